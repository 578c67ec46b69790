// Per-element device math for the mixed-FEM Newton step on gfx950.
//
// Everything here runs inside one lane (one tetrahedron per lane), on values
// held in registers. Matrices are column-major arrays (a[r + 3*c]) so that
// vec(F) matches the reference's column-major Eigen vectorisation
// (reference src/mesh/tet_mesh.cpp:16-27, tests/test_dsvd.cpp:226).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace mfemhip {

#define MFEM_DEV __device__ __forceinline__

template <typename T> MFEM_DEV T tsqrt(T x);
template <> MFEM_DEV double tsqrt<double>(double x) { return sqrt(x); }
template <> MFEM_DEV float tsqrt<float>(float x) { return sqrtf(x); }
template <typename T> MFEM_DEV T tabs(T x) { return x < T(0) ? -x : x; }

// ---------------------------------------------------------------------------
// Rotation fitting: maximise tr(R F^T) over SO(3) by Newton's method with the
// reference's warm start, tolerance and backtracking rule
// (reference src/svd/newton_procrustes.h:108-189 with A = I, B = F).
//
// With Y = R F^T the gradient and Hessian used by the reference reduce to
//   g = (Y21 - Y12, Y02 - Y20, Y10 - Y01),
//   H = tr(Y) I - (Y + Y^T)/2.
// The step is omega = -H^{-1} g; the inner loop keeps shrinking omega by 0.6
// while -tr(dR Y) > -tr(R Y) (the reference compares against tr(R Y), not
// tr(Y)) and |omega| > tol; the last trial rotation is applied.
MFEM_DEV void solve3_sym(const double H[6], const double b[3], double x[3]) {
  // H packed as (00,11,22,01,02,12); adjugate solve
  double a = H[0], d = H[1], f = H[2], bb = H[3], c = H[4], e = H[5];
  double A00 = d * f - e * e, A01 = c * e - bb * f, A02 = bb * e - c * d;
  double A11 = a * f - c * c, A12 = bb * c - a * e, A22 = a * d - bb * bb;
  double det = a * A00 + bb * A01 + c * A02;
  double inv = 1.0 / det;
  x[0] = (A00 * b[0] + A01 * b[1] + A02 * b[2]) * inv;
  x[1] = (A01 * b[0] + A11 * b[1] + A12 * b[2]) * inv;
  x[2] = (A02 * b[0] + A12 * b[1] + A22 * b[2]) * inv;
}

MFEM_DEV void matmul3(const double A[9], const double B[9], double C[9]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      C[r + 3 * c] = A[r] * B[3 * c] + A[r + 3] * B[1 + 3 * c] + A[r + 6] * B[2 + 3 * c];
}

// Rodrigues' formula (reference src/svd/newton_procrustes.h:87-106).
MFEM_DEV void rodrigues(const double w[3], double dR[9]) {
  double angle = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  // below 1e-8 the rotation is the identity: sin and 1 - cos are taken as
  // 0 (and the axis as 0), for which the formula below gives exactly I;
  // selecting instead of branching keeps dR in registers
  const bool small = fabs(angle) < 1e-8;
  double ia = small ? 0.0 : 1.0 / angle;
  double kx = w[0] * ia, ky = w[1] * ia, kz = w[2] * ia;
  double sn, cs;
  sincos(angle, &sn, &cs);
  sn = small ? 0.0 : sn;
  double oc = small ? 0.0 : 1.0 - cs;
  // K = [0 -kz ky; kz 0 -kx; -ky kx 0]; K^2 = k k^T - I
  dR[0] = 1.0 + oc * (kx * kx - 1.0);
  dR[4] = 1.0 + oc * (ky * ky - 1.0);
  dR[8] = 1.0 + oc * (kz * kz - 1.0);
  dR[1] = sn * kz + oc * kx * ky;   // (1,0)
  dR[3] = -sn * kz + oc * kx * ky;  // (0,1)
  dR[2] = -sn * ky + oc * kx * kz;  // (2,0)
  dR[6] = sn * ky + oc * kx * kz;   // (0,2)
  dR[5] = sn * kx + oc * ky * kz;   // (2,1)
  dR[7] = -sn * kx + oc * ky * kz;  // (1,2)
}

// R (in: warm start, out: rotation), F column-major. Returns iterations.
MFEM_DEV int procrustes(double R[9], const double F[9], double tol, int max_iter) {
  double Y[9];
  // Y = R F^T
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      Y[r + 3 * c] = R[r] * F[c] + R[r + 3] * F[c + 3] + R[r + 6] * F[c + 6];
  int itr = 0;
  do {
    double g[3] = {Y[5] - Y[7], Y[6] - Y[2], Y[1] - Y[3]};
    if (sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) < tol) break;
    double tr = Y[0] + Y[4] + Y[8];
    double H[6] = {tr - Y[0], tr - Y[4], tr - Y[8], -0.5 * (Y[3] + Y[1]),
                   -0.5 * (Y[6] + Y[2]), -0.5 * (Y[7] + Y[5])};
    double w[3];
    solve3_sym(H, g, w);
    w[0] = -w[0]; w[1] = -w[1]; w[2] = -w[2];
    // E0 = -tr(R Y)
    double trRY = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) trRY += R[i + 3 * k] * Y[k + 3 * i];
    double E0 = -trRY, E1;
    double dR[9];
    do {
      rodrigues(w, dR);
      double t = 0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) t += dR[i + 3 * k] * Y[k + 3 * i];
      E1 = -t;
      w[0] *= 0.6; w[1] *= 0.6; w[2] *= 0.6;
    } while (E1 > E0 && sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) > tol);
    double T1[9];
    matmul3(dR, R, T1);
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = T1[i];
    matmul3(dR, Y, T1);
#pragma unroll
    for (int i = 0; i < 9; ++i) Y[i] = T1[i];
    ++itr;
  } while (itr < max_iter);
  return itr;
}

// s = sym(R^T F) as (S00,S11,S22,S01,S02,S12) (reference mixed_stretch.cpp:25-29)
MFEM_DEV void stretch_of(const double R[9], const double F[9], double s[6], double S[9]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      S[r + 3 * c] = R[3 * r] * F[3 * c] + R[3 * r + 1] * F[3 * c + 1] + R[3 * r + 2] * F[3 * c + 2];
  s[0] = S[0]; s[1] = S[4]; s[2] = S[8];
  s[3] = 0.5 * (S[1] + S[3]);
  s[4] = 0.5 * (S[2] + S[6]);
  s[5] = 0.5 * (S[5] + S[7]);
}

// ds/dvec(F) (6 x 9, column-major dsdF[i + 6*c]) at the fitted rotation,
// the implicit derivative of the reference (newton_procrustes.h:161-187,
// mixed_stretch.cpp:31-41): for the unit perturbation E_ij of F,
//   w  = H^{-1} (e_i x r_j),   H = tr(Y) I - sym(Y),  Y = R F^T,
//   dR = -[w]x R,
//   dS = R^T (E_ij - dR S) = R^T (E_ij + [w]x F)   (R S = F).
template <typename Real>
MFEM_DEV void dsdF_of(const double R[9], const double F[9], Real dsdF[54]) {
  double Y[9];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      Y[r + 3 * c] = R[r] * F[c] + R[r + 3] * F[c + 3] + R[r + 6] * F[c + 6];
  double tr = Y[0] + Y[4] + Y[8];
  double H[6] = {tr - Y[0], tr - Y[4], tr - Y[8], -0.5 * (Y[3] + Y[1]),
                 -0.5 * (Y[6] + Y[2]), -0.5 * (Y[7] + Y[5])};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      // d = e_i x r_j, r_j = column j of R
      double rj0 = R[3 * j], rj1 = R[3 * j + 1], rj2 = R[3 * j + 2];
      double d[3];
      if (i == 0) { d[0] = 0; d[1] = -rj2; d[2] = rj1; }
      else if (i == 1) { d[0] = rj2; d[1] = 0; d[2] = -rj0; }
      else { d[0] = -rj1; d[1] = rj0; d[2] = 0; }
      double w[3];
      solve3_sym(H, d, w);
      // M = E_ij + [w]x F
      double M[9];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double f0 = F[3 * c], f1 = F[3 * c + 1], f2 = F[3 * c + 2];
        M[3 * c + 0] = w[1] * f2 - w[2] * f1;
        M[3 * c + 1] = w[2] * f0 - w[0] * f2;
        M[3 * c + 2] = w[0] * f1 - w[1] * f0;
      }
      M[i + 3 * j] += 1.0;
      // dS = R^T M
      double dS[9];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
          dS[r + 3 * c] = R[3 * r] * M[3 * c] + R[3 * r + 1] * M[3 * c + 1] + R[3 * r + 2] * M[3 * c + 2];
      int col = i + 3 * j;
      dsdF[0 + 6 * col] = (Real)dS[0];
      dsdF[1 + 6 * col] = (Real)dS[4];
      dsdF[2 + 6 * col] = (Real)dS[8];
      dsdF[3 + 6 * col] = (Real)(0.5 * (dS[1] + dS[3]));
      dsdF[4 + 6 * col] = (Real)(0.5 * (dS[2] + dS[6]));
      dsdF[5 + 6 * col] = (Real)(0.5 * (dS[5] + dS[7]));
    }
  }
}

// ---------------------------------------------------------------------------
// Materials on s = (S00,S11,S22,S01,S02,S12), Voigt off-diagonals counted
// twice (reference src/energies/*.cpp 6-vector API; closed forms of
// scripts/derivs.m). Symmetric 6x6 matrices are packed upper-triangular,
// row by row: index sym6(i,j), i <= j.
enum { MAT_SNH = 0, MAT_NH = 1, MAT_COROT = 2, MAT_FCR = 3, MAT_ARAP = 4, MAT_FUNG = 5 };

MFEM_DEV constexpr int sym6(int i, int j) {
  return i <= j ? (i * 6 - (i * (i - 1)) / 2 + (j - i)) : (j * 6 - (j * (j - 1)) / 2 + (i - j));
}

// det S and its first/second derivatives with respect to s
MFEM_DEV void det_s6(const double s[6], double& J, double dJ[6]) {
  double a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5];
  J = a * b * c + 2.0 * d * e * f - a * f * f - b * e * e - c * d * d;
  dJ[0] = b * c - f * f;
  dJ[1] = a * c - e * e;
  dJ[2] = a * b - d * d;
  dJ[3] = 2.0 * (e * f - c * d);
  dJ[4] = 2.0 * (d * f - b * e);
  dJ[5] = 2.0 * (d * e - a * f);
}

// Fung exponent constant (reference src/energies/fung.cpp:7)
constexpr double kFungC = 4.0;

// energy only (reference energies: stable_neohookean.cpp, neohookean.cpp,
// corotational.cpp, fixed_corotational.cpp, arap.cpp, fung.cpp; closed forms
// in scripts/derivs.m)
MFEM_DEV double material_energy(int type, double mu, double la, const double s[6]) {
  double d0 = s[0] - 1.0, d1 = s[1] - 1.0, d2 = s[2] - 1.0;
  double off = 2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  if (type == MAT_ARAP) return 0.5 * mu * (d0 * d0 + d1 * d1 + d2 * d2 + off);
  if (type == MAT_COROT) {
    double t = d0 + d1 + d2;
    return 0.5 * la * t * t + mu * (d0 * d0 + d1 * d1 + d2 * d2 + off);
  }
  double J, dJ[6];
  det_s6(s, J, dJ);
  if (type == MAT_FCR) return 0.5 * la * (J - 1.0) * (J - 1.0) + mu * (d0 * d0 + d1 * d1 + d2 * d2 + off);
  double trS2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + off;
  if (type == MAT_NH) {
    const double lJ = log(J);
    return 0.5 * mu * (trS2 - 3.0) - mu * lJ + 0.5 * la * lJ * lJ;
  }
  if (type == MAT_FUNG) {
    const double A = 1.0 + (mu + mu * kFungC) / la;
    return 0.5 * mu * (trS2 - 3.0) + 0.5 * la * (J - A) * (J - A) +
           0.5 * mu * (exp(0.5 * kFungC * (trS2 - 3.0)) - 1.0);
  }
  return 0.5 * mu * (trS2 - 3.0) - mu * (J - 1.0) + 0.5 * la * (J - 1.0) * (J - 1.0);
}

// gradient g (FP64) and Hessian H (packed 21, FP64, not PSD-projected)
MFEM_DEV void material_grad_hess(int type, double mu, double la, const double s[6],
                                 double g[6], double H[21]) {
  const double vg[6] = {1, 1, 1, 2, 2, 2};
#pragma unroll
  for (int i = 0; i < 21; ++i) H[i] = 0.0;
  if (type == MAT_ARAP) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      g[i] = mu * vg[i] * (s[i] - (i < 3 ? 1.0 : 0.0));
      H[sym6(i, i)] = mu * vg[i];
    }
    return;
  }
  if (type == MAT_COROT) {
    double t = s[0] + s[1] + s[2] - 3.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      g[i] = 2.0 * mu * vg[i] * (s[i] - (i < 3 ? 1.0 : 0.0)) + (i < 3 ? la * t : 0.0);
      H[sym6(i, i)] = 2.0 * mu * vg[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) H[sym6(i, j)] += la;
    return;
  }
  double J, dJ[6];
  det_s6(s, J, dJ);
  double k;         // coefficient of d2J
  double cjj = la;  // coefficient of dJ dJ^T
  if (type == MAT_FCR) {
    k = la * (J - 1.0);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      g[i] = 2.0 * mu * vg[i] * (s[i] - (i < 3 ? 1.0 : 0.0)) + k * dJ[i];
      H[sym6(i, i)] = 2.0 * mu * vg[i];
    }
  } else if (type == MAT_NH) {
    // 0.5 mu (I2 - 3) - mu log J + 0.5 la (log J)^2 (symmetric closed form)
    const double lJ = log(J);
    k = (la * lJ - mu) / J;
    cjj = (la + mu - la * lJ) / (J * J);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      g[i] = mu * vg[i] * s[i] + k * dJ[i];
      H[sym6(i, i)] = mu * vg[i];
    }
  } else if (type == MAT_FUNG) {
    // 0.5 mu (I2 - 3) + 0.5 la (J - A)^2 + 0.5 mu (exp(c (I2 - 3) / 2) - 1)
    double trS2 = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) trS2 += vg[i] * s[i] * s[i];
    const double ex = exp(0.5 * kFungC * (trS2 - 3.0));
    const double A = 1.0 + (mu + mu * kFungC) / la;
    const double a = mu + 0.5 * kFungC * mu * ex, b = 0.5 * kFungC * kFungC * mu * ex;
    k = la * (J - A);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      g[i] = a * vg[i] * s[i] + k * dJ[i];
      H[sym6(i, i)] = a * vg[i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j) H[sym6(i, j)] += b * vg[i] * s[i] * vg[j] * s[j];
  } else {  // SNH
    k = la * (J - 1.0) - mu;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      g[i] = mu * vg[i] * s[i] + k * dJ[i];
      H[sym6(i, i)] = mu * vg[i];
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) H[sym6(i, j)] += cjj * dJ[i] * dJ[j];
  // + k * d2J (d2J is linear in s)
  double a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5];
  H[sym6(0, 1)] += k * c;
  H[sym6(0, 2)] += k * b;
  H[sym6(1, 2)] += k * a;
  H[sym6(0, 5)] += k * (-2.0 * f);
  H[sym6(1, 4)] += k * (-2.0 * e);
  H[sym6(2, 3)] += k * (-2.0 * d);
  H[sym6(3, 3)] += k * (-2.0 * c);
  H[sym6(4, 4)] += k * (-2.0 * b);
  H[sym6(5, 5)] += k * (-2.0 * a);
  H[sym6(3, 4)] += k * (2.0 * f);
  H[sym6(3, 5)] += k * (2.0 * e);
  H[sym6(4, 5)] += k * (2.0 * d);
}

// ---------------------------------------------------------------------------
// 2D (reference DIM = 2 paths). Symmetric 3x3 matrices are packed
// upper-triangular row by row: sym3(i, j) in 0..5.
MFEM_DEV constexpr int sym3(int i, int j) {
  return i <= j ? (i * 3 - (i * (i - 1)) / 2 + (j - i)) : (j * 3 - (j * (j - 1)) / 2 + (i - j));
}

// 2D polar fit (mixed_stretch.cpp:42-77 with svd_eigen.h:7-31 and
// dsvd.cpp:138-191): R = the rotation by atan2(F10 - F01, F00 + F11), which
// is the flip-corrected U V^T; S = R^T F, s = (S00, S11, (S10 + S01) / 2);
// dsdF (3 x 4 column-major, entry (i, c) at i + 3c) = rows {0, 3, (1 + 2)/2}
// of (I (x) R^T)(I - (S^T (x) I) dR/dF), with dR/dF dropped when
// |sigma_1^2 - sigma_0^2| < 1e-4 as the reference's dsvd does. F, R and S
// are 2x2 column-major.
MFEM_DEV void polar2d(const double F[4], double R[4], double S[4], double s[3], double* dsdF) {
  const double a = F[0] + F[3], b = F[1] - F[2];
  const double th = atan2(b, a), c = cos(th), sn = sin(th);
  R[0] = c; R[1] = sn; R[2] = -sn; R[3] = c;
  // S = R^T F
  S[0] = c * F[0] + sn * F[1];
  S[1] = -sn * F[0] + c * F[1];
  S[2] = c * F[2] + sn * F[3];
  S[3] = -sn * F[2] + c * F[3];
  s[0] = S[0];
  s[1] = S[3];
  s[2] = 0.5 * (S[1] + S[2]);
  if (!dsdF) return;
  const double n2 = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  const double det = F[0] * F[3] - F[2] * F[1];
  const double gap = sqrt(fmax(n2 * n2 - 4.0 * det * det, 0.0));
  double dRdF[16];  // (i, j) at i + 4j: d vec(R)_i / d vec(F)_j
#pragma unroll
  for (int i = 0; i < 16; ++i) dRdF[i] = 0.0;
  if (gap >= 1e-4) {
    const double r2 = a * a + b * b;
    const double dth[4] = {-b / r2, a / r2, -a / r2, -b / r2};
    const double dR[4] = {-sn, c, -c, -sn};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) dRdF[i + 4 * j] = dR[i] * dth[j];
  }
  // K = I - (S^T (x) I) dRdF; (S^T (x) I)(bi*2 + i, bj*2 + i) = S(bj, bi)
  double K[16];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int bi = i / 2, ii = i % 2;
      double v = 0.0;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) v += S[bj + 2 * bi] * dRdF[(bj * 2 + ii) + 4 * j];
      K[i + 4 * j] = (i == j ? 1.0 : 0.0) - v;
    }
  // J = (I (x) R^T) K; (I (x) R^T)(blk*2 + i, blk*2 + j) = R(j, i)
  double J[16];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int blk = i / 2, ii = i % 2;
      J[i + 4 * j] = R[0 + 2 * ii] * K[(blk * 2 + 0) + 4 * j] + R[1 + 2 * ii] * K[(blk * 2 + 1) + 4 * j];
    }
#pragma unroll
  for (int col = 0; col < 4; ++col) {
    dsdF[0 + 3 * col] = J[0 + 4 * col];
    dsdF[1 + 3 * col] = J[3 + 4 * col];
    dsdF[2 + 3 * col] = 0.5 * (J[1 + 4 * col] + J[2 + 4 * col]);
  }
}

// Materials on the 2D 3-vector s = (S00, S11, S01) (stable_neohookean.cpp:
// 236-277, arap.cpp:67-98, corotational.cpp:62-96, fixed_corotational.cpp:
// 93-132), with q = 1 - det S. SNH's Hessian is PSD-projected by the caller;
// the host admits no other type on this path.
MFEM_DEV double material2d_energy(int type, double mu, double la, const double s[3]) {
  const double S1 = s[0], S2 = s[1], S3 = s[2];
  const double q = 1.0 - S1 * S2 + S3 * S3;
  const double dev = (S1 - 1) * (S1 - 1) + (S2 - 1) * (S2 - 1) + 2.0 * S3 * S3;
  if (type == MAT_ARAP) return 0.5 * mu * dev;
  if (type == MAT_COROT) return 0.5 * la * (S1 + S2 - 2.0) * (S1 + S2 - 2.0) + mu * dev;
  if (type == MAT_FCR) return 0.5 * la * q * q + mu * dev;
  return 0.5 * la * q * q + 0.5 * mu * (S1 * S1 + S2 * S2 + 2.0 * S3 * S3 - 2.0) + mu * q;
}

MFEM_DEV void material2d_grad_hess(int type, double mu, double la, const double s[3], double g[3],
                                   double H[6]) {
  const double S1 = s[0], S2 = s[1], S3 = s[2];
  const double q = 1.0 - S1 * S2 + S3 * S3;
  const double dq[3] = {-S2, -S1, 2.0 * S3};
#pragma unroll
  for (int i = 0; i < 6; ++i) H[i] = 0.0;
  if (type == MAT_ARAP) {
    g[0] = mu * (S1 - 1); g[1] = mu * (S2 - 1); g[2] = 2.0 * mu * S3;
    H[sym3(0, 0)] = mu; H[sym3(1, 1)] = mu; H[sym3(2, 2)] = 2.0 * mu;
    return;
  }
  if (type == MAT_COROT) {
    const double t = S1 + S2 - 2.0;
    g[0] = la * t + 2.0 * mu * (S1 - 1); g[1] = la * t + 2.0 * mu * (S2 - 1); g[2] = 4.0 * mu * S3;
    H[sym3(0, 0)] = H[sym3(1, 1)] = la + 2.0 * mu;
    H[sym3(0, 1)] = la;
    H[sym3(2, 2)] = 4.0 * mu;
    return;
  }
  // SNH: (la q + mu) dq + mu diag(1,1,2) s; FCR: la q dq + mu diag(2,2,4)(s - (1,1,0))
  const bool fcr = type == MAT_FCR;
  const double k = fcr ? la * q : la * q + mu;
  const double v[3] = {1, 1, 2};
#pragma unroll
  for (int i = 0; i < 3; ++i)
    g[i] = k * dq[i] + (fcr ? 2.0 * mu * v[i] * (s[i] - (i < 2 ? 1.0 : 0.0)) : mu * v[i] * s[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) H[sym3(i, j)] = la * dq[i] * dq[j];
  // d2q: (0,1) = -1, (2,2) = 2
  H[sym3(0, 1)] += -k;
  H[sym3(2, 2)] += 2.0 * k;
#pragma unroll
  for (int i = 0; i < 3; ++i) H[sym3(i, i)] += (fcr ? 2.0 : 1.0) * mu * v[i];
}

MFEM_DEV bool material_needs_psd(int type) {
  // SNH (reference stable_neohookean.cpp:91-93) and NH/Fung project; ARAP,
  // Corot and FCR do not.
  return type == MAT_SNH || type == MAT_NH || type == MAT_FUNG;
}

// ---------------------------------------------------------------------------
// Symmetric NxN helpers on full column-major storage.
// Cholesky test of A - tol*I: true iff every pivot is positive.
template <typename Real, int N>
MFEM_DEV bool chol_pd(const Real* A, Real tol) {
  Real L[N * (N + 1) / 2];
  int idx = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
#pragma unroll
    for (int i = j; i < N; ++i) {
      Real s = A[i + N * j] - (i == j ? tol : Real(0));
      // L stored column-wise packed: L(i,j) at off(j) + (i - j)
#pragma unroll
      for (int k = 0; k < j; ++k) {
        int offj = k * N - (k * (k - 1)) / 2;
        s -= L[offj + (i - k)] * L[offj + (j - k)];
      }
      int offj = j * N - (j * (j - 1)) / 2;
      if (i == j) {
        if (!(s > Real(0))) return false;
        L[offj] = tsqrt<Real>(s);
      } else {
        L[offj + (i - j)] = s / L[offj];
      }
      ++idx;
    }
  }
  return true;
}

// Cyclic Jacobi eigen-decomposition (Golub & Van Loan Alg. 8.5.3) followed by
// clamping eigenvalues below tol to tol: A <- Q max(W, tol) Q^T. A is full,
// column-major, symmetric. This is the PSD projection applied at the
// reference's sim::simple_psd_fix call sites; a matrix that passes the
// Cholesky test above is left unchanged.
template <typename Real, int N>
MFEM_DEV void jacobi_psd_project(Real* A, Real tol) {
  Real Q[N * N];
#pragma unroll
  for (int i = 0; i < N * N; ++i) Q[i] = (i % (N + 1) == 0) ? Real(1) : Real(0);
  const Real eps = sizeof(Real) == 8 ? Real(1e-30) : Real(1e-14);
  for (int sweep = 0; sweep < 30; ++sweep) {
    Real off = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        Real v = A[i + N * j] * A[i + N * j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= eps * tot) break;
#pragma unroll
    for (int p = 0; p < N - 1; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        Real apq = A[p + N * q];
        if (apq == Real(0)) continue;
        Real theta = (A[q + N * q] - A[p + N * p]) / (Real(2) * apq);
        Real t = (theta >= Real(0) ? Real(1) : Real(-1)) /
                 (tabs(theta) + tsqrt<Real>(theta * theta + Real(1)));
        Real c = Real(1) / tsqrt<Real>(t * t + Real(1)), s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          Real akp = A[k + N * p], akq = A[k + N * q];
          A[k + N * p] = c * akp - s * akq;
          A[k + N * q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          Real apk = A[p + N * k], aqk = A[q + N * k];
          A[p + N * k] = c * apk - s * aqk;
          A[q + N * k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          Real qkp = Q[k + N * p], qkq = Q[k + N * q];
          Q[k + N * p] = c * qkp - s * qkq;
          Q[k + N * q] = s * qkp + c * qkq;
        }
      }
  }
  Real w[N];
#pragma unroll
  for (int i = 0; i < N; ++i) w[i] = A[i + N * i] < tol ? tol : A[i + N * i];
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = j; i < N; ++i) {
      Real s = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) s += Q[i + N * k] * w[k] * Q[j + N * k];
      A[i + N * j] = s;
      A[j + N * i] = s;
    }
}

// Packed symmetric storage, upper triangle row by row: (i, j), i <= j, at
// i*N - i(i-1)/2 + (j - i) (sym6 for N = 6).
template <int N>
MFEM_DEV constexpr int sym_pk(int i, int j) {
  return i <= j ? i * N - (i * (i - 1)) / 2 + (j - i) : j * N - (j * (j - 1)) / 2 + (i - j);
}

// chol_pd on packed symmetric storage.
template <typename Real, int N>
MFEM_DEV bool chol_pd_packed(const Real* A, Real tol) {
  Real L[N * (N + 1) / 2];
#pragma unroll
  for (int j = 0; j < N; ++j) {
#pragma unroll
    for (int i = j; i < N; ++i) {
      Real s = A[sym_pk<N>(j, i)] - (i == j ? tol : Real(0));
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[sym_pk<N>(k, i)] * L[sym_pk<N>(k, j)];
      if (i == j) {
        if (!(s > Real(0))) return false;
        L[sym_pk<N>(j, j)] = tsqrt<Real>(s);
      } else {
        L[sym_pk<N>(j, i)] = s / L[sym_pk<N>(j, j)];
      }
    }
  }
  return true;
}

// jacobi_psd_project on packed symmetric storage: the same cyclic sweeps and
// rotation angles, with each rotation applied to both sides at once (Golub &
// Van Loan Alg. 8.5.2: a_pq <- 0, a_pp <- a_pp - t a_pq, a_qq <- a_qq + t a_pq,
// the other entries of rows p and q rotated). Holding 21 entries instead of
// 36 keeps the FP64 projection within half the register file.
template <typename Real, int N>
MFEM_DEV void jacobi_psd_project_packed(Real* A, Real tol) {
  Real Q[N * N];
#pragma unroll
  for (int i = 0; i < N * N; ++i) Q[i] = (i % (N + 1) == 0) ? Real(1) : Real(0);
  const Real eps = sizeof(Real) == 8 ? Real(1e-30) : Real(1e-14);
  for (int sweep = 0; sweep < 30; ++sweep) {
    Real off = 0, dia = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      dia += A[sym_pk<N>(i, i)] * A[sym_pk<N>(i, i)];
#pragma unroll
      for (int j = i + 1; j < N; ++j) off += A[sym_pk<N>(i, j)] * A[sym_pk<N>(i, j)];
    }
    off *= Real(2);
    if (off <= eps * (off + dia)) break;
#pragma unroll
    for (int p = 0; p < N - 1; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const Real apq = A[sym_pk<N>(p, q)];
        if (apq == Real(0)) continue;
        const Real app = A[sym_pk<N>(p, p)], aqq = A[sym_pk<N>(q, q)];
        const Real theta = (aqq - app) / (Real(2) * apq);
        const Real t = (theta >= Real(0) ? Real(1) : Real(-1)) /
                       (tabs(theta) + tsqrt<Real>(theta * theta + Real(1)));
        const Real c = Real(1) / tsqrt<Real>(t * t + Real(1)), s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if (k == p || k == q) continue;
          const Real akp = A[sym_pk<N>(k, p)], akq = A[sym_pk<N>(k, q)];
          A[sym_pk<N>(k, p)] = c * akp - s * akq;
          A[sym_pk<N>(k, q)] = s * akp + c * akq;
        }
        A[sym_pk<N>(p, p)] = app - t * apq;
        A[sym_pk<N>(q, q)] = aqq + t * apq;
        A[sym_pk<N>(p, q)] = Real(0);
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const Real qkp = Q[k + N * p], qkq = Q[k + N * q];
          Q[k + N * p] = c * qkp - s * qkq;
          Q[k + N * q] = s * qkp + c * qkq;
        }
      }
  }
  Real w[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const Real d = A[sym_pk<N>(i, i)];
    w[i] = d < tol ? tol : d;
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i; j < N; ++j) {
      Real s = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) s += Q[i + N * k] * w[k] * Q[j + N * k];
      A[sym_pk<N>(i, j)] = s;
    }
}

// Inverse of an SPD matrix via Cholesky (full column-major in/out).
template <typename Real, int N>
MFEM_DEV void spd_inverse(const Real* A, Real* X) {
  Real L[N * N];
#pragma unroll
  for (int i = 0; i < N * N; ++i) L[i] = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    Real d = A[j + N * j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[j + N * k] * L[j + N * k];
    d = tsqrt<Real>(d);
    L[j + N * j] = d;
    Real id = Real(1) / d;
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      Real s = A[i + N * j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[i + N * k] * L[j + N * k];
      L[i + N * j] = s * id;
    }
  }
  // X = L^{-T} L^{-1}: solve column by column
#pragma unroll
  for (int c = 0; c < N; ++c) {
    Real y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      Real s = (i == c) ? Real(1) : Real(0);
#pragma unroll
      for (int k = 0; k < i; ++k) s -= L[i + N * k] * y[k];
      y[i] = s / L[i + N * i];
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
      Real s = y[i];
#pragma unroll
      for (int k = i + 1; k < N; ++k) s -= L[k + N * i] * X[k + N * c];
      X[i + N * c] = s / L[i + N * i];
    }
  }
}

}  // namespace mfemhip
