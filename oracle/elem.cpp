// ORACLE — test infrastructure only. Not part of the product library.
//
// Per-element FP64 math of the reference, restated without Eigen:
//   * newton_procrustes + rodrigues   (reference src/svd/newton_procrustes.h:87-189)
//   * polar_svd<3,6>                  (reference src/variables/mixed_stretch.cpp:14-42)
//   * material models on the 6-vector s, the 9-vector F and the 4-vector F
//     (reference src/energies/stable_neohookean.cpp, arap.cpp,
//      corotational.cpp, fixed_corotational.cpp; closed forms of
//      scripts/derivs.m:20-34,66-78)
// The polynomial derivatives below are derived here from the closed forms
// (via det S and its derivatives), not transcribed from the generated code.
#include "oracle.h"
#include "omath.h"

namespace orc {

// Cross-product generator matrices [e_x]x, [e_y]x, [e_z]x
// (reference src/svd/newton_procrustes.h:11-30) and their symmetric
// products (:32-66).
static M3 cross_mat(int a) {
  M3 m;
  if (a == 0) { m(1, 2) = -1; m(2, 1) = 1; }
  if (a == 1) { m(0, 2) = 1; m(2, 0) = -1; }
  if (a == 2) { m(0, 1) = -1; m(1, 0) = 1; }
  return m;
}

struct ProcrustesConsts {
  M3 cp[3];
  M3 cpp[3][3];  // 0.5*(cp_a cp_b + cp_b cp_a)
  Mat<9, 3> skew_to_full;
  ProcrustesConsts() {
    for (int a = 0; a < 3; ++a) cp[a] = cross_mat(a);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) cpp[a][b] = 0.5 * (cp[a] * cp[b] + cp[b] * cp[a]);
    // vec([w]x) = skew_to_full * w, column-major vec (:69-85)
    for (int a = 0; a < 3; ++a) {
      V3 e; e[a] = 1;
      M3 S = cross_mat(a);
      Vec<9> v = flatten(S);
      for (int i = 0; i < 9; ++i) skew_to_full(i, a) = v[i];
    }
  }
};
static const ProcrustesConsts PC;

// tr(A*B) without forming the product.
static inline double trace_prod(const M3& A, const M3& B) {
  double t = 0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) t += A(i, k) * B(k, i);
  return t;
}

// Reference src/svd/newton_procrustes.h:87-106
static M3 rodrigues(const V3& omega) {
  double angle = omega.norm();
  if (std::fabs(angle) < 1e-8) return M3::Identity();
  V3 axis = (1.0 / angle) * omega;
  M3 K;
  K(0, 1) = -axis[2]; K(0, 2) = axis[1];
  K(1, 0) = axis[2];  K(1, 2) = -axis[0];
  K(2, 0) = -axis[1]; K(2, 1) = axis[0];
  return M3::Identity() + std::sin(angle) * K + (1.0 - std::cos(angle)) * (K * K);
}

// Reference src/svd/newton_procrustes.h:108-189. Maximises tr(R A B^T) over
// rotations by Newton's method on SO(3), warm-started from R. Replicates the
// reference's quirks: E0 is -tr(R*Y) although Y already contains R, the
// tolerance is absolute on |g|, and the last trial rotation of the inner
// backtracking loop is applied. Returns 0, or -1 if |H| < 1e-6 at the
// gradient evaluation (the reference prints "H FAILURE" and exits there).
int newton_procrustes(M3& R, const M3& A, const M3& B, bool compute_gradients,
                      Mat<9, 9>* dRdF, double tol, int max_iter, int* iters) {
  M3 Y = R * A * B.T();
  V3 g;
  M3 H;
  M3 dR;
  int itr = 0;
  do {
    for (int a = 0; a < 3; ++a) g[a] = -trace_prod(PC.cp[a], Y);
    if (g.norm() < tol) break;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) H(a, b) = -trace_prod(PC.cpp[a][b], Y);
    V3 omega = -colpiv_qr_solve<3, 1>(H, g);
    double E0 = -trace_prod(R, Y);
    double E1 = E0 + 1.0;
    do {
      dR = rodrigues(omega);
      E1 = -trace_prod(dR, Y);
      omega *= 0.6;
    } while (E1 > E0 && omega.norm() > tol);
    R = dR * R;
    Y = dR * Y;
    ++itr;
  } while (itr < max_iter);
  if (iters) *iters = itr;

  if (compute_gradients) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) H(a, b) = -trace_prod(PC.cpp[a][b], Y);
    if (H.norm() < 1e-6) return -1;
    Mat<3, 9> dF;
    M3 RA = R * A;
    for (int a = 0; a < 3; ++a) {
      Vec<9> v = flatten(-(PC.cp[a] * RA));
      for (int i = 0; i < 9; ++i) dF(a, i) = v[i];
    }
    Mat<3, 9> HdF = colpiv_qr_solve<3, 9>(H, dF);
    *dRdF = -(flatten_multiply_right<3>(R) * PC.skew_to_full * HdF);
  }
  return 0;
}

// Reference src/variables/mixed_stretch.cpp:14-42 (D == 3 branch).
int polar_svd3(M3& R, Vec<6>& s, const M3& F, bool compute_gradients,
               Mat<6, 9>* dsdF, int* iters) {
  Mat<9, 9> dRdF;
  int rc = newton_procrustes(R, M3::Identity(), F, compute_gradients, &dRdF,
                             1e-6, 100, iters);
  M3 S = R.T() * F;
  s[0] = S(0, 0); s[1] = S(1, 1); s[2] = S(2, 2);
  s[3] = 0.5 * (S(1, 0) + S(0, 1));
  s[4] = 0.5 * (S(2, 0) + S(0, 2));
  s[5] = 0.5 * (S(2, 1) + S(1, 2));
  if (compute_gradients && dsdF) {
    Mat<9, 9> J = flatten_multiply<3>(R.T()) *
                  (Mat<9, 9>::Identity() - flatten_multiply_right<3>(S) * dRdF);
    for (int c = 0; c < 9; ++c) {
      (*dsdF)(0, c) = J(0, c);
      (*dsdF)(1, c) = J(4, c);
      (*dsdF)(2, c) = J(8, c);
      (*dsdF)(3, c) = 0.5 * (J(1, c) + J(3, c));
      (*dsdF)(4, c) = 0.5 * (J(2, c) + J(6, c));
      (*dsdF)(5, c) = 0.5 * (J(5, c) + J(7, c));
    }
  }
  return rc;
}

// ---------------------------------------------------------------------------
// 2D polar fit of mixed_stretch.cpp:42-77 (DIM = 2). The reference takes
// R = U V^T from JacobiSVD with the flip fix of svd_eigen.h:7-31; for any F
// this is the rotation by theta = atan2(F10 - F01, F00 + F11), the rotation
// that makes R^T F symmetric with a non-negative trace. S = R^T F,
// s = (S00, S11, (S10 + S01) / 2); dsdF holds rows {0, 3, (1 + 2) / 2} of
// J = (I (x) R^T)(I - (S^T (x) I) dRdF). dRdF follows dsvd.cpp:138-191: zero
// when |sigma_1^2 - sigma_0^2| < 1e-4 (its "no angular velocity" guard),
// else the exact derivative of the rotation, vec(dR/dtheta) dtheta/dF.
int polar2d(Mat<2, 2>& R, Vec<3>& s, const Mat<2, 2>& F, bool compute_gradients,
            Mat<3, 4>* dsdF) {
  const double a = F(0, 0) + F(1, 1), b = F(1, 0) - F(0, 1);
  const double th = std::atan2(b, a), c = std::cos(th), sn = std::sin(th);
  R(0, 0) = c; R(1, 0) = sn; R(0, 1) = -sn; R(1, 1) = c;
  Mat<2, 2> S = R.T() * F;
  s[0] = S(0, 0); s[1] = S(1, 1); s[2] = 0.5 * (S(1, 0) + S(0, 1));
  if (compute_gradients && dsdF) {
    Mat<4, 4> dRdF;
    double n2 = 0;
    for (int i = 0; i < 4; ++i) n2 += F[i] * F[i];
    const double det = F(0, 0) * F(1, 1) - F(0, 1) * F(1, 0);
    const double gap = std::sqrt(std::max(n2 * n2 - 4.0 * det * det, 0.0));  // s1^2 - s0^2
    if (gap >= 1e-4) {
      const double r2 = a * a + b * b;
      const double dth[4] = {-b / r2, a / r2, -a / r2, -b / r2};  // d theta / vec F
      const double dR[4] = {-sn, c, -c, -sn};                     // vec dR/dtheta
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) dRdF(i, j) = dR[i] * dth[j];
    }
    Mat<4, 4> J = flatten_multiply<2>(R.T()) *
                  (Mat<4, 4>::Identity() - flatten_multiply_right<2>(S) * dRdF);
    for (int col = 0; col < 4; ++col) {
      (*dsdF)(0, col) = J(0, col);
      (*dsdF)(1, col) = J(3, col);
      (*dsdF)(2, col) = 0.5 * (J(1, col) + J(2, col));
    }
  }
  return 0;
}

// Materials on the 2D symmetric 3-vector s = (S00, S11, S01), the 3-vector
// API of stable_neohookean.cpp:236-277 (PSD fix), arap.cpp:67-98,
// corotational.cpp:62-96 and fixed_corotational.cpp:93-132 (closed forms of
// their symbolic expressions, q = 1 - det S = 1 - S00 S11 + S01^2). Neo-Hookean
// and Fung have no 3-vector API (material_model.h, base "unimplemented"):
// false.
bool material_s3(int type, double mu, double la, const Vec<3>& s, bool psd_fix, double* E,
                 Vec<3>* g, Mat<3, 3>* H) {
  const double S1 = s[0], S2 = s[1], S3 = s[2];
  const double q = 1.0 - S1 * S2 + S3 * S3;
  Vec<3> dq; dq[0] = -S2; dq[1] = -S1; dq[2] = 2.0 * S3;
  Mat<3, 3> d2q; d2q(0, 1) = d2q(1, 0) = -1.0; d2q(2, 2) = 2.0;
  const double v[3] = {1, 1, 2};
  double e = 0;
  Vec<3> gg;
  Mat<3, 3> HH;
  switch (type) {
    case MAT_SNH: {
      e = 0.5 * la * q * q + 0.5 * mu * (S1 * S1 + S2 * S2 + 2.0 * S3 * S3 - 2.0) + mu * q;
      for (int i = 0; i < 3; ++i) gg[i] = (la * q + mu) * dq[i] + mu * v[i] * s[i];
      HH = la * (dq * dq.T()) + (la * q + mu) * d2q;
      for (int i = 0; i < 3; ++i) HH(i, i) += mu * v[i];
      if (psd_fix) simple_psd_fix<3>(HH);
      break;
    }
    case MAT_ARAP: {
      e = 0.5 * mu * ((S1 - 1) * (S1 - 1) + (S2 - 1) * (S2 - 1) + 2.0 * S3 * S3);
      gg[0] = mu * (S1 - 1); gg[1] = mu * (S2 - 1); gg[2] = 2.0 * mu * S3;
      HH(0, 0) = mu; HH(1, 1) = mu; HH(2, 2) = 2.0 * mu;
      break;
    }
    case MAT_COROT: {
      const double t = S1 + S2 - 2.0;
      e = 0.5 * la * t * t + mu * ((S1 - 1) * (S1 - 1) + (S2 - 1) * (S2 - 1) + 2.0 * S3 * S3);
      gg[0] = la * t + 2.0 * mu * (S1 - 1); gg[1] = la * t + 2.0 * mu * (S2 - 1);
      gg[2] = 4.0 * mu * S3;
      HH(0, 0) = HH(1, 1) = la + 2.0 * mu; HH(0, 1) = HH(1, 0) = la; HH(2, 2) = 4.0 * mu;
      break;
    }
    case MAT_FCR: {
      e = 0.5 * la * q * q + mu * ((S1 - 1) * (S1 - 1) + (S2 - 1) * (S2 - 1) + 2.0 * S3 * S3);
      gg[0] = la * q * dq[0] + 2.0 * mu * (S1 - 1);
      gg[1] = la * q * dq[1] + 2.0 * mu * (S2 - 1);
      gg[2] = la * q * dq[2] + 4.0 * mu * S3;
      HH = la * (dq * dq.T()) + (la * q) * d2q;
      HH(0, 0) += 2.0 * mu; HH(1, 1) += 2.0 * mu; HH(2, 2) += 4.0 * mu;
      break;
    }
    default:
      return false;
  }
  if (E) *E = e;
  if (g) *g = gg;
  if (H) *H = HH;
  return true;
}

// ---------------------------------------------------------------------------
// Materials on the symmetric 6-vector s = (S00,S11,S22,S01,S02,S12).
// J = det S, with gradient dJ and Hessian d2J with respect to s.
static void det_s6(const Vec<6>& s, double& J, Vec<6>& dJ, Mat<6, 6>& d2J) {
  double a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5];
  J = a * b * c + 2.0 * d * e * f - a * f * f - b * e * e - c * d * d;
  dJ[0] = b * c - f * f;
  dJ[1] = a * c - e * e;
  dJ[2] = a * b - d * d;
  dJ[3] = 2.0 * e * f - 2.0 * c * d;
  dJ[4] = 2.0 * d * f - 2.0 * b * e;
  dJ[5] = 2.0 * d * e - 2.0 * a * f;
  d2J = Mat<6, 6>();
  auto set = [&](int i, int j, double v) { d2J(i, j) = v; d2J(j, i) = v; };
  set(0, 1, c); set(0, 2, b); set(1, 2, a);
  set(0, 5, -2.0 * f); set(1, 4, -2.0 * e); set(2, 3, -2.0 * d);
  set(3, 3, -2.0 * c); set(4, 4, -2.0 * b); set(5, 5, -2.0 * a);
  set(3, 4, 2.0 * f); set(3, 5, 2.0 * e); set(4, 5, 2.0 * d);
}

static const double kVoigt[6] = {1, 1, 1, 2, 2, 2};

void material_s6(int type, double mu, double la, const Vec<6>& s, bool psd_fix,
                 double* E, Vec<6>* g, Mat<6, 6>* H) {
  double J; Vec<6> dJ; Mat<6, 6> d2J;
  det_s6(s, J, dJ, d2J);
  // I2/2 = 0.5*tr(S^T S) and its derivatives
  double trS2 = 0;
  for (int i = 0; i < 6; ++i) trS2 += kVoigt[i] * s[i] * s[i];
  Vec<6> dI; for (int i = 0; i < 6; ++i) dI[i] = kVoigt[i] * s[i];
  Mat<6, 6> D = diag<6>(kVoigt);
  double e = 0; Vec<6> gg; Mat<6, 6> HH;
  switch (type) {
    case MAT_SNH: {  // derivs.m:29 with I2,I3 on S; stable_neohookean.cpp:8-95
      e = 0.5 * mu * (trS2 - 3.0) - mu * (J - 1.0) + 0.5 * la * (J - 1.0) * (J - 1.0);
      gg = mu * dI + (la * (J - 1.0) - mu) * dJ;
      HH = mu * D + (la * (J - 1.0) - mu) * d2J + la * (dJ * dJ.T());
      if (psd_fix) simple_psd_fix<6>(HH);
      break;
    }
    case MAT_NH: {  // derivs.m:50: 0.5 mu (I2 - 3) - mu log J + 0.5 la (log J)^2 (neohookean.cpp:8-101)
      // The Hessian is the symmetric closed form; neohookean.cpp:61 (H(0,2))
      // carries a typo (S2*3 for S2*S3) that makes the reference's matrix
      // asymmetric, which this restatement does not reproduce.
      const double lJ = std::log(J);
      e = 0.5 * mu * (trS2 - 3.0) - mu * lJ + 0.5 * la * lJ * lJ;
      const double k1 = (la * lJ - mu) / J, k2 = (la + mu - la * lJ) / (J * J);
      gg = mu * dI + k1 * dJ;
      HH = mu * D + k2 * (dJ * dJ.T()) + k1 * d2J;
      if (psd_fix) simple_psd_fix<6>(HH);
      break;
    }
    case MAT_FUNG: {  // derivs.m:41-42 with c = 4 (fung.cpp:7-97)
      const double cf = 4.0;
      const double ex = std::exp(0.5 * cf * (trS2 - 3.0));
      const double A = 1.0 + (mu + mu * cf) / la;
      e = 0.5 * mu * (trS2 - 3.0) + 0.5 * la * (J - A) * (J - A) + 0.5 * mu * (ex - 1.0);
      const double a = mu + 0.5 * cf * mu * ex;
      gg = a * dI + (la * (J - A)) * dJ;
      HH = a * D + (0.5 * cf * cf * mu * ex) * (dI * dI.T()) + la * (dJ * dJ.T()) + (la * (J - A)) * d2J;
      if (psd_fix) simple_psd_fix<6>(HH);
      break;
    }
    case MAT_ARAP: {  // arap.cpp:7-65 (no PSD fix; constant Hessian)
      Vec<6> d = s; d[0] -= 1; d[1] -= 1; d[2] -= 1;
      for (int i = 0; i < 6; ++i) e += kVoigt[i] * d[i] * d[i];
      e *= 0.5 * mu;
      for (int i = 0; i < 6; ++i) gg[i] = mu * kVoigt[i] * d[i];
      HH = mu * D;
      break;
    }
    case MAT_COROT: {  // corotational.cpp:8-56
      double t = s[0] + s[1] + s[2] - 3.0;
      Vec<6> d = s; d[0] -= 1; d[1] -= 1; d[2] -= 1;
      double q = 0; for (int i = 0; i < 6; ++i) q += kVoigt[i] * d[i] * d[i];
      e = 0.5 * la * t * t + mu * q;
      for (int i = 0; i < 6; ++i) gg[i] = 2.0 * mu * kVoigt[i] * d[i] + (i < 3 ? la * t : 0.0);
      HH = (2.0 * mu) * D;
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) HH(i, j) += la;
      break;
    }
    case MAT_FCR: {  // fixed_corotational.cpp:8-90 (no PSD fix)
      Vec<6> d = s; d[0] -= 1; d[1] -= 1; d[2] -= 1;
      double q = 0; for (int i = 0; i < 6; ++i) q += kVoigt[i] * d[i] * d[i];
      e = 0.5 * la * (J - 1.0) * (J - 1.0) + mu * q;
      for (int i = 0; i < 6; ++i) gg[i] = 2.0 * mu * kVoigt[i] * d[i];
      gg += (la * (J - 1.0)) * dJ;
      HH = (2.0 * mu) * D + (la * (J - 1.0)) * d2J + la * (dJ * dJ.T());
      break;
    }
    default: {  // unsupported here: report NaN so callers notice
      e = NAN;
      for (int i = 0; i < 6; ++i) gg[i] = NAN;
      for (int i = 0; i < 36; ++i) HH[i] = NAN;
    }
  }
  if (E) *E = e;
  if (g) *g = gg;
  if (H) *H = HH;
}

// ---------------------------------------------------------------------------
// Plain (displacement-only) SNH on the column-major 9-vector F
// (reference src/energies/stable_neohookean.cpp:97-234; the Hessian is
// PSD-projected with tol 0.0 at :232).
static inline int lc(int i, int j, int k) {  // Levi-Civita
  if (i == j || j == k || i == k) return 0;
  return ((i == 0 && j == 1 && k == 2) || (i == 1 && j == 2 && k == 0) ||
          (i == 2 && j == 0 && k == 1)) ? 1 : -1;
}

void snh_f9(double mu, double la, const Vec<9>& f, double* E, Vec<9>* g,
            Mat<9, 9>* H) {
  M3 F; for (int i = 0; i < 9; ++i) F[i] = f[i];
  double J = det3(F);
  // cofactor: dJ/dF_ij
  M3 C;
  C(0, 0) = F(1, 1) * F(2, 2) - F(1, 2) * F(2, 1);
  C(0, 1) = F(1, 2) * F(2, 0) - F(1, 0) * F(2, 2);
  C(0, 2) = F(1, 0) * F(2, 1) - F(1, 1) * F(2, 0);
  C(1, 0) = F(0, 2) * F(2, 1) - F(0, 1) * F(2, 2);
  C(1, 1) = F(0, 0) * F(2, 2) - F(0, 2) * F(2, 0);
  C(1, 2) = F(0, 1) * F(2, 0) - F(0, 0) * F(2, 1);
  C(2, 0) = F(0, 1) * F(1, 2) - F(0, 2) * F(1, 1);
  C(2, 1) = F(0, 2) * F(1, 0) - F(0, 0) * F(1, 2);
  C(2, 2) = F(0, 0) * F(1, 1) - F(0, 1) * F(1, 0);
  Vec<9> c = flatten(C);
  if (E) *E = 0.5 * mu * (F.squaredNorm() - 3.0) - mu * (J - 1.0) +
              0.5 * la * (J - 1.0) * (J - 1.0);
  double k = la * (J - 1.0) - mu;
  if (g) *g = mu * f + k * c;
  if (H) {
    Mat<9, 9> HH = mu * Mat<9, 9>::Identity() + la * (c * c.T());
    // d2J/dF_ij dF_kl = sum_mn eps_ikm eps_jln F_mn
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
      for (int kk = 0; kk < 3; ++kk) for (int l = 0; l < 3; ++l) {
        double v = 0;
        for (int m = 0; m < 3; ++m) for (int n = 0; n < 3; ++n)
          v += lc(i, kk, m) * lc(j, l, n) * F(m, n);
        HH(i + 3 * j, kk + 3 * l) += k * v;
      }
    simple_psd_fix<9>(HH, 0.0);
    *H = HH;
  }
}

// Plain Neo-Hookean on the 9-vector F (reference neohookean.cpp:96-231, the
// symbolic derivatives of derivs.m:50 on F): Psi = mu/2 (|F|^2 - 3) - mu log J
// + la/2 (log J)^2; g = mu F + a c with a = (la log J - mu) / J and c = dJ/dF;
// H = mu I + (la - la log J + mu) / J^2 c c^T + a d2J/dF2, PSD-projected with
// the default tolerance (:230).
void nh_f9(double mu, double la, const Vec<9>& f, double* E, Vec<9>* g, Mat<9, 9>* H) {
  M3 F; for (int i = 0; i < 9; ++i) F[i] = f[i];
  double J = det3(F);
  M3 C;  // cofactor, dJ/dF
  C(0, 0) = F(1, 1) * F(2, 2) - F(1, 2) * F(2, 1);
  C(0, 1) = F(1, 2) * F(2, 0) - F(1, 0) * F(2, 2);
  C(0, 2) = F(1, 0) * F(2, 1) - F(1, 1) * F(2, 0);
  C(1, 0) = F(0, 2) * F(2, 1) - F(0, 1) * F(2, 2);
  C(1, 1) = F(0, 0) * F(2, 2) - F(0, 2) * F(2, 0);
  C(1, 2) = F(0, 1) * F(2, 0) - F(0, 0) * F(2, 1);
  C(2, 0) = F(0, 1) * F(1, 2) - F(0, 2) * F(1, 1);
  C(2, 1) = F(0, 2) * F(1, 0) - F(0, 0) * F(1, 2);
  C(2, 2) = F(0, 0) * F(1, 1) - F(0, 1) * F(1, 0);
  Vec<9> c = flatten(C);
  const double lj = std::log(J);
  if (E) *E = 0.5 * mu * (F.squaredNorm() - 3.0) - mu * lj + 0.5 * la * lj * lj;
  const double a = (la * lj - mu) / J;
  if (g) *g = mu * f + a * c;
  if (H) {
    Mat<9, 9> HH = mu * Mat<9, 9>::Identity() + ((la - la * lj + mu) / (J * J)) * (c * c.T());
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
      for (int kk = 0; kk < 3; ++kk) for (int l = 0; l < 3; ++l) {
        double v = 0;
        for (int m = 0; m < 3; ++m) for (int n = 0; n < 3; ++n)
          v += lc(i, kk, m) * lc(j, l, n) * F(m, n);
        HH(i + 3 * j, kk + 3 * l) += a * v;
      }
    simple_psd_fix<9>(HH);
    *H = HH;
  }
}

// The 9-vector (displacement-only) API of each material model: SNH and NH
// implement it; Fung returns 0, 0, I (fung.cpp:101-107); Corot, FCR and ARAP
// print "unimplemented" and return uninitialised values
// (material_model.h:43-62), so they are rejected here (false).
bool material_f9(int type, double mu, double la, const Vec<9>& f, double* E, Vec<9>* g,
                 Mat<9, 9>* H) {
  switch (type) {
    case MAT_SNH: snh_f9(mu, la, f, E, g, H); return true;
    case MAT_NH: nh_f9(mu, la, f, E, g, H); return true;
    case MAT_FUNG:
      if (E) *E = 0.0;
      if (g) *g = Vec<9>();
      if (H) *H = Mat<9, 9>::Identity();
      return true;
    default: return false;
  }
}

// Plain 2D SNH on the column-major 4-vector F (stable_neohookean.cpp:279-328;
// no PSD projection).
void snh_f4(double mu, double la, const Vec<4>& f, double* E, Vec<4>* g,
            Mat<4, 4>* H) {
  double F00 = f[0], F10 = f[1], F01 = f[2], F11 = f[3];
  double J = F00 * F11 - F01 * F10;
  Vec<4> c; c[0] = F11; c[1] = -F01; c[2] = -F10; c[3] = F00;
  if (E) *E = 0.5 * mu * (f.squaredNorm() - 2.0) - mu * (J - 1.0) +
              0.5 * la * (J - 1.0) * (J - 1.0);
  double k = la * (J - 1.0) - mu;
  if (g) *g = mu * f + k * c;
  if (H) {
    Mat<4, 4> HH = mu * Mat<4, 4>::Identity() + la * (c * c.T());
    HH(0, 3) += k; HH(3, 0) += k; HH(1, 2) -= k; HH(2, 1) -= k;
    *H = HH;
  }
}

}  // namespace orc
