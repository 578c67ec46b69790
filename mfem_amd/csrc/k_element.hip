// Element kernels of the Newton step: one tetrahedron per lane, structure-of-
// arrays loads (stride nT) so that each load instruction of a wavefront is
// coalesced; vertex positions are gathered through L2 (xyz interleaved).
#include <hip/hip_runtime.h>
#include <type_traits>
#include "device_math.h"
#include "kernels.h"
#include "reduce.h"

namespace mfemhip {

namespace {

constexpr int kElemBlock = 128;

// Gather the 4 corner positions of tet e and form F = sum_k x_k (x) dphi_k.
__device__ __forceinline__ void gather_F(const TetMeshDev& m, int e, const double* __restrict__ x,
                                         int v[4], double xc[12], double F[9]) {
  int4 t = m.T[e];
  v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) xc[3 * k + r] = x[3 * v[k] + r];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double d = m.dX[(3 * k + c) * m.nT + e];
#pragma unroll
      for (int r = 0; r < 3; ++r) F[r + 3 * c] += xc[3 * k + r] * d;
    }
}

__device__ __forceinline__ void unpack_sym6(const double H[21], double A[36]) {
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i + 6 * j] = H[sym6(i, j)];
}

// General 6x6 inverse by LU with partial pivoting (the reference inverts the
// projected Hessian with Eigen's inverse(), mixed_stretch.cpp:162).
template <typename Real>
__device__ __forceinline__ void lu_inverse6(const Real* Ain, Real* X) {
  Real A[36];
  int piv[6];
#pragma unroll
  for (int i = 0; i < 36; ++i) A[i] = Ain[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) piv[i] = i;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int p = k;
    Real best = tabs(A[k + 6 * k]);
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
      if (tabs(A[i + 6 * k]) > best) { best = tabs(A[i + 6 * k]); p = i; }
    // the swap with row p, written as one predicated swap per candidate
    // row so that every index is a compile-time constant (a run-time row
    // index would put A in scratch memory)
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
      if (p == i) {
#pragma unroll
        for (int j = 0; j < 6; ++j) { Real t = A[k + 6 * j]; A[k + 6 * j] = A[i + 6 * j]; A[i + 6 * j] = t; }
        int t = piv[k]; piv[k] = piv[i]; piv[i] = t;
      }
    Real id = Real(1) / A[k + 6 * k];
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      A[i + 6 * k] *= id;
      Real l = A[i + 6 * k];
#pragma unroll
      for (int j = k + 1; j < 6; ++j) A[i + 6 * j] -= l * A[k + 6 * j];
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    Real y[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      Real s = (piv[i] == c) ? Real(1) : Real(0);
#pragma unroll
      for (int k = 0; k < i; ++k) s -= A[i + 6 * k] * y[k];
      y[i] = s;
    }
#pragma unroll
    for (int i = 5; i >= 0; --i) {
      Real s = y[i];
#pragma unroll
      for (int k = i + 1; k < 6; ++k) s -= A[i + 6 * k] * X[k + 6 * c];
      X[i + 6 * c] = s / A[i + 6 * i];
    }
  }
}

// x = A^-1 b for a general 6x6 A by LU with partial pivoting (the factor of
// lu_inverse6, applied to one right-hand side instead of forming A^-1).
__device__ __forceinline__ void lu_solve6(const double* Ain, const double* b, double* x) {
  double A[36], y[6];
#pragma unroll
  for (int i = 0; i < 36; ++i) A[i] = Ain[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) y[i] = b[i];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int p = k;
    double best = fabs(A[k + 6 * k]);
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
      if (fabs(A[i + 6 * k]) > best) { best = fabs(A[i + 6 * k]); p = i; }
    // predicated swaps with compile-time indices (as in lu_inverse6)
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
      if (p == i) {
#pragma unroll
        for (int j = 0; j < 6; ++j) { double t = A[k + 6 * j]; A[k + 6 * j] = A[i + 6 * j]; A[i + 6 * j] = t; }
        double t = y[k]; y[k] = y[i]; y[i] = t;
      }
    const double id = 1.0 / A[k + 6 * k];
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const double l = A[i + 6 * k] * id;
#pragma unroll
      for (int j = k + 1; j < 6; ++j) A[i + 6 * j] -= l * A[k + 6 * j];
      y[i] -= l * y[k];
    }
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) s -= A[i + 6 * k] * x[k];
    x[i] = s / A[i + 6 * i];
  }
}

// Projected, h^2-scaled material Hessian (full 6x6) from s, in FP64 in every
// precision mode: the projection clamps eigenvalues to 1e-6 against entries
// of order lambda + mu (up to ~1e7), a spread FP32 cannot hold, and the local
// solve needs this matrix and its inverse to be inverses of each other to the
// working precision (ds = Hpsi^-1 (Hpsi v) cancels exactly only then).
__device__ __forceinline__ void projected_hessian(int type, double mu, double la,
                                                  const double s[6], double h2, double g[6],
                                                  double Hs[36]) {
  double gg[6], HH[21];
  material_grad_hess(type, mu, la, s, gg, HH);
  if (material_needs_psd(type) && !chol_pd_packed<double, 6>(HH, 1e-6))
    jacobi_psd_project_packed<double, 6>(HH, 1e-6);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) Hs[i + 6 * j] = h2 * HH[sym6(i, j)];
#pragma unroll
  for (int i = 0; i < 6; ++i) g[i] = h2 * gg[i];
}

__device__ __forceinline__ constexpr int blk_index(int k, int l) {
  // upper pair (k <= l) of the 4 tet corners -> 0..9
  return k * 4 - (k * (k - 1)) / 2 + (l - k);
}

template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_mixed_polar_material(
    TetMeshDev m, const double* __restrict__ x, double* __restrict__ R, double* __restrict__ S,
    const double* __restrict__ s, Real* __restrict__ dsdF, HessOut<Real> ho,
    double* __restrict__ g, double* __restrict__ gl, double h2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.nT) return;
  const int nT = m.nT;
  int v[4];
  double xc[12], F[9];
  gather_F(m, e, x, v, xc, F);
  double Rl[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Rl[i] = R[i * nT + e];
  procrustes(Rl, F, 1e-6, 100);
#pragma unroll
  for (int i = 0; i < 9; ++i) R[i * nT + e] = Rl[i];
  double Sm[9], sv[6];
  stretch_of(Rl, F, sv, Sm);
#pragma unroll
  for (int i = 0; i < 6; ++i) S[i * nT + e] = sv[i];
  {
    Real d[54];
    dsdF_of<Real>(Rl, F, d);
#pragma unroll
    for (int i = 0; i < 54; ++i) dsdF[i * nT + e] = d[i];
  }
  double se[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) se[i] = s[i * nT + e];
  double gg[6], Hs[36];
  projected_hessian(m.mat[e], m.mu[e], m.la[e], se, h2, gg, Hs);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) {
      const int k = sym6(i, j) * nT + e;
      ho.Hpsi[k] = (Real)Hs[i + 6 * j];
      if (ho.Hloc) ho.Hloc[k] = Hs[i + 6 * j];
    }
  if (ho.Hinv || ho.Hlinv) {  // uniform over the launch
    double Hi[36];
    lu_inverse6<double>(Hs, Hi);
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j) {
        const int k = sym6(i, j) * nT + e;
        const double hi = 0.5 * (Hi[i + 6 * j] + Hi[j + 6 * i]);
        if (ho.Hinv) ho.Hinv[k] = (Real)hi;
        if (ho.Hlinv) ho.Hlinv[k] = hi;
      }
  }
  const double syminv[6] = {1, 1, 1, 0.5, 0.5, 0.5};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double a = gg[i];
#pragma unroll
    for (int j = 0; j < 6; ++j) a += Hs[i + 6 * j] * (sv[j] - se[j]);
    g[i * nT + e] = gg[i];
    gl[i * nT + e] = syminv[i] * a;
  }
}

// Element outputs of k_mixed_blocks_rhs leave through LDS: tet e's block
// record (90 values) starts 90 values after tet e-1's, so a store from every
// lane would touch 64 cache lines per instruction. Each wave stages a round
// of its 64 tets' values (an odd pitch keeps the LDS banks apart) and writes
// the round back as contiguous runs.
constexpr int kStagePitch = 37;  // widest round: 36 values per tet

template <int N, int PITCH, typename V, int TETS = kWave>
__device__ __forceinline__ void stage_flush(const V* stage, int lane, int e0, int nT, V* dst,
                                            int stride, int off) {
  __syncthreads();
  for (int idx = lane; idx < TETS * N; idx += kWave) {
    const int t = idx / N, j = idx - t * N;
    if (e0 + t < nT) dst[(size_t)(e0 + t) * stride + off + j] = stage[t * PITCH + j];
  }
  __syncthreads();
}

// B = dsdF * Jloc (6 x 12): B[i][3k+r] = sum_c dsdF[i][r+3c] dX(k,c), and
// C = Hpsi * B (6 x 12), both column-major (entry (i, q) at i + 6 q).
template <typename Real>
__device__ __forceinline__ void mixed_element_BC(const TetMeshDev& m, int ec,
                                                 const Real* __restrict__ dsdF,
                                                 const Real* __restrict__ Hpsi, Real (&B)[72],
                                                 Real (&C)[72]) {
  const int nT = m.nT;
  {
    Real d[54];
#pragma unroll
    for (int i = 0; i < 54; ++i) d[i] = dsdF[i * nT + ec];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      Real dx0 = (Real)m.dX[(3 * k + 0) * nT + ec];
      Real dx1 = (Real)m.dX[(3 * k + 1) * nT + ec];
      Real dx2 = (Real)m.dX[(3 * k + 2) * nT + ec];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int i = 0; i < 6; ++i)
          B[i + 6 * (3 * k + r)] = d[i + 6 * (r + 0)] * dx0 + d[i + 6 * (r + 3)] * dx1 +
                                   d[i + 6 * (r + 6)] * dx2;
    }
  }
  Real H[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) {
      Real h = Hpsi[sym6(i, j) * nT + ec];
      H[i + 6 * j] = h;
      H[j + 6 * i] = h;
    }
#pragma unroll
  for (int q = 0; q < 12; ++q)
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      Real a = 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) a += H[i + 6 * j] * B[j + 6 * q];
      C[i + 6 * q] = a;
    }
}

// Element residual -vol B^T (Sym gl) - M_e (x - x_tilde - h^2 g) of the
// wave's tets, staged through the wave's LDS (`wave_raw`, at least
// 64 x 13 doubles) and written as contiguous runs.
template <typename Real>
__device__ __forceinline__ void mixed_element_fe(const TetMeshDev& m, int ec, int lane, int e0,
                                                 const double* __restrict__ x, const Inertia& in,
                                                 const double* __restrict__ gl, const Real (&B)[72],
                                                 double vol, unsigned char* wave_raw,
                                                 double* __restrict__ fe) {
  const int nT = m.nT;
  double w[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) w[i] = (i < 3 ? 1.0 : 2.0) * gl[i * nT + ec];
  int4 t = m.T[ec];
  int v[4] = {t.x, t.y, t.z, t.w};
  double d[12], dsum[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double dd = x[3 * v[k] + r] - in.x_tilde[3 * v[k] + r] - in.h2g[r];
      d[3 * k + r] = dd;
      dsum[r] += dd;
    }
  const double mc = m.density * vol / 20.0;
  constexpr int kFePitch = 13;
  double* fst = reinterpret_cast<double*>(wave_raw);
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    double a = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) a += (double)B[i + 6 * j] * w[i];
    fst[lane * kFePitch + j] = -vol * a - mc * (d[j] + dsum[j % 3]);
  }
  stage_flush<12, kFePitch>(fst, lane, e0, nT, fe, 12, 0);
}

// FP32 records are staged whole and leave as contiguous runs, 32 tets at a
// time (lanes 0-31, then 32-63): 32 records are 11520 B, exactly 90 lines of
// 128 B, so every run covers whole lines. Staged in parts of a record, the
// runs of a part covered partial cache lines, and the L2 wrote such lines
// back more than once (WRITE_SIZE 1.3x the record bytes at the 960k-tet
// beam); 64 records at a time need twice the LDS (fewer workgroups per CU).
// FP64 records leave in three rounds of 36, 27 and 27 values per tet.
constexpr int kFullPitch = 91;  // floats per staged FP32 record (odd: no bank conflicts)
constexpr int kHalfWave = kWave / 2;

template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_mixed_blocks_rhs(
    TetMeshDev m, const double* __restrict__ x, Inertia in, const Real* __restrict__ dsdF,
    const Real* __restrict__ Hpsi, const double* __restrict__ gl, Real* __restrict__ blocks,
    double* __restrict__ fe) {
  constexpr bool kWhole = sizeof(Real) == sizeof(float);
  constexpr int kPitch = kWhole ? kFullPitch : kStagePitch;
  constexpr int kStageTets = kWhole ? kHalfWave : kWave;  // tets staged per wave at a time
  constexpr int kWaveBytes = kStageTets * kPitch * sizeof(Real);
  __shared__ __align__(16) unsigned char stage_raw[(kElemBlock / kWave) * kWaveBytes];
  static_assert(kWave * 13 * sizeof(double) <= kWaveBytes, "f_e stage exceeds the wave's LDS");
  const int lane = threadIdx.x & (kWave - 1);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int e0 = e - lane;  // first tet of this wave
  const int nT = m.nT;
  // lanes past the end compute on the last tet and store nothing, so that
  // every lane reaches the barriers of the rounds
  const int ec = e < nT ? e : nT - 1;
  unsigned char* wave_raw = stage_raw + (size_t)(threadIdx.x / kWave) * kWaveBytes;
  Real* stage = reinterpret_cast<Real*>(wave_raw);
  const double vol = m.vol[ec];
  Real B[72], C[72];
  mixed_element_BC<Real>(m, ec, dsdF, Hpsi, B, C);
  const Real rv = (Real)vol;
  // the lane's staged record
  const int srow = (kWhole ? (lane & (kHalfWave - 1)) : lane) * kPitch;
  // block (k, l) = vol B_k^T C_l into staged slot `slot`
  auto block_out = [&](int k, int l, int slot) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        Real a = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) a += B[i + 6 * (3 * k + r)] * C[i + 6 * (3 * l + q)];
        stage[srow + 9 * slot + 3 * r + q] = rv * a;
      }
  };
  // record order blk_index(k, l): (0,0..3) | (1,1..3) | (2,2), (2,3), (3,3)
  if constexpr (kWhole) {
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      if (lane / kHalfWave == h) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
          for (int l = k; l < 4; ++l) block_out(k, l, blk_index(k, l));
      }
      stage_flush<90, kPitch, Real, kHalfWave>(stage, lane, e0 + kHalfWave * h, nT, blocks, 90,
                                               0);
    }
  } else {
#pragma unroll
    for (int l = 0; l < 4; ++l) block_out(0, l, l);
    stage_flush<36, kPitch>(stage, lane, e0, nT, blocks, 90, 0);
#pragma unroll
    for (int l = 1; l < 4; ++l) block_out(1, l, l - 1);
    stage_flush<27, kPitch>(stage, lane, e0, nT, blocks, 90, 36);
    block_out(2, 2, 0);
    block_out(2, 3, 1);
    block_out(3, 3, 2);
    stage_flush<27, kPitch>(stage, lane, e0, nT, blocks, 90, 63);
  }
  mixed_element_fe<Real>(m, ec, lane, e0, x, in, gl, B, vol, wave_raw, fe);
}

// The same element blocks with the products B_k^T C_l on the matrix cores
// (selectable; the vector-unit kernel above is the default: at the 960k-tet
// beam it takes 236 us per launch against 255 us for this one, both writing
// the records as whole contiguous runs):
// v_mfma_f32_4x4x1f32 runs 16 independent 4x4 outer products per wave, one
// per 4-lane group. The 12 x 12 product B^T C of a tet is cut into 4 x 4
// tiles; its 6 upper tiles (I <= J) hold every entry of the 10 blocks
// (k <= l) except the lower entries of the diagonal blocks that fall into a
// lower tile, which are taken from their transposes. A wave stages B and C of
// 16 tets at a time in LDS and runs, per upper tile, 6 accumulating MFMAs
// (k = 0..5) for those 16 tets, lane 4t + m supplying B_t(k, 4I + m) and
// C_t(k, 4J + m) and receiving column 4J + m of the tile (rows in its four
// accumulators). Accumulation runs over k in order in FP32, as on the
// vector units; the results may differ from them in the last bit.
constexpr int kMfmaTets = 16;              // tets per MFMA round of a wave
constexpr int kMfmaBCPitch = 73;           // floats per tet of staged B (and C)
constexpr int kMfmaOutPitch = 91;          // floats per tet of staged blocks
constexpr int kMfmaWaveFloats = kMfmaTets * (2 * kMfmaBCPitch + kMfmaOutPitch);

__global__ __launch_bounds__(kElemBlock) void k_mixed_blocks_rhs_mfma(
    TetMeshDev m, const double* __restrict__ x, Inertia in, const float* __restrict__ dsdF,
    const float* __restrict__ Hpsi, const double* __restrict__ gl, float* __restrict__ blocks,
    double* __restrict__ fe) {
  __shared__ __align__(16) float lds[(kElemBlock / kWave) * kMfmaWaveFloats];
  static_assert(kWave * 13 * sizeof(double) <= kMfmaWaveFloats * sizeof(float),
                "f_e stage exceeds the wave's LDS");
  const int lane = threadIdx.x & (kWave - 1);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int e0 = e - lane;
  const int nT = m.nT;
  const int ec = e < nT ? e : nT - 1;
  float* wl = lds + (threadIdx.x / kWave) * kMfmaWaveFloats;
  float* sB = wl;
  float* sC = wl + kMfmaTets * kMfmaBCPitch;
  float* sOut = wl + 2 * kMfmaTets * kMfmaBCPitch;
  const double vol = m.vol[ec];
  float B[72], C[72];
  mixed_element_BC<float>(m, ec, dsdF, Hpsi, B, C);
  const int t = lane >> 2, mm = lane & 3;  // MFMA block (tet of the round) and lane in it
  for (int round = 0; round < kWave / kMfmaTets; ++round) {
    // stage B and C (row k = 0..5, column 0..11) of this round's 16 tets
    if ((lane >> 4) == round) {
      const int tt = lane & (kMfmaTets - 1);
#pragma unroll
      for (int i = 0; i < 72; ++i) {
        const int k = i % 6, col = i / 6;
        sB[tt * kMfmaBCPitch + 12 * k + col] = B[i];
        sC[tt * kMfmaBCPitch + 12 * k + col] = C[i];
      }
    }
    __syncthreads();
    const int et = e0 + round * kMfmaTets + t;  // tet of this lane's MFMA block
    const float rv = (float)m.vol[et < nT ? et : nT - 1];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
      for (int J = I; J < 3; ++J) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const float a = sB[t * kMfmaBCPitch + 12 * k + 4 * I + mm];
          const float b = sC[t * kMfmaBCPitch + 12 * k + 4 * J + mm];
          acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, acc, 0, 0, 0);
        }
        // acc[r] = (B^T C)(4I + r, 4J + mm)
        const int col = 4 * J + mm, l = col / 3, q = col % 3;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 4 * I + r, k = row / 3, rr = row % 3;
          const float v = rv * acc[r];
          if (k < l || (k == l && I <= J))
            sOut[t * kMfmaOutPitch + 9 * blk_index(k, l) + 3 * rr + q] = v;
          // the transposed entry of a diagonal block when its tile is lower
          if (k == l && (row / 4) < (col / 4) && rr != q)
            sOut[t * kMfmaOutPitch + 9 * blk_index(k, k) + 3 * q + rr] = v;
        }
      }
    __syncthreads();
    const int eb = e0 + round * kMfmaTets;
    for (int idx = lane; idx < kMfmaTets * 90; idx += kWave) {
      const int tt = idx / 90, j = idx - tt * 90;
      if (eb + tt < nT) blocks[(size_t)eb * 90 + idx] = sOut[tt * kMfmaOutPitch + j];
    }
    __syncthreads();
  }
  mixed_element_fe<float>(m, ec, lane, e0, x, in, gl, B, vol,
                          reinterpret_cast<unsigned char*>(wl), fe);
}

template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_mixed_local_solve(
    TetMeshDev m, const double* __restrict__ dx, const Real* __restrict__ dsdF,
    const double* __restrict__ Hpsi, const double* __restrict__ Hinv, const double* __restrict__ g,
    const double* __restrict__ gl, double* __restrict__ lam, double* __restrict__ ds,
    ReduceSlot max_ds) {
  __shared__ double sm[kElemBlock / kWave];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double local_max = 0.0;
  if (e < m.nT) {
    const int nT = m.nT;
    int4 t = m.T[e];
    int v[4] = {t.x, t.y, t.z, t.w};
    double dc[12];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int f = m.free_map[v[k]];
#pragma unroll
      for (int r = 0; r < 3; ++r) dc[3 * k + r] = f >= 0 ? dx[3 * f + r] : 0.0;
    }
    double Fd[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Fd[i] = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double dd = m.dX[(3 * k + c) * nT + e];
#pragma unroll
        for (int r = 0; r < 3; ++r) Fd[r + 3 * c] += dc[3 * k + r] * dd;
      }
    double y[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double a = 0;
#pragma unroll
      for (int c = 0; c < 9; ++c) a += (double)dsdF[(i + 6 * c) * nT + e] * Fd[c];
      y[i] = a;
    }
    const double syminv[6] = {1, 1, 1, 0.5, 0.5, 0.5};
    const double sym[6] = {1, 1, 1, 2, 2, 2};
    double l[6], rr[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double a = 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) a += Hpsi[sym6(i, j) * nT + e] * y[j];
      l[i] = -gl[i * nT + e] - syminv[i] * a;
      rr[i] = sym[i] * l[i] + g[i * nT + e];
    }
    // ds = -Hpsi^-1 rr: the stored inverse, or (Hinv null) an LU solve
    // with the same FP64 matrix
    double sol[6];
    if (Hinv) {
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double a = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j) a += Hinv[sym6(i, j) * nT + e] * rr[j];
        sol[i] = a;
      }
    } else {
      double Hs[36];
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) Hs[i + 6 * j] = Hpsi[sym6(i, j) * nT + e];
      lu_solve6(Hs, rr, sol);
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      lam[i * nT + e] = l[i];
      ds[i * nT + e] = -sol[i];
      local_max = fmax(local_max, fabs(sol[i]));
    }
  }
  block_partial<kElemBlock, 1>(local_max, max_ds.partials, sm);
}

// Line-search energy of tet e at (x + alpha dx, s + alpha ds): inertia share
// (consistent mass), h^2 vol Psi(s) and the constraint term -vol lambda^T
// Sym (S(x) - s), with S(x) from a polar fit warm-started from R (the stored
// R is not changed).
__device__ __forceinline__ double mixed_energy_tet(const TetMeshDev& m, int e,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ dx, double alpha,
                                                   const Inertia& in, const double* __restrict__ R,
                                                   const double* __restrict__ s,
                                                   const double* __restrict__ ds,
                                                   const double* __restrict__ lam, double h2) {
  const int nT = m.nT;
  int4 t = m.T[e];
  int v[4] = {t.x, t.y, t.z, t.w};
  double q[12], dsum[3] = {0, 0, 0}, dsq = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int f = m.free_map[v[k]];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double qq = x[3 * v[k] + r] + (f >= 0 ? alpha * dx[3 * f + r] : 0.0);
      q[3 * k + r] = qq;
      double d = qq - in.x_tilde[3 * v[k] + r] - in.h2g[r];
      dsum[r] += d;
      dsq += d * d;
    }
  }
  const double vol = m.vol[e];
  double e_in = 0.5 * (m.density * vol / 20.0) *
                (dsq + dsum[0] * dsum[0] + dsum[1] * dsum[1] + dsum[2] * dsum[2]);
  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double d = m.dX[(3 * k + c) * nT + e];
#pragma unroll
      for (int r = 0; r < 3; ++r) F[r + 3 * c] += q[3 * k + r] * d;
    }
  double Rl[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Rl[i] = R[i * nT + e];
  procrustes(Rl, F, 1e-6, 100);
  double Sm[9], sv[6], sa[6];
  stretch_of(Rl, F, sv, Sm);
  double ec = 0;
  const double sym[6] = {1, 1, 1, 2, 2, 2};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    sa[i] = s[i * nT + e] + alpha * ds[i * nT + e];
    ec += lam[i * nT + e] * sym[i] * (sv[i] - sa[i]);
  }
  double psi = material_energy(m.mat[e], m.mu[e], m.la[e], sa);
  return e_in + h2 * vol * psi - vol * ec;
}

__global__ __launch_bounds__(kElemBlock) void k_mixed_energy(
    TetMeshDev m, const double* __restrict__ x, const double* __restrict__ dx, double alpha,
    Inertia in, const double* __restrict__ R, const double* __restrict__ s,
    const double* __restrict__ ds, const double* __restrict__ lam, double h2, ReduceSlot out) {
  __shared__ double sm[kElemBlock / kWave];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double en = 0.0;
  if (e < m.nT) en = mixed_energy_tet(m, e, x, dx, alpha, in, R, s, ds, lam, h2);
  block_partial<kElemBlock, 0>(en, out.partials, sm);
}

// E(alpha0) and E(alpha1) of the line search's first two trials in one
// launch: a workgroup holds the 128 tets of one k_mixed_energy workgroup
// twice, waves 0-1 at alpha0 and waves 2-3 at alpha1, so the second pair
// reads the tets' data from the CU's caches, and each half is reduced as
// k_mixed_energy reduces its workgroup: both sums are bit-identical to two
// separate launches.
__global__ __launch_bounds__(2 * kElemBlock) void k_mixed_energy2(
    TetMeshDev m, const double* __restrict__ x, const double* __restrict__ dx, double alpha0,
    double alpha1, Inertia in, const double* __restrict__ R, const double* __restrict__ s,
    const double* __restrict__ ds, const double* __restrict__ lam, double h2, ReduceSlot out0,
    ReduceSlot out1) {
  static_assert(kElemBlock == 2 * kWave, "each half of the workgroup is two waves");
  __shared__ double sm[2 * kElemBlock / kWave];
  const int half = threadIdx.x / kElemBlock;
  const int e = blockIdx.x * kElemBlock + (threadIdx.x - half * kElemBlock);
  double en = 0.0;
  if (e < m.nT) en = mixed_energy_tet(m, e, x, dx, half ? alpha1 : alpha0, in, R, s, ds, lam, h2);
  en = wave_sum(en);
  if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = en;
  __syncthreads();
  if (threadIdx.x == 0) out0.partials[blockIdx.x] = sm[0] + sm[1];
  if (threadIdx.x == kElemBlock) out1.partials[blockIdx.x] = sm[2] + sm[3];
}

// Batched simulations: energy of every tet with the step length of its
// simulation (tets of simulations that are off get 0).
__global__ __launch_bounds__(kElemBlock) void k_mixed_energy_batch(
    TetMeshDev m, const double* __restrict__ x, const double* __restrict__ dx,
    const int* __restrict__ tet_sim, const double* __restrict__ alpha_sim,
    const int* __restrict__ sim_on, Inertia in, const double* __restrict__ R,
    const double* __restrict__ s, const double* __restrict__ ds, const double* __restrict__ lam,
    double h2, double* __restrict__ e_tet) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.nT) return;
  const int b = tet_sim[e];
  e_tet[e] = sim_on[b] ? mixed_energy_tet(m, e, x, dx, alpha_sim[b], in, R, s, ds, lam, h2) : 0.0;
}

// ---------------------------------------------------------------------------
// Plain SNH on F (reference stable_neohookean.cpp:97-234): energy, gradient,
// and the 9x9 Hessian mu I + la c c^T + (la (J-1) - mu) d2J/dF2, projected
// with tolerance 0 (:232).
__device__ __forceinline__ int eps3(int i, int j, int k) {
  return (i == j || j == k || i == k) ? 0 : (((j - i + 3) % 3 == 1) ? 1 : -1);
}

__device__ __forceinline__ void snh_f9_cof(const double F[9], double& J, double c[9]) {
  c[0] = F[4] * F[8] - F[7] * F[5];
  c[3] = F[7] * F[2] - F[1] * F[8];
  c[6] = F[1] * F[5] - F[4] * F[2];
  c[1] = F[6] * F[5] - F[3] * F[8];
  c[4] = F[0] * F[8] - F[6] * F[2];
  c[7] = F[3] * F[2] - F[0] * F[5];
  c[2] = F[3] * F[7] - F[6] * F[4];
  c[5] = F[6] * F[1] - F[0] * F[7];
  c[8] = F[0] * F[4] - F[3] * F[1];
  J = F[0] * c[0] + F[3] * c[3] + F[6] * c[6];
}

__device__ __forceinline__ double snh_f9_energy(double mu, double la, const double F[9]) {
  double J, c[9];
  snh_f9_cof(F, J, c);
  double n2 = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) n2 += F[i] * F[i];
  return 0.5 * mu * (n2 - 3.0) - mu * (J - 1.0) + 0.5 * la * (J - 1.0) * (J - 1.0);
}

template <typename Real>
__device__ __forceinline__ void snh_f9_grad_hess(double mu, double la, const double F[9],
                                                 double g[9], Real H[81]) {
  double J, c[9];
  snh_f9_cof(F, J, c);
  double k = la * (J - 1.0) - mu;
#pragma unroll
  for (int i = 0; i < 9; ++i) g[i] = mu * F[i] + k * c[i];
#pragma unroll
  for (int p = 0; p < 9; ++p)
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      int i = p % 3, j = p / 3, kk = q % 3, l = q / 3;
      double v = (p == q ? mu : 0.0) + la * c[p] * c[q];
      if (i != kk && j != l) {
        int mm = 3 - i - kk, n = 3 - j - l;
        v += k * (double)(eps3(i, kk, mm) * eps3(j, l, n)) * F[mm + 3 * n];
      }
      H[p + 9 * q] = (Real)v;
    }
}

// Neo-Hookean on F (reference neohookean.cpp:96-231): Psi = mu/2 (|F|^2 - 3)
// - mu log J + la/2 (log J)^2, g = mu F + a c with a = (la log J - mu) / J,
// H = mu I + (la - la log J + mu) / J^2 c c^T + a d2J/dF2, projected with the
// default tolerance (:230).
__device__ __forceinline__ double nh_f9_energy(double mu, double la, const double F[9]) {
  double J, c[9];
  snh_f9_cof(F, J, c);
  double n2 = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) n2 += F[i] * F[i];
  const double lj = log(J);
  return 0.5 * mu * (n2 - 3.0) - mu * lj + 0.5 * la * lj * lj;
}

template <typename Real>
__device__ __forceinline__ void nh_f9_grad_hess(double mu, double la, const double F[9],
                                                double g[9], Real H[81]) {
  double J, c[9];
  snh_f9_cof(F, J, c);
  const double lj = log(J);
  const double a = (la * lj - mu) / J;
  const double cc = (la - la * lj + mu) / (J * J);
#pragma unroll
  for (int i = 0; i < 9; ++i) g[i] = mu * F[i] + a * c[i];
#pragma unroll
  for (int p = 0; p < 9; ++p)
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      int i = p % 3, j = p / 3, kk = q % 3, l = q / 3;
      double v = (p == q ? mu : 0.0) + cc * c[p] * c[q];
      if (i != kk && j != l) {
        int mm = 3 - i - kk, n = 3 - j - l;
        v += a * (double)(eps3(i, kk, mm) * eps3(j, l, n)) * F[mm + 3 * n];
      }
      H[p + 9 * q] = (Real)v;
    }
}

// The 9-vector API of material `type` (SNH, NH, and Fung's constant 0, 0, I,
// fung.cpp:101-107; the host admits no other type on this path), with each
// model's PSD projection.
__device__ __forceinline__ double f9_energy(int type, double mu, double la, const double F[9]) {
  if (type == MAT_NH) return nh_f9_energy(mu, la, F);
  if (type == MAT_FUNG) return 0.0;
  return snh_f9_energy(mu, la, F);
}

template <typename Real>
__device__ __forceinline__ void f9_grad_hess(int type, double mu, double la, const double F[9],
                                             double g[9], Real H[81]) {
  if (type == MAT_FUNG) {
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 81; ++i) H[i] = (i % 10 == 0) ? Real(1) : Real(0);
  } else if (type == MAT_NH) {
    nh_f9_grad_hess<Real>(mu, la, F, g, H);
    if (!chol_pd<Real, 9>(H, Real(1e-6))) jacobi_psd_project<Real, 9>(H, Real(1e-6));
  } else {
    snh_f9_grad_hess<Real>(mu, la, F, g, H);
    if (!chol_pd<Real, 9>(H, Real(0))) jacobi_psd_project<Real, 9>(H, Real(0));
  }
}

template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_plain_update(
    TetMeshDev m, const double* __restrict__ x, Inertia in, double h2, Real* __restrict__ blocks,
    double* __restrict__ fe) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.nT) return;
  const int nT = m.nT;
  int v[4];
  double xc[12], F[9];
  gather_F(m, e, x, v, xc, F);
  double g[9];
  Real H[81];
  f9_grad_hess<Real>(m.mat[e], m.mu[e], m.la[e], F, g, H);
  const double vol = m.vol[e];
  const Real sc = (Real)(vol * h2);
  Real dxr[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) dxr[i] = (Real)m.dX[i * nT + e];
  Real* out = blocks + (size_t)e * 90;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    // C = H Jloc_l (9 x 3): C[p][q] = sum_c2 H[p][q + 3 c2] dX(l, c2)
    Real C[27];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int p = 0; p < 9; ++p)
        C[p + 9 * q] = H[p + 9 * (q + 0)] * dxr[3 * l + 0] + H[p + 9 * (q + 3)] * dxr[3 * l + 1] +
                       H[p + 9 * (q + 6)] * dxr[3 * l + 2];
#pragma unroll
    for (int k = 0; k <= l; ++k) {
      Real* o = out + 9 * blk_index(k, l);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          o[3 * r + q] = sc * (dxr[3 * k + 0] * C[(r + 0) + 9 * q] + dxr[3 * k + 1] * C[(r + 3) + 9 * q] +
                               dxr[3 * k + 2] * C[(r + 6) + 9 * q]);
    }
  }
  double d[12], dsum[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double dd = xc[3 * k + r] - in.x_tilde[3 * v[k] + r] - in.h2g[r];
      d[3 * k + r] = dd;
      dsum[r] += dd;
    }
  const double mc = m.density * vol / 20.0;
  double* fo = fe + (size_t)e * 12;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double a = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) a += m.dX[(3 * k + c) * nT + e] * g[r + 3 * c];
      fo[3 * k + r] = -vol * h2 * a - mc * (d[3 * k + r] + dsum[r]);
    }
}

__global__ __launch_bounds__(kElemBlock) void k_plain_energy(
    TetMeshDev m, const double* __restrict__ x, const double* __restrict__ dx, double alpha,
    Inertia in, double h2, ReduceSlot out) {
  __shared__ double sm[kElemBlock / kWave];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double en = 0.0;
  if (e < m.nT) {
    const int nT = m.nT;
    int4 t = m.T[e];
    int v[4] = {t.x, t.y, t.z, t.w};
    double q[12], dsum[3] = {0, 0, 0}, dsq = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int f = m.free_map[v[k]];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double qq = x[3 * v[k] + r] + (f >= 0 ? alpha * dx[3 * f + r] : 0.0);
        q[3 * k + r] = qq;
        double d = qq - in.x_tilde[3 * v[k] + r] - in.h2g[r];
        dsum[r] += d;
        dsq += d * d;
      }
    }
    const double vol = m.vol[e];
    double e_in = 0.5 * (m.density * vol / 20.0) *
                  (dsq + dsum[0] * dsum[0] + dsum[1] * dsum[1] + dsum[2] * dsum[2]);
    double F[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double d = m.dX[(3 * k + c) * nT + e];
#pragma unroll
        for (int r = 0; r < 3; ++r) F[r + 3 * c] += q[3 * k + r] * d;
      }
    en = e_in + h2 * vol * f9_energy(m.mat[e], m.mu[e], m.la[e], F);
  }
  block_partial<kElemBlock, 0>(en, out.partials, sm);
}

// ---------------------------------------------------------------------------
// 2D displacement-only SNH on the 4-vector F (reference
// stable_neohookean.cpp:279-328, no PSD projection) over triangles (corners
// x, y, z of T; dphi/dX in rows 3k + c, k < 3, c < 2; vol = the reference's
// signed doublearea, tri2d_mesh.cpp:32-34). Blocks and residual use the 3D
// layout with zero z coupling, so the z rows of K hold only the mass and dz
// stays zero.
__device__ __forceinline__ void snh_f4_cof(const double F[4], double& J, double c[4]) {
  // F column-major (F00, F10, F01, F11); c = dJ/dF
  J = F[0] * F[3] - F[2] * F[1];
  c[0] = F[3];
  c[1] = -F[2];
  c[2] = -F[1];
  c[3] = F[0];
}

__device__ __forceinline__ double snh_f4_energy(double mu, double la, const double F[4]) {
  double J, c[4];
  snh_f4_cof(F, J, c);
  const double n2 = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  return 0.5 * mu * (n2 - 2.0) - mu * (J - 1.0) + 0.5 * la * (J - 1.0) * (J - 1.0);
}

__device__ __forceinline__ void gather_F2(const TetMeshDev& m, int e, const double* __restrict__ x,
                                          const double* __restrict__ dx, double alpha, int v[3],
                                          double xc[9], double F[4]) {
  const int4 t = m.T[e];
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int f = dx ? m.free_map[v[k]] : -1;
#pragma unroll
    for (int r = 0; r < 3; ++r)
      xc[3 * k + r] = x[3 * v[k] + r] + (f >= 0 ? alpha * dx[3 * f + r] : 0.0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) F[i] = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double d = m.dX[(3 * k + c) * m.nT + e];
#pragma unroll
      for (int r = 0; r < 2; ++r) F[r + 2 * c] += xc[3 * k + r] * d;
    }
}

template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_plain2d_update(
    TetMeshDev m, const double* __restrict__ x, Inertia in, double h2, Real* __restrict__ blocks,
    double* __restrict__ fe) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.nT) return;
  const int nT = m.nT;
  int v[3];
  double xc[9], F[4];
  gather_F2(m, e, x, nullptr, 0.0, v, xc, F);
  double J, c[4];
  snh_f4_cof(F, J, c);
  const double mu = m.mu[e], la = m.la[e];
  const double k = la * (J - 1.0) - mu;
  double g[4], H[16];
#pragma unroll
  for (int i = 0; i < 4; ++i) g[i] = mu * F[i] + k * c[i];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) H[p + 4 * q] = (p == q ? mu : 0.0) + la * c[p] * c[q];
  H[0 + 4 * 3] += k;
  H[3 + 4 * 0] += k;
  H[1 + 4 * 2] -= k;
  H[2 + 4 * 1] -= k;
  const double vol = m.vol[e];
  const double sc = vol * h2;
  double d2[6];  // dphi_k/dX_c, k < 3, c < 2
#pragma unroll
  for (int kk = 0; kk < 3; ++kk)
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) d2[2 * kk + cc] = m.dX[(3 * kk + cc) * nT + e];
  Real* out = blocks + (size_t)e * 90;
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int a = 0; a <= l; ++a) {
      Real* o = out + 9 * blk_index(a, l);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double val = 0.0;
          if (r < 2 && q < 2) {
#pragma unroll
            for (int c1 = 0; c1 < 2; ++c1)
#pragma unroll
              for (int c2 = 0; c2 < 2; ++c2)
                val += d2[2 * a + c1] * H[(r + 2 * c1) + 4 * (q + 2 * c2)] * d2[2 * l + c2];
            val *= sc;
          }
          o[3 * r + q] = (Real)val;
        }
    }
  // residual: -vol h^2 Jloc^T g - M_e (x - x_tilde - h^2 g_ext), M_e = rho vol / 12 (1 + delta)
  double d[9], dsum[3] = {0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 3; ++kk)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double dd = xc[3 * kk + r] - in.x_tilde[3 * v[kk] + r] - in.h2g[r];
      d[3 * kk + r] = dd;
      dsum[r] += dd;
    }
  const double mc = m.density * vol / 12.0;
  double* fo = fe + (size_t)e * 12;
#pragma unroll
  for (int kk = 0; kk < 3; ++kk)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double a = 0.0;
      if (r < 2) a = d2[2 * kk] * g[r] + d2[2 * kk + 1] * g[r + 2];
      fo[3 * kk + r] = -vol * h2 * a - mc * (d[3 * kk + r] + dsum[r]);
    }
#pragma unroll
  for (int r = 0; r < 3; ++r) fo[9 + r] = 0.0;
}

// 2D mixed stretch (the DIM = 2 instantiation of mixed_stretch.cpp:134-332):
// s = (S00, S11, S01), Sym = diag(1, 1, 2); per triangle R (4), S, s, g, gl
// (3 each), dsdF (3 x 4 at (i + 3c)), Hpsi and its inverse (6 packed), all
// structure of arrays with stride nT in the 3D arrays' leading rows.
__device__ __forceinline__ double sym3w(int i) { return i < 2 ? 1.0 : 2.0; }

template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_mixed2d_polar_material(
    TetMeshDev m, const double* __restrict__ x, double* __restrict__ R, double* __restrict__ S,
    const double* __restrict__ s, Real* __restrict__ dsdF, HessOut<Real> ho,
    double* __restrict__ g, double* __restrict__ gl, double h2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.nT) return;
  const int nT = m.nT;
  int v[3];
  double xc[9], F[4];
  gather_F2(m, e, x, nullptr, 0.0, v, xc, F);
  double Rl[4], Sm[4], sv[3], d[12];
  polar2d(F, Rl, Sm, sv, d);
#pragma unroll
  for (int i = 0; i < 4; ++i) R[i * nT + e] = Rl[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) S[i * nT + e] = sv[i];
#pragma unroll
  for (int i = 0; i < 12; ++i) dsdF[i * nT + e] = (Real)d[i];
  double se[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) se[i] = s[i * nT + e];
  const int type = m.mat[e];
  double gg[3], HH[6];
  material2d_grad_hess(type, m.mu[e], m.la[e], se, gg, HH);
  // projection and inverse in FP64 in every precision mode (see
  // projected_hessian above)
  double A[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i + 3 * j] = HH[sym3(i, j)];
  if (type == MAT_SNH && !chol_pd<double, 3>(A, 1e-6)) jacobi_psd_project<double, 3>(A, 1e-6);
  double Hs[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Hs[i] = h2 * A[i];
  // inverse of the symmetric 3x3 by cofactors
  const double c00 = Hs[4] * Hs[8] - Hs[7] * Hs[5], c01 = Hs[7] * Hs[2] - Hs[1] * Hs[8],
               c02 = Hs[1] * Hs[5] - Hs[4] * Hs[2];
  const double det = Hs[0] * c00 + Hs[3] * c01 + Hs[6] * c02;
  double Hi[9];
  Hi[0] = c00 / det; Hi[1] = c01 / det; Hi[2] = c02 / det;
  Hi[3] = (Hs[6] * Hs[5] - Hs[3] * Hs[8]) / det;
  Hi[4] = (Hs[0] * Hs[8] - Hs[6] * Hs[2]) / det;
  Hi[5] = (Hs[3] * Hs[2] - Hs[0] * Hs[5]) / det;
  Hi[6] = (Hs[3] * Hs[7] - Hs[6] * Hs[4]) / det;
  Hi[7] = (Hs[6] * Hs[1] - Hs[0] * Hs[7]) / det;
  Hi[8] = (Hs[0] * Hs[4] - Hs[3] * Hs[1]) / det;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      const int k = sym3(i, j) * nT + e;
      const double hi = 0.5 * (Hi[i + 3 * j] + Hi[j + 3 * i]);
      ho.Hpsi[k] = (Real)Hs[i + 3 * j];
      if (ho.Hinv) ho.Hinv[k] = (Real)hi;
      if (ho.Hloc) ho.Hloc[k] = Hs[i + 3 * j];
      if (ho.Hlinv) ho.Hlinv[k] = hi;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double a = h2 * gg[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) a += Hs[i + 3 * j] * (sv[j] - se[j]);
    g[i * nT + e] = h2 * gg[i];
    gl[i * nT + e] = (i < 2 ? 1.0 : 0.5) * a;
  }
}

// B = dsdF Jloc (3 x 9 over the 3 corners' xyz; z columns zero):
// B(i, 3k + r) = sum_c dsdF(i, r + 2c) dX(k, c) for r < 2.
template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_mixed2d_blocks_rhs(
    TetMeshDev m, const double* __restrict__ x, Inertia in, const Real* __restrict__ dsdF,
    const Real* __restrict__ Hpsi, const double* __restrict__ gl, Real* __restrict__ blocks,
    double* __restrict__ fe) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.nT) return;
  const int nT = m.nT;
  const double vol = m.vol[e];
  double d2[6];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 2; ++c) d2[2 * k + c] = m.dX[(3 * k + c) * nT + e];
  double B[27];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double a = 0.0;
        if (r < 2)
#pragma unroll
          for (int c = 0; c < 2; ++c) a += (double)dsdF[(i + 3 * (r + 2 * c)) * nT + e] * d2[2 * k + c];
        B[i + 3 * (3 * k + r)] = a;
      }
  double H[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) H[i + 3 * j] = (double)Hpsi[sym3(i, j) * nT + e];
  Real* out = blocks + (size_t)e * 90;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    double C[9];  // H B_l (3 x 3)
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double a = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) a += H[i + 3 * j] * B[j + 3 * (3 * l + q)];
        C[i + 3 * q] = a;
      }
#pragma unroll
    for (int k = 0; k <= l; ++k) {
      Real* o = out + 9 * blk_index(k, l);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double a = 0;
#pragma unroll
          for (int i = 0; i < 3; ++i) a += B[i + 3 * (3 * k + r)] * C[i + 3 * q];
          o[3 * r + q] = (Real)(vol * a);
        }
    }
  }
  // residual: -vol B^T (Sym gl) - M_e (x - x_tilde - h^2 g), M_e = rho vol / 12 (1 + delta)
  double w[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) w[i] = sym3w(i) * gl[i * nT + e];
  const int4 t = m.T[e];
  const int v[3] = {t.x, t.y, t.z};
  double dd[9], dsum[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double a = x[3 * v[k] + r] - in.x_tilde[3 * v[k] + r] - in.h2g[r];
      dd[3 * k + r] = a;
      dsum[r] += a;
    }
  const double mc = m.density * vol / 12.0;
  double* fo = fe + (size_t)e * 12;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    double a = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) a += B[i + 3 * j] * w[i];
    fo[j] = -vol * a - mc * (dd[j] + dsum[j % 3]);
  }
#pragma unroll
  for (int j = 9; j < 12; ++j) fo[j] = 0.0;
}

template <typename Real>
__global__ __launch_bounds__(kElemBlock) void k_mixed2d_local_solve(
    TetMeshDev m, const double* __restrict__ dx, const Real* __restrict__ dsdF,
    const double* __restrict__ Hpsi, const double* __restrict__ Hinv, const double* __restrict__ g,
    const double* __restrict__ gl, double* __restrict__ lam, double* __restrict__ ds,
    ReduceSlot max_ds) {
  __shared__ double sm[kElemBlock / kWave];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double local_max = 0.0;
  if (e < m.nT) {
    const int nT = m.nT;
    const int4 t = m.T[e];
    const int v[3] = {t.x, t.y, t.z};
    double Fd[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int f = m.free_map[v[k]];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double d = m.dX[(3 * k + c) * nT + e];
#pragma unroll
        for (int r = 0; r < 2; ++r) Fd[r + 2 * c] += (f >= 0 ? dx[3 * f + r] : 0.0) * d;
      }
    }
    double y[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double a = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) a += (double)dsdF[(i + 3 * c) * nT + e] * Fd[c];
      y[i] = a;
    }
    double l[3], rr[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double a = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) a += Hpsi[sym3(i, j) * nT + e] * y[j];
      l[i] = -gl[i * nT + e] - (i < 2 ? 1.0 : 0.5) * a;
      rr[i] = sym3w(i) * l[i] + g[i * nT + e];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double a = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) a += Hinv[sym3(i, j) * nT + e] * rr[j];
      lam[i * nT + e] = l[i];
      ds[i * nT + e] = -a;
      local_max = fmax(local_max, fabs(a));
    }
  }
  block_partial<kElemBlock, 1>(local_max, max_ds.partials, sm);
}

// Line-search energy of triangle e at (x + alpha dx, s + alpha ds): inertia,
// h^2 vol Psi(s) and -vol lambda^T Sym (S(x) - s), S(x) by the 2D polar fit.
__global__ __launch_bounds__(kElemBlock) void k_mixed2d_energy(
    TetMeshDev m, const double* __restrict__ x, const double* __restrict__ dx, double alpha,
    Inertia in, const double* __restrict__ s, const double* __restrict__ ds,
    const double* __restrict__ lam, double h2, ReduceSlot out) {
  __shared__ double sm[kElemBlock / kWave];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double en = 0.0;
  if (e < m.nT) {
    const int nT = m.nT;
    int v[3];
    double xc[9], F[4];
    gather_F2(m, e, x, dx, alpha, v, xc, F);
    double dsum[3] = {0, 0, 0}, dsq = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double dd = xc[3 * k + r] - in.x_tilde[3 * v[k] + r] - in.h2g[r];
        dsum[r] += dd;
        dsq += dd * dd;
      }
    const double vol = m.vol[e];
    const double e_in = 0.5 * (m.density * vol / 12.0) *
                        (dsq + dsum[0] * dsum[0] + dsum[1] * dsum[1] + dsum[2] * dsum[2]);
    double Rl[4], Sm[4], st[3];
    polar2d(F, Rl, Sm, st, nullptr);
    double si[3], c = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      si[i] = s[i * nT + e] + alpha * ds[i * nT + e];
      c += lam[i * nT + e] * sym3w(i) * (st[i] - si[i]);
    }
    en = e_in + h2 * vol * material2d_energy(m.mat[e], m.mu[e], m.la[e], si) - vol * c;
  }
  block_partial<kElemBlock, 0>(en, out.partials, sm);
}

__global__ __launch_bounds__(kElemBlock) void k_plain2d_energy(
    TetMeshDev m, const double* __restrict__ x, const double* __restrict__ dx, double alpha,
    Inertia in, double h2, ReduceSlot out) {
  __shared__ double sm[kElemBlock / kWave];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double en = 0.0;
  if (e < m.nT) {
    int v[3];
    double xc[9], F[4];
    gather_F2(m, e, x, dx, alpha, v, xc, F);
    double dsum[3] = {0, 0, 0}, dsq = 0;
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double dd = xc[3 * kk + r] - in.x_tilde[3 * v[kk] + r] - in.h2g[r];
        dsum[r] += dd;
        dsq += dd * dd;
      }
    const double vol = m.vol[e];
    const double e_in = 0.5 * (m.density * vol / 12.0) *
                        (dsq + dsum[0] * dsum[0] + dsum[1] * dsum[1] + dsum[2] * dsum[2]);
    en = e_in + h2 * vol * snh_f4_energy(m.mu[e], m.la[e], F);
  }
  block_partial<kElemBlock, 0>(en, out.partials, sm);
}

// ---------------------------------------------------------------------------
// Batched element evaluations (points in AoS, column-major per point).
__global__ void k_batch_polar(int n, const double* __restrict__ F, double* __restrict__ R,
                              double* __restrict__ s, double* __restrict__ dsdF,
                              int* __restrict__ iters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double Fl[9], Rl[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) { Fl[k] = F[9 * i + k]; Rl[k] = R[9 * i + k]; }
  int it = procrustes(Rl, Fl, 1e-6, 100);
  double Sm[9], sv[6];
  stretch_of(Rl, Fl, sv, Sm);
#pragma unroll
  for (int k = 0; k < 9; ++k) R[9 * i + k] = Rl[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) s[6 * i + k] = sv[k];
  if (dsdF) {
    double d[54];
    dsdF_of<double>(Rl, Fl, d);
#pragma unroll
    for (int k = 0; k < 54; ++k) dsdF[54 * i + k] = d[k];
  }
  if (iters) iters[i] = it;
}

__global__ void k_batch_material(int n, const int* __restrict__ type, const double* __restrict__ mu,
                                 const double* __restrict__ la, const double* __restrict__ s,
                                 int psd, double* __restrict__ E, double* __restrict__ g,
                                 double* __restrict__ H) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sv[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sv[k] = s[6 * i + k];
  if (E) E[i] = material_energy(type[i], mu[i], la[i], sv);
  double gg[6], HH[21];
  material_grad_hess(type[i], mu[i], la[i], sv, gg, HH);
  if (psd && material_needs_psd(type[i]) && !chol_pd_packed<double, 6>(HH, 1e-6))
    jacobi_psd_project_packed<double, 6>(HH, 1e-6);
  double A[36];
  unpack_sym6(HH, A);
  if (g)
#pragma unroll
    for (int k = 0; k < 6; ++k) g[6 * i + k] = gg[k];
  if (H)
#pragma unroll
    for (int k = 0; k < 36; ++k) H[36 * i + k] = A[k];
}

__global__ void k_batch_snh_f9(int n, const double* __restrict__ mu, const double* __restrict__ la,
                               const double* __restrict__ F, double* __restrict__ E,
                               double* __restrict__ g, double* __restrict__ H) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double Fl[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Fl[k] = F[9 * i + k];
  if (E) E[i] = snh_f9_energy(mu[i], la[i], Fl);
  double gg[9], HH[81];
  snh_f9_grad_hess<double>(mu[i], la[i], Fl, gg, HH);
  if (!chol_pd<double, 9>(HH, 0.0)) jacobi_psd_project<double, 9>(HH, 0.0);
  if (g)
#pragma unroll
    for (int k = 0; k < 9; ++k) g[9 * i + k] = gg[k];
  if (H)
#pragma unroll
    for (int k = 0; k < 81; ++k) H[81 * i + k] = HH[k];
}

inline int grid_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

int reduce_grid(long n) { return n > 0 ? grid_for(n, kElemBlock) : 1; }

template <typename Real>
void launch_mixed_polar_material(const TetMeshDev& m, const double* x_full, double* R, double* S,
                                 const double* s, Real* dsdF, const HessOut<Real>& ho, double* g,
                                 double* gl, double h2, hipStream_t st) {
  if (m.nT == 0) return;
  hipLaunchKernelGGL(k_mixed_polar_material<Real>, dim3(grid_for(m.nT, kElemBlock)),
                     dim3(kElemBlock), 0, st, m, x_full, R, S, s, dsdF, ho, g, gl, h2);
}

template <typename Real>
void launch_mixed_blocks_rhs(const TetMeshDev& m, const double* x_full, const Inertia& in,
                             const Real* dsdF, const Real* Hpsi, const double* gl, Real* blocks,
                             double* fe, hipStream_t st, bool mfma) {
  if (m.nT == 0) return;
  if constexpr (std::is_same_v<Real, float>) {
    if (mfma) {
      hipLaunchKernelGGL(k_mixed_blocks_rhs_mfma, dim3(grid_for(m.nT, kElemBlock)),
                         dim3(kElemBlock), 0, st, m, x_full, in, dsdF, Hpsi, gl, blocks, fe);
      return;
    }
  }
  hipLaunchKernelGGL(k_mixed_blocks_rhs<Real>, dim3(grid_for(m.nT, kElemBlock)), dim3(kElemBlock),
                     0, st, m, x_full, in, dsdF, Hpsi, gl, blocks, fe);
}

template <typename Real>
void launch_mixed_local_solve(const TetMeshDev& m, const double* dx, const Real* dsdF,
                              const double* Hpsi, const double* Hinv, const double* g,
                              const double* gl, double* lam, double* ds, ReduceSlot max_ds,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_mixed_local_solve<Real>, dim3(reduce_grid(m.nT)), dim3(kElemBlock), 0, st,
                     m, dx, dsdF, Hpsi, Hinv, g, gl, lam, ds, max_ds);
  launch_finalize(max_ds.partials, reduce_grid(m.nT), 1, max_ds.result, st);
}

void launch_mixed_energy_batch(const TetMeshDev& m, const double* x_full, const double* dx,
                               const int* tet_sim, const double* alpha_sim, const int* sim_on,
                               const Inertia& in, const double* R, const double* s,
                               const double* ds, const double* lam, double h2, double* e_tet,
                               hipStream_t st) {
  if (m.nT == 0) return;
  hipLaunchKernelGGL(k_mixed_energy_batch, dim3(grid_for(m.nT, kElemBlock)), dim3(kElemBlock), 0,
                     st, m, x_full, dx, tet_sim, alpha_sim, sim_on, in, R, s, ds, lam, h2, e_tet);
}

void launch_mixed_energy(const TetMeshDev& m, const double* x_full, const double* dx, double alpha,
                         const Inertia& in, const double* R, const double* s, const double* ds,
                         const double* lam, double h2, ReduceSlot out, hipStream_t st) {
  hipLaunchKernelGGL(k_mixed_energy, dim3(reduce_grid(m.nT)), dim3(kElemBlock), 0, st, m, x_full,
                     dx, alpha, in, R, s, ds, lam, h2, out);
  launch_finalize(out.partials, reduce_grid(m.nT), 0, out.result, st);
}

void launch_mixed_energy2(const TetMeshDev& m, const double* x_full, const double* dx,
                          double alpha0, double alpha1, const Inertia& in, const double* R,
                          const double* s, const double* ds, const double* lam, double h2,
                          ReduceSlot out0, ReduceSlot out1, hipStream_t st) {
  hipLaunchKernelGGL(k_mixed_energy2, dim3(reduce_grid(m.nT)), dim3(2 * kElemBlock), 0, st, m,
                     x_full, dx, alpha0, alpha1, in, R, s, ds, lam, h2, out0, out1);
  launch_finalize(out0.partials, reduce_grid(m.nT), 0, out0.result, st);
  launch_finalize(out1.partials, reduce_grid(m.nT), 0, out1.result, st);
}

template <typename Real>
void launch_plain_update(const TetMeshDev& m, const double* x_full, const Inertia& in, double h2,
                         Real* blocks, double* fe, hipStream_t st) {
  if (m.nT == 0) return;
  hipLaunchKernelGGL(k_plain_update<Real>, dim3(grid_for(m.nT, kElemBlock)), dim3(kElemBlock), 0,
                     st, m, x_full, in, h2, blocks, fe);
}

void launch_plain_energy(const TetMeshDev& m, const double* x_full, const double* dx, double alpha,
                         const Inertia& in, double h2, ReduceSlot out, hipStream_t st) {
  hipLaunchKernelGGL(k_plain_energy, dim3(reduce_grid(m.nT)), dim3(kElemBlock), 0, st, m, x_full,
                     dx, alpha, in, h2, out);
  launch_finalize(out.partials, reduce_grid(m.nT), 0, out.result, st);
}

template <typename Real>
void launch_plain2d_update(const TetMeshDev& m, const double* x_full, const Inertia& in,
                           double h2, Real* blocks, double* fe, hipStream_t st) {
  if (m.nT == 0) return;
  hipLaunchKernelGGL(k_plain2d_update<Real>, dim3(grid_for(m.nT, kElemBlock)), dim3(kElemBlock), 0,
                     st, m, x_full, in, h2, blocks, fe);
}

void launch_plain2d_energy(const TetMeshDev& m, const double* x_full, const double* dx,
                           double alpha, const Inertia& in, double h2, ReduceSlot out,
                           hipStream_t st) {
  const int g = reduce_grid(m.nT);
  hipLaunchKernelGGL(k_plain2d_energy, dim3(g), dim3(kElemBlock), 0, st, m, x_full, dx, alpha, in,
                     h2, out);
  launch_finalize(out.partials, g, 0, out.result, st);
}

template <typename Real>
void launch_mixed2d_polar_material(const TetMeshDev& m, const double* x_full, double* R, double* S,
                                   const double* s, Real* dsdF, const HessOut<Real>& ho,
                                   double* g, double* gl, double h2, hipStream_t st) {
  if (m.nT == 0) return;
  hipLaunchKernelGGL(k_mixed2d_polar_material<Real>, dim3(grid_for(m.nT, kElemBlock)),
                     dim3(kElemBlock), 0, st, m, x_full, R, S, s, dsdF, ho, g, gl, h2);
}

template <typename Real>
void launch_mixed2d_blocks_rhs(const TetMeshDev& m, const double* x_full, const Inertia& in,
                               const Real* dsdF, const Real* Hpsi, const double* gl, Real* blocks,
                               double* fe, hipStream_t st) {
  if (m.nT == 0) return;
  hipLaunchKernelGGL(k_mixed2d_blocks_rhs<Real>, dim3(grid_for(m.nT, kElemBlock)), dim3(kElemBlock),
                     0, st, m, x_full, in, dsdF, Hpsi, gl, blocks, fe);
}

template <typename Real>
void launch_mixed2d_local_solve(const TetMeshDev& m, const double* dx, const Real* dsdF,
                                const double* Hpsi, const double* Hinv, const double* g,
                                const double* gl, double* lam, double* ds, ReduceSlot max_ds,
                                hipStream_t st) {
  const int gsz = reduce_grid(m.nT);
  hipLaunchKernelGGL(k_mixed2d_local_solve<Real>, dim3(gsz), dim3(kElemBlock), 0, st, m, dx, dsdF,
                     Hpsi, Hinv, g, gl, lam, ds, max_ds);
  launch_finalize(max_ds.partials, gsz, 1, max_ds.result, st);
}

void launch_mixed2d_energy(const TetMeshDev& m, const double* x_full, const double* dx,
                           double alpha, const Inertia& in, const double* s, const double* ds,
                           const double* lam, double h2, ReduceSlot out, hipStream_t st) {
  const int gsz = reduce_grid(m.nT);
  hipLaunchKernelGGL(k_mixed2d_energy, dim3(gsz), dim3(kElemBlock), 0, st, m, x_full, dx, alpha,
                     in, s, ds, lam, h2, out);
  launch_finalize(out.partials, gsz, 0, out.result, st);
}

void launch_batch_polar(int n, const double* F, double* R, double* s, double* dsdF, int* iters,
                        hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_batch_polar, dim3(grid_for(n, 64)), dim3(64), 0, st, n, F, R, s, dsdF, iters);
}

void launch_batch_material(int n, const int* type, const double* mu, const double* la,
                           const double* s, int psd, double* E, double* g, double* H,
                           hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_batch_material, dim3(grid_for(n, 64)), dim3(64), 0, st, n, type, mu, la, s,
                     psd, E, g, H);
}

void launch_batch_snh_f9(int n, const double* mu, const double* la, const double* F, double* E,
                         double* g, double* H, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_batch_snh_f9, dim3(grid_for(n, 64)), dim3(64), 0, st, n, mu, la, F, E, g, H);
}

#define MFEM_INST(Real)                                                                          \
  template void launch_mixed_polar_material<Real>(const TetMeshDev&, const double*, double*,    \
                                                  double*, const double*, Real*,                \
                                                  const HessOut<Real>&, double*, double*, double, \
                                                  hipStream_t);                                 \
  template void launch_mixed_blocks_rhs<Real>(const TetMeshDev&, const double*, const Inertia&, \
                                              const Real*, const Real*, const double*, Real*,   \
                                              double*, hipStream_t, bool);                      \
  template void launch_mixed_local_solve<Real>(const TetMeshDev&, const double*, const Real*,   \
                                               const double*, const double*, const double*,     \
                                               const double*, double*, double*, ReduceSlot,     \
                                               hipStream_t);                                    \
  template void launch_plain_update<Real>(const TetMeshDev&, const double*, const Inertia&,     \
                                          double, Real*, double*, hipStream_t);                 \
  template void launch_plain2d_update<Real>(const TetMeshDev&, const double*, const Inertia&,   \
                                            double, Real*, double*, hipStream_t);               \
  template void launch_mixed2d_polar_material<Real>(const TetMeshDev&, const double*, double*,  \
                                                    double*, const double*, Real*,              \
                                                    const HessOut<Real>&, double*, double*,     \
                                                    double, hipStream_t);                       \
  template void launch_mixed2d_blocks_rhs<Real>(const TetMeshDev&, const double*, const Inertia&, \
                                                const Real*, const Real*, const double*, Real*,   \
                                                double*, hipStream_t);                            \
  template void launch_mixed2d_local_solve<Real>(const TetMeshDev&, const double*, const Real*,   \
                                                 const double*, const double*, const double*,     \
                                                 const double*, double*, double*, ReduceSlot,     \
                                                 hipStream_t);
MFEM_INST(double)
MFEM_INST(float)
#undef MFEM_INST

}  // namespace mfemhip
