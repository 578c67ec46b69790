// "eigen-pcg-IC": CG with an incomplete-Cholesky preconditioner.
//
// The reference pairs Eigen's ConjugateGradient with IncompleteCholesky
// (reference linear_solver_factory.cpp:87-97). Its scaling and shift loop are
// kept here: K~ = S K S with S = diag(|row|_2^-1/2), a diagonal shift of
// 1e-3 - min diag(K~) when that is not positive, and on a non-positive pivot
// a restart with max(1e-3, 2 shift) (the host drives the restarts). The
// ordering and fill rule become IC(0) on the 3x3 vertex blocks of K in greedy
// multicolour order: vertices of one colour share no block, so one colour is
// one parallel step of the factorization and of each triangular solve.
//
//   L_ij = (K~_ij - sum_k L_ik L_jk^T) L_jj^-T   for j before i (k before j,
//                                               k in both rows' lower lists)
//   L_ii L_ii^T = K~_ii + shift I - sum_k L_ik L_ik^T
//   z = S L^-T L^-1 S r
//
// Per block row the lower list holds the blocks of earlier colours (sorted
// by position in the colour order) and the upper list the later ones, as
// (row j, index of L_ji in row j's lower list), so both sweeps gather and
// nothing is written twice. In 2D the z rows of K hold only the mass and
// decouple; the pivot and shift tests look at the first `dim` rows.
//
// One CG iteration follows Eigen's loop (x0 = 0, stop at |r|^2 < tol^2 |b|^2):
//   k_ic_spmv (q = K p, partial p.q), k_ic_fin (alpha), k_ic_update (x, r,
//   partial r.r), k_ic_fin (stop test), forward and backward sweeps per
//   colour, k_ic_rz (partial r.z), k_ic_fin (beta), k_ic_dir (p = z + beta p)
// with the scalars in one IcCtrl on the device; once `done` is set every
// later kernel returns at once.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "reduce.h"

namespace mfemhip {

namespace {

constexpr int kVecBS = 256;
constexpr int kSweepBS = 64;
constexpr int kFinBS = 1024;

// XCD-contiguous split of n items (workgroup b on XCD b % 8 takes the
// (b % 8)-th eighth); n for lanes past the end of their range.
__device__ __forceinline__ int xcd_item(int n) {
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3;
  const long long start = (long long)n * xcd / 8, end = (long long)n * (xcd + 1) / 8;
  const long long i = start + (long long)lb * blockDim.x + threadIdx.x;
  return i < end ? (int)i : n;
}

inline int grid_for(int n, int bs) {
  if (n <= 0) return 0;
  const int per_xcd = (n + 7) / 8;
  return 8 * ((per_xcd + bs - 1) / bs);
}

template <typename MT>
__device__ __forceinline__ void load_block(const MT* __restrict__ K, long k, double B[9]) {
#pragma unroll
  for (int e = 0; e < 9; ++e) B[e] = (double)K[9 * k + e];
}

// 3x3 Cholesky of the symmetric D (row-major, lower part read). Returns the
// inverse of the factor (lower triangular, row-major) in Dinv and the number
// of failed pivots among the first `dim`; a failed pivot of a decoupled z row
// (2D) is replaced by 1.
__device__ __forceinline__ int chol3_inv(const double D[9], int dim, double Dinv[9]) {
  int fail = 0;
  double l00 = D[0];
  if (!(l00 > 0)) { fail += 1; l00 = 1; }
  l00 = sqrt(l00);
  const double l10 = D[3] / l00, l20 = D[6] / l00;
  double l11 = D[4] - l10 * l10;
  if (!(l11 > 0)) { fail += 1; l11 = 1; }
  l11 = sqrt(l11);
  const double l21 = (D[7] - l20 * l10) / l11;
  double l22 = D[8] - l20 * l20 - l21 * l21;
  if (!(l22 > 0)) { if (dim > 2) fail += 1; l22 = 1; }
  l22 = sqrt(l22);
  // inverse of [[l00,0,0],[l10,l11,0],[l20,l21,l22]]
  const double i00 = 1 / l00, i11 = 1 / l11, i22 = 1 / l22;
  const double i10 = -l10 * i00 * i11;
  const double i21 = -l21 * i11 * i22;
  const double i20 = -(l20 * i00 + l21 * i10) * i22;
  Dinv[0] = i00; Dinv[1] = 0;   Dinv[2] = 0;
  Dinv[3] = i10; Dinv[4] = i11; Dinv[5] = 0;
  Dinv[6] = i20; Dinv[7] = i21; Dinv[8] = i22;
  return fail;
}

// ---------------------------------------------------------------- factor
// Row scaling and the scaled diagonal's minimum (over the first dim rows),
// as a per-workgroup partial of -min.
template <typename MT>
__global__ __launch_bounds__(kVecBS) void k_ic_scale(IcFactorDev f, const MT* __restrict__ K,
                                                     double* __restrict__ parts) {
  __shared__ double sm[kVecBS / kWave];
  const int i = xcd_item(f.nfree);
  double neg_min = -INFINITY;
  if (i < f.nfree) {
    double s[3] = {0, 0, 0};
    for (int k = f.rowptr[i]; k < f.rowptr[i + 1]; ++k) {
      double B[9];
      load_block(K, k, B);
#pragma unroll
      for (int r = 0; r < 3; ++r) s[r] += B[3 * r] * B[3 * r] + B[3 * r + 1] * B[3 * r + 1] +
                                          B[3 * r + 2] * B[3 * r + 2];
    }
    double D[9];
    load_block(K, f.diag[i], D);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double q = sqrt(sqrt(s[r]));
      const double sc = q > 2.2250738585072014e-308 ? 1.0 / q : 1.0;
      f.scale[3 * i + r] = sc;
      if (r < f.dim) neg_min = fmax(neg_min, -D[4 * r] * sc * sc);
    }
  }
  block_partial<kVecBS, 1>(neg_min, parts, sm);
}

// One colour of the factorization: one block row per lane.
template <typename MT>
__global__ __launch_bounds__(kSweepBS) void k_ic_factor(IcFactorDev f, const MT* __restrict__ K,
                                                        double shift, int off, int count) {
  const int t = xcd_item(count);
  if (t >= count) return;
  const int i = f.color_rows[off + t];
  const double* si = f.scale + 3 * i;
  const int a0 = f.lo_ptr[i], a1 = f.lo_ptr[i + 1];
  for (int a = a0; a < a1; ++a) {
    const int j = f.lo_col[a];
    const double* sj = f.scale + 3 * j;
    double B[9];
    load_block(K, f.lo_k[a], B);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) B[3 * r + c] *= si[r] * sj[c];
    // sum over k in lower(i) (before a) and lower(j), both sorted by position
    int ia = a0, ja = f.lo_ptr[j];
    const int jend = f.lo_ptr[j + 1];
    while (ia < a && ja < jend) {
      const int pi = f.lo_pos[ia], pj = f.lo_pos[ja];
      if (pi == pj) {
        const double* Li = f.L + 9 * (long)ia;
        const double* Lj = f.L + 9 * (long)ja;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            B[3 * r + c] -= Li[3 * r] * Lj[3 * c] + Li[3 * r + 1] * Lj[3 * c + 1] +
                            Li[3 * r + 2] * Lj[3 * c + 2];
        ++ia;
        ++ja;
      } else if (pi < pj) {
        ++ia;
      } else {
        ++ja;
      }
    }
    // L_ij = B Dinv_j^T
    const double* Dj = f.Dinv + 9 * (long)j;
    double* Lij = f.L + 9 * (long)a;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        Lij[3 * r + c] = B[3 * r] * Dj[3 * c] + B[3 * r + 1] * Dj[3 * c + 1] +
                         B[3 * r + 2] * Dj[3 * c + 2];
  }
  double D[9];
  load_block(K, f.diag[i], D);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) D[3 * r + c] *= si[r] * si[c];
  D[0] += shift;
  D[4] += shift;
  D[8] += shift;
  for (int a = a0; a < a1; ++a) {
    const double* La = f.L + 9 * (long)a;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        D[3 * r + c] -= La[3 * r] * La[3 * c] + La[3 * r + 1] * La[3 * c + 1] +
                        La[3 * r + 2] * La[3 * c + 2];
  }
  double Di[9];
  const int fail = chol3_inv(D, f.dim, Di);
#pragma unroll
  for (int e = 0; e < 9; ++e) f.Dinv[9 * (long)i + e] = Di[e];
  if (fail) f.fail[0] = 1;  // every failing lane stores the same value
}

// ---------------------------------------------------------------- solve
// y_i = Dinv_i (S r_i - sum_j L_ij y_j) over the rows of one colour
__global__ __launch_bounds__(kSweepBS) void k_ic_forward(IcFactorDev f, IcVectors v, int off,
                                                         int count) {
  if (v.ctrl->done) return;
  const int t = xcd_item(count);
  if (t >= count) return;
  const int i = f.color_rows[off + t];
  double s[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) s[r] = f.scale[3 * i + r] * v.r[3 * i + r];
  for (int a = f.lo_ptr[i]; a < f.lo_ptr[i + 1]; ++a) {
    const double* La = f.L + 9 * (long)a;
    const double* yj = v.y + 3 * (long)f.lo_col[a];
    const double y0 = yj[0], y1 = yj[1], y2 = yj[2];
#pragma unroll
    for (int r = 0; r < 3; ++r) s[r] -= La[3 * r] * y0 + La[3 * r + 1] * y1 + La[3 * r + 2] * y2;
  }
  const double* Di = f.Dinv + 9 * (long)i;
  v.y[3 * i] = Di[0] * s[0];
  v.y[3 * i + 1] = Di[3] * s[0] + Di[4] * s[1];
  v.y[3 * i + 2] = Di[6] * s[0] + Di[7] * s[1] + Di[8] * s[2];
}

// zt_i = Dinv_i^T (y_i - sum_j L_ji^T zt_j) over the rows of one colour
__global__ __launch_bounds__(kSweepBS) void k_ic_backward(IcFactorDev f, IcVectors v, int off,
                                                          int count) {
  if (v.ctrl->done) return;
  const int t = xcd_item(count);
  if (t >= count) return;
  const int i = f.color_rows[off + t];
  double s[3] = {v.y[3 * i], v.y[3 * i + 1], v.y[3 * i + 2]};
  for (int b = f.up_ptr[i]; b < f.up_ptr[i + 1]; ++b) {
    const double* Lb = f.L + 9 * (long)f.up_l[b];
    const double* zj = v.zt + 3 * (long)f.up_col[b];
    const double z0 = zj[0], z1 = zj[1], z2 = zj[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) s[c] -= Lb[c] * z0 + Lb[3 + c] * z1 + Lb[6 + c] * z2;
  }
  const double* Di = f.Dinv + 9 * (long)i;
  const double z2 = Di[8] * s[2];
  const double z1 = Di[4] * s[1] + Di[7] * s[2];
  const double z0 = Di[0] * s[0] + Di[3] * s[1] + Di[6] * s[2];
  v.zt[3 * i] = z0;
  v.zt[3 * i + 1] = z1;
  v.zt[3 * i + 2] = z2;
}

// ---------------------------------------------------------------- CG
// x = 0, r = b, partial b.b
__global__ __launch_bounds__(kVecBS) void k_ic_init(int n3, IcVectors v) {
  __shared__ double sm[kVecBS / kWave];
  const int k = xcd_item(n3);
  double s = 0;
  if (k < n3) {
    const double b = v.b[k];
    v.x[k] = 0;
    v.r[k] = b;
    s = b * b;
  }
  block_partial<kVecBS, 0>(s, v.parts, sm);
}

// q = K p (one block row per lane), partial p.q
template <typename MT>
__global__ __launch_bounds__(kVecBS) void k_ic_spmv(IcFactorDev f, const MT* __restrict__ K,
                                                    IcVectors v) {
  __shared__ double sm[kVecBS / kWave];
  if (v.ctrl->done) return;
  const int i = xcd_item(f.nfree);
  double s = 0;
  if (i < f.nfree) {
    double q[3] = {0, 0, 0};
    for (int k = f.rowptr[i]; k < f.rowptr[i + 1]; ++k) {
      const double* pj = v.p + 3 * (long)f.col[k];
      const double p0 = pj[0], p1 = pj[1], p2 = pj[2];
      const MT* B = K + 9 * (long)k;
#pragma unroll
      for (int r = 0; r < 3; ++r)
        q[r] += (double)B[3 * r] * p0 + (double)B[3 * r + 1] * p1 + (double)B[3 * r + 2] * p2;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      v.q[3 * i + r] = q[r];
      s += v.p[3 * i + r] * q[r];
    }
  }
  block_partial<kVecBS, 0>(s, v.parts, sm);
}

// x += alpha p, r -= alpha q, partial r.r
__global__ __launch_bounds__(kVecBS) void k_ic_update(int n3, IcVectors v) {
  __shared__ double sm[kVecBS / kWave];
  if (v.ctrl->done) return;
  const int k = xcd_item(n3);
  const double alpha = v.ctrl->alpha;
  double s = 0;
  if (k < n3) {
    v.x[k] += alpha * v.p[k];
    const double r = v.r[k] - alpha * v.q[k];
    v.r[k] = r;
    s = r * r;
  }
  block_partial<kVecBS, 0>(s, v.parts, sm);
}

// partial r.z with z = S zt
__global__ __launch_bounds__(kVecBS) void k_ic_rz(int n3, IcFactorDev f, IcVectors v) {
  __shared__ double sm[kVecBS / kWave];
  if (v.ctrl->done) return;
  const int k = xcd_item(n3);
  double s = 0;
  if (k < n3) s = v.r[k] * f.scale[k] * v.zt[k];
  block_partial<kVecBS, 0>(s, v.parts, sm);
}

// p = z + beta p (p = z on the first pass)
__global__ __launch_bounds__(kVecBS) void k_ic_dir(int n3, IcFactorDev f, IcVectors v) {
  if (v.ctrl->done) return;
  const int k = xcd_item(n3);
  if (k >= n3) return;
  const double z = f.scale[k] * v.zt[k];
  v.p[k] = v.ctrl->first ? z : z + v.ctrl->beta * v.p[k];
}

// The scalar steps of Eigen's loop on the summed partials.
enum IcFinStage { IC_FIN_INIT = 0, IC_FIN_ALPHA, IC_FIN_RR, IC_FIN_BETA, IC_FIN_MINDIAG };

__global__ __launch_bounds__(kFinBS) void k_ic_fin(const double* __restrict__ parts, int nparts,
                                                   IcCtrl* ctrl, int stage) {
  __shared__ double sm[kFinBS / kWave];
  if (stage == IC_FIN_MINDIAG) {
    const double m = reduce_array<kFinBS, 1>(parts, nparts, 1, 0, sm);
    if (threadIdx.x == 0) ctrl->mindiag = -m;
    return;
  }
  if (ctrl->done) return;
  const double s = reduce_array<kFinBS, 0>(parts, nparts, 1, 0, sm);
  if (threadIdx.x != 0) return;
  IcCtrl c = *ctrl;
  switch (stage) {
    case IC_FIN_INIT:
      c.b2 = s;
      c.rr = s;
      c.threshold = fmax(c.tol2 * s, 2.2250738585072014e-308);
      c.iter = 0;
      c.first = 1;
      if (s == 0 || s < c.threshold) c.done = 1;
      break;
    case IC_FIN_ALPHA:
      c.alpha = c.abs_new / s;
      break;
    case IC_FIN_RR:
      c.rr = s;
      if (s < c.threshold) c.done = 1;
      break;
    case IC_FIN_BETA:
      if (c.first) {
        c.abs_new = s;
        if (c.max_iters <= 0) c.done = 1;
      } else {
        c.beta = s / c.abs_new;
        c.abs_new = s;
        c.iter += 1;
        if (c.iter >= c.max_iters) c.done = 1;
      }
      break;
  }
  *ctrl = c;
}

// Clears `first` after the first direction kernel has read it.
__global__ void k_ic_clear_first(IcCtrl* ctrl) {
  if (threadIdx.x == 0 && !ctrl->done) ctrl->first = 0;
}

__global__ void k_ic_ctrl_init(IcCtrl* ctrl, double tol2, int max_iters) {
  if (threadIdx.x != 0) return;
  IcCtrl c{};
  c.tol2 = tol2;
  c.max_iters = max_iters;
  *ctrl = c;
}

}  // namespace

int ic_partials(int nfree) { return grid_for(3 * nfree, kVecBS); }

template <typename MT>
void launch_ic_scale(const IcFactorDev& f, const MT* K, double* parts, IcCtrl* ctrl,
                     hipStream_t st) {
  const int g = grid_for(f.nfree, kVecBS);
  if (g == 0) return;
  hipLaunchKernelGGL(k_ic_scale<MT>, dim3(g), dim3(kVecBS), 0, st, f, K, parts);
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, parts, g, ctrl,
                     (int)IC_FIN_MINDIAG);
}

template <typename MT>
void launch_ic_factor(const IcFactorDev& f, const MT* K, double shift, const int* color_ptr,
                      int ncolors, hipStream_t st) {
  for (int c = 0; c < ncolors; ++c) {
    const int off = color_ptr[c], count = color_ptr[c + 1] - color_ptr[c];
    const int g = grid_for(count, kSweepBS);
    if (g == 0) continue;
    hipLaunchKernelGGL(k_ic_factor<MT>, dim3(g), dim3(kSweepBS), 0, st, f, K, shift, off, count);
  }
}

static void precond_and_direction(const IcFactorDev& f, const IcVectors& v, const int* color_ptr,
                                  int ncolors, hipStream_t st) {
  const int n3 = 3 * f.nfree;
  const int gv = grid_for(n3, kVecBS);
  for (int c = 0; c < ncolors; ++c) {
    const int off = color_ptr[c], count = color_ptr[c + 1] - color_ptr[c];
    const int g = grid_for(count, kSweepBS);
    if (g) hipLaunchKernelGGL(k_ic_forward, dim3(g), dim3(kSweepBS), 0, st, f, v, off, count);
  }
  for (int c = ncolors - 1; c >= 0; --c) {
    const int off = color_ptr[c], count = color_ptr[c + 1] - color_ptr[c];
    const int g = grid_for(count, kSweepBS);
    if (g) hipLaunchKernelGGL(k_ic_backward, dim3(g), dim3(kSweepBS), 0, st, f, v, off, count);
  }
  hipLaunchKernelGGL(k_ic_rz, dim3(gv), dim3(kVecBS), 0, st, n3, f, v);
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, gv, v.ctrl,
                     (int)IC_FIN_BETA);
  hipLaunchKernelGGL(k_ic_dir, dim3(gv), dim3(kVecBS), 0, st, n3, f, v);
  hipLaunchKernelGGL(k_ic_clear_first, dim3(1), dim3(64), 0, st, v.ctrl);
}

void launch_ic_cg_init(const IcFactorDev& f, const IcVectors& v, double tol, int max_iters,
                       const int* color_ptr, int ncolors, hipStream_t st) {
  const int n3 = 3 * f.nfree;
  const int gv = grid_for(n3, kVecBS);
  hipLaunchKernelGGL(k_ic_ctrl_init, dim3(1), dim3(64), 0, st, v.ctrl, tol * tol, max_iters);
  if (gv == 0) return;
  hipLaunchKernelGGL(k_ic_init, dim3(gv), dim3(kVecBS), 0, st, n3, v);
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, gv, v.ctrl,
                     (int)IC_FIN_INIT);
  precond_and_direction(f, v, color_ptr, ncolors, st);
}

template <typename MT>
void launch_ic_cg_iteration(const IcFactorDev& f, const MT* K, const IcVectors& v,
                            const int* color_ptr, int ncolors, hipStream_t st) {
  const int n3 = 3 * f.nfree;
  const int gv = grid_for(n3, kVecBS);
  const int gr = grid_for(f.nfree, kVecBS);
  if (gv == 0) return;
  hipLaunchKernelGGL(k_ic_spmv<MT>, dim3(gr), dim3(kVecBS), 0, st, f, K, v);
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, gr, v.ctrl,
                     (int)IC_FIN_ALPHA);
  hipLaunchKernelGGL(k_ic_update, dim3(gv), dim3(kVecBS), 0, st, n3, v);
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, gv, v.ctrl,
                     (int)IC_FIN_RR);
  precond_and_direction(f, v, color_ptr, ncolors, st);
}

template void launch_ic_scale<double>(const IcFactorDev&, const double*, double*, IcCtrl*,
                                      hipStream_t);
template void launch_ic_scale<float>(const IcFactorDev&, const float*, double*, IcCtrl*,
                                     hipStream_t);
template void launch_ic_factor<double>(const IcFactorDev&, const double*, double, const int*, int,
                                       hipStream_t);
template void launch_ic_factor<float>(const IcFactorDev&, const float*, double, const int*, int,
                                      hipStream_t);
template void launch_ic_cg_iteration<double>(const IcFactorDev&, const double*, const IcVectors&,
                                             const int*, int, hipStream_t);
template void launch_ic_cg_iteration<float>(const IcFactorDev&, const float*, const IcVectors&,
                                            const int*, int, hipStream_t);

}  // namespace mfemhip
