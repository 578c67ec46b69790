// "eigen-pcg-IC": CG with an incomplete-Cholesky preconditioner.
//
// The reference pairs Eigen's ConjugateGradient with IncompleteCholesky
// (reference linear_solver_factory.cpp:87-97). Its scaling and shift loop are
// kept here: K~ = S K S with S = diag(|row|_2^-1/2), a diagonal shift of
// 1e-3 - min diag(K~) when that is not positive, and on a non-positive pivot
// a restart with max(1e-3, 2 shift) (the host drives the restarts). The
// ordering and fill rule become IC(0) on the 3x3 vertex blocks of K in
// multicolour order (smallest-last colouring, built on the host): vertices of
// one colour share no block, so one colour is one parallel step of the
// factorization and of each triangular solve.
//
//   L_ij = (K~_ij - sum_k L_ik L_jk^T) L_jj^-T   for j before i (k before j,
//                                               k in both rows' lower lists)
//   L_ii L_ii^T = K~_ii + shift I - sum_k L_ik L_ik^T
//   z = S L^-T L^-1 S r
//
// Layout: rows are numbered by position p in the colour order (colour, then
// index). Per position the lower list holds the blocks of earlier colours
// (sorted by position) and the upper list the later ones, as (row j, index
// of L_ji in the lower lists), so both sweeps gather and nothing is written
// twice. L, the lists and the inverted diagonal factors are stored by
// position, so the rows of one colour read one contiguous stretch. Vectors
// stay in row order. In 2D the z rows of K hold only the mass and decouple;
// the pivot and shift tests look at the first `dim` rows.
//
// One CG iteration follows Eigen's loop (x0 = 0, stop at |r|^2 < tol^2 |b|^2)
// in 2 C + 4 launches for C colours:
//   k_ic_spmv      p = z + beta p_old (ping-pong), q = K p, partial p.q
//   k_ic_fin       alpha = |r.z| / p.q
//   k_ic_forward   per colour: x += alpha p, r -= alpha q, partial r.r;
//                  y = Dinv (S r - sum L y)
//   k_ic_fin       stop test on r.r
//   k_ic_backward  per colour (reversed): zt = Dinv^T (y - sum L^T zt),
//                  z = S zt, partial r.z
//   k_ic_fin       beta, iteration count
// with the scalars in one IcCtrl on the device; once `done` is set every
// later kernel returns at once. The sweeps and the SpMV give each block row
// kLanes lanes, which split its blocks and sum with lane shuffles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kernels.h"
#include "reduce.h"

namespace mfemhip {

namespace {

constexpr int kVecBS = 256;
constexpr int kRowBS = 256;    // sweeps and SpMV: kRowBS / kLanes rows per workgroup
constexpr int kLanes = 4;      // lanes per block row
constexpr int kFactorBS = 64;  // factorization: one row per lane
constexpr int kFinBS = 1024;

// XCD-contiguous split of n items (workgroup b on XCD b % 8 takes the
// (b % 8)-th eighth); n for lanes past the end of their range.
__device__ __forceinline__ int xcd_item(int n) {
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3;
  const long long start = (long long)n * xcd / 8, end = (long long)n * (xcd + 1) / 8;
  const long long i = start + (long long)lb * blockDim.x + threadIdx.x;
  return i < end ? (int)i : n;
}

// The same split over rows with kLanes lanes each: the lanes of one row are
// adjacent and in one workgroup. Returns n for rows past the end.
__device__ __forceinline__ int xcd_row(int n) {
  constexpr int rows_per_block = kRowBS / kLanes;
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3;
  const long long start = (long long)n * xcd / 8, end = (long long)n * (xcd + 1) / 8;
  const long long i = start + (long long)lb * rows_per_block + threadIdx.x / kLanes;
  return i < end ? (int)i : n;
}

inline int grid_for(int n, int bs) {
  if (n <= 0) return 0;
  const int per_xcd = (n + 7) / 8;
  return 8 * ((per_xcd + bs - 1) / bs);
}
inline int row_grid(int n) { return grid_for(n, kRowBS / kLanes); }

__device__ __forceinline__ double lane_sum(double v) {
#pragma unroll
  for (int off = 1; off < kLanes; off <<= 1) v += __shfl_xor(v, off, kWave);
  return v;
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

// One colour of the factorization (positions off .. off + count - 1), one
// row per lane.
template <typename MT>
__global__ __launch_bounds__(kFactorBS) void k_ic_factor(IcFactorDev f, const MT* __restrict__ K,
                                                         double shift, int off, int count) {
  const int t = xcd_item(count);
  if (t >= count) return;
  const int p = off + t;
  const int i = f.color_rows[p];
  const double* si = f.scale + 3 * i;
  const int a0 = f.lo_ptr[p], a1 = f.lo_ptr[p + 1];
  for (int a = a0; a < a1; ++a) {
    const int j = f.lo_col[a], q = f.lo_pos[a];
    const double* sj = f.scale + 3 * j;
    double B[9];
    load_block(K, f.lo_k[a], B);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) B[3 * r + c] *= si[r] * sj[c];
    // sum over k in lower(i) (before a) and lower(j), both sorted by position
    int ia = a0, ja = f.lo_ptr[q];
    const int jend = f.lo_ptr[q + 1];
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
    const double* Dj = f.Dinv + 9 * (long)q;
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
  for (int e = 0; e < 9; ++e) f.Dinv[9 * (long)p + e] = Di[e];
  if (fail) f.fail[0] = 1;  // every failing lane stores the same value
}

// ---------------------------------------------------------------- CG
// x = 0, r = b, p_old = 0, partial b.b
__global__ __launch_bounds__(kVecBS) void k_ic_init(int n3, IcVectors v, double* p_old) {
  __shared__ double sm[kVecBS / kWave];
  const int k = xcd_item(n3);
  double s = 0;
  if (k < n3) {
    const double b = v.b[k];
    v.x[k] = 0;
    v.r[k] = b;
    p_old[k] = 0;
    s = b * b;
  }
  block_partial<kVecBS, 0>(s, v.parts, sm);
}

// One colour of L^-1 (positions off .. off + count - 1). With `update`, the
// CG update of the colour's rows first: x += alpha p, r -= alpha q, and a
// partial of r.r.
__global__ __launch_bounds__(kRowBS) void k_ic_forward(IcFactorDev f, IcVectors v, const double* p,
                                                       double* parts, int off, int count,
                                                       int update) {
  __shared__ double sm[kRowBS / kWave];
  if (v.ctrl->done) return;
  const int t = xcd_row(count);
  const int lane = threadIdx.x & (kLanes - 1);
  double rr = 0;
  if (t < count) {
    const int pos = off + t;
    const int i = f.color_rows[pos];
    // the row's own operands, loaded ahead of the gather chain
    double ri[3] = {0, 0, 0}, qi[3] = {0, 0, 0}, pi[3] = {0, 0, 0}, sc[3] = {0, 0, 0}, Di[6] = {};
    if (lane == 0) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        ri[r] = v.r[3 * i + r];
        sc[r] = f.scale[3 * i + r];
        if (update) {
          qi[r] = v.q[3 * i + r];
          pi[r] = p[3 * i + r];
        }
      }
      const double* D = f.Dinv + 9 * (long)pos;
      Di[0] = D[0]; Di[1] = D[3]; Di[2] = D[4]; Di[3] = D[6]; Di[4] = D[7]; Di[5] = D[8];
    }
    double acc[3] = {0, 0, 0};
    for (int a = f.lo_ptr[pos] + lane; a < f.lo_ptr[pos + 1]; a += kLanes) {
      const double* La = f.L + 9 * (long)a;
      const double* yj = v.y + 3 * (long)f.lo_col[a];
      const double y0 = yj[0], y1 = yj[1], y2 = yj[2];
#pragma unroll
      for (int r = 0; r < 3; ++r) acc[r] += La[3 * r] * y0 + La[3 * r + 1] * y1 + La[3 * r + 2] * y2;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] = lane_sum(acc[r]);
    if (lane == 0) {
      if (update) {
        const double alpha = v.ctrl->alpha;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          v.x[3 * i + r] += alpha * pi[r];
          ri[r] -= alpha * qi[r];
          v.r[3 * i + r] = ri[r];
          rr += ri[r] * ri[r];
        }
      }
      double s[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) s[r] = sc[r] * ri[r] - acc[r];
      v.y[3 * i] = Di[0] * s[0];
      v.y[3 * i + 1] = Di[1] * s[0] + Di[2] * s[1];
      v.y[3 * i + 2] = Di[3] * s[0] + Di[4] * s[1] + Di[5] * s[2];
    }
  }
  if (update) block_partial<kRowBS, 0>(rr, parts, sm);
}

// One colour of L^-T (reversed order): zt = Dinv^T (y - sum_j L_ji^T zt_j),
// z = S zt, and a partial of r.z.
__global__ __launch_bounds__(kRowBS) void k_ic_backward(IcFactorDev f, IcVectors v, double* parts,
                                                        int off, int count) {
  __shared__ double sm[kRowBS / kWave];
  if (v.ctrl->done) return;
  const int t = xcd_row(count);
  const int lane = threadIdx.x & (kLanes - 1);
  double rz = 0;
  if (t < count) {
    const int pos = off + t;
    const int i = f.color_rows[pos];
    // the row's own operands, loaded ahead of the gather chain
    double yi[3] = {0, 0, 0}, ri[3] = {0, 0, 0}, sc[3] = {0, 0, 0}, Di[6] = {};
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        yi[c] = v.y[3 * i + c];
        ri[c] = v.r[3 * i + c];
        sc[c] = f.scale[3 * i + c];
      }
      const double* D = f.Dinv + 9 * (long)pos;
      Di[0] = D[0]; Di[1] = D[3]; Di[2] = D[4]; Di[3] = D[6]; Di[4] = D[7]; Di[5] = D[8];
    }
    double acc[3] = {0, 0, 0};
    for (int b = f.up_ptr[pos] + lane; b < f.up_ptr[pos + 1]; b += kLanes) {
      const double* Lb = f.L + 9 * (long)f.up_l[b];
      const double* zj = v.zt + 3 * (long)f.up_col[b];
      const double z0 = zj[0], z1 = zj[1], z2 = zj[2];
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] += Lb[c] * z0 + Lb[3 + c] * z1 + Lb[6 + c] * z2;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = lane_sum(acc[c]);
    if (lane == 0) {
      double s[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) s[c] = yi[c] - acc[c];
      // Dinv^T s with Dinv lower triangular (packed rows 00; 10 11; 20 21 22)
      double z[3];
      z[2] = Di[5] * s[2];
      z[1] = Di[2] * s[1] + Di[4] * s[2];
      z[0] = Di[0] * s[0] + Di[1] * s[1] + Di[3] * s[2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v.zt[3 * i + c] = z[c];
        const double zs = sc[c] * z[c];
        v.z[3 * i + c] = zs;
        rz += ri[c] * zs;
      }
    }
  }
  block_partial<kRowBS, 0>(rz, parts, sm);
}

// p = z + beta p_old, q = K p, partial p.q (kLanes lanes per block row)
template <typename MT>
__global__ __launch_bounds__(kRowBS) void k_ic_spmv(IcFactorDev f, const MT* __restrict__ K,
                                                    IcVectors v, const double* __restrict__ p_old,
                                                    double* __restrict__ p_new) {
  __shared__ double sm[kRowBS / kWave];
  if (v.ctrl->done) return;
  const double beta = v.ctrl->beta;
  const int i = xcd_row(f.nfree);
  const int lane = threadIdx.x & (kLanes - 1);
  double pq = 0;
  if (i < f.nfree) {
    double acc[3] = {0, 0, 0};
    for (int k = f.rowptr[i] + lane; k < f.rowptr[i + 1]; k += kLanes) {
      const long j = f.col[k];
      const double p0 = v.z[3 * j] + beta * p_old[3 * j];
      const double p1 = v.z[3 * j + 1] + beta * p_old[3 * j + 1];
      const double p2 = v.z[3 * j + 2] + beta * p_old[3 * j + 2];
      const MT* B = K + 9 * (long)k;
#pragma unroll
      for (int r = 0; r < 3; ++r)
        acc[r] += (double)B[3 * r] * p0 + (double)B[3 * r + 1] * p1 + (double)B[3 * r + 2] * p2;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] = lane_sum(acc[r]);
    if (lane == 0) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double pr = v.z[3 * i + r] + beta * p_old[3 * i + r];
        p_new[3 * i + r] = pr;
        v.q[3 * i + r] = acc[r];
        pq += pr * acc[r];
      }
    }
  }
  block_partial<kRowBS, 0>(pq, v.parts, sm);
}

// The scalar steps of Eigen's loop on the summed partials.
enum IcFinStage { IC_FIN_INIT = 0, IC_FIN_ALPHA, IC_FIN_RR, IC_FIN_BETA_FIRST, IC_FIN_BETA,
                  IC_FIN_MINDIAG };

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
      if (s == 0 || s < c.threshold) c.done = 1;
      break;
    case IC_FIN_ALPHA:
      c.alpha = c.abs_new / s;
      break;
    case IC_FIN_RR:
      c.rr = s;
      if (s < c.threshold) c.done = 1;
      break;
    case IC_FIN_BETA_FIRST:
      c.abs_new = s;
      c.beta = 0;
      if (c.max_iters <= 0) c.done = 1;
      break;
    case IC_FIN_BETA:
      c.beta = s / c.abs_new;
      c.abs_new = s;
      c.iter += 1;
      if (c.iter >= c.max_iters) c.done = 1;
      break;
  }
  *ctrl = c;
}

__global__ void k_ic_ctrl_init(IcCtrl* ctrl, double tol2, int max_iters) {
  if (threadIdx.x != 0) return;
  IcCtrl c{};
  c.tol2 = tol2;
  c.max_iters = max_iters;
  *ctrl = c;
}

// Partials of the sweeps of all colours, side by side: the total count and
// the offset of colour c.
int sweep_partials(const int* color_ptr, int ncolors, int* offsets) {
  int n = 0;
  for (int c = 0; c < ncolors; ++c) {
    if (offsets) offsets[c] = n;
    n += row_grid(color_ptr[c + 1] - color_ptr[c]);
  }
  return n;
}

void forward_sweeps(const IcFactorDev& f, const IcVectors& v, const double* p,
                    const int* color_ptr, int ncolors, int update, hipStream_t st) {
  int off = 0;
  for (int c = 0; c < ncolors; ++c) {
    const int pos = color_ptr[c], count = color_ptr[c + 1] - color_ptr[c];
    const int g = row_grid(count);
    if (g) hipLaunchKernelGGL(k_ic_forward, dim3(g), dim3(kRowBS), 0, st, f, v, p, v.parts + off,
                              pos, count, update);
    off += g;
  }
  if (update)
    hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, off, v.ctrl,
                       (int)IC_FIN_RR);
}

void backward_sweeps(const IcFactorDev& f, const IcVectors& v, const int* color_ptr, int ncolors,
                     int first, hipStream_t st) {
  std::vector<int> offs(ncolors);
  const int total = sweep_partials(color_ptr, ncolors, offs.data());
  for (int c = ncolors - 1; c >= 0; --c) {
    const int pos = color_ptr[c], count = color_ptr[c + 1] - color_ptr[c];
    const int g = row_grid(count);
    if (g) hipLaunchKernelGGL(k_ic_backward, dim3(g), dim3(kRowBS), 0, st, f, v,
                              v.parts + offs[c], pos, count);
  }
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, total, v.ctrl,
                     (int)(first ? IC_FIN_BETA_FIRST : IC_FIN_BETA));
}

}  // namespace

int ic_partials(int nfree, const int* color_ptr, int ncolors) {
  return std::max({grid_for(3 * nfree, kVecBS), grid_for(nfree, kVecBS), row_grid(nfree),
                   sweep_partials(color_ptr, ncolors, nullptr), 1});
}

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
    const int g = grid_for(count, kFactorBS);
    if (g == 0) continue;
    hipLaunchKernelGGL(k_ic_factor<MT>, dim3(g), dim3(kFactorBS), 0, st, f, K, shift, off, count);
  }
}

void launch_ic_cg_init(const IcFactorDev& f, const IcVectors& v, double tol, int max_iters,
                       const int* color_ptr, int ncolors, hipStream_t st) {
  const int n3 = 3 * f.nfree;
  const int gv = grid_for(n3, kVecBS);
  hipLaunchKernelGGL(k_ic_ctrl_init, dim3(1), dim3(64), 0, st, v.ctrl, tol * tol, max_iters);
  if (gv == 0) return;
  // the first iteration reads p[0] as p_old
  hipLaunchKernelGGL(k_ic_init, dim3(gv), dim3(kVecBS), 0, st, n3, v, v.p[0]);
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, gv, v.ctrl,
                     (int)IC_FIN_INIT);
  forward_sweeps(f, v, nullptr, color_ptr, ncolors, 0, st);
  backward_sweeps(f, v, color_ptr, ncolors, 1, st);
}

template <typename MT>
void launch_ic_cg_iteration(const IcFactorDev& f, const MT* K, const IcVectors& v, int k,
                            const int* color_ptr, int ncolors, hipStream_t st) {
  if (f.nfree == 0) return;
  double* p_old = v.p[k & 1];
  double* p_new = v.p[(k + 1) & 1];
  const int gr = row_grid(f.nfree);
  hipLaunchKernelGGL(k_ic_spmv<MT>, dim3(gr), dim3(kRowBS), 0, st, f, K, v, p_old, p_new);
  hipLaunchKernelGGL(k_ic_fin, dim3(1), dim3(kFinBS), 0, st, v.parts, gr, v.ctrl,
                     (int)IC_FIN_ALPHA);
  forward_sweeps(f, v, p_new, color_ptr, ncolors, 1, st);
  backward_sweeps(f, v, color_ptr, ncolors, 0, st);
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
                                             int, const int*, int, hipStream_t);
template void launch_ic_cg_iteration<float>(const IcFactorDev&, const float*, const IcVectors&, int,
                                            const int*, int, hipStream_t);

}  // namespace mfemhip
