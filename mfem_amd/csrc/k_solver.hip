// Deterministic assembly gathers, BSR(3x3) SpMV and Jacobi-preconditioned
// conjugate gradients for gfx950.
//
// Assembly follows the reference's scatter-map design (reference
// src/utils/sparse_utils.cpp:19-203): every nonzero 3x3 block of the free-DOF
// system owns a precomputed, sorted list of (tet, local pair) slots and sums
// them in that order -- no atomics, bitwise reproducible.
//
// CG follows Eigen's ConjugateGradient as the reference uses it
// (linear_solver_factory.cpp:144-154, eigen_iterative_solver.h:20-29):
// x0 = 0, Jacobi preconditioner, stop when |r|^2 < max(tol^2 |b|^2, DBL_MIN)
// or after max_iters iterations. Scalars live on the device; every iteration
// is three kernels that exit at once when the solve has converged, so the
// host only polls the control block every few iterations.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cfloat>
#include <utility>
#include <hip/hip_ext.h>
#include "kernels.h"
#include "reduce.h"

namespace mfemhip {

namespace {

constexpr int kBS = 256;

__device__ __forceinline__ constexpr int blk_index(int k, int l) {
  return k * 4 - (k * (k - 1)) / 2 + (l - k);
}

template <typename Real>
__global__ __launch_bounds__(kBS) void k_assemble_bsr(int nnzb, const int* __restrict__ slot_ptr,
                                                      const int* __restrict__ slots,
                                                      const double* __restrict__ mass,
                                                      const Real* __restrict__ blocks,
                                                      Real* __restrict__ K) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nnzb) return;
  double acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = 0.0;
  const int p1 = slot_ptr[b + 1];
  for (int p = slot_ptr[b]; p < p1; ++p) {
    const int sl = slots[p];
    const int e = sl >> 4, a = (sl >> 2) & 3, c = sl & 3;
    if (a <= c) {
      const Real* src = blocks + (size_t)e * 90 + 9 * blk_index(a, c);
#pragma unroll
      for (int i = 0; i < 9; ++i) acc[i] += (double)src[i];
    } else {
      const Real* src = blocks + (size_t)e * 90 + 9 * blk_index(c, a);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[3 * r + q] += (double)src[3 * q + r];
    }
  }
  const double m = mass[b];
  acc[0] += m;
  acc[4] += m;
  acc[8] += m;
  Real* o = K + (size_t)9 * b;
#pragma unroll
  for (int i = 0; i < 9; ++i) o[i] = (Real)acc[i];
}

template <typename Real, typename VT>
__global__ void k_jacobi_diag(int nfree, const int* __restrict__ diag_block, const Real* __restrict__ K,
                              VT* __restrict__ invdiag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfree) return;
  const Real* d = K + (size_t)9 * diag_block[i];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double v = (double)d[4 * r];
    invdiag[3 * i + r] = (VT)(v != 0.0 ? 1.0 / v : 1.0);
  }
}

__global__ void k_assemble_rhs(int nfree, const int* __restrict__ inc_ptr, const int* __restrict__ inc,
                               const double* __restrict__ fe, double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfree) return;
  double a0 = 0, a1 = 0, a2 = 0;
  const int p1 = inc_ptr[i + 1];
  for (int p = inc_ptr[i]; p < p1; ++p) {
    const int ek = inc[p];
    const double* f = fe + (size_t)(ek >> 2) * 12 + 3 * (ek & 3);
    a0 += f[0];
    a1 += f[1];
    a2 += f[2];
  }
  rhs[3 * i] = a0;
  rhs[3 * i + 1] = a1;
  rhs[3 * i + 2] = a2;
}


// ---------------------------------------------------------------------------
// BSR 3x3 SpMV, in several shapes picked at run time by an autotuner:
//   G  lanes cooperate on one block row; lane l takes blocks l, l+G, ... and
//      the 3 partial row sums are combined with cross-lane shuffles;
//   U  blocks per lane per loop trip (more loads in flight);
//   NT non-temporal value loads (the matrix is streamed once per SpMV, so it
//      need not displace the vectors from the caches).
// Consecutive lanes read consecutive 3x3 blocks, so a wavefront streams the
// value array contiguously. Products accumulate in the matrix precision
// (FP32 SpMV for the mixed mode); `xv(c, k)` gives component k at block
// column c.
struct SpmvShape {
  int G, U, NT;
};
constexpr SpmvShape kShapes[] = {{8, 1, 0}, {4, 1, 0}, {16, 1, 0}, {8, 2, 0},
                                 {8, 1, 1}, {4, 2, 1}, {16, 1, 1}, {8, 2, 1}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <bool NT, typename MT>
__device__ __forceinline__ MT ldv(const MT* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}

template <int G, int U, bool NT, typename MT, typename XF>
__device__ __forceinline__ void bsr_row(int row, int lane, const int* __restrict__ rowptr,
                                        const int* __restrict__ col, const MT* __restrict__ val,
                                        XF xv, MT& y0, MT& y1, MT& y2) {
  MT a0 = 0, a1 = 0, a2 = 0;
  const int p0 = rowptr[row], p1 = rowptr[row + 1];
  int p = p0 + lane;
  if constexpr (U == 2) {
    for (; p + G < p1; p += 2 * G) {
      const int c = col[p], c2 = col[p + G];
      const MT* v = val + (size_t)9 * p;
      const MT* w = val + (size_t)9 * (p + G);
      MT vv[9], ww[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) { vv[k] = ldv<NT>(v + k); ww[k] = ldv<NT>(w + k); }
      const MT x0 = xv(c, 0), x1 = xv(c, 1), x2 = xv(c, 2);
      const MT z0 = xv(c2, 0), z1 = xv(c2, 1), z2 = xv(c2, 2);
      a0 += vv[0] * x0 + vv[1] * x1 + vv[2] * x2 + ww[0] * z0 + ww[1] * z1 + ww[2] * z2;
      a1 += vv[3] * x0 + vv[4] * x1 + vv[5] * x2 + ww[3] * z0 + ww[4] * z1 + ww[5] * z2;
      a2 += vv[6] * x0 + vv[7] * x1 + vv[8] * x2 + ww[6] * z0 + ww[7] * z1 + ww[8] * z2;
    }
  }
  for (; p < p1; p += G) {
    const int c = col[p];
    const MT* v = val + (size_t)9 * p;
    MT vv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) vv[k] = ldv<NT>(v + k);
    const MT x0 = xv(c, 0), x1 = xv(c, 1), x2 = xv(c, 2);
    a0 += vv[0] * x0 + vv[1] * x1 + vv[2] * x2;
    a1 += vv[3] * x0 + vv[4] * x1 + vv[5] * x2;
    a2 += vv[6] * x0 + vv[7] * x1 + vv[8] * x2;
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, G);
    a1 += __shfl_xor(a1, off, G);
    a2 += __shfl_xor(a2, off, G);
  }
  y0 = a0;
  y1 = a1;
  y2 = a2;
}

template <int G, int U, bool NT, typename MT, typename VT>
__global__ __launch_bounds__(kBS) void k_spmv(int nrows, const int* __restrict__ rowptr,
                                              const int* __restrict__ col, const MT* __restrict__ val,
                                              const VT* __restrict__ x, VT* __restrict__ y) {
  const int row = (blockIdx.x * kBS + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (row >= nrows) return;
  MT y0, y1, y2;
  bsr_row<G, U, NT>(row, lane, rowptr, col, val, [&](int c, int k) { return (MT)x[3 * c + k]; },
                    y0, y1, y2);
  if (lane == 0) {
    y[3 * row] = (VT)y0;
    y[3 * row + 1] = (VT)y1;
    y[3 * row + 2] = (VT)y2;
  }
}

// r = b - K x with per-block partials of |r|^2. The products use the stored
// matrix values promoted to FP64, so the defect-correction residual is FP64.
template <int G, typename MT>
__global__ __launch_bounds__(kBS) void k_residual(int nrows, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col,
                                                  const MT* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ r, double* partials) {
  __shared__ double sm[kBS / kWave];
  const int row = (blockIdx.x * kBS + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  double acc = 0;
  if (row < nrows) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int p = rowptr[row] + lane; p < rowptr[row + 1]; p += G) {
      const int c = col[p];
      const MT* v = val + (size_t)9 * p;
      const double x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
      a0 += (double)v[0] * x0 + (double)v[1] * x1 + (double)v[2] * x2;
      a1 += (double)v[3] * x0 + (double)v[4] * x1 + (double)v[5] * x2;
      a2 += (double)v[6] * x0 + (double)v[7] * x1 + (double)v[8] * x2;
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
      a0 += __shfl_xor(a0, off, G);
      a1 += __shfl_xor(a1, off, G);
      a2 += __shfl_xor(a2, off, G);
    }
    if (lane == 0) {
      double r0 = b[3 * row] - a0, r1 = b[3 * row + 1] - a1, rr2 = b[3 * row + 2] - a2;
      r[3 * row] = r0;
      r[3 * row + 1] = r1;
      r[3 * row + 2] = rr2;
      acc = r0 * r0 + r1 * r1 + rr2 * rr2;
    }
  }
  block_partial<kBS, 0>(acc, partials, sm);
}

__global__ __launch_bounds__(kBS) void k_finalize(const double* __restrict__ partials, int n,
                                                  int op, double* result) {
  __shared__ double sm[kBS / kWave];
  double r = op == 0 ? reduce_array<kBS, 0>(partials, n, 1, 0, sm)
                     : reduce_array<kBS, 1>(partials, n, 1, 0, sm);
  if (threadIdx.x == 0) *result = r;
}

// ---------------------------------------------------------------------------
// CG, two kernels per iteration `it`. z and p are stored interleaved as
// pairs zp = {z, p}, double-buffered: iteration `it` reads zp[it & 1] and
// builds zp[(it+1) & 1].
//   k_cg_spmv:  reduce the (r.r, r.z) partials of the previous update;
//               stop if |r|^2 < threshold or it == max_iters (Eigen's loop
//               test); beta = rz / rz_prev (0 at it == 0);
//               p_new = z + beta p_old, formed on the fly for the gathered
//               columns (one paired load gives z and p_old) and stored by
//               the row owner into zp_next.p; q = K p_new; partial p.q
//   k_cg_update: alpha = rz / (p.q); x += alpha p; r -= alpha q;
//               zp_next.z = D^-1 r; partials of r.r and r.z
// Every block of a kernel reduces the previous kernel's partials itself, in
// the same fixed order, so all blocks take the same branch without any
// inter-block communication inside a kernel.

template <typename VT> struct Pair;
template <> struct Pair<float> { using T = float2; };
template <> struct Pair<double> { using T = double2; };

// z + beta p, one rounding (the same expression for gathered and owned rows)
template <typename VT>
__device__ __forceinline__ VT pdir(VT z, VT beta, VT p) {
  return fma(beta, p, z);
}

// Sums of the two halves [a_0..a_{n-1}, b_0..b_{n-1}] in one pass.
template <int BS>
__device__ __forceinline__ void reduce_pair(const double* __restrict__ parts, int n, double& a,
                                            double& b, double* sm) {
  double s0 = 0, s1 = 0;
  for (int i = threadIdx.x; i < n; i += BS) {
    s0 += parts[i];
    s1 += parts[n + i];
  }
  a = block_reduce_all<BS, 0>(s0, sm);
  b = block_reduce_all<BS, 0>(s1, sm);
}

template <int G, int U, bool NT, typename MT, typename VT>
__global__ __launch_bounds__(kBS) void k_cg_spmv(
    int it, int nrows, const int* __restrict__ rowptr, const int* __restrict__ col,
    const MT* __restrict__ val, const typename Pair<VT>::T* __restrict__ zp,
    VT* __restrict__ zp_next, VT* __restrict__ q, PcgCtrl* ctrl,
    const double* __restrict__ parts_rz, int np_rz, double* __restrict__ parts_pq) {
  __shared__ double sm[kBS / kWave];
  if (ctrl->done) return;
  double rr, rz;
  reduce_pair<kBS>(parts_rz, np_rz, rr, rz, sm);
  const bool converged = rr < ctrl->threshold;
  const bool exhausted = it >= ctrl->max_iters;
  if (converged || exhausted) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctrl->rr = rr;
      ctrl->iter = converged ? (it > 0 ? it - 1 : 0) : it;
      ctrl->stop_it = it;  // SpMV launches 0 .. it-1 did work
      ctrl->done = 1;
    }
    return;
  }
  const double beta = it == 0 ? 0.0 : rz / ctrl->rz[(it - 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctrl->rz[it & 1] = rz;
    ctrl->rr = rr;
  }
  const VT bv = (VT)beta;
  const int lane = threadIdx.x & (G - 1);
  const int rows_per_pass = gridDim.x * (kBS / G);
  double acc = 0;
  for (int row = (blockIdx.x * kBS + threadIdx.x) / G; row < nrows; row += rows_per_pass) {
    MT y0, y1, y2;
    bsr_row<G, U, NT>(row, lane, rowptr, col, val,
                      [&](int c, int k) {
                        const auto v = zp[3 * c + k];
                        return (MT)pdir(v.x, bv, v.y);
                      },
                      y0, y1, y2);
    if (lane == 0) {
      const auto v0 = zp[3 * row], v1 = zp[3 * row + 1], v2 = zp[3 * row + 2];
      const VT pn0 = pdir(v0.x, bv, v0.y), pn1 = pdir(v1.x, bv, v1.y), pn2 = pdir(v2.x, bv, v2.y);
      // p component of the next pair (the z component is written by the update)
      zp_next[2 * (3 * row) + 1] = pn0;
      zp_next[2 * (3 * row + 1) + 1] = pn1;
      zp_next[2 * (3 * row + 2) + 1] = pn2;
      const VT q0 = (VT)y0, q1 = (VT)y1, q2 = (VT)y2;
      q[3 * row] = q0;
      q[3 * row + 1] = q1;
      q[3 * row + 2] = q2;
      acc += (double)pn0 * (double)q0 + (double)pn1 * (double)q1 + (double)pn2 * (double)q2;
    }
  }
  block_partial<kBS, 0>(acc, parts_pq, sm);
}

template <typename VT>
__global__ __launch_bounds__(kBS) void k_cg_update(int it, int n,
                                                   typename Pair<VT>::T* __restrict__ zp_next,
                                                   const VT* __restrict__ q,
                                                   const VT* __restrict__ invdiag,
                                                   double* __restrict__ x, VT* __restrict__ r,
                                                   const PcgCtrl* ctrl,
                                                   const double* __restrict__ parts_pq, int np_pq,
                                                   double* __restrict__ parts_rz) {
  __shared__ double sm[kBS / kWave];
  if (ctrl->done) return;
  const double pq = reduce_array<kBS, 0>(parts_pq, np_pq, 1, 0, sm);
  const double alpha = ctrl->rz[it & 1] / pq;
  const VT av = (VT)alpha;
  double rr = 0, rz = 0;
  for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
    auto v = zp_next[i];  // v.y = p_new
    x[i] += alpha * (double)v.y;
    const VT ri = r[i] - av * q[i];
    r[i] = ri;
    const VT zi = invdiag[i] * ri;
    v.x = zi;
    zp_next[i] = v;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  const double a = block_reduce_all<kBS, 0>(rr, sm);
  const double b = block_reduce_all<kBS, 0>(rz, sm);
  if (threadIdx.x == 0) {
    parts_rz[blockIdx.x] = a;
    parts_rz[gridDim.x + blockIdx.x] = b;
  }
}

// r = b, x = 0, zp[0] = {D^-1 b, 0}; partials of (b.b, r.z).
template <typename VT>
__global__ __launch_bounds__(kBS) void k_cg_init(int n, const double* __restrict__ b,
                                                 const VT* __restrict__ invdiag,
                                                 double* __restrict__ x, VT* __restrict__ r,
                                                 typename Pair<VT>::T* __restrict__ zp0,
                                                 double* __restrict__ parts_rz) {
  __shared__ double sm[kBS / kWave];
  double rr = 0, rz = 0;
  for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
    const double bi = b[i];
    x[i] = 0.0;
    const VT ri = (VT)bi;
    r[i] = ri;
    const VT zi = invdiag[i] * ri;
    typename Pair<VT>::T v;
    v.x = zi;
    v.y = (VT)0;
    zp0[i] = v;
    rr += bi * bi;
    rz += (double)ri * (double)zi;
  }
  const double a = block_reduce_all<kBS, 0>(rr, sm);
  const double c = block_reduce_all<kBS, 0>(rz, sm);
  if (threadIdx.x == 0) {
    parts_rz[blockIdx.x] = a;
    parts_rz[gridDim.x + blockIdx.x] = c;
  }
}

// |b|^2 and the stopping threshold max(tol^2 |b|^2, DBL_MIN).
__global__ __launch_bounds__(kBS) void k_cg_setup(PcgCtrl* ctrl, const double* __restrict__ parts_rz,
                                                  int np, double tol, int max_iters) {
  __shared__ double sm[kBS / kWave];
  const double b2 = reduce_array<kBS, 0>(parts_rz, np, 1, 0, sm);
  if (threadIdx.x == 0) {
    ctrl->b2 = b2;
    ctrl->rr = b2;
    ctrl->threshold = fmax(tol * tol * b2, DBL_MIN);
    ctrl->iter = 0;
    ctrl->max_iters = max_iters;
    ctrl->rz[0] = ctrl->rz[1] = 0.0;
    ctrl->stop_it = 0;
    ctrl->done = (b2 == 0.0) ? 1 : 0;  // x = 0 (Eigen returns at once)
  }
}

inline int grid(long n, int bs) { return n > 0 ? (int)((n + bs - 1) / bs) : 1; }

constexpr int kResidualG = 8;

// Dispatch a templated SpMV-shaped kernel over the shape table.
template <template <int, int, bool> class K, typename... A>
void dispatch_shape(int variant, A&&... args) {
  switch (variant) {
#define MFEM_SHAPE(i)                                                                         \
  case i:                                                                                     \
    K<kShapes[i].G, kShapes[i].U, (bool)kShapes[i].NT>::run(std::forward<A>(args)...); \
    break;
    MFEM_SHAPE(0) MFEM_SHAPE(1) MFEM_SHAPE(2) MFEM_SHAPE(3)
    MFEM_SHAPE(4) MFEM_SHAPE(5) MFEM_SHAPE(6) MFEM_SHAPE(7)
#undef MFEM_SHAPE
    default:
      K<8, 1, false>::run(std::forward<A>(args)...);
  }
}

template <typename MT, typename VT>
struct SpmvRun {
  template <int G, int U, bool NT>
  struct R {
    static void run(int nrows, const int* rowptr, const int* col, const MT* val, const VT* x, VT* y,
                    hipStream_t st) {
      hipLaunchKernelGGL((k_spmv<G, U, NT, MT, VT>), dim3(grid((long)nrows * G, kBS)), dim3(kBS), 0,
                         st, nrows, rowptr, col, val, x, y);
    }
  };
};

template <typename MT, typename VT>
struct CgSpmvRun {
  template <int G, int U, bool NT>
  struct R {
    static void run(int it, int nrows, const PcgBuffers<MT, VT>& B, int gv, hipStream_t st,
                    hipEvent_t ev0, hipEvent_t ev1) {
      using P2 = typename Pair<VT>::T;
      const int gs = std::min(grid((long)nrows * G, kBS), 2048);
      const P2* zp = reinterpret_cast<const P2*>(B.zp[it & 1]);
      VT* zp_next = B.zp[(it + 1) & 1];
      if (ev0)
        hipExtLaunchKernelGGL((k_cg_spmv<G, U, NT, MT, VT>), dim3(gs), dim3(kBS), 0, st, ev0, ev1, 0,
                              it, nrows, B.rowptr, B.col, B.K, zp, zp_next, B.q, B.ctrl, B.parts_rz,
                              gv, B.parts_pq);
      else
        hipLaunchKernelGGL((k_cg_spmv<G, U, NT, MT, VT>), dim3(gs), dim3(kBS), 0, st, it, nrows,
                           B.rowptr, B.col, B.K, zp, zp_next, B.q, B.ctrl, B.parts_rz, gv,
                           B.parts_pq);
    }
  };
};

}  // namespace

int spmv_variant_count() { return kNumShapes; }
void spmv_variant_shape(int v, int* G, int* U, int* NT) {
  const SpmvShape s = (v >= 0 && v < kNumShapes) ? kShapes[v] : kShapes[0];
  *G = s.G;
  *U = s.U;
  *NT = s.NT;
}

int pcg_spmv_blocks(int nrows) { return 2048; }  // upper bound over all shapes
int pcg_vec_blocks(int n) { return std::min(grid(n, kBS), 256); }

void launch_finalize(const double* partials, int n, int op, double* result, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, st, partials, n, op, result);
}

template <typename Real>
void launch_assemble_bsr(int nnzb, const int* slot_ptr, const int* slots, const double* mass,
                         const Real* blocks, Real* K, hipStream_t st) {
  if (nnzb == 0) return;
  hipLaunchKernelGGL(k_assemble_bsr<Real>, dim3(grid(nnzb, kBS)), dim3(kBS), 0, st, nnzb, slot_ptr,
                     slots, mass, blocks, K);
}

template <typename Real, typename VT>
void launch_jacobi_diag(int nfree, const int* diag_block, const Real* K, VT* invdiag,
                        hipStream_t st) {
  if (nfree == 0) return;
  hipLaunchKernelGGL((k_jacobi_diag<Real, VT>), dim3(grid(nfree, kBS)), dim3(kBS), 0, st, nfree,
                     diag_block, K, invdiag);
}

void launch_assemble_rhs(int nfree, const int* inc_ptr, const int* inc, const double* fe,
                         double* rhs, hipStream_t st) {
  if (nfree == 0) return;
  hipLaunchKernelGGL(k_assemble_rhs, dim3(grid(nfree, kBS)), dim3(kBS), 0, st, nfree, inc_ptr, inc,
                     fe, rhs);
}

template <typename MT, typename VT>
void launch_spmv(int nrows, const int* rowptr, const int* col, const MT* val, const VT* x, VT* y,
                 hipStream_t st, int variant) {
  if (nrows == 0) return;
  dispatch_shape<SpmvRun<MT, VT>::template R>(variant, nrows, rowptr, col, val, x, y, st);
}

template <typename MT>
void launch_residual(int nrows, const int* rowptr, const int* col, const MT* val, const double* x,
                     const double* b, double* r, ReduceSlot r2, hipStream_t st) {
  const int g = grid((long)nrows * kResidualG, kBS);
  hipLaunchKernelGGL((k_residual<kResidualG, MT>), dim3(g), dim3(kBS), 0, st, nrows, rowptr, col, val,
                     x, b, r, r2.partials);
  launch_finalize(r2.partials, g, 0, r2.result, st);
}

template <typename MT, typename VT>
void launch_pcg_init(const PcgBuffers<MT, VT>& B, double tol, int max_iters, hipStream_t st) {
  using P2 = typename Pair<VT>::T;
  const int gv = pcg_vec_blocks(B.n);
  hipLaunchKernelGGL(k_cg_init<VT>, dim3(gv), dim3(kBS), 0, st, B.n, B.b, B.invdiag, B.x, B.r,
                     reinterpret_cast<P2*>(B.zp[0]), B.parts_rz);
  hipLaunchKernelGGL(k_cg_setup, dim3(1), dim3(kBS), 0, st, B.ctrl, B.parts_rz, gv, tol, max_iters);
}

template <typename MT, typename VT>
void launch_pcg_iteration(const PcgBuffers<MT, VT>& B, int it, hipStream_t st, hipEvent_t ev0,
                          hipEvent_t ev1) {
  using P2 = typename Pair<VT>::T;
  const int nrows = B.n / 3;
  const int gv = pcg_vec_blocks(B.n);
  int G, U, NT;
  spmv_variant_shape(B.variant, &G, &U, &NT);
  const int gs = std::min(grid((long)nrows * G, kBS), 2048);
  dispatch_shape<CgSpmvRun<MT, VT>::template R>(B.variant, it, nrows, B, gv, st, ev0, ev1);
  hipLaunchKernelGGL(k_cg_update<VT>, dim3(gv), dim3(kBS), 0, st, it, B.n,
                     reinterpret_cast<P2*>(B.zp[(it + 1) & 1]), B.q, B.invdiag, B.x, B.r, B.ctrl,
                     B.parts_pq, gs, B.parts_rz);
}

template void launch_assemble_bsr<double>(int, const int*, const int*, const double*, const double*,
                                          double*, hipStream_t);
template void launch_assemble_bsr<float>(int, const int*, const int*, const double*, const float*,
                                         float*, hipStream_t);
template void launch_jacobi_diag<double, double>(int, const int*, const double*, double*, hipStream_t);
template void launch_jacobi_diag<float, float>(int, const int*, const float*, float*, hipStream_t);
template void launch_spmv<double, double>(int, const int*, const int*, const double*, const double*,
                                          double*, hipStream_t, int);
template void launch_spmv<float, float>(int, const int*, const int*, const float*, const float*,
                                        float*, hipStream_t, int);
template void launch_residual<double>(int, const int*, const int*, const double*, const double*,
                                      const double*, double*, ReduceSlot, hipStream_t);
template void launch_residual<float>(int, const int*, const int*, const float*, const double*,
                                     const double*, double*, ReduceSlot, hipStream_t);
template void launch_pcg_init<double, double>(const PcgBuffers<double, double>&, double, int,
                                              hipStream_t);
template void launch_pcg_init<float, float>(const PcgBuffers<float, float>&, double, int,
                                            hipStream_t);
template void launch_pcg_iteration<double, double>(const PcgBuffers<double, double>&, int,
                                                   hipStream_t, hipEvent_t, hipEvent_t);
template void launch_pcg_iteration<float, float>(const PcgBuffers<float, float>&, int, hipStream_t,
                                                 hipEvent_t, hipEvent_t);

}  // namespace mfemhip
