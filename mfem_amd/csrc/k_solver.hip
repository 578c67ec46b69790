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
#include <cfloat>
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
// BSR 3x3 SpMV: G lanes cooperate on one block row; lane l takes blocks
// l, l+G, ... of the row, then the 3 partial row sums are combined with
// cross-lane shuffles. Consecutive lanes read consecutive 3x3 blocks.
template <int G, typename MT, typename VT>
__device__ __forceinline__ void bsr_row(int row, int lane, const int* __restrict__ rowptr,
                                        const int* __restrict__ col, const MT* __restrict__ val,
                                        const VT* __restrict__ x, double& y0, double& y1,
                                        double& y2) {
  double a0 = 0, a1 = 0, a2 = 0;
  const int p1 = rowptr[row + 1];
  for (int p = rowptr[row] + lane; p < p1; p += G) {
    const int c = col[p];
    const MT* v = val + (size_t)9 * p;
    const double x0 = (double)x[3 * c], x1 = (double)x[3 * c + 1], x2 = (double)x[3 * c + 2];
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
  y0 = a0;
  y1 = a1;
  y2 = a2;
}

template <int G, typename MT, typename VT>
__global__ __launch_bounds__(kBS) void k_spmv(int nrows, const int* __restrict__ rowptr,
                                              const int* __restrict__ col, const MT* __restrict__ val,
                                              const VT* __restrict__ x, VT* __restrict__ y) {
  const int row = (blockIdx.x * kBS + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (row >= nrows) return;
  double y0, y1, y2;
  bsr_row<G>(row, lane, rowptr, col, val, x, y0, y1, y2);
  if (lane == 0) {
    y[3 * row] = (VT)y0;
    y[3 * row + 1] = (VT)y1;
    y[3 * row + 2] = (VT)y2;
  }
}

template <int G, typename MT>
__global__ __launch_bounds__(kBS) void k_residual(int nrows, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col,
                                                  const MT* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ r, ReduceSlot r2) {
  __shared__ double sm[kBS / kWave];
  const int row = (blockIdx.x * kBS + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  double acc = 0;
  if (row < nrows) {
    double y0, y1, y2;
    bsr_row<G>(row, lane, rowptr, col, val, x, y0, y1, y2);
    if (lane == 0) {
      double r0 = b[3 * row] - y0, r1 = b[3 * row + 1] - y1, rr2 = b[3 * row + 2] - y2;
      r[3 * row] = r0;
      r[3 * row + 1] = r1;
      r[3 * row + 2] = rr2;
      acc = r0 * r0 + r1 * r1 + rr2 * rr2;
    }
  }
  block_reduce_finalize<kBS, 0>(acc, r2, sm);
}

// q = K p, and alpha = rz / (p . q) finalised by the last block.
template <int G, typename MT, typename VT>
__global__ __launch_bounds__(kBS) void k_pcg_spmv(int nrows, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col,
                                                  const MT* __restrict__ val,
                                                  const VT* __restrict__ p, VT* __restrict__ q,
                                                  PcgCtrl* ctrl, ReduceSlot red) {
  __shared__ double sm[kBS / kWave];
  if (ctrl->done) return;
  const int row = (blockIdx.x * kBS + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  double acc = 0;
  if (row < nrows) {
    double y0, y1, y2;
    bsr_row<G>(row, lane, rowptr, col, val, p, y0, y1, y2);
    if (lane == 0) {
      const VT q0 = (VT)y0, q1 = (VT)y1, q2 = (VT)y2;
      q[3 * row] = q0;
      q[3 * row + 1] = q1;
      q[3 * row + 2] = q2;
      acc = (double)p[3 * row] * (double)q0 + (double)p[3 * row + 1] * (double)q1 +
            (double)p[3 * row + 2] * (double)q2;
    }
  }
  if (block_reduce_finalize<kBS, 0>(acc, red, sm) && threadIdx.x == 0) {
    ctrl->pq = *red.result;
    ctrl->alpha = ctrl->rz / ctrl->pq;
  }
}

// Two-value deterministic reduction finalised by one block (rr and rz).
template <int BS>
__device__ __forceinline__ bool block_reduce2_finalize(double v0, double v1, double* partials,
                                                       unsigned* counter, double* sm,
                                                       double& t0, double& t1) {
  __shared__ int is_last;
  double a = block_reduce<BS, 0>(v0, sm);
  double b = block_reduce<BS, 0>(v1, sm);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
    __threadfence();
    unsigned ticket = atomicAdd(counter, 1u);
    is_last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double s0 = 0, s1 = 0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += BS) {
    s0 += __hip_atomic_load(&partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s1 += __hip_atomic_load(&partials[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  t0 = block_reduce<BS, 0>(s0, sm);
  t1 = block_reduce<BS, 0>(s1, sm);
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

// x += alpha p; r -= alpha q; z = D^{-1} r; then (last block) the stopping
// test on |r|^2, beta = rz_new / rz and the iteration count.
template <typename VT>
__global__ __launch_bounds__(kBS) void k_pcg_update(int n, const VT* __restrict__ p,
                                                    const VT* __restrict__ q,
                                                    const VT* __restrict__ invdiag,
                                                    double* __restrict__ x, VT* __restrict__ r,
                                                    VT* __restrict__ z, PcgCtrl* ctrl,
                                                    ReduceSlot red) {
  __shared__ double sm[kBS / kWave];
  if (ctrl->done) return;
  const double alpha = ctrl->alpha;
  const int i = blockIdx.x * kBS + threadIdx.x;
  double rr = 0, rz = 0;
  if (i < n) {
    x[i] += alpha * (double)p[i];
    const VT ri = (VT)((double)r[i] - alpha * (double)q[i]);
    r[i] = ri;
    const VT zi = invdiag[i] * ri;
    z[i] = zi;
    rr = (double)ri * (double)ri;
    rz = (double)ri * (double)zi;
  }
  double t0, t1;
  if (block_reduce2_finalize<kBS>(rr, rz, red.partials, red.counter, sm, t0, t1) &&
      threadIdx.x == 0) {
    ctrl->rr = t0;
    if (t0 < ctrl->threshold) {
      ctrl->done = 1;
    } else {
      ctrl->rz_new = t1;
      ctrl->beta = t1 / ctrl->rz;
      ctrl->rz = t1;
      ctrl->iter += 1;
      if (ctrl->iter >= ctrl->max_iters) ctrl->done = 1;
    }
  }
}

// p = z + beta p (skipped when the stopping test fired or the iteration
// budget ran out; in the latter case Eigen also leaves p unused).
template <typename VT>
__global__ __launch_bounds__(kBS) void k_pcg_dir(int n, const VT* __restrict__ z, VT* __restrict__ p,
                                                 const PcgCtrl* ctrl) {
  if (ctrl->done) return;
  const double beta = ctrl->beta;
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) p[i] = (VT)((double)z[i] + beta * (double)p[i]);
}

// r = b, x = 0, z = p = D^{-1} b; reductions of r.z, r.r (= b.b).
template <typename VT>
__global__ __launch_bounds__(kBS) void k_pcg_init(int n, const double* __restrict__ b,
                                                  const VT* __restrict__ invdiag,
                                                  double* __restrict__ x, VT* __restrict__ r,
                                                  VT* __restrict__ z, VT* __restrict__ p,
                                                  ReduceSlot red_rr, ReduceSlot red_rz) {
  __shared__ double sm[kBS / kWave];
  const int i = blockIdx.x * kBS + threadIdx.x;
  double rr = 0, rz = 0;
  if (i < n) {
    const double bi = b[i];
    x[i] = 0.0;
    const VT ri = (VT)bi;
    r[i] = ri;
    const VT zi = invdiag[i] * ri;
    z[i] = zi;
    p[i] = zi;
    rr = bi * bi;
    rz = (double)ri * (double)zi;
  }
  block_reduce_finalize<kBS, 0>(rr, red_rr, sm);
  block_reduce_finalize<kBS, 0>(rz, red_rz, sm);
}

__global__ void k_pcg_init_ctrl(PcgCtrl* ctrl, const double* rr, const double* rz, double tol,
                                int max_iters) {
  const double b2 = *rr;
  ctrl->b2 = b2;
  ctrl->rr = b2;
  ctrl->rz = *rz;
  ctrl->threshold = fmax(tol * tol * b2, DBL_MIN);
  ctrl->iter = 0;
  ctrl->max_iters = max_iters;
  ctrl->alpha = 0;
  ctrl->beta = 0;
  ctrl->pq = 0;
  ctrl->done = (b2 == 0.0 || b2 < ctrl->threshold || max_iters <= 0) ? 1 : 0;
}

inline int grid(long n, int bs) { return (int)((n + bs - 1) / bs); }

constexpr int kG = 8;  // lanes per BSR block row

}  // namespace

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
                 hipStream_t st) {
  if (nrows == 0) return;
  hipLaunchKernelGGL((k_spmv<kG, MT, VT>), dim3(grid((long)nrows * kG, kBS)), dim3(kBS), 0, st,
                     nrows, rowptr, col, val, x, y);
}

template <typename MT>
void launch_residual(int nrows, const int* rowptr, const int* col, const MT* val, const double* x,
                     const double* b, double* r, ReduceSlot r2, hipStream_t st) {
  hipLaunchKernelGGL((k_residual<kG, MT>), dim3(grid((long)nrows * kG, kBS) > 0 ? grid((long)nrows * kG, kBS) : 1),
                     dim3(kBS), 0, st, nrows, rowptr, col, val, x, b, r, r2);
}

template <typename MT, typename VT>
void launch_pcg_init(const PcgBuffers<MT, VT>& B, double tol, int max_iters, hipStream_t st) {
  int g = B.n > 0 ? grid(B.n, kBS) : 1;
  hipLaunchKernelGGL(k_pcg_init<VT>, dim3(g), dim3(kBS), 0, st, B.n, B.b, B.invdiag, B.x, B.r, B.z,
                     B.p, B.red[0], B.red[1]);
  hipLaunchKernelGGL(k_pcg_init_ctrl, dim3(1), dim3(1), 0, st, B.ctrl, B.red[0].result,
                     B.red[1].result, tol, max_iters);
}

template <typename MT, typename VT>
void launch_pcg_iteration(const PcgBuffers<MT, VT>& B, hipStream_t st, hipEvent_t ev0,
                          hipEvent_t ev1) {
  const int nrows = B.n / 3;
  if (ev0) (void)hipEventRecord(ev0, st);
  hipLaunchKernelGGL((k_pcg_spmv<kG, MT, VT>), dim3(grid((long)nrows * kG, kBS)), dim3(kBS), 0, st,
                     nrows, B.rowptr, B.col, B.K, B.p, B.q, B.ctrl, B.red[0]);
  if (ev1) (void)hipEventRecord(ev1, st);
  hipLaunchKernelGGL(k_pcg_update<VT>, dim3(grid(B.n, kBS)), dim3(kBS), 0, st, B.n, B.p, B.q,
                     B.invdiag, B.x, B.r, B.z, B.ctrl, B.red[1]);
  hipLaunchKernelGGL(k_pcg_dir<VT>, dim3(grid(B.n, kBS)), dim3(kBS), 0, st, B.n, B.z, B.p, B.ctrl);
}

template void launch_assemble_bsr<double>(int, const int*, const int*, const double*, const double*,
                                          double*, hipStream_t);
template void launch_assemble_bsr<float>(int, const int*, const int*, const double*, const float*,
                                         float*, hipStream_t);
template void launch_jacobi_diag<double, double>(int, const int*, const double*, double*, hipStream_t);
template void launch_jacobi_diag<float, float>(int, const int*, const float*, float*, hipStream_t);
template void launch_spmv<double, double>(int, const int*, const int*, const double*, const double*,
                                          double*, hipStream_t);
template void launch_spmv<float, float>(int, const int*, const int*, const float*, const float*,
                                        float*, hipStream_t);
template void launch_residual<double>(int, const int*, const int*, const double*, const double*,
                                      const double*, double*, ReduceSlot, hipStream_t);
template void launch_residual<float>(int, const int*, const int*, const float*, const double*,
                                     const double*, double*, ReduceSlot, hipStream_t);
template void launch_pcg_init<double, double>(const PcgBuffers<double, double>&, double, int,
                                              hipStream_t);
template void launch_pcg_init<float, float>(const PcgBuffers<float, float>&, double, int,
                                            hipStream_t);
template void launch_pcg_iteration<double, double>(const PcgBuffers<double, double>&, hipStream_t,
                                                   hipEvent_t, hipEvent_t);
template void launch_pcg_iteration<float, float>(const PcgBuffers<float, float>&, hipStream_t,
                                                 hipEvent_t, hipEvent_t);

}  // namespace mfemhip
