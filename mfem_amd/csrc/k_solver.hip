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
// is one kernel that exits at once when the solve has stopped, so the host
// only polls the control block every few iterations.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cfloat>
#include <stdexcept>
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

// Item of this thread when [0, n) is split into 8 contiguous ranges, one per
// XCD (workgroups are dispatched round-robin over the XCDs, so workgroup b
// runs on XCD b % 8), each walked in order by that XCD's workgroups; -1 past
// the range. Rows that share tets then share one L2. The grid is
// xcd_grid(n, bs).
__device__ __forceinline__ int xcd_item(int n, int block) {
  const int xcd = block & 7, lb = block >> 3;
  const long long start = (long long)n * xcd / 8, end = (long long)n * (xcd + 1) / 8;
  const long long i = start + (long long)lb * blockDim.x + threadIdx.x;
  return i < end ? (int)i : -1;
}
__device__ __forceinline__ int xcd_item(int n) { return xcd_item(n, (int)blockIdx.x); }
inline int xcd_grid(int n, int bs) {
  const int per_xcd = (n + 7) / 8;
  return 8 * std::max(1, (per_xcd + bs - 1) / bs);
}

// Four lanes per K block: lane q sums entries q, q + 4, q + 8 (row-major
// 3x3) over the block's slots, so the four lanes of a block read the same
// 36 bytes of each element block with at most 3 loads each, and the diagonal
// entries (0, 4, 8) with the mass term all fall to lane 0. Every entry is
// summed in slot order, as with one lane per block. Only the blocks j >= i
// are summed; each is also written transposed into block (j, i), whose
// slots are the same element blocks transposed in the same order, so both
// halves are what summing them separately gives, and each element block is
// read once instead of twice.
template <typename Real>
__device__ __forceinline__ void assemble_bsr_item(int t, const int* __restrict__ upper,
                                                  const int* __restrict__ tpos,
                                                  const int* __restrict__ slot_ptr,
                                                  const int* __restrict__ slots,
                                                  const double* __restrict__ mass,
                                                  const Real* __restrict__ blocks,
                                                  Real* __restrict__ K) {
  const int q = t & 3;
  const int b = upper[t >> 2];
  const int bt = tpos[t >> 2];
  const int ne = q == 0 ? 3 : 2;
  double acc[3] = {0.0, 0.0, 0.0};
  const int p1 = slot_ptr[b + 1];
  // slots in batches of kAsmBatch: all slot ids, then all element values,
  // are in flight together (the sums still run in slot order)
  constexpr int kAsmBatch = 8;
  for (int p = slot_ptr[b]; p < p1; p += kAsmBatch) {
    int sl[kAsmBatch];
#pragma unroll
    for (int u = 0; u < kAsmBatch; ++u) sl[u] = p + u < p1 ? slots[p + u] : -1;
    Real val[kAsmBatch][3];
#pragma unroll
    for (int u = 0; u < kAsmBatch; ++u) {
      if (sl[u] < 0) continue;
      const int e = sl[u] >> 4, a = (sl[u] >> 2) & 3, c = sl[u] & 3;
      const Real* src = blocks + (size_t)e * 90 + 9 * (a <= c ? blk_index(a, c) : blk_index(c, a));
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int o = q + 4 * j;                             // entry 3r + s of the K block
        const int src_o = a <= c ? o : 3 * (o % 3) + o / 3;  // transposed for a > c
        val[u][j] = j < ne ? src[src_o] : (Real)0;
      }
    }
#pragma unroll
    for (int u = 0; u < kAsmBatch; ++u)
      if (sl[u] >= 0)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          if (j < ne) acc[j] += (double)val[u][j];
  }
  Real* out = K + (size_t)9 * b;
  Real* out_t = K + (size_t)9 * bt;
  const double m = q == 0 ? mass[b] : 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if (j < ne) {
      const int o = q + 4 * j;
      const Real v = (Real)(acc[j] + m);
      out[o] = v;
      if (bt != b) out_t[3 * (o % 3) + o / 3] = v;
    }
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

// Rows are visited in the order `rows` lists them (each XCD a contiguous
// part of the list).
__device__ __forceinline__ void assemble_rhs_row(int i, const int* __restrict__ inc_ptr,
                                                 const int* __restrict__ inc,
                                                 const double* __restrict__ fe,
                                                 double* __restrict__ rhs) {
  double a0 = 0, a1 = 0, a2 = 0;
  const int p1 = inc_ptr[i + 1];
  // incidences in batches: ids, then element values, in flight together
  // (the sums still run in incidence order)
  constexpr int kRhsBatch = 8;
  for (int p = inc_ptr[i]; p < p1; p += kRhsBatch) {
    int ek[kRhsBatch];
#pragma unroll
    for (int u = 0; u < kRhsBatch; ++u) ek[u] = p + u < p1 ? inc[p + u] : -1;
    double f[kRhsBatch][3];
#pragma unroll
    for (int u = 0; u < kRhsBatch; ++u)
      if (ek[u] >= 0) {
        const double* src = fe + (size_t)(ek[u] >> 2) * 12 + 3 * (ek[u] & 3);
        f[u][0] = src[0];
        f[u][1] = src[1];
        f[u][2] = src[2];
      }
#pragma unroll
    for (int u = 0; u < kRhsBatch; ++u)
      if (ek[u] >= 0) {
        a0 += f[u][0];
        a1 += f[u][1];
        a2 += f[u][2];
      }
  }
  rhs[3 * i] = a0;
  rhs[3 * i + 1] = a1;
  rhs[3 * i + 2] = a2;
}

// Both gathers in one launch: workgroups [0, nb_bsr) assemble K, the rest
// the right-hand side. Neither is bound by bandwidth (each waits on its
// id -> value load chains), so the rhs gather runs in the K gather's
// shadow. nb_bsr is a multiple of 8, so both parts keep their XCD mapping.
template <typename Real>
__global__ __launch_bounds__(kBS) void k_assemble_bsr_rhs(
    int nub, const int* __restrict__ upper, const int* __restrict__ tpos,
    const int* __restrict__ slot_ptr, const int* __restrict__ slots,
    const double* __restrict__ mass, const Real* __restrict__ blocks, Real* __restrict__ K,
    int nb_bsr, int nfree, const int* __restrict__ rows, const int* __restrict__ inc_ptr,
    const int* __restrict__ inc, const double* __restrict__ fe, double* __restrict__ rhs) {
  if ((int)blockIdx.x < nb_bsr) {
    const int t = xcd_item(4 * nub, (int)blockIdx.x);
    if (t >= 0) assemble_bsr_item<Real>(t, upper, tpos, slot_ptr, slots, mass, blocks, K);
  } else {
    const int item = xcd_item(nfree, (int)blockIdx.x - nb_bsr);
    if (item >= 0) assemble_rhs_row(rows[item], inc_ptr, inc, fe, rhs);
  }
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

// One block row with every lane's U blocks in flight at once: the U column
// ids and the U 3x3 value blocks are loaded together, then the U gathers, so
// a row of up to U*G blocks costs one column-id round trip and one gather
// round trip. Lanes past the row end are masked off.
template <int G, int U, typename MT, typename XF>
__device__ __forceinline__ void bsr_row_batched(int row, int lane, const int* __restrict__ rowptr,
                                                const int* __restrict__ col,
                                                const MT* __restrict__ val, XF xv, MT& y0, MT& y1,
                                                MT& y2) {
  MT a0 = 0, a1 = 0, a2 = 0;
  const int p0 = rowptr[row], p1 = rowptr[row + 1];
  for (int p = p0 + lane; p < p1; p += U * G) {
    int c[U];
    MT v[U][9];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = p + u * G;
      c[u] = q < p1 ? col[q] : -1;
      if (q < p1) {
#pragma unroll
        for (int k = 0; k < 9; ++k) v[u][k] = val[(size_t)9 * q + k];
      }
    }
    MT xs[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (c[u] >= 0) {
        xs[u][0] = xv(c[u], 0);
        xs[u][1] = xv(c[u], 1);
        xs[u][2] = xv(c[u], 2);
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (c[u] >= 0) {
        a0 += v[u][0] * xs[u][0] + v[u][1] * xs[u][1] + v[u][2] * xs[u][2];
        a1 += v[u][3] * xs[u][0] + v[u][4] * xs[u][1] + v[u][5] * xs[u][2];
        a2 += v[u][6] * xs[u][0] + v[u][7] * xs[u][1] + v[u][8] * xs[u][2];
      }
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
// CG, one kernel per iteration.
//
// Eigen's loop (x0 = 0, r0 = b, z = D^-1 r, p0 = z0):
//   q = K p; alpha = rz / p.q; x += alpha p; r -= alpha q;
//   stop if |r|^2 < threshold; z = D^-1 r; beta = rz_new / rz; p = z + beta p.
// Kernel `it` reads the state {r, q, p, D^-1} of iteration it-1 and the five
// partial sums its predecessor wrote, then
//   1. reduces the partials (every block, in the same fixed order) to p.q,
//      z.q, q.D^-1 q, r.r and r.z of iteration it-1;
//   2. applies Eigen's stopping test to r_{it-1} (one kernel late: x already
//      holds x_{it-1}), or stops when it-1 == max_iters;
//   3. alpha = rz / p.q and beta from the r.z of the coming residual,
//      (r - alpha q).D^-1(r - alpha q) = rz - 2 alpha z.q + alpha^2 q.D^-1 q;
//   4. forms r = r - alpha q, z = D^-1 r, p = z + beta p on the fly for every
//      gathered block column and computes q = K p; the row owner applies
//      x += alpha p_{it-1} and stores the new state;
//   5. writes the five partial sums of the new state.
// Step 3 starts from the exact r.z of r_{it-1} (partial 4), so the scalar
// recurrence carries the rounding of one step only. Kernel 0 runs the same
// code with alpha = beta = 0 on the initial state {b, 0, 0, D^-1}.
//
// Rows are split into 8 contiguous ranges, one per XCD (workgroups are
// dispatched round-robin over the 8 XCDs), so the block columns a range
// gathers stay in that XCD's L2.
template <typename T>
struct alignas(4 * sizeof(T)) CgDof {
  T r, q, p, d;
};

constexpr int kCusPerXcd = 32;  // MI355X: 8 XCDs x 32 CUs

struct CgShape {
  int G, U, BS;  // lanes per block row, blocks per lane in flight, workgroup size
};
constexpr CgShape kCgShapes[] = {{4, 4, 256},  {4, 4, 512},  {8, 2, 256},  {8, 2, 1024},
                                 {4, 2, 256},  {4, 2, 1024}, {16, 1, 1024}, {8, 1, 256}};
constexpr int kNumCgShapes = sizeof(kCgShapes) / sizeof(kCgShapes[0]);

// r = r - alpha q, z = D^-1 r, p = z + beta p (identical for owned and
// gathered rows)
template <typename T>
__device__ __forceinline__ void cg_form(const CgDof<T>& s, T alpha, T beta, T& r, T& z, T& p) {
  r = fma(-alpha, s.q, s.r);
  z = s.d * r;
  p = fma(beta, s.p, z);
}

// Profiling: the span of the previous CG kernel, from the per-block
// (start, end) wall-clock stamps it wrote. Called by all threads of block 0.
template <int BS>
__device__ __forceinline__ void cg_account(const unsigned long long* __restrict__ stamps, int nb,
                                           PcgCtrl* ctrl, double* sm) {
  double lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < nb; i += BS) {
    lo = fmin(lo, (double)stamps[2 * i]);
    hi = fmax(hi, (double)stamps[2 * i + 1]);
  }
  lo = -block_reduce_all<BS, 1>(-lo, sm);
  hi = block_reduce_all<BS, 1>(hi, sm);
  if (threadIdx.x == 0) {
    ctrl->kern_ticks += hi - lo;
    ctrl->kern_launches += 1;
  }
}

// Sum of the kCgParts partial arrays (partial i of quantity k at
// parts[k * stride + i], i < nb) by one workgroup, with all of a thread's
// loads issued before the first add.
template <int BS>
__device__ __forceinline__ void cg_reduce_parts(const double* __restrict__ parts, int stride,
                                                int nb, double (&v)[kCgParts], double* sm) {
  constexpr int R = BS >= 1024 ? 1 : 4;  // loads in flight per thread (register budget)
#pragma unroll
  for (int k = 0; k < kCgParts; ++k) v[k] = 0.0;
  for (int i0 = threadIdx.x; i0 < nb; i0 += R * BS) {
    double t[R][kCgParts];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = i0 + j * BS;
#pragma unroll
      for (int k = 0; k < kCgParts; ++k) t[j][k] = i < nb ? parts[(size_t)k * stride + i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
      for (int k = 0; k < kCgParts; ++k) v[k] += t[j][k];
  }
  block_sum_n<BS, kCgParts>(v, sm);
}

// A control block into the host's mapped copy: every field, then (after a
// system-scope fence) done, so that a host that sees done set reads the
// final fields. Called when a solve stops, and by the batched setup for
// every system (done then tells whether the system starts stopped).
__device__ __forceinline__ void publish_stop(const PcgCtrl& c, PcgCtrl* mirror) {
  mirror->rr = c.rr;
  mirror->b2 = c.b2;
  mirror->threshold = c.threshold;
  mirror->kern_ticks = c.kern_ticks;
  mirror->iter = c.iter;
  mirror->max_iters = c.max_iters;
  mirror->stop_it = c.stop_it;
  mirror->kern_launches = c.kern_launches;
  __threadfence_system();
  __hip_atomic_store(&mirror->done, c.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One pass of a lane over a block row in the CG kernel: the U column ids,
// value blocks and gathered states {r, q, p, D^-1} of the block columns. The
// loads depend on neither alpha nor beta, so the first pass of a workgroup is
// issued before the partial sums of the previous iteration are reduced, and
// the two memory round trips overlap.
template <int U, typename T>
struct CgPass {
  int c[U];
  T v[U][9];
  CgDof<T> g[U][3];
};

template <int G, int U, typename T>
__device__ __forceinline__ void cg_load_pass(int p, int p1, const int* __restrict__ col,
                                             const T* __restrict__ val,
                                             const CgDof<T>* __restrict__ s_old,
                                             CgPass<U, T>& P) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int q = p + u * G;
    P.c[u] = q < p1 ? col[q] : -1;
    if (q < p1) {
#pragma unroll
      for (int k = 0; k < 9; ++k) P.v[u][k] = val[(size_t)9 * q + k];
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (P.c[u] >= 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) P.g[u][k] = s_old[3 * P.c[u] + k];
    }
}

// y += K_pass p with p = D^-1 (r - alpha q) + beta p of every gathered column
template <int U, typename T>
__device__ __forceinline__ void cg_apply_pass(const CgPass<U, T>& P, T a, T bt, T& y0, T& y1,
                                              T& y2) {
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (P.c[u] >= 0) {
      T x[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        T r, z;
        cg_form(P.g[u][k], a, bt, r, z, x[k]);
      }
      y0 += P.v[u][0] * x[0] + P.v[u][1] * x[1] + P.v[u][2] * x[2];
      y1 += P.v[u][3] * x[0] + P.v[u][4] * x[1] + P.v[u][5] * x[2];
      y2 += P.v[u][6] * x[0] + P.v[u][7] * x[1] + P.v[u][8] * x[2];
    }
}

template <int G, int U, int BS, typename T>
__global__ __launch_bounds__(BS) void k_cg_iter(
    int it, int nrows, const int* __restrict__ rowptr, const int* __restrict__ col,
    const T* __restrict__ val, const CgDof<T>* __restrict__ s_old, CgDof<T>* __restrict__ s_new,
    double* __restrict__ x, PcgCtrl* ctrl, PcgCtrl* mirror, const double* __restrict__ parts_old,
    double* __restrict__ parts_new, const unsigned long long* __restrict__ stamps_old,
    unsigned long long* __restrict__ stamps_new) {
  __shared__ double sm[kCgParts * (BS / kWave)];
  __shared__ int s_done;
  const unsigned long long t_start = (unsigned long long)wall_clock64();
  // one read of the stop flag per workgroup: workgroup 0 may set it during
  // this launch, and waves that read it at different times would otherwise
  // split, leaving the block sums below with the slots of exited waves unset
  if (threadIdx.x == 0) s_done = ctrl->done;
  __syncthreads();
  if (s_done) return;
  const int nb = gridDim.x;
  // progress for the host, which keeps a few kernels queued ahead of it
  if (mirror && blockIdx.x == 0 && threadIdx.x == 0)
    __hip_atomic_store(&mirror->started, it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  const int lane = threadIdx.x & (G - 1);
  // XCD x (= blockIdx % 8) owns rows [nrows x / 8, nrows (x+1) / 8); its nbx
  // workgroups split that range into equal contiguous chunks.
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, nbx = nb >> 3;
  const long long r0 = (long long)nrows * xcd / 8, r1 = (long long)nrows * (xcd + 1) / 8;
  const int c0 = (int)(r0 + (r1 - r0) * lb / nbx);
  const int c1 = (int)(r0 + (r1 - r0) * (lb + 1) / nbx);
  // first pass of this lane's first row, in flight during the reduction
  int row = c0 + (int)threadIdx.x / G;
  int p0 = 0, p1 = 0;
  CgPass<U, T> P;
  CgDof<T> own;
  if (row < c1) {
    p0 = rowptr[row];
    p1 = rowptr[row + 1];
    cg_load_pass<G, U>(p0 + lane, p1, col, val, s_old, P);
    if (lane < 3) own = s_old[3 * row + lane];
  }
  double alpha = 0.0, beta = 0.0;
  if (it > 0) {
    double v[kCgParts];
    cg_reduce_parts<BS>(parts_old, nb, nb, v, sm);
    if (stamps_old && blockIdx.x == 0) cg_account<BS>(stamps_old, nb, ctrl, sm);
    const double pq = v[0], zq = v[1], qdq = v[2], rr = v[3], rz = v[4];
    const int j = it - 1;  // iteration of the partials
    const bool converged = rr < ctrl->threshold;
    if (converged || j >= ctrl->max_iters) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->rr = rr;
        ctrl->iter = converged ? (j > 0 ? j - 1 : 0) : j;
        ctrl->stop_it = it;
        ctrl->done = 1;
        if (mirror) publish_stop(*ctrl, mirror);
      }
      return;
    }
    alpha = rz / pq;
    beta = (rz - 2.0 * alpha * zq + alpha * alpha * qdq) / rz;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->rr = rr;
  }
  const T a = (T)alpha, bt = (T)beta;
  double acc[kCgParts];
#pragma unroll
  for (int k = 0; k < kCgParts; ++k) acc[k] = 0.0;
  for (bool first = true; row < c1; row += BS / G, first = false) {
    if (!first) {
      p0 = rowptr[row];
      p1 = rowptr[row + 1];
      cg_load_pass<G, U>(p0 + lane, p1, col, val, s_old, P);
      if (lane < 3) own = s_old[3 * row + lane];
    }
    T y0 = 0, y1 = 0, y2 = 0;
    cg_apply_pass<U>(P, a, bt, y0, y1, y2);
    for (int p = p0 + lane + U * G; p < p1; p += U * G) {
      cg_load_pass<G, U>(p, p1, col, val, s_old, P);
      cg_apply_pass<U>(P, a, bt, y0, y1, y2);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
      y0 += __shfl_xor(y0, off, G);
      y1 += __shfl_xor(y1, off, G);
      y2 += __shfl_xor(y2, off, G);
    }
    if (lane < 3) {
      const int i = 3 * row + lane;
      const CgDof<T> s = own;
      T r, z, p;
      cg_form(s, a, bt, r, z, p);
      x[i] += alpha * (double)s.p;
      const T q = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
      CgDof<T> o;
      o.r = r;
      o.q = q;
      o.p = p;
      o.d = s.d;
      s_new[i] = o;
      const double dq = (double)q, dz = (double)z, dr = (double)r;
      acc[0] += (double)p * dq;
      acc[1] += dz * dq;
      acc[2] += dq * ((double)s.d * dq);
      acc[3] += dr * dr;
      acc[4] += dr * dz;
    }
  }
  block_sum_n<BS, kCgParts>(acc, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kCgParts; ++k) parts_new[k * nb + blockIdx.x] = acc[k];
    if (stamps_new) {
      stamps_new[2 * blockIdx.x] = t_start;
      stamps_new[2 * blockIdx.x + 1] = (unsigned long long)wall_clock64();
    }
  }
}

// Batched CG: many independent systems (the diagonal blocks of a block-
// diagonal K, one per simulation) advance one iteration per launch. The
// workgroups [wg0[b], wg0[b+1]) belong to system b and split its block rows
// [row0[b], row0[b+1]) into equal chunks; each system has its own control
// block, scalars and Eigen stopping test (the single-system kernel above,
// with the partial-sum reduction restricted to the system's workgroups).
//
// Graph replay: with it_in set, the iteration index is read from *it_in
// instead of the argument, and workgroup 0 writes it + 1 to *it_out (another
// word than it_in, so no workgroup of this launch reads what it writes), so
// that a captured pair of launches advances two iterations per replay.
template <int G, int U, int BS, typename T>
__global__ __launch_bounds__(BS) void k_cg_iter_batch(
    int it_arg, const int* __restrict__ rowptr, const int* __restrict__ col,
    const T* __restrict__ val, const CgDof<T>* __restrict__ s_old, CgDof<T>* __restrict__ s_new,
    double* __restrict__ x, PcgCtrl* ctrls, const double* __restrict__ parts_old,
    double* __restrict__ parts_new, int nwg, const int* __restrict__ wg_sys,
    const int* __restrict__ wg0, const int* __restrict__ row0, const int* __restrict__ active,
    const int* it_in, int* it_out, PcgCtrl* mirror, PcgCtrl* progress) {
  __shared__ double sm[kCgParts * (BS / kWave)];
  const int it = it_in ? *it_in : it_arg;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it_out) *it_out = it + 1;
    // progress for the host, which keeps a few kernels queued ahead of it
    if (progress)
      __hip_atomic_store(&progress->started, it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  // the workgroup this block runs: itself, or the active list's entry
  const int wg = active ? active[blockIdx.x] : (int)blockIdx.x;
  const int b = wg_sys[wg];
  PcgCtrl* ctrl = ctrls + b;
  // one read of the stop flag per workgroup (see k_cg_iter)
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = ctrl->done;
  __syncthreads();
  if (s_done) return;
  const int w0 = wg0[b], nw = wg0[b + 1] - w0;
  double alpha = 0.0, beta = 0.0;
  if (it > 0) {
    double v[kCgParts];
    cg_reduce_parts<BS>(parts_old + w0, nwg, nw, v, sm);
    const double pq = v[0], zq = v[1], qdq = v[2], rr = v[3], rz = v[4];
    const int j = it - 1;
    const bool converged = rr < ctrl->threshold;
    if (converged || j >= ctrl->max_iters) {
      if (wg == w0 && threadIdx.x == 0) {
        ctrl->rr = rr;
        ctrl->iter = converged ? (j > 0 ? j - 1 : 0) : j;
        ctrl->stop_it = it;
        ctrl->done = 1;
        if (mirror) publish_stop(*ctrl, mirror + b);
      }
      return;
    }
    alpha = rz / pq;
    beta = (rz - 2.0 * alpha * zq + alpha * alpha * qdq) / rz;
    if (wg == w0 && threadIdx.x == 0) ctrl->rr = rr;
  }
  const T a = (T)alpha, bt = (T)beta;
  const int lane = threadIdx.x & (G - 1);
  const int lb = wg - w0;
  const long long r0 = row0[b], r1 = row0[b + 1];
  const int c0 = (int)(r0 + (r1 - r0) * lb / nw);
  const int c1 = (int)(r0 + (r1 - r0) * (lb + 1) / nw);
  double acc[kCgParts];
#pragma unroll
  for (int k = 0; k < kCgParts; ++k) acc[k] = 0.0;
  for (int row = c0 + (int)threadIdx.x / G; row < c1; row += BS / G) {
    T y0, y1, y2;
    bsr_row_batched<G, U>(row, lane, rowptr, col, val,
                          [&](int c, int k) {
                            T r, z, p;
                            cg_form(s_old[3 * c + k], a, bt, r, z, p);
                            return p;
                          },
                          y0, y1, y2);
    if (lane < 3) {
      const int i = 3 * row + lane;
      const CgDof<T> st = s_old[i];
      T r, z, p;
      cg_form(st, a, bt, r, z, p);
      x[i] += alpha * (double)st.p;
      const T q = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
      CgDof<T> o;
      o.r = r;
      o.q = q;
      o.p = p;
      o.d = st.d;
      s_new[i] = o;
      const double dq = (double)q, dz = (double)z, dr = (double)r;
      acc[0] += (double)p * dq;
      acc[1] += dz * dq;
      acc[2] += dq * ((double)st.d * dq);
      acc[3] += dr * dr;
      acc[4] += dr * dz;
    }
  }
  block_sum_n<BS, kCgParts>(acc, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kCgParts; ++k) parts_new[(size_t)k * nwg + wg] = acc[k];
  }
}

// Per-system control blocks: threshold max(tol_b^2 |b_b|^2, DBL_MIN) from the
// per-system |b|^2; systems that are off (or have b = 0) start done.
__global__ void k_cg_setup_batch(int nsys, PcgCtrl* ctrls, const double* __restrict__ b2,
                                 const double* __restrict__ tol, const int* __restrict__ max_iters,
                                 const int* __restrict__ on, PcgCtrl* mirror) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nsys) return;
  PcgCtrl c;
  c.b2 = b2[b];
  c.rr = c.b2;
  c.threshold = fmax(tol[b] * tol[b] * c.b2, DBL_MIN);
  c.kern_ticks = 0.0;
  c.iter = 0;
  c.max_iters = max_iters[b];
  c.stop_it = 0;
  c.kern_launches = 0;
  c.started = 0;
  c.done = (!on[b] || c.b2 == 0.0) ? 1 : 0;
  ctrls[b] = c;
  if (mirror) publish_stop(c, mirror + b);
}

// x = 0, state = {b, 0, 0, D^-1}; partials of b.b
template <typename T>
__global__ __launch_bounds__(kBS) void k_cg_init(int n, const double* __restrict__ b,
                                                 const T* __restrict__ invdiag,
                                                 double* __restrict__ x, CgDof<T>* __restrict__ s0,
                                                 double* __restrict__ parts) {
  __shared__ double sm[kBS / kWave];
  double bb = 0;
  for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) {
    const double bi = b[i];
    x[i] = 0.0;
    CgDof<T> o;
    o.r = (T)bi;
    o.q = (T)0;
    o.p = (T)0;
    o.d = invdiag[i];
    s0[i] = o;
    bb += bi * bi;
  }
  block_partial<kBS, 0>(bb, parts, sm);
}

// |b|^2 and the stopping threshold max(tol^2 |b|^2, DBL_MIN).
__global__ __launch_bounds__(kBS) void k_cg_setup(PcgCtrl* ctrl, const double* __restrict__ parts,
                                                  int np, double tol, int max_iters,
                                                  PcgCtrl* mirror) {
  __shared__ double sm[kBS / kWave];
  const double b2 = reduce_array<kBS, 0>(parts, np, 1, 0, sm);
  if (threadIdx.x == 0) {
    ctrl->b2 = b2;
    ctrl->rr = b2;
    ctrl->threshold = fmax(tol * tol * b2, DBL_MIN);
    ctrl->kern_ticks = 0.0;
    ctrl->iter = 0;
    ctrl->max_iters = max_iters;
    ctrl->stop_it = 0;
    ctrl->kern_launches = 0;
    ctrl->started = -1;
    ctrl->done = (b2 == 0.0) ? 1 : 0;  // x = 0 (Eigen returns at once)
    if (mirror && ctrl->done) publish_stop(*ctrl, mirror);
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

}  // namespace

int spmv_variant_count() { return kNumShapes; }
void spmv_variant_shape(int v, int* G, int* U, int* NT) {
  const SpmvShape s = (v >= 0 && v < kNumShapes) ? kShapes[v] : kShapes[0];
  *G = s.G;
  *U = s.U;
  *NT = s.NT;
}

int cg_variant_count() { return kNumCgShapes; }
void cg_variant_shape(int v, int* G, int* U, int* BS) {
  const CgShape s = (v >= 0 && v < kNumCgShapes) ? kCgShapes[v] : kCgShapes[0];
  *G = s.G;
  *U = s.U;
  *BS = s.BS;
}

int cg_blocks_per_cu(int variant, int n, int per_cu) {
  const CgShape s = (variant >= 0 && variant < kNumCgShapes) ? kCgShapes[variant] : kCgShapes[0];
  // no more workgroups than chunks of at least one full pass of rows
  const long rows_per_xcd = (n / 3 + 7) / 8;
  const long max_nbx = std::max<long>(1, rows_per_xcd / (s.BS / s.G));
  const long nbx = std::min<long>((long)per_cu * kCusPerXcd, max_nbx);
  return (int)std::min<long>(8 * nbx, kCgMaxBlocks);
}

int pcg_default_blocks(int n) { return cg_blocks_per_cu(0, n, 2); }

void launch_finalize(const double* partials, int n, int op, double* result, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, st, partials, n, op, result);
}

template <typename Real, typename VT>
void launch_jacobi_diag(int nfree, const int* diag_block, const Real* K, VT* invdiag,
                        hipStream_t st) {
  if (nfree == 0) return;
  hipLaunchKernelGGL((k_jacobi_diag<Real, VT>), dim3(grid(nfree, kBS)), dim3(kBS), 0, st, nfree,
                     diag_block, K, invdiag);
}

// Dense column-major copy of the BSR matrix (the direct-solver path):
// A[(3 col + c) lda + 3 row + r] = K_b(r, c), one block row per lane.
template <typename Real>
__global__ __launch_bounds__(kBS) void k_bsr_to_dense(int nrows, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col,
                                                      const Real* __restrict__ K,
                                                      double* __restrict__ A, long lda) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= nrows) return;
  for (int b = rowptr[i]; b < rowptr[i + 1]; ++b) {
    const long j = col[b];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) A[(3 * j + c) * lda + 3L * i + r] = (double)K[9L * b + 3 * r + c];
  }
}

template <typename Real>
void launch_bsr_to_dense(int nrows, const int* rowptr, const int* col, const Real* K, double* A,
                         long lda, hipStream_t st) {
  if (nrows == 0) return;
  hipLaunchKernelGGL(k_bsr_to_dense<Real>, dim3(grid(nrows, kBS)), dim3(kBS), 0, st, nrows, rowptr,
                     col, K, A, lda);
}
template void launch_bsr_to_dense<float>(int, const int*, const int*, const float*, double*, long,
                                         hipStream_t);
template void launch_bsr_to_dense<double>(int, const int*, const int*, const double*, double*,
                                          long, hipStream_t);

template <typename Real>
void launch_assemble_bsr_rhs(int nub, const int* upper, const int* tpos, const int* slot_ptr,
                             const int* slots, const double* mass, const Real* blocks, Real* K,
                             int nfree, const int* rows, const int* inc_ptr, const int* inc,
                             const double* fe, double* rhs, hipStream_t st) {
  const int nb_bsr = nub > 0 ? xcd_grid(4 * nub, kBS) : 0;
  const int nb_rhs = nfree > 0 ? xcd_grid(nfree, kBS) : 0;
  if (nb_bsr + nb_rhs == 0) return;
  hipLaunchKernelGGL(k_assemble_bsr_rhs<Real>, dim3(nb_bsr + nb_rhs), dim3(kBS), 0, st, nub,
                     upper, tpos, slot_ptr, slots, mass, blocks, K, nb_bsr, nfree, rows, inc_ptr,
                     inc, fe, rhs);
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

template <typename T>
void launch_pcg_init(const PcgBuffers<T>& B, double tol, int max_iters, hipStream_t st) {
  const int gv = std::min(grid(B.n, kBS), 256);
  hipLaunchKernelGGL(k_cg_init<T>, dim3(gv), dim3(kBS), 0, st, B.n, B.b, B.invdiag, B.x,
                     reinterpret_cast<CgDof<T>*>(B.state[0]), B.parts[1]);
  hipLaunchKernelGGL(k_cg_setup, dim3(1), dim3(kBS), 0, st, B.ctrl, B.parts[1], gv, tol, max_iters,
                     B.mirror);
}

template <typename T>
void launch_pcg_iteration(const PcgBuffers<T>& B, int it, hipStream_t st) {
  if (B.blocks < 8 || B.blocks > kCgMaxBlocks || (B.blocks & 7))
    throw std::invalid_argument("CG grid must be a multiple of 8 in [8, 4096]");
  const auto* so = reinterpret_cast<const CgDof<T>*>(B.state[it & 1]);
  auto* sn = reinterpret_cast<CgDof<T>*>(B.state[(it + 1) & 1]);
  const double* po = B.parts[(it + 1) & 1];
  double* pn = B.parts[it & 1];
  const unsigned long long* to = B.stamps[(it + 1) & 1];
  unsigned long long* tn = B.stamps[it & 1];
  const int nrows = B.n / 3;
  switch (B.variant) {
#define MFEM_CG(i)                                                                             \
  case i:                                                                                      \
    hipLaunchKernelGGL((k_cg_iter<kCgShapes[i].G, kCgShapes[i].U, kCgShapes[i].BS, T>),          \
                       dim3(B.blocks), dim3(kCgShapes[i].BS), 0, st, it, nrows, B.rowptr, B.col, \
                       B.K, so, sn, B.x, B.ctrl, B.mirror, po, pn, to, tn);                    \
    break;
    MFEM_CG(0) MFEM_CG(1) MFEM_CG(2) MFEM_CG(3) MFEM_CG(4) MFEM_CG(5) MFEM_CG(6) MFEM_CG(7)
#undef MFEM_CG
    default:
      throw std::invalid_argument("CG kernel shape out of range");
  }
}

template <typename T>
void launch_pcg_batch_init(const BatchCgBuffers<T>& B, double* scratch, hipStream_t st) {
  const int gv = std::min(grid(B.n, kBS), 256);
  hipLaunchKernelGGL(k_cg_init<T>, dim3(gv), dim3(kBS), 0, st, B.n, B.b, B.invdiag, B.x,
                     reinterpret_cast<CgDof<T>*>(B.state[0]), scratch);
}

void launch_pcg_batch_setup(int nsys, PcgCtrl* ctrls, const double* b2, const double* tol,
                            const int* max_iters, const int* on, hipStream_t st,
                            PcgCtrl* mirror) {
  if (nsys == 0) return;
  hipLaunchKernelGGL(k_cg_setup_batch, dim3(grid(nsys, kBS)), dim3(kBS), 0, st, nsys, ctrls, b2, tol,
                     max_iters, on, mirror);
}

template <typename T>
void launch_pcg_batch_iteration(const BatchCgBuffers<T>& B, int it, hipStream_t st,
                                const int* it_in, int* it_out) {
  if (B.nwg <= 0) return;
  if (B.active && (B.nactive <= 0 || B.nactive > B.nwg))
    throw std::invalid_argument("batched CG: active workgroup count out of range");
  const int grid = B.active ? B.nactive : B.nwg;
  // with it_in set, `it` only selects the buffer parity
  const auto* so = reinterpret_cast<const CgDof<T>*>(B.state[it & 1]);
  auto* sn = reinterpret_cast<CgDof<T>*>(B.state[(it + 1) & 1]);
  const double* po = B.parts[(it + 1) & 1];
  double* pn = B.parts[it & 1];
  switch (B.variant) {
#define MFEM_CGB(i)                                                                             \
  case i:                                                                                      \
    hipLaunchKernelGGL((k_cg_iter_batch<kCgShapes[i].G, kCgShapes[i].U, kCgShapes[i].BS, T>),    \
                       dim3(grid), dim3(kCgShapes[i].BS), 0, st, it, B.rowptr, B.col, B.K, so,   \
                       sn, B.x, B.ctrls, po, pn, B.nwg, B.wg_sys, B.wg0, B.row0, B.active, it_in, \
                       it_out, B.mirror, B.progress);                                          \
    break;
    MFEM_CGB(0) MFEM_CGB(1) MFEM_CGB(2) MFEM_CGB(3) MFEM_CGB(4) MFEM_CGB(5) MFEM_CGB(6) MFEM_CGB(7)
#undef MFEM_CGB
    default:
      throw std::invalid_argument("CG kernel shape out of range");
  }
}

namespace {
__global__ void k_set_int(int* p, int v) { *p = v; }
}  // namespace

void launch_set_int(int* p, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p, v);
}

template void launch_pcg_batch_init<double>(const BatchCgBuffers<double>&, double*, hipStream_t);
template void launch_pcg_batch_init<float>(const BatchCgBuffers<float>&, double*, hipStream_t);
template void launch_pcg_batch_iteration<double>(const BatchCgBuffers<double>&, int, hipStream_t,
                                                 const int*, int*);
template void launch_pcg_batch_iteration<float>(const BatchCgBuffers<float>&, int, hipStream_t,
                                                const int*, int*);

template void launch_assemble_bsr_rhs<double>(int, const int*, const int*, const int*,
                                              const int*, const double*, const double*, double*,
                                              int, const int*, const int*, const int*,
                                              const double*, double*, hipStream_t);
template void launch_assemble_bsr_rhs<float>(int, const int*, const int*, const int*,
                                             const int*, const double*, const float*, float*,
                                             int, const int*, const int*, const int*,
                                             const double*, double*, hipStream_t);
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
template void launch_pcg_init<double>(const PcgBuffers<double>&, double, int, hipStream_t);
template void launch_pcg_init<float>(const PcgBuffers<float>&, double, int, hipStream_t);
template void launch_pcg_iteration<double>(const PcgBuffers<double>&, int, hipStream_t);
template void launch_pcg_iteration<float>(const PcgBuffers<float>&, int, hipStream_t);

}  // namespace mfemhip
