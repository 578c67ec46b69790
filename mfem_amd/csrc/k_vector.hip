// Vector kernels of the timestep: line-search acceptance, BDF history
// (reference src/time_integrators/BDF.cpp:7-28) and small reductions.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "reduce.h"

namespace mfemhip {

namespace {

constexpr int kBS = 256;

struct Hist {
  const double* p[6];
  double a[6];
};

__global__ __launch_bounds__(kBS) void k_accept_step(int nV, const int* __restrict__ free_map,
                                                     double* __restrict__ x,
                                                     const double* __restrict__ dx, long ns,
                                                     double* __restrict__ s,
                                                     const double* __restrict__ ds, double alpha) {
  const long i = (long)blockIdx.x * kBS + threadIdx.x;
  if (i < nV) {
    const int f = free_map[i];
    if (f >= 0) {
#pragma unroll
      for (int r = 0; r < 3; ++r) x[3 * i + r] += alpha * dx[3 * f + r];
    }
  }
  for (long k = i; k < ns; k += (long)gridDim.x * kBS) s[k] += alpha * ds[k];
}

__global__ __launch_bounds__(kBS) void k_bdf_xtilde(int n, int order, Hist xh, Hist vh, double dt,
                                                    double* __restrict__ xt) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  double wx = 0, wv = 0;
  for (int j = 0; j < order; ++j) {
    wx += xh.a[j] * xh.p[j][i];
    wv += xh.a[j] * vh.p[j][i];
  }
  xt[i] = wx + dt * wv;
}

// x_new / v_new may alias the oldest history slot: each lane reads its
// element of every slot before writing.
__global__ __launch_bounds__(kBS) void k_bdf_update(int n, int order, Hist xh, double dt,
                                                    const double* __restrict__ x, double* x_new,
                                                    double* v_new) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  double wx = 0;
  for (int j = 0; j < order; ++j) wx += xh.a[j] * xh.p[j][i];
  const double xi = x[i];
  x_new[i] = xi;
  v_new[i] = (xi - wx) / dt;
}

__global__ __launch_bounds__(kBS) void k_maxabs(int n, const double* __restrict__ v, ReduceSlot out) {
  __shared__ double sm[kBS / kWave];
  double m = 0;
  for (int i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) m = fmax(m, fabs(v[i]));
  block_partial<kBS, 1>(m, out.partials, sm);
}

__global__ void k_fill(double* p, double v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_axpy(long n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += alpha * x[i];
}

inline int grid(long n, int bs) { return n > 0 ? (int)((n + bs - 1) / bs) : 1; }

}  // namespace

void launch_accept_step(const TetMeshDev& m, double* x_full, const double* dx, double* s,
                        const double* ds, double alpha, int with_s, hipStream_t st) {
  long ns = with_s ? 6L * m.nT : 0;
  hipLaunchKernelGGL(k_accept_step, dim3(grid(m.nV, kBS)), dim3(kBS), 0, st, m.nV, m.free_map,
                     x_full, dx, ns, s, ds, alpha);
}

void launch_bdf_xtilde(int n, int order, const double* const* xh, const double* const* vh,
                       const double* a, double dt, double* x_tilde, hipStream_t st) {
  Hist X{}, Vh{};
  for (int j = 0; j < order; ++j) {
    X.p[j] = xh[j];
    X.a[j] = a[j];
    Vh.p[j] = vh[j];
    Vh.a[j] = a[j];
  }
  hipLaunchKernelGGL(k_bdf_xtilde, dim3(grid(n, kBS)), dim3(kBS), 0, st, n, order, X, Vh, dt,
                     x_tilde);
}

void launch_bdf_update(int n, int order, const double* const* xh, const double* a, double dt,
                       const double* x_full, double* x_new, double* v_new, hipStream_t st) {
  Hist X{};
  for (int j = 0; j < order; ++j) {
    X.p[j] = xh[j];
    X.a[j] = a[j];
  }
  hipLaunchKernelGGL(k_bdf_update, dim3(grid(n, kBS)), dim3(kBS), 0, st, n, order, X, dt, x_full,
                     x_new, v_new);
}

void launch_reduce_maxabs(int n, const double* v, ReduceSlot out, hipStream_t st) {
  int g = grid(n, kBS);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_maxabs, dim3(g), dim3(kBS), 0, st, n, v, out);
  launch_finalize(out.partials, g, 1, out.result, st);
}

// x[3 ids[k] + r] = pos[3k + r]: prescribed positions of pinned vertices
__global__ __launch_bounds__(kBS) void k_scatter3(int n, const int* __restrict__ ids,
                                                  const double* __restrict__ pos,
                                                  double* __restrict__ x) {
  const int k = blockIdx.x * kBS + threadIdx.x;
  if (k >= n) return;
  const long v = ids[k];
#pragma unroll
  for (int r = 0; r < 3; ++r) x[3 * v + r] = pos[3 * k + r];
}

void launch_scatter3(int n, const int* ids, const double* pos, double* x, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scatter3, dim3(grid(n, kBS)), dim3(kBS), 0, st, n, ids, pos, x);
}

void launch_fill(double* p, double v, long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid(n, kBS)), dim3(kBS), 0, st, p, v, n);
}

void launch_axpy(long n, double alpha, const double* x, double* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy, dim3(grid(n, kBS)), dim3(kBS), 0, st, n, alpha, x, y);
}

namespace {

// Stage 1 of the per-segment reductions: workgroup (b, p) reduces the p-th
// of kSegSlices equal slices of segment b (every component) into
// part[b * kSegSlices + p].
__global__ __launch_bounds__(kBS) void k_seg_reduce_slices(const double* __restrict__ v, int ncomp,
                                                           long comp_stride,
                                                           const int* __restrict__ seg, int mode,
                                                           double* __restrict__ part) {
  __shared__ double sm[kBS / kWave];
  const int b = blockIdx.x / kSegSlices, p = blockIdx.x % kSegSlices;
  const long lo0 = seg[b], len = seg[b + 1] - lo0;
  const int lo = (int)(lo0 + len * p / kSegSlices), hi = (int)(lo0 + len * (p + 1) / kSegSlices);
  double acc = 0.0;
  for (int c = 0; c < ncomp; ++c) {
    const double* vc = v + (size_t)c * comp_stride;
    for (int i = lo + (int)threadIdx.x; i < hi; i += kBS) {
      const double a = vc[i];
      if (mode == 0) acc += a;
      else if (mode == 1) acc += a * a;
      else acc = fmax(acc, fabs(a));
    }
  }
  const double r = mode == 2 ? block_reduce_all<kBS, 1>(acc, sm) : block_reduce_all<kBS, 0>(acc, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Stage 2: one thread per segment combines its slices in slice order.
__global__ __launch_bounds__(kBS) void k_seg_reduce_combine(int nseg, const double* __restrict__ part,
                                                            int mode, double* __restrict__ out) {
  const int b = blockIdx.x * kBS + threadIdx.x;
  if (b >= nseg) return;
  double acc = part[(size_t)b * kSegSlices];
  for (int p = 1; p < kSegSlices; ++p) {
    const double a = part[(size_t)b * kSegSlices + p];
    acc = mode == 2 ? fmax(acc, a) : acc + a;
  }
  out[b] = acc;
}

__global__ __launch_bounds__(kBS) void k_accept_batch(int nV, const int* __restrict__ free_map,
                                                      const int* __restrict__ vert_sim, int nT,
                                                      const int* __restrict__ tet_sim,
                                                      const double* __restrict__ alpha_sim,
                                                      double* __restrict__ x,
                                                      const double* __restrict__ dx,
                                                      double* __restrict__ s,
                                                      const double* __restrict__ ds) {
  const long i = (long)blockIdx.x * kBS + threadIdx.x;
  for (long v = i; v < nV; v += (long)gridDim.x * kBS) {
    const int f = free_map[v];
    const double a = alpha_sim[vert_sim[v]];
    if (f >= 0 && a != 0.0) {
#pragma unroll
      for (int r = 0; r < 3; ++r) x[3 * v + r] += a * dx[3 * f + r];
    }
  }
  if (s) {
    const long ns = 6L * nT;
    for (long k = i; k < ns; k += (long)gridDim.x * kBS) {
      const double a = alpha_sim[tet_sim[k % nT]];
      if (a != 0.0) s[k] += a * ds[k];
    }
  }
}

}  // namespace

void launch_seg_reduce(int nseg, const double* v, int ncomp, long comp_stride, const int* seg,
                       int mode, double* out, double* part, hipStream_t st) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_seg_reduce_slices, dim3(nseg * kSegSlices), dim3(kBS), 0, st, v, ncomp,
                     comp_stride, seg, mode, part);
  hipLaunchKernelGGL(k_seg_reduce_combine, dim3(grid(nseg, kBS)), dim3(kBS), 0, st, nseg, part,
                     mode, out);
}

void launch_accept_batch(int nV, const int* free_map, const int* vert_sim, int nT,
                         const int* tet_sim, const double* alpha_sim, double* x_full,
                         const double* dx, double* s, const double* ds, hipStream_t st) {
  const long n = std::max<long>(nV, s ? 6L * nT : 0L);
  if (n == 0) return;
  const int g = (int)std::min<long>((n + kBS - 1) / kBS, 65535);
  hipLaunchKernelGGL(k_accept_batch, dim3(g), dim3(kBS), 0, st, nV, free_map, vert_sim, nT, tet_sim,
                     alpha_sim, x_full, dx, s, ds);
}

}  // namespace mfemhip
