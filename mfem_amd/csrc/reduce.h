// Deterministic reductions for gfx950 (64-lane wavefronts), without atomics
// or device-scope fences.
//
// A reducing kernel writes one partial per block (fixed tree inside the
// block); the kernel boundary makes the partials visible, and either a
// one-block finalize kernel or every block of the consuming kernel sums them
// in index order (each lane a fixed strided subset, then the same fixed
// tree). The result therefore does not depend on block scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace mfemhip {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// op: 0 = sum, 1 = max. Called by every thread of the block; the result is
// returned to every thread.
template <int BS, int OP>
__device__ __forceinline__ double block_reduce_all(double v, double* sm) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  v = OP == 0 ? wave_sum(v) : wave_max(v);
  __syncthreads();  // sm may still be read by a previous call
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  double r = sm[0];
#pragma unroll
  for (int w = 1; w < BS / kWave; ++w) r = OP == 0 ? r + sm[w] : fmax(r, sm[w]);
  return r;
}

// Sums of N values over the block in one LDS round; sm holds N * BS / 64
// doubles. Every thread gets the N sums.
template <int BS, int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double* sm) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) sm[k * (BS / kWave) + wid] = v[k];
  }
  __syncthreads();
  // one quantity at a time, so large workgroups (16 waves) do not hold all
  // N * waves values in registers at once
#pragma unroll 1
  for (int k = 0; k < N; ++k) {
    double r = sm[k * (BS / kWave)];
#pragma unroll
    for (int w = 1; w < BS / kWave; ++w) r += sm[k * (BS / kWave) + w];
    v[k] = r;
  }
}

// Block partial: partials[blockIdx.x] = reduction of v over the block.
template <int BS, int OP>
__device__ __forceinline__ void block_partial(double v, double* partials, double* sm) {
  double r = block_reduce_all<BS, OP>(v, sm);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// Fixed-order reduction of n values (stride `stride`, offset `off`) by one
// block; result returned to every thread.
template <int BS, int OP>
__device__ __forceinline__ double reduce_array(const double* a, int n, int stride, int off,
                                               double* sm) {
  double acc = OP == 0 ? 0.0 : -INFINITY;
  for (int i = threadIdx.x; i < n; i += BS) {
    double v = a[(size_t)i * stride + off];
    acc = OP == 0 ? acc + v : fmax(acc, v);
  }
  return block_reduce_all<BS, OP>(acc, sm);
}

// Launch the one-block finalize of `n` partials into *result.
void launch_finalize(const double* partials, int n, int op, double* result, hipStream_t st);

}  // namespace mfemhip
