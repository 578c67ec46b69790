// Deterministic block reductions for gfx950 (64-lane wavefronts).
//
// Each block reduces its values in a fixed tree, writes one partial, and
// counts itself in with an integer atomic. The last block to arrive sums all
// partials in index order (each lane a fixed strided subset, then the same
// fixed tree), so the result does not depend on block scheduling. No float
// atomics are used anywhere.
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

// op: 0 = sum, 1 = max. Must be called by every thread of the block.
template <int BS, int OP>
__device__ __forceinline__ double block_reduce(double v, double* sm) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  v = OP == 0 ? wave_sum(v) : wave_max(v);
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) {
    r = sm[0];
#pragma unroll
    for (int w = 1; w < BS / kWave; ++w) r = OP == 0 ? r + sm[w] : fmax(r, sm[w]);
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// Writes this block's partial; the last block reduces all partials into
// *out.result. Returns true in every thread of the block that finalised.
template <int BS, int OP>
__device__ __forceinline__ bool block_reduce_finalize(double v, ReduceSlot out, double* sm) {
  __shared__ int is_last;
  double r = block_reduce<BS, OP>(v, sm);
  if (threadIdx.x == 0) {
    out.partials[blockIdx.x] = r;
    __threadfence();
    unsigned ticket = atomicAdd(out.counter, 1u);
    is_last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double a = OP == 0 ? 0.0 : -INFINITY;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += BS) {
    double p = __hip_atomic_load(&out.partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a = OP == 0 ? a + p : fmax(a, p);
  }
  double tot = block_reduce<BS, OP>(a, sm);
  if (threadIdx.x == 0) {
    *out.result = tot;
    *out.counter = 0u;
  }
  return true;
}

}  // namespace mfemhip
