// Device-side reduction helpers shared by the HIP kernels of the FEM step.
#pragma once

#include <hip/hip_runtime.h>

#include "sdx_kernels.h"

namespace sdx {

// ------------------------------------------------------------------ reductions
//
// Cross-block sums use spread atomic accumulators instead of a last-block
// reduction: on a multi-XCD CDNA device an agent-scope fence per block (L2
// write-back / invalidate) costs far more than the kernel's own traffic,
// whereas FP64 atomic adds spread over kSlots separate 128-byte lines cost
// almost nothing.  Consumers (the next kernel, or the host) sum the slots in
// a fixed order; kernel boundaries order producers and consumers.

template <class A>
__device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum NV values over the block; result valid in thread 0.
template <class A, int NV>
__device__ __forceinline__ void block_sum(A (&v)[NV]) {
  __shared__ A sh[NV][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][wid] = v[k];
  __syncthreads();
  if (wid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      A t = lane < nw ? sh[k][lane] : A(0);
      v[k] = wave_sum(t);
    }
  }
  __syncthreads();
}

// Accumulator set `set` of environment e: kSlots lines of kSlotStride values.
template <class A>
__device__ __forceinline__ A* acc_ptr(A* acc, int set, int nEnv, int e) {
  return acc + ((size_t)(set * nEnv + e) * kSlots) * kSlotStride;
}

// Block-sum v (one barrier: wave sums, then thread 0 adds the per-wave
// values) and add it into slot (tile_in_env % kSlots) of `dst`.  Uses its own
// LDS, so a kernel may call it more than once.
template <class A, int NV>
__device__ __forceinline__ void acc_add(A (&v)[NV], A* dst, int tile_in_env) {
  __shared__ A sh[NV][16];
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][threadIdx.x >> 6] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    A* d = dst + (tile_in_env & (kSlots - 1)) * kSlotStride;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      A t = sh[k][0];
      for (int w = 1; w < nw; ++w) t += sh[k][w];
      atomicAdd(d + k, t);
    }
  }
}

template <class A, int NV>
__device__ __forceinline__ void acc_read(const A* src, double (&out)[NV]) {
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = 0.0;
  for (int sl = 0; sl < kSlots; ++sl)
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] += double(src[sl * kSlotStride + k]);
}

// Zero one accumulator set of one environment (kSlots * kSlotStride values).
template <class A>
__device__ __forceinline__ void acc_zero(A* dst) {
  for (int k = threadIdx.x; k < kSlots * kSlotStride; k += blockDim.x) dst[k] = A(0);
}

// Workgroups are dealt to the 8 XCDs round-robin; block b of a grid of n
// gets work item xcd_tile(b, n) so that each XCD takes a contiguous run of
// the items and the neighbouring data they gather stays in its own L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, rem = n % kXcd;
  const int x = b % kXcd, i = b / kXcd;
  return x * q + (x < rem ? x : rem) + i;
}

}  // namespace sdx
