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

// Cross-lane sums on the VALU (DPP lane moves) rather than through the LDS
// crossbar (ds_bpermute).  quad_perm [1,0,3,2] and [2,3,0,1] pair each lane
// with its xor-1 and xor-2 partner, row_half_mirror reverses every group of 8
// lanes and row_mirror every row of 16, so after the four steps each lane of
// a row holds the sum of the row (every lane adds the same two values at each
// step, so all lanes of a group end bit-identical).  Every lane of the wave
// must be active where these are called.
enum : int { kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140 };

template <int kCtrl>
__device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_mov_dpp(v, kCtrl, 0xF, 0xF, false);
}
template <int kCtrl>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(dpp_move<kCtrl>(__float_as_int(v)));
}
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_move<kCtrl>(int(b)), hi = dpp_move<kCtrl>(int(b >> 32));
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Sum over each aligned group of L lanes (L = 2, 4, 8 or 16), in every lane of the group.
template <int L, class A>
__device__ __forceinline__ A group_sum(A v) {
  static_assert(L == 2 || L == 4 || L == 8 || L == 16, "DPP sums groups of 2 to 16 lanes");
  v += dpp_move<kDppXor1>(v);
  if constexpr (L >= 4) v += dpp_move<kDppXor2>(v);
  if constexpr (L >= 8) v += dpp_move<kDppHalfMirror>(v);
  if constexpr (L >= 16) v += dpp_move<kDppMirror>(v);
  return v;
}

__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b), lane), hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Sum over the 64 lanes of the wave, in every lane: the 4 row sums read from
// lanes 0, 16, 32 and 48.
template <class A>
__device__ __forceinline__ A wave_sum(A v) {
  v = group_sum<16>(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
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
