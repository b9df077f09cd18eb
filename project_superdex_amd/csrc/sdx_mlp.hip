// HIP kernels of the learned CG warm start (gfx950).
//
// Inference: one persistent workgroup per CU keeps the BF16 weights of all
// four layers in LDS (112 KB) and streams 32-vertex tiles through its four
// waves.  Each wave runs its tile through the layers with
// v_mfma_f32_16x16x16_bf16 (FP32 accumulation): the activations of a layer
// leave the accumulators, get bias + GELU, and are written to the wave's LDS
// slab as BF16 rows, which the next layer reads back as its A operand.
//
// Training: FP32 SIMT kernels (tiled GEMM with split-K, bias + GELU, GELU
// backward, column sums, MSE gradient, Adam).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sdx_device.h"
#include "sdx_mlp.h"

namespace sdx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// GELU(z) = z Phi(z) with the normal CDF Phi written through erf.
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.7071067811865476f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.7071067811865476f));
  const float pdf = 0.3989422804014327f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// ------------------------------------------------------------------ packing

__global__ void k_mlp_pack(const float* __restrict__ p, uint16_t* __restrict__ w1, uint16_t* __restrict__ w2,
                           uint16_t* __restrict__ w3, uint16_t* __restrict__ w4, float* __restrict__ b1,
                           float* __restrict__ b2, float* __restrict__ b3, float* __restrict__ b4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kMlpHid * kMlpInPad) {
    const int n = i / kMlpInPad, k = i % kMlpInPad;
    w1[i] = k < kMlpIn ? f2bf(p[kOffW1 + n * kMlpIn + k]) : uint16_t(0);
  }
  if (i < kMlpHid * kMlpHid) {
    w2[i] = f2bf(p[kOffW2 + i]);
    w3[i] = f2bf(p[kOffW3 + i]);
  }
  if (i < 16 * kMlpHid) {
    const int n = i / kMlpHid, k = i % kMlpHid;
    w4[i] = n < kMlpOut ? f2bf(p[kOffW4 + n * kMlpHid + k]) : uint16_t(0);
  }
  if (i < kMlpHid) {
    b1[i] = p[kOffB1 + i];
    b2[i] = p[kOffB2 + i];
    b3[i] = p[kOffB3 + i];
  }
  if (i < 16) b4[i] = i < kMlpOut ? p[kOffB4 + i] : 0.0f;
}

void launch_mlp_pack(hipStream_t s, const float* params, uint16_t* w1, uint16_t* w2, uint16_t* w3, uint16_t* w4,
                     float* b1, float* b2, float* b3, float* b4) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((kMlpHid * kMlpHid + 255) / 256), dim3(256), 0, s, params, w1, w2, w3, w4, b1,
                     b2, b3, b4);
}

// ------------------------------------------------------------------ inference
//
// One workgroup of kInferWaves waves per CU holds the BF16 weights in LDS;
// each wave runs tiles of 32 vertices through the layers.  A layer computes
// C[u][v] = sum_k W[u][k] act[v][k] on the matrix cores with W as the A
// operand (rows = units) and the activations as the B operand (columns =
// vertices), so the accumulator of lane l holds 4 consecutive units
// 4 (l / 16) + e of vertex l % 16 (v_mfma_f32_16x16x*_bf16 output layout) and
// bias + GELU + BF16 packing end in one 8-byte LDS store per 4 units.  Both
// operands of an MFMA hold k = K (l / 16) + j in lane l, slot j (K = 4 or 8
// slots); any one lane-to-k map used for both operands gives the same sum.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

constexpr int kLd = kMlpHid + 8;  // LDS row stride (BF16) of W2, W3, W4 and the activations
constexpr int kInferWaves = 8;
constexpr int kRowsPerWave = 32;

// GELU(z) = z Phi(z) = z (1 + erf(z / sqrt 2)) / 2 for inference, with erf
// from Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.26:
// erf(a) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-a^2),
// t = 1 / (1 + p a), a >= 0, |error| <= 1.5e-7 (far below BF16 resolution).
__device__ __forceinline__ float gelu_infer(float z) {
  const float a = fabsf(z) * 0.7071067811865476f;
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, a, 1.0f));
  const float poly =
      t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f), 0.254829592f);
  const float e = __builtin_amdgcn_exp2f(-(a * a) * 1.4426950408889634f);  // exp(-a^2)
  const float erf_v = copysignf(fmaf(-poly, e, 1.0f), z);
  const float hz = 0.5f * z;
  return fmaf(hz, erf_v, hz);
}

// Two FP32 values to packed BF16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

// Bias + GELU of the 128 x 32 accumulator block (units x vertices), stored
// as BF16 activation rows act[v][u].
__device__ __forceinline__ void store_act(f32x4 (&acc)[8][2], const float* __restrict__ bias,
                                          uint16_t* __restrict__ act, int lr, int lg) {
#pragma unroll
  for (int ut = 0; ut < 8; ++ut) {
    const f32x4 b = *reinterpret_cast<const f32x4*>(bias + ut * 16 + 4 * lg);
#pragma unroll
    for (int vt = 0; vt < 2; ++vt) {
      const f32x4 z = acc[ut][vt] + b;
      uint2 q;
      q.x = pack_bf16(gelu_infer(z[0]), gelu_infer(z[1]));
      q.y = pack_bf16(gelu_infer(z[2]), gelu_infer(z[3]));
      *reinterpret_cast<uint2*>(act + (vt * 16 + lr) * kLd + ut * 16 + 4 * lg) = q;
    }
  }
}

// Order one wave's LDS writes before its own later reads of other lanes' data.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8_t lds_frag8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const s16x8*>(p));
}

// One hidden layer (K = 128): acc[ut][vt] = W[ut block] act[vt block]^T.
__device__ __forceinline__ void hidden_layer(const uint16_t* __restrict__ act, const uint16_t* __restrict__ W,
                                             f32x4 (&acc)[8][2], int lr, int lg) {
#pragma unroll
  for (int ut = 0; ut < 8; ++ut)
#pragma unroll
    for (int vt = 0; vt < 2; ++vt) acc[ut][vt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < kMlpHid / 32; ++ks) {
    const bf16x8_t b0 = lds_frag8(act + lr * kLd + ks * 32 + 8 * lg);
    const bf16x8_t b1 = lds_frag8(act + (16 + lr) * kLd + ks * 32 + 8 * lg);
#pragma unroll
    for (int ut = 0; ut < 8; ++ut) {
      const bf16x8_t w = lds_frag8(W + (ut * 16 + lr) * kLd + ks * 32 + 8 * lg);
      acc[ut][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, b0, acc[ut][0], 0, 0, 0);
      acc[ut][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, b1, acc[ut][1], 0, 0, 0);
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kInferWaves * 64) k_mlp_infer(MlpBf16 w, const float* __restrict__ x, int n,
                                                               const int32_t* __restrict__ row_env,
                                                               const double* __restrict__ out_scale,
                                                               const uint8_t* __restrict__ pinned,
                                                               float* __restrict__ y, T* __restrict__ yt) {
  __shared__ __align__(16) uint16_t sW1[kMlpHid * kMlpInPad];
  __shared__ __align__(16) uint16_t sW2[kMlpHid * kLd];
  __shared__ __align__(16) uint16_t sW3[kMlpHid * kLd];
  __shared__ __align__(16) uint16_t sW4[16 * kLd];
  __shared__ __align__(16) float sB[3 * kMlpHid + 16];
  __shared__ __align__(16) uint16_t sAct[kInferWaves][kRowsPerWave * kLd];
  for (int i = threadIdx.x; i < kMlpHid * kMlpInPad; i += blockDim.x) sW1[i] = w.w1[i];
  for (int i = threadIdx.x; i < kMlpHid * kMlpHid; i += blockDim.x) {
    const int r = i / kMlpHid, c = i % kMlpHid;
    sW2[r * kLd + c] = w.w2[i];
    sW3[r * kLd + c] = w.w3[i];
  }
  for (int i = threadIdx.x; i < 16 * kMlpHid; i += blockDim.x) sW4[(i / kMlpHid) * kLd + i % kMlpHid] = w.w4[i];
  for (int i = threadIdx.x; i < kMlpHid; i += blockDim.x) {
    sB[i] = w.b1[i];
    sB[kMlpHid + i] = w.b2[i];
    sB[2 * kMlpHid + i] = w.b3[i];
  }
  if (threadIdx.x < 16) sB[3 * kMlpHid + threadIdx.x] = w.b4[threadIdx.x];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lg = lane >> 4;
  uint16_t* act = sAct[wave];
  const int n_tiles = (n + kRowsPerWave - 1) / kRowsPerWave;
  f32x4 acc[8][2];
  for (int tile = blockIdx.x * kInferWaves + wave; tile < n_tiles; tile += gridDim.x * kInferWaves) {
    const int row0 = tile * kRowsPerWave;
    // layer 1 (K = 16): the B operand comes straight from the FP32 feature rows
#pragma unroll
    for (int vt = 0; vt < 2; ++vt) {
      const int r = row0 + vt * 16 + lr;
      uint2 q = make_uint2(0u, 0u);
      if (r < n) {
        const float4 f = *reinterpret_cast<const float4*>(x + (size_t)r * kMlpInPad + 4 * lg);
        q.x = pack_bf16(f.x, f.y);
        q.y = pack_bf16(f.z, f.w);
      }
      const bf16x4 bx = __builtin_bit_cast(bf16x4, q);
#pragma unroll
      for (int ut = 0; ut < 8; ++ut) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(sW1 + (ut * 16 + lr) * kMlpInPad + 4 * lg);
        acc[ut][vt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      }
    }
    store_act(acc, sB, act, lr, lg);
    wave_lds_sync();
    hidden_layer(act, sW2, acc, lr, lg);
    wave_lds_sync();
    store_act(acc, sB + kMlpHid, act, lr, lg);
    wave_lds_sync();
    hidden_layer(act, sW3, acc, lr, lg);
    wave_lds_sync();
    store_act(acc, sB + 2 * kMlpHid, act, lr, lg);
    wave_lds_sync();
    // output layer (3 units, padded to 16): lanes 0..15 hold units 0..3 of vertex lr
    f32x4 o[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int ks = 0; ks < kMlpHid / 32; ++ks) {
      const bf16x8_t wv = lds_frag8(sW4 + lr * kLd + ks * 32 + 8 * lg);
#pragma unroll
      for (int vt = 0; vt < 2; ++vt)
        o[vt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv, lds_frag8(act + (vt * 16 + lr) * kLd + ks * 32 + 8 * lg),
                                                        o[vt], 0, 0, 0);
    }
    wave_lds_sync();  // the next tile overwrites act
    if (lg == 0) {
#pragma unroll
      for (int vt = 0; vt < 2; ++vt) {
        const int r = row0 + vt * 16 + lr;
        if (r >= n) continue;
        const float sc = out_scale ? float(out_scale[row_env[r]]) : 1.0f;
#pragma unroll
        for (int e = 0; e < kMlpOut; ++e) {
          float v = (o[vt][e] + sB[3 * kMlpHid + e]) * sc;
          if (pinned && pinned[3 * r + e]) v = 0.0f;
          if (yt)
            yt[3 * r + e] = T(v);
          else
            y[3 * r + e] = v;
        }
      }
    }
  }
}

template <class T>
void launch_mlp_infer(hipStream_t s, const MlpBf16& w, const float* x, int n, const int32_t* row_env,
                      const double* out_scale, const uint8_t* pinned, float* y, T* yt) {
  if (n <= 0) return;
  const int tiles = (n + kRowsPerWave - 1) / kRowsPerWave;
  const int grid = std::max(1, std::min(256, (tiles + kInferWaves - 1) / kInferWaves));
  hipLaunchKernelGGL(k_mlp_infer<T>, dim3(grid), dim3(kInferWaves * 64), 0, s, w, x, n, row_env, out_scale, pinned,
                     y, yt);
}

// ------------------------------------------------------------------ training

constexpr int kGemmTile = 64, kGemmK = 16;

// C = opA . opB over k in [z * kchunk, (z + 1) * kchunk); with split-K the
// partial tiles are added atomically into C.
__global__ void __launch_bounds__(256) k_gemm_f32(int M, int N, int K, const float* __restrict__ A, int lda, int tA,
                                                  const float* __restrict__ B, int ldb, int tB, float* __restrict__ C,
                                                  int ldc, int accumulate, int kchunk, int atomic) {
  __shared__ float sA[kGemmK][kGemmTile + 4];
  __shared__ float sB[kGemmK][kGemmTile + 4];
  const int m0 = blockIdx.y * kGemmTile, n0 = blockIdx.x * kGemmTile;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  float c[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) c[i][j] = 0.0f;
  for (int k0 = kb; k0 < ke; k0 += kGemmK) {
    for (int q = threadIdx.x; q < kGemmTile * kGemmK; q += 256) {
      int mm, kk;
      if (tA) {  // consecutive threads along m
        mm = q % kGemmTile;
        kk = q / kGemmTile;
      } else {  // along k
        mm = q / kGemmK;
        kk = q % kGemmK;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.0f;
      if (gm < M && gk < ke) v = tA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      sA[kk][mm] = v;
    }
    for (int q = threadIdx.x; q < kGemmTile * kGemmK; q += 256) {
      int nn, kk;
      if (tB) {
        nn = q / kGemmK;
        kk = q % kGemmK;
      } else {
        nn = q % kGemmTile;
        kk = q / kGemmTile;
      }
      const int gn = n0 + nn, gk = k0 + kk;
      float v = 0.0f;
      if (gn < N && gk < ke) v = tB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      sB[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGemmK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sB[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) c[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gm = m0 + ty * 4 + i, gn = n0 + tx * 4 + j;
      if (gm >= M || gn >= N) continue;
      float* d = C + (size_t)gm * ldc + gn;
      if (atomic)
        atomicAdd(d, c[i][j]);
      else
        *d = accumulate ? *d + c[i][j] : c[i][j];
    }
}

__global__ void k_zero_f32(int M, int N, float* C, int ldc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M * N) C[(size_t)(i / N) * ldc + i % N] = 0.0f;
}

void launch_gemm_f32(hipStream_t s, int M, int N, int K, const float* A, int lda, bool transA, const float* B,
                     int ldb, bool transB, float* C, int ldc, bool accumulate) {
  if (M <= 0 || N <= 0) return;
  const int gx = (N + kGemmTile - 1) / kGemmTile, gy = (M + kGemmTile - 1) / kGemmTile;
  // split long reductions (weight gradients over the batch) across workgroups
  int splits = 1;
  while (gx * gy * splits < 256 && K / (splits * 2) >= 256) splits *= 2;
  const int kchunk = ((K + splits - 1) / splits + kGemmK - 1) / kGemmK * kGemmK;
  if (splits > 1 && !accumulate)
    hipLaunchKernelGGL(k_zero_f32, dim3((M * N + 255) / 256), dim3(256), 0, s, M, N, C, ldc);
  hipLaunchKernelGGL(k_gemm_f32, dim3(gx, gy, splits), dim3(256), 0, s, M, N, K, A, lda, transA ? 1 : 0, B, ldb,
                     transB ? 1 : 0, C, ldc, accumulate ? 1 : 0, kchunk, splits > 1 ? 1 : 0);
}

__global__ void k_bias_gelu(int M, int N, float* __restrict__ Z, const float* __restrict__ b, float* __restrict__ H) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  const float z = Z[i] + b[i % N];
  Z[i] = z;
  if (H) H[i] = gelu(z);
}

void launch_bias_gelu(hipStream_t s, int M, int N, float* Z, const float* b, float* H) {
  const int64_t n = (int64_t)M * N;
  if (n == 0) return;
  hipLaunchKernelGGL(k_bias_gelu, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, M, N, Z, b, H);
}

__global__ void k_gelu_backward(int64_t n, const float* __restrict__ Z, const float* __restrict__ dH,
                                float* __restrict__ dZ) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dZ[i] = dH[i] * gelu_grad(Z[i]);
}

void launch_gelu_backward(hipStream_t s, int64_t n, const float* Z, const float* dH, float* dZ) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gelu_backward, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, Z, dH, dZ);
}

__global__ void __launch_bounds__(256) k_colsum(int M, int N, const float* __restrict__ dZ, float* __restrict__ db) {
  const int c = blockIdx.x;
  float v = 0.0f;
  for (int m = threadIdx.x; m < M; m += 256) v += dZ[(size_t)m * N + c];
  float vv[1] = {v};
  block_sum<float, 1>(vv);
  if (threadIdx.x == 0) db[c] = vv[0];
}

void launch_colsum(hipStream_t s, int M, int N, const float* dZ, float* db) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_colsum, dim3(N), dim3(256), 0, s, M, N, dZ, db);
}

__global__ void __launch_bounds__(256) k_mse_grad(int M, const float* __restrict__ Y, const float* __restrict__ T,
                                                  float* __restrict__ dY, double* loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double part[1] = {0.0};
  if (i < M * kMlpOut) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d / float(M);
    part[0] = double(d) * double(d) / double(M);
  }
  block_sum<double, 1>(part);
  if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

void launch_mse_grad(hipStream_t s, int M, const float* Y, const float* T, float* dY, double* loss) {
  if (M == 0) return;
  hipLaunchKernelGGL(k_mse_grad, dim3((M * kMlpOut + 255) / 256), dim3(256), 0, s, M, Y, T, dY, loss);
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void launch_adam(hipStream_t s, int n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                 float eps, int t) {
  const float c1 = 1.0f - powf(b1, float(t)), c2 = 1.0f - powf(b2, float(t));
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps, c1, c2);
}

__global__ void k_gather_rows(int rows, int width, const float* __restrict__ src, int64_t n, int64_t mult,
                              int64_t off, int64_t base, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * width) return;
  const int64_t r = i / width, c = i % width;
  // mult < 2^31 and base + r < n < 2^31, so the product fits in 64 bits
  const int64_t row = int64_t((uint64_t(mult) * uint64_t(base + r) + uint64_t(off)) % uint64_t(n));
  out[i] = src[row * width + c];
}

void launch_gather_rows(hipStream_t s, int rows, int width, const float* src, int64_t n, int64_t mult, int64_t off,
                        int64_t base, float* out) {
  const int64_t tot = (int64_t)rows * width;
  if (tot == 0) return;
  hipLaunchKernelGGL(k_gather_rows, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, s, rows, width, src, n, mult,
                     off, base, out);
}

// ------------------------------------------------------------------ features

template <class T>
__global__ void __launch_bounds__(kVecThreads) k_ws_scaled_residual(MeshView m, Tiles tv, const T* __restrict__ b,
                                                                    const T* __restrict__ vals,
                                                                    const T* __restrict__ dprev, float* __restrict__ y,
                                                                    double* acc) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int e = tv.env[tile];
  const int i = tv.row0[tile] + threadIdx.x;
  double part[3] = {0.0, 0.0, 0.0};
  if (i < tv.row1[tile]) {
    const int sl = m.diag_slot[i];
    const size_t nB = m.nB;
    const double d = (double(vals[sl]) + double(vals[4 * nB + sl]) + double(vals[8 * nB + sl])) / 3.0;
    const double inv = d > 0.0 ? 1.0 / d : 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double yc = double(b[3 * i + c]) * inv;
      const double pc = double(dprev[3 * i + c]);
      y[3 * i + c] = float(yc);
      part[0] += yc * yc;
      part[1] += pc * pc;
    }
    part[2] = 1.0;
  }
  acc_add<double, 3>(part, acc_ptr(acc, 0, m.nEnv, e), tile - tv.first[e]);
}

template <class T>
void launch_ws_scaled_residual(hipStream_t s, const MeshView& m, const Tiles& tv, const T* b, const T* vals,
                               const T* dprev, float* y, double* acc) {
  if (tv.n == 0) return;
  hipLaunchKernelGGL(k_ws_scaled_residual<T>, dim3(tv.n), dim3(kVecThreads), 0, s, m, tv, b, vals, dprev, y, acc);
}

__global__ void k_ws_env_scale(int nEnv, const double* __restrict__ acc, double* __restrict__ scale) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEnv) return;
  double q[3];
  acc_read<double, 3>(acc_ptr(acc, 0, nEnv, e), q);
  const double n = q[2] > 0.0 ? q[2] : 1.0;
  const double v = sqrt((q[1] > q[0] ? q[1] : q[0]) / n);
  scale[e] = v > 0.0 ? v : 1.0;
}

void launch_ws_env_scale(hipStream_t s, int nEnv, const double* acc, double* scale) {
  if (nEnv == 0) return;
  hipLaunchKernelGGL(k_ws_env_scale, dim3((nEnv + 63) / 64), dim3(64), 0, s, nEnv, acc, scale);
}

// Four lanes per vertex: the 1-ring loads of a row are split over the lanes
// and issued together (a row holds about 13 neighbours); lane 0 writes the
// feature row.  A 256-row vector tile takes four workgroups.
template <class T>
__global__ void __launch_bounds__(256) k_ws_features(MeshView m, Tiles tv, const T* __restrict__ vals,
                                                     const float* __restrict__ y, const T* __restrict__ dprev,
                                                     const double* __restrict__ scale, float* __restrict__ F) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = b >> 2, quarter = b & 3;
  const int e = tv.env[tile];
  const int i = tv.row0[tile] + quarter * 64 + (int(threadIdx.x) >> 2);
  const int lane = threadIdx.x & 3;
  const bool valid = i < tv.row1[tile];
  float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
  int cnt = 0;
  if (valid) {
    const int k0 = m.row_ptr[i], k1 = m.row_ptr[i + 1];
    for (int k = k0 + lane; k < k1; k += 16) {
      int jj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) jj[u] = k + 4 * u < k1 ? m.col[k + 4 * u] : i;  // i: skipped
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (jj[u] == i) continue;
        n0 += y[3 * jj[u]];
        n1 += y[3 * jj[u] + 1];
        n2 += y[3 * jj[u] + 2];
        ++cnt;
      }
    }
  }
#pragma unroll
  for (int off = 1; off < 4; off <<= 1) {
    n0 += __shfl_xor(n0, off, 4);
    n1 += __shfl_xor(n1, off, 4);
    n2 += __shfl_xor(n2, off, 4);
    cnt += __shfl_xor(cnt, off, 4);
  }
  if (!valid || lane != 0) return;
  const float invs = float(1.0 / scale[e]);
  const int sl = m.diag_slot[i];
  const size_t nB = m.nB;
  const float d00 = float(vals[sl]), d11 = float(vals[4 * nB + sl]), d22 = float(vals[8 * nB + sl]);
  const float d01 = float(vals[1 * nB + sl]), d02 = float(vals[2 * nB + sl]), d12 = float(vals[5 * nB + sl]);
  const float d = (d00 + d11 + d22) / 3.0f;
  const float invd = d > 0.0f ? 1.0f / d : 0.0f;
  const float invc = cnt ? invs / float(cnt) : 0.0f;
  float4* out = reinterpret_cast<float4*>(F + (size_t)i * kMlpInPad);
  out[0] = make_float4(y[3 * i] * invs, y[3 * i + 1] * invs, y[3 * i + 2] * invs, d00 * invd);
  out[1] = make_float4(d11 * invd, d22 * invd, d01 * invd, d02 * invd);
  out[2] = make_float4(d12 * invd, float(dprev[3 * i]) * invs, float(dprev[3 * i + 1]) * invs,
                       float(dprev[3 * i + 2]) * invs);
  out[3] = make_float4(n0 * invc, n1 * invc, n2 * invc, 0.0f);
}

template <class T>
void launch_ws_features(hipStream_t s, const MeshView& m, const Tiles& tv, const T* vals, const float* y,
                        const T* dprev, const double* scale, float* F) {
  if (tv.n == 0) return;
  static_assert(kVecThreads == 256, "four workgroups of 64 rows per vector tile");
  hipLaunchKernelGGL(k_ws_features<T>, dim3(unsigned(tv.n) * 4u), dim3(256), 0, s, m, tv, vals, y, dprev, scale, F);
}

template <class T>
__global__ void k_ws_collect(int nEnv, const int32_t* __restrict__ env_v0, const int32_t* __restrict__ mask,
                             int per_env, uint32_t salt, const float* __restrict__ F, const T* __restrict__ dx,
                             const double* __restrict__ scale, int64_t base, float* __restrict__ XF,
                             float* __restrict__ XT) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)nEnv * per_env) return;
  const int e = int(q / per_env), j = int(q % per_env);
  const int64_t slot = base + q;
  const int v0 = env_v0[e], ne = env_v0[e + 1] - v0;
  int v = -1;
  if (mask[e] && ne > 0) {
    const int stride = ne / per_env > 0 ? ne / per_env : 1;
    uint32_t h = salt * 2654435761u ^ (uint32_t(e) * 2246822519u);
    h ^= h >> 15;
    v = v0 + int((h % uint32_t(stride) + uint32_t(j) * uint32_t(stride)) % uint32_t(ne));
  }
  const float invs = v >= 0 ? float(1.0 / scale[e]) : 0.0f;
  for (int c = 0; c < kMlpInPad; ++c) XF[slot * kMlpInPad + c] = v >= 0 ? F[(size_t)v * kMlpInPad + c] : 0.0f;
  for (int c = 0; c < kMlpOut; ++c) XT[slot * kMlpOut + c] = v >= 0 ? float(dx[3 * v + c]) * invs : 0.0f;
}

template <class T>
void launch_ws_collect(hipStream_t s, int nEnv, const int32_t* env_v0, const int32_t* mask, int per_env,
                       uint32_t salt, const float* F, const T* dx, const double* scale, int64_t base, float* XF,
                       float* XT) {
  const int64_t n = (int64_t)nEnv * per_env;
  if (n == 0) return;
  hipLaunchKernelGGL(k_ws_collect<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, nEnv, env_v0, mask, per_env,
                     salt, F, dx, scale, base, XF, XT);
}

// Least |b - sum_i c_i A g_i|^2 = |b|^2 - 2 c.h + c^T G c over the guesses of
// `mask` (G the Gram matrix of the A g_i, h_i = b.A g_i): solves the normal
// equations of the subset; false when they are (nearly) singular.
__device__ bool ws_fit(const double G[3][3], const double h[3], int mask, double c[3]) {
  int idx[3], n = 0;
  for (int i = 0; i < 3; ++i) {
    c[i] = 0.0;
    if (mask >> i & 1) idx[n++] = i;
  }
  double M[3][3], y[3];
  for (int i = 0; i < n; ++i) {
    y[i] = h[idx[i]];
    for (int j = 0; j < n; ++j) M[i][j] = G[idx[i]][idx[j]];
  }
  double diag = 1.0;
  for (int i = 0; i < n; ++i) {
    if (!(M[i][i] > 0.0)) return false;
    diag *= M[i][i];
  }
  if (n == 1) {
    c[idx[0]] = y[0] / M[0][0];
  } else if (n == 2) {
    const double det = M[0][0] * M[1][1] - M[0][1] * M[1][0];
    if (!(det > 1e-10 * diag)) return false;
    c[idx[0]] = (y[0] * M[1][1] - y[1] * M[0][1]) / det;
    c[idx[1]] = (y[1] * M[0][0] - y[0] * M[1][0]) / det;
  } else {  // symmetric 3 x 3: c = adj(M) y / det(M)
    const double a00 = M[1][1] * M[2][2] - M[1][2] * M[2][1], a01 = M[0][2] * M[2][1] - M[0][1] * M[2][2],
                 a02 = M[0][1] * M[1][2] - M[0][2] * M[1][1], a11 = M[0][0] * M[2][2] - M[0][2] * M[2][0],
                 a12 = M[0][2] * M[1][0] - M[0][0] * M[1][2], a22 = M[0][0] * M[1][1] - M[0][1] * M[1][0];
    const double det = M[0][0] * a00 + M[0][1] * a01 + M[0][2] * a02;
    if (!(det > 1e-10 * diag)) return false;
    c[idx[0]] = (a00 * y[0] + a01 * y[1] + a02 * y[2]) / det;
    c[idx[1]] = (a01 * y[0] + a11 * y[1] + a12 * y[2]) / det;
    c[idx[2]] = (a02 * y[0] + a12 * y[1] + a22 * y[2]) / det;
  }
  return isfinite(c[0]) && isfinite(c[1]) && isfinite(c[2]);
}

__global__ void k_ws_decide_multi(int nEnv, const int32_t* __restrict__ solve, const double* __restrict__ acc,
                                  int32_t* __restrict__ warm, double* __restrict__ ratio,
                                  double* __restrict__ ratio_hist, double* __restrict__ alpha,
                                  double* __restrict__ beta, double* __restrict__ gamma) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEnv) return;
  alpha[e] = 1.0;
  beta[e] = 0.0;
  gamma[e] = 0.0;
  if (!solve[e]) {
    warm[e] = 0;
    ratio[e] = 0.0;
    ratio_hist[e] = 0.0;
    return;
  }
  double q[10];  // |b|^2, b.Au, b.Av, b.Aw, |Au|^2, |Av|^2, |Aw|^2, Au.Av, Au.Aw, Av.Aw
  acc_read<double, 10>(acc_ptr(acc, 0, nEnv, e), q);
  const double bb = q[0];
  const double h[3] = {q[1], q[2], q[3]};
  const double G[3][3] = {{q[4], q[7], q[8]}, {q[7], q[5], q[9]}, {q[8], q[9], q[6]}};
  // every nonempty subset of the three guesses; the least residual wins
  // (best_h: the same over the subsets without u, the previous steps alone)
  double best = bb, best_h = bb, bc[3] = {0.0, 0.0, 0.0};
  for (int mask = 1; mask < 8; ++mask) {
    double c[3];
    if (!ws_fit(G, h, mask, c)) continue;
    double r2 = bb;
    for (int i = 0; i < 3; ++i) {
      r2 -= 2.0 * c[i] * h[i];
      for (int j = 0; j < 3; ++j) r2 += c[i] * c[j] * G[i][j];
    }
    if (!(mask & 1) && r2 < best_h) best_h = r2;
    if (r2 < best) {
      best = r2;
      for (int i = 0; i < 3; ++i) bc[i] = c[i];
    }
  }
  if (best < 0.0) best = 0.0;
  if (best_h < 0.0) best_h = 0.0;
  const double r = bb > 0.0 ? sqrt(best / bb) : 1.0;
  const bool keep = bb > 0.0 && r < kWarmKeep;
  warm[e] = keep ? 1 : 0;
  ratio[e] = r;
  ratio_hist[e] = bb > 0.0 ? sqrt(best_h / bb) : 1.0;
  if (keep) {
    alpha[e] = bc[0];
    beta[e] = bc[1];
    gamma[e] = bc[2];
  }
}

void launch_ws_decide_multi(hipStream_t s, int nEnv, const int32_t* solve, const double* acc, int32_t* warm,
                            double* ratio, double* ratio_hist, double* alpha, double* beta, double* gamma) {
  if (nEnv == 0) return;
  hipLaunchKernelGGL(k_ws_decide_multi, dim3((nEnv + 63) / 64), dim3(64), 0, s, nEnv, solve, acc, warm, ratio,
                     ratio_hist, alpha, beta, gamma);
}

#define SDX_MLP_INST(T)                                                                                          \
  template void launch_mlp_infer<T>(hipStream_t, const MlpBf16&, const float*, int, const int32_t*,             \
                                    const double*, const uint8_t*, float*, T*);                                 \
  template void launch_ws_scaled_residual<T>(hipStream_t, const MeshView&, const Tiles&, const T*, const T*,    \
                                             const T*, float*, double*);                                        \
  template void launch_ws_features<T>(hipStream_t, const MeshView&, const Tiles&, const T*, const float*,       \
                                      const T*, const double*, float*);                                         \
  template void launch_ws_collect<T>(hipStream_t, int, const int32_t*, const int32_t*, int, uint32_t,          \
                                     const float*, const T*, const double*, int64_t, float*, float*);

SDX_MLP_INST(float)
SDX_MLP_INST(double)

}  // namespace sdx
