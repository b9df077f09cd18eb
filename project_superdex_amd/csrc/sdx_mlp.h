// Per-vertex warm-start network of the CG: 15 -> 128 -> 128 -> 128 -> 3,
// GELU (erf form), no counterpart in the reference (its PCG starts from zero,
// pcg.h:109-130).  Inference runs in BF16 on the matrix cores (FP32
// accumulation); training runs in FP32 on hand-written kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdx_kernels.h"

namespace sdx {

constexpr int kMlpIn = 15;     // features per vertex
constexpr int kMlpInPad = 16;  // feature row stride (last entry zero)
constexpr int kMlpHid = 128;
constexpr int kMlpOut = 3;
// FP32 parameter vector: W1[128][15] b1[128] W2[128][128] b2[128]
// W3[128][128] b3[128] W4[3][128] b4[3]  (W[out][in], y = W x + b)
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHid * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpHid;
constexpr int kOffW4 = kOffB3 + kMlpHid;
constexpr int kOffB4 = kOffW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kOffB4 + kMlpOut;  // 35,459

// BF16 weights for the matrix-core kernel: W1p[128][16], W2p, W3p [128][128],
// W4p[16][128] (outputs padded to 16), biases FP32 (b4 padded to 16).
struct MlpBf16 {
  const uint16_t* w1;
  const uint16_t* w2;
  const uint16_t* w3;
  const uint16_t* w4;
  const float* b1;
  const float* b2;
  const float* b3;
  const float* b4;
};
constexpr size_t kBf16W1 = size_t(kMlpHid) * kMlpInPad;
constexpr size_t kBf16W2 = size_t(kMlpHid) * kMlpHid;
constexpr size_t kBf16W4 = size_t(16) * kMlpHid;

// Packed BF16 weights from the FP32 parameters (round to nearest even).
void launch_mlp_pack(hipStream_t s, const float* params, uint16_t* w1, uint16_t* w2, uint16_t* w3, uint16_t* w4,
                     float* b1, float* b2, float* b3, float* b4);
// y[n][3] = net(x[n][16]) with BF16 matrix-core layers.  out_scale (may be
// NULL) multiplies row i by out_scale[row_env[i]]; pinned (may be NULL)
// zeroes pinned DoFs; the result goes to y (float) or yt (T) as given.
template <class T>
void launch_mlp_infer(hipStream_t s, const MlpBf16& w, const float* x, int n, const int32_t* row_env,
                      const double* out_scale, const uint8_t* pinned, float* y, T* yt);

// ---- FP32 training kernels ----
// C[M][N] = sum_k opA(m,k) opB(k,n) (+ C if accumulate); opA(m,k) = A[m*lda+k]
// or A[k*lda+m] (transA), opB(k,n) = B[k*ldb+n] or B[n*ldb+k] (transB).
void launch_gemm_f32(hipStream_t s, int M, int N, int K, const float* A, int lda, bool transA, const float* B,
                     int ldb, bool transB, float* C, int ldc, bool accumulate);
// Z[M][N] += b[N]; H = gelu(Z)  (H may be NULL: bias only)
void launch_bias_gelu(hipStream_t s, int M, int N, float* Z, const float* b, float* H);
// dZ = dH * gelu'(Z)
void launch_gelu_backward(hipStream_t s, int64_t n, const float* Z, const float* dH, float* dZ);
// db[N] = sum over rows of dZ[M][N]
void launch_colsum(hipStream_t s, int M, int N, const float* dZ, float* db);
// dY = 2 (Y - T) / M over [M][3]; loss (FP64 accumulator, zeroed by the
// caller) += sum (Y - T)^2 / M
void launch_mse_grad(hipStream_t s, int M, const float* Y, const float* T, float* dY, double* loss);
// Adam (Kingma & Ba 2015) with bias correction for step t (1-based).
void launch_adam(hipStream_t s, int n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                 float eps, int t);
// Gather rows: out[r] = src[(mult * (base + r) + off) % n] for r < rows.
void launch_gather_rows(hipStream_t s, int rows, int width, const float* src, int64_t n, int64_t mult, int64_t off,
                        int64_t base, float* out);

// ---- features of the warm start (world side) ----
// Per vertex: y = b / d with d the mean diagonal of the vertex's diagonal
// block; per-environment sums |y|^2, |dprev|^2 and vertex counts into set 0
// of acc (zeroed by the caller).
template <class T>
void launch_ws_scaled_residual(hipStream_t s, const MeshView& m, const Tiles& tv, const T* b, const T* vals,
                               const T* dprev, float* y, double* acc);
// Per-environment output scale from the sums: scale[e] = sqrt(max(|dprev|^2,
// |y|^2) / n_e) (1 where both vanish).
void launch_ws_env_scale(hipStream_t s, int nEnv, const double* acc, double* scale);
// Feature rows F[i][16]: y_i / s (3), the 6 unique entries of the diagonal
// block / d_i, dprev_i / s (3), mean of y_j / s over the 1-ring (3), 0.
template <class T>
void launch_ws_features(hipStream_t s, const MeshView& m, const Tiles& tv, const T* vals, const float* y,
                        const T* dprev, const double* scale, float* F);
// Training samples of environment e: `per_env` vertices at a stride through
// its rows starting at an offset derived from `salt`; feature row and target
// dx / scale[e] written at slot base + e * per_env + j (j < per_env).
template <class T>
void launch_ws_collect(hipStream_t s, int nEnv, const int32_t* env_v0, const int32_t* mask, int per_env,
                       uint32_t salt, const float* F, const T* dx, const double* scale, int64_t base, float* XF,
                       float* XT);
// A warm-start guess is kept only if it leaves less than kWarmKeep |b| of the
// residual.  Measured on batched indentation and plate presses, guesses that
// leave 40% of the residual cost CG as many iterations as they save (the
// per-vertex error they carry is rough), so only guesses that remove 80% of it
// are kept.
constexpr double kWarmKeep = 0.2;
// Per environment with solve[e], for three guesses u (the network's), v and w
// (the two previous Newton steps), from the sums of launch_residual64_multi in
// set 0 of acc: the combination alpha u + beta v + gamma w with the least
// residual over every nonempty subset of the three (normal equations; subsets
// whose equations are nearly singular are skipped); ratio[e] = that residual /
// |b| (0 when not tried), ratio_hist[e] the same over the subsets without u
// (what the previous steps alone leave), warm[e] = 1 when ratio[e] <
// kWarmKeep, else the environment starts from zero.
void launch_ws_decide_multi(hipStream_t s, int nEnv, const int32_t* solve, const double* acc, int32_t* warm,
                            double* ratio, double* ratio_hist, double* alpha, double* beta, double* gamma);

}  // namespace sdx
