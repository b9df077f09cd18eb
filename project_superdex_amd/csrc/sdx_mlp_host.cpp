// Host side of the learned CG warm start (see sdx_mlp_host.h).
#include "sdx_mlp_host.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <stdexcept>

#include "sdx_world.h"

namespace sdx {

namespace {

template <class P>
void dmalloc(P** p, size_t bytes) {
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e == hipErrorOutOfMemory) throw std::bad_alloc();
  hip_check(e, "hipMalloc");
}

template <class P>
void dfree(P*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace

Mlp::Mlp(int device, uint64_t seed) : device_(device) {
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
  const size_t pb = sizeof(float) * kMlpParams;
  dmalloc(&params_, pb);
  dmalloc(&grad_, pb);
  dmalloc(&m_, pb);
  dmalloc(&v_, pb);
  dmalloc(&wpack_, sizeof(uint16_t) * (kBf16W1 + 2 * kBf16W2 + kBf16W4));
  dmalloc(&bpack_, sizeof(float) * (3 * kMlpHid + 16));
  dmalloc(&loss_, sizeof(double));
  bf16_.w1 = wpack_;
  bf16_.w2 = wpack_ + kBf16W1;
  bf16_.w3 = wpack_ + kBf16W1 + kBf16W2;
  bf16_.w4 = wpack_ + kBf16W1 + 2 * kBf16W2;
  bf16_.b1 = bpack_;
  bf16_.b2 = bpack_ + kMlpHid;
  bf16_.b3 = bpack_ + 2 * kMlpHid;
  bf16_.b4 = bpack_ + 3 * kMlpHid;
  // He-style initialisation: W ~ N(0, 2 / fan_in) for the GELU layers,
  // N(0, 1 / fan_in) for the output layer, biases zero.
  std::vector<float> p(kMlpParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int rows, int cols, double var) {
    std::normal_distribution<double> nd(0.0, std::sqrt(var / cols));
    for (int i = 0; i < rows * cols; ++i) p[size_t(off + i)] = float(nd(rng));
  };
  fill(kOffW1, kMlpHid, kMlpIn, 2.0);
  fill(kOffW2, kMlpHid, kMlpHid, 2.0);
  fill(kOffW3, kMlpHid, kMlpHid, 2.0);
  fill(kOffW4, kMlpOut, kMlpHid, 1.0);
  set_params(p.data());
}

Mlp::~Mlp() {
  for (float** q : {&params_, &grad_, &m_, &v_, &bpack_, &X_, &T_, &Z1_, &H1_, &Z2_, &H2_, &Z3_, &H3_, &Y_, &dY_,
                    &dH_, &dZ_})
    dfree(*q);
  dfree(wpack_);
  dfree(loss_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Mlp::repack() {
  launch_mlp_pack(stream_, params_, wpack_, wpack_ + kBf16W1, wpack_ + kBf16W1 + kBf16W2,
                  wpack_ + kBf16W1 + 2 * kBf16W2, bpack_, bpack_ + kMlpHid, bpack_ + 2 * kMlpHid,
                  bpack_ + 3 * kMlpHid);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Mlp::set_params(const float* host) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipMemcpyAsync(params_, host, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, stream_), "hipMemcpy");
  hip_check(hipMemsetAsync(m_, 0, sizeof(float) * kMlpParams, stream_), "hipMemsetAsync");
  hip_check(hipMemsetAsync(v_, 0, sizeof(float) * kMlpParams, stream_), "hipMemsetAsync");
  adam_t_ = 0;
  repack();
}

void Mlp::get_params(float* host) {
  hip_check(hipMemcpyAsync(host, params_, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, stream_), "hipMemcpy");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Mlp::ensure(int rows) {
  if (rows <= cap_) return;
  for (float** q : {&X_, &T_, &Z1_, &H1_, &Z2_, &H2_, &Z3_, &H3_, &Y_, &dY_, &dH_, &dZ_}) dfree(*q);
  cap_ = rows;
  const size_t r = size_t(rows);
  dmalloc(&X_, sizeof(float) * r * kMlpInPad);
  dmalloc(&T_, sizeof(float) * r * kMlpOut);
  for (float** q : {&Z1_, &H1_, &Z2_, &H2_, &Z3_, &H3_, &dH_, &dZ_}) dmalloc(q, sizeof(float) * r * kMlpHid);
  dmalloc(&Y_, sizeof(float) * r * kMlpOut);
  dmalloc(&dY_, sizeof(float) * r * kMlpOut);
}

void Mlp::forward_fp32(const float* X, int B) {
  const float* p = params_;
  launch_gemm_f32(stream_, B, kMlpHid, kMlpIn, X, kMlpInPad, false, p + kOffW1, kMlpIn, true, Z1_, kMlpHid, false);
  launch_bias_gelu(stream_, B, kMlpHid, Z1_, p + kOffB1, H1_);
  launch_gemm_f32(stream_, B, kMlpHid, kMlpHid, H1_, kMlpHid, false, p + kOffW2, kMlpHid, true, Z2_, kMlpHid, false);
  launch_bias_gelu(stream_, B, kMlpHid, Z2_, p + kOffB2, H2_);
  launch_gemm_f32(stream_, B, kMlpHid, kMlpHid, H2_, kMlpHid, false, p + kOffW3, kMlpHid, true, Z3_, kMlpHid, false);
  launch_bias_gelu(stream_, B, kMlpHid, Z3_, p + kOffB3, H3_);
  launch_gemm_f32(stream_, B, kMlpOut, kMlpHid, H3_, kMlpHid, false, p + kOffW4, kMlpHid, true, Y_, kMlpOut, false);
  launch_bias_gelu(stream_, B, kMlpOut, Y_, p + kOffB4, nullptr);
}

// Gradients of the loss whose dL/dY is in dY_, into grad_ (parameter layout).
void Mlp::backward(const float* X, int B) {
  const float* p = params_;
  float* g = grad_;
  // output layer: dW4 = dY^T H3, db4, dH3 = dY W4
  launch_gemm_f32(stream_, kMlpOut, kMlpHid, B, dY_, kMlpOut, true, H3_, kMlpHid, false, g + kOffW4, kMlpHid, false);
  launch_colsum(stream_, B, kMlpOut, dY_, g + kOffB4);
  launch_gemm_f32(stream_, B, kMlpHid, kMlpOut, dY_, kMlpOut, false, p + kOffW4, kMlpHid, false, dH_, kMlpHid, false);
  // layer 3
  launch_gelu_backward(stream_, int64_t(B) * kMlpHid, Z3_, dH_, dZ_);
  launch_gemm_f32(stream_, kMlpHid, kMlpHid, B, dZ_, kMlpHid, true, H2_, kMlpHid, false, g + kOffW3, kMlpHid, false);
  launch_colsum(stream_, B, kMlpHid, dZ_, g + kOffB3);
  launch_gemm_f32(stream_, B, kMlpHid, kMlpHid, dZ_, kMlpHid, false, p + kOffW3, kMlpHid, false, dH_, kMlpHid, false);
  // layer 2
  launch_gelu_backward(stream_, int64_t(B) * kMlpHid, Z2_, dH_, dZ_);
  launch_gemm_f32(stream_, kMlpHid, kMlpHid, B, dZ_, kMlpHid, true, H1_, kMlpHid, false, g + kOffW2, kMlpHid, false);
  launch_colsum(stream_, B, kMlpHid, dZ_, g + kOffB2);
  launch_gemm_f32(stream_, B, kMlpHid, kMlpHid, dZ_, kMlpHid, false, p + kOffW2, kMlpHid, false, dH_, kMlpHid, false);
  // layer 1
  launch_gelu_backward(stream_, int64_t(B) * kMlpHid, Z1_, dH_, dZ_);
  launch_gemm_f32(stream_, kMlpHid, kMlpIn, B, dZ_, kMlpHid, true, X, kMlpInPad, false, g + kOffW1, kMlpIn, false);
  launch_colsum(stream_, B, kMlpHid, dZ_, g + kOffB1);
}

void Mlp::forward(const float* hx, int64_t n, bool bf16, float* hy) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  const int chunk = 65536;
  ensure(chunk);
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    const int B = int(std::min<int64_t>(chunk, n - r0));
    hip_check(hipMemcpyAsync(X_, hx + r0 * kMlpInPad, sizeof(float) * size_t(B) * kMlpInPad, hipMemcpyHostToDevice,
                             stream_),
              "hipMemcpyAsync");
    if (bf16)
      launch_mlp_infer<float>(stream_, bf16_, X_, B, nullptr, nullptr, nullptr, Y_, nullptr);
    else
      forward_fp32(X_, B);
    hip_check(hipMemcpyAsync(hy + r0 * kMlpOut, Y_, sizeof(float) * size_t(B) * kMlpOut, hipMemcpyDeviceToHost,
                             stream_),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }
}

double Mlp::loss_grad(const float* hx, const float* ht, int n, float* hgrad) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  ensure(n);
  hip_check(hipMemcpyAsync(X_, hx, sizeof(float) * size_t(n) * kMlpInPad, hipMemcpyHostToDevice, stream_), "memcpy");
  hip_check(hipMemcpyAsync(T_, ht, sizeof(float) * size_t(n) * kMlpOut, hipMemcpyHostToDevice, stream_), "memcpy");
  hip_check(hipMemsetAsync(loss_, 0, sizeof(double), stream_), "hipMemsetAsync");
  forward_fp32(X_, n);
  launch_mse_grad(stream_, n, Y_, T_, dY_, loss_);
  backward(X_, n);
  double loss = 0.0;
  hip_check(hipMemcpyAsync(&loss, loss_, sizeof(double), hipMemcpyDeviceToHost, stream_), "memcpy");
  if (hgrad)
    hip_check(hipMemcpyAsync(hgrad, grad_, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, stream_), "memcpy");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  return loss;
}

std::vector<double> Mlp::train(const float* dX, const float* dT, int64_t n, const TrainConfig& c) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  if (n <= 0 || c.batch <= 0 || c.epochs <= 0) throw std::invalid_argument("empty training set or schedule");
  if (n >= (int64_t(1) << 31)) throw std::invalid_argument("at most 2^31 - 1 training rows");
  const int B = int(std::min<int64_t>(c.batch, n));
  ensure(B);
  std::mt19937_64 rng(c.seed);
  std::vector<double> out;
  const int64_t batches = n / B;
  for (int ep = 0; ep < c.epochs; ++ep) {
    // a fresh affine permutation of the rows per epoch: r -> (mult r + off) mod n
    uint64_t mult = 1 + rng() % uint64_t(std::min<int64_t>(n, (int64_t(1) << 31) - 1));
    while (gcd64(mult, uint64_t(n)) != 1) mult = 1 + (mult % uint64_t(n));
    const uint64_t off = rng() % uint64_t(n);
    hip_check(hipMemsetAsync(loss_, 0, sizeof(double), stream_), "hipMemsetAsync");
    for (int64_t b = 0; b < std::max<int64_t>(batches, 1); ++b) {
      launch_gather_rows(stream_, B, kMlpInPad, dX, n, int64_t(mult), int64_t(off), b * B, X_);
      launch_gather_rows(stream_, B, kMlpOut, dT, n, int64_t(mult), int64_t(off), b * B, T_);
      forward_fp32(X_, B);
      launch_mse_grad(stream_, B, Y_, T_, dY_, loss_);
      backward(X_, B);
      launch_adam(stream_, kMlpParams, params_, grad_, m_, v_, c.lr, c.beta1, c.beta2, c.eps, ++adam_t_);
    }
    double loss = 0.0;
    hip_check(hipMemcpyAsync(&loss, loss_, sizeof(double), hipMemcpyDeviceToHost, stream_), "memcpy");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    out.push_back(loss / double(std::max<int64_t>(batches, 1)));
  }
  repack();
  return out;
}

}  // namespace sdx
