// Host side of the learned CG warm start: parameters, BF16 packing, FP32
// training (forward, backward, Adam) on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "sdx_mlp.h"

namespace sdx {

struct TrainConfig {
  int epochs = 10;
  int batch = 4096;
  float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  uint64_t seed = 1;
};

class Mlp {
 public:
  Mlp(int device, uint64_t seed);
  ~Mlp();
  Mlp(const Mlp&) = delete;
  Mlp& operator=(const Mlp&) = delete;

  int device() const { return device_; }
  void set_params(const float* host);
  void get_params(float* host);
  // BF16 weights for inference (packed after every parameter change).
  const MlpBf16& weights() const { return bf16_; }
  // y[n][3] for host rows x[n][16]: matrix-core BF16 path or FP32 path.
  void forward(const float* hx, int64_t n, bool bf16, float* hy);
  // Loss (mean over rows of |y - t|^2) and its gradient for host rows.
  double loss_grad(const float* hx, const float* ht, int n, float* hgrad);
  // Adam training on device rows X[n][16], T[n][3]; mean loss per epoch.
  std::vector<double> train(const float* dX, const float* dT, int64_t n, const TrainConfig& cfg);

 private:
  void ensure(int rows);
  void forward_fp32(const float* X, int rows);  // Z1..Z3, H1..H3, Y
  void backward(const float* X, int rows);      // grad_ from dY_
  void repack();

  int device_;
  hipStream_t stream_ = nullptr;
  float *params_ = nullptr, *grad_ = nullptr, *m_ = nullptr, *v_ = nullptr;
  uint16_t* wpack_ = nullptr;
  float* bpack_ = nullptr;
  MlpBf16 bf16_{};
  int cap_ = 0;
  float *X_ = nullptr, *T_ = nullptr, *Z1_ = nullptr, *H1_ = nullptr, *Z2_ = nullptr, *H2_ = nullptr, *Z3_ = nullptr,
        *H3_ = nullptr, *Y_ = nullptr, *dY_ = nullptr, *dH_ = nullptr, *dZ_ = nullptr;
  double* loss_ = nullptr;
  int adam_t_ = 0;
};

}  // namespace sdx
