// mlp.h -- internal: the surrogate multilayer perceptron (BF16 MFMA training).
//
// A four-layer perceptron x (p_in) -> 256 -> 256 -> 256 -> [mu (p_out),
// log-variance (p_out)] with SiLU activations, trained by Adam on the
// heteroscedastic Gaussian negative log-likelihood
//   L = mean_b sum_a 0.5 [ (y_ba - mu_ba)^2 exp(-lv_ba) + lv_ba ].
// Matrix products run on the gfx950 BF16 matrix cores (v_mfma_f32_16x16x32_bf16)
// with float32 accumulation; weights and activations enter the products as
// bfloat16 (round to nearest even), master weights and optimizer state are
// float32.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "plan.h"

namespace cusna_impl {

constexpr int kMlpLayers = 4;

class Mlp {
 public:
  // p_in, p_out <= 64; width a multiple of 64 up to 256; batch a multiple of 64.
  cusna_status init(int p_in, int p_out, int width, int batch, uint64_t seed, hipStream_t s);
  // Fresh Glorot-uniform weights, zero biases and optimizer state.
  cusna_status reset(uint64_t seed);

  int p_in() const { return p_in_; }
  int p_out() const { return p_out_; }
  int batch() const { return B_; }
  size_t param_count() const { return n_params_; }
  // flat layout: for each layer l, W_l [K_l][N_l] row-major then b_l [N_l];
  // K_0 = p_in padded to 64, N_3 = 2 p_out padded to 64 (padding stays zero)
  int K(int l) const { return K_[l]; }
  int N(int l) const { return N_[l]; }
  size_t w_offset(int l) const { return off_w_[l]; }
  size_t b_offset(int l) const { return off_b_[l]; }

  cusna_status set_params(const float* host_flat);
  cusna_status get_params(float* host_flat);

  // One forward + backward pass over the batch x (B x p_in), y (B x p_out),
  // both float32 on the device; gradients land in the flat gradient buffer,
  // outputs (B x N_3, float32) and per-sample losses stay on the device.
  cusna_status forward_backward(const float* x_dev, const float* y_dev);
  // Forward only (outputs).
  cusna_status forward(const float* x_dev);
  // Train on n samples (device, float32, x n x p_in, y n x p_out): `steps`
  // Adam steps on minibatches drawn with replacement by a counter hash; the
  // learning rate decays linearly from lr to lr / 10.
  cusna_status train(const float* x_all, const float* y_all, int64_t n, int steps, float lr,
                     uint64_t seed);
  // Mean loss of the last forward_backward (device -> host, synchronises).
  cusna_status last_loss(double* loss);

  const float* outputs() const { return O_; }
  const float* grads() const { return grads_; }
  hipStream_t stream() const { return s_; }

 private:
  // One launch of the slice kernel over the batch: mode 0 forward, 1 forward
  // and backward of x, y (B rows), 2 the same on rows drawn by hash from the
  // n samples x, y (step `step` of the stream `seed`).
  cusna_status slice(int mode, const float* x, const float* y, int64_t n, uint64_t seed,
                     long long step);
  // Weight and bias gradients from the slice kernel's transposed activations
  // and gradients; with adam, one Adam step at learning rate lr.
  cusna_status wgrad(bool adam, float lr);
  cusna_status cast_weights();

  int p_in_ = 0, p_out_ = 0, width_ = 0, B_ = 0;
  int K_[kMlpLayers] = {}, N_[kMlpLayers] = {};
  size_t off_w_[kMlpLayers] = {}, off_b_[kMlpLayers] = {};
  size_t n_params_ = 0;
  long long t_ = 0;  // Adam step count
  hipStream_t s_ = nullptr;
  DeviceBlock D_;
  float *params_ = nullptr, *m_ = nullptr, *v_ = nullptr, *grads_ = nullptr;
  __bf16* Wb_[kMlpLayers] = {};   // W_l [K][N] (backward operands)
  __bf16* Wtb_[kMlpLayers] = {};  // W_l^T [N][K] (forward operands)
  __bf16* XT_ = nullptr;          // inputs, transposed (K_0 x B)
  __bf16* HT_[kMlpLayers - 1] = {};  // hidden activations, transposed (N_l x B)
  __bf16* GT_[kMlpLayers] = {};      // pre-activation (output) gradients, transposed
  float* O_ = nullptr;     // outputs (B x N_3)
  float* loss_ = nullptr;  // per sample
};

}  // namespace cusna_impl
