// Learned step selector (additive; no reference counterpart): a 16 -> 64 ->
// 64 -> 2 MLP with FP32 weights and GELU, trained and evaluated on the GPU by
// this library's own kernels, plus device feature extraction from a CSR
// matrix (Gershgorin and 1-/inf-norm estimates, nnz-per-row statistics, t).
#pragma once

#include <hip/hip_runtime.h>

#include "mek/csr_host.hpp"

namespace mek {

class Mlp {
 public:
  explicit Mlp(unsigned seed, int device = 0);
  ~Mlp();
  Mlp(const Mlp&) = delete;
  Mlp& operator=(const Mlp&) = delete;
  static int param_count();
  void set_params(const float* p);  // also resets the Adam state
  void get_params(float* p) const;
  void forward(const float* X, int B, float* Y);                      // X: B x 16, Y: B x 2
  float loss_grad(const float* X, const float* T, int B, float* grads);  // MSE and its gradient
  // Mini-batch Adam over N samples; returns the MSE over all N afterwards.
  float train(const float* X, const float* T, int N, int batch, int epochs, float lr, unsigned seed);

 private:
  void ensure_batch(int B);
  void fwd_device(const float* x, int B);
  void bwd_device(const float* x, const float* t, int B);
  int device_ = 0;
  int cap_ = 0;
  int step_ = 0;
  hipStream_t stream_ = nullptr;
  float *d_p_ = nullptr, *d_g_ = nullptr, *d_m_ = nullptr, *d_v_ = nullptr, *d_loss_ = nullptr;
  float *d_x_ = nullptr, *d_t_ = nullptr, *d_z1_ = nullptr, *d_a1_ = nullptr, *d_z2_ = nullptr;
  float *d_a2_ = nullptr, *d_z3_ = nullptr, *d_y_ = nullptr, *d_dy_ = nullptr, *d_d2_ = nullptr;
  float *d_d1_ = nullptr, *d_data_x_ = nullptr, *d_data_t_ = nullptr;
  int* d_idx_ = nullptr;
};

// The 16 selector features of (A, t, tol), computed on the device.
void features_csr(const CSRHost& A, double t, double tol, float* out16, int device = 0);

}  // namespace mek
