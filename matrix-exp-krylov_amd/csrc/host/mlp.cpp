// Learned step selector host side: the MLP's parameters, Adam state and
// training loop on the device (kernels in learn.hip), and device feature
// extraction for a CSRHost.
#include "mek/mlp.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "../kernels/learn.hpp"
#include "hip_check.hpp"

namespace mek {

namespace {
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kNumFeatures * kHidden;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kHidden * kNumOutputs;
}  // namespace

int Mlp::param_count() { return kMlpParams; }

Mlp::Mlp(unsigned seed, int device) : device_(device) {
  MEK_HIP(hipSetDevice(device_));
  MEK_HIP(hipStreamCreate(&stream_));
  for (float** p : {&d_p_, &d_g_, &d_m_, &d_v_}) MEK_HIP(hipMalloc(p, sizeof(float) * kMlpParams));
  MEK_HIP(hipMalloc(&d_loss_, sizeof(float)));
  MEK_HIP(hipMemset(d_m_, 0, sizeof(float) * kMlpParams));
  MEK_HIP(hipMemset(d_v_, 0, sizeof(float) * kMlpParams));
  // Glorot-uniform weights, zero biases
  std::mt19937 rng(seed);
  std::vector<float> p(kMlpParams, 0.0f);
  auto fill = [&](int off, int fan_in, int fan_out) {
    const float a = std::sqrt(6.0f / (fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < fan_in * fan_out; ++i) p[off + i] = u(rng);
  };
  fill(kW1, kNumFeatures, kHidden);
  fill(kW2, kHidden, kHidden);
  fill(kW3, kHidden, kNumOutputs);
  set_params(p.data());
}

Mlp::~Mlp() {
  (void)hipSetDevice(device_);
  for (float* p : {d_p_, d_g_, d_m_, d_v_, d_loss_, d_x_, d_t_, d_z1_, d_a1_, d_z2_, d_a2_, d_z3_, d_y_,
                   d_dy_, d_d2_, d_d1_, d_data_x_, d_data_t_})
    if (p) (void)hipFree(p);
  if (d_idx_) (void)hipFree(d_idx_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Mlp::set_params(const float* p) {
  MEK_HIP(hipMemcpy(d_p_, p, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  step_ = 0;
  MEK_HIP(hipMemset(d_m_, 0, sizeof(float) * kMlpParams));
  MEK_HIP(hipMemset(d_v_, 0, sizeof(float) * kMlpParams));
}

void Mlp::get_params(float* p) const {
  MEK_HIP(hipMemcpy(p, d_p_, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
}

void Mlp::ensure_batch(int B) {
  if (B <= cap_) return;
  for (float* p : {d_x_, d_t_, d_z1_, d_a1_, d_z2_, d_a2_, d_z3_, d_y_, d_dy_, d_d2_, d_d1_})
    if (p) (void)hipFree(p);
  if (d_idx_) (void)hipFree(d_idx_);
  cap_ = B;
  MEK_HIP(hipMalloc(&d_x_, sizeof(float) * B * kNumFeatures));
  MEK_HIP(hipMalloc(&d_t_, sizeof(float) * B * kNumOutputs));
  for (float** p : {&d_z1_, &d_a1_, &d_z2_, &d_a2_, &d_d2_, &d_d1_}) MEK_HIP(hipMalloc(p, sizeof(float) * B * kHidden));
  for (float** p : {&d_z3_, &d_y_, &d_dy_}) MEK_HIP(hipMalloc(p, sizeof(float) * B * kNumOutputs));
  MEK_HIP(hipMalloc(&d_idx_, sizeof(int) * B));
}

void Mlp::fwd_device(const float* x, int B) {
  mlp_dense_fwd(x, d_p_ + kW1, d_p_ + kB1, d_z1_, d_a1_, B, kNumFeatures, kHidden, 1, stream_);
  mlp_dense_fwd(d_a1_, d_p_ + kW2, d_p_ + kB2, d_z2_, d_a2_, B, kHidden, kHidden, 1, stream_);
  mlp_dense_fwd(d_a2_, d_p_ + kW3, d_p_ + kB3, d_z3_, d_y_, B, kHidden, kNumOutputs, 0, stream_);
}

void Mlp::bwd_device(const float* x, const float* t, int B) {
  mlp_mse(d_y_, t, d_dy_, d_loss_, B * kNumOutputs, stream_);
  mlp_dense_grad(d_a2_, d_dy_, d_g_ + kW3, d_g_ + kB3, B, kHidden, kNumOutputs, stream_);
  mlp_dense_bwd_in(d_dy_, d_p_ + kW3, d_z2_, d_d2_, B, kHidden, kNumOutputs, 1, stream_);
  mlp_dense_grad(d_a1_, d_d2_, d_g_ + kW2, d_g_ + kB2, B, kHidden, kHidden, stream_);
  mlp_dense_bwd_in(d_d2_, d_p_ + kW2, d_z1_, d_d1_, B, kHidden, kHidden, 1, stream_);
  mlp_dense_grad(x, d_d1_, d_g_ + kW1, d_g_ + kB1, B, kNumFeatures, kHidden, stream_);
}

void Mlp::forward(const float* X, int B, float* Y) {
  MEK_HIP(hipSetDevice(device_));
  ensure_batch(B);
  MEK_HIP(hipMemcpyAsync(d_x_, X, sizeof(float) * B * kNumFeatures, hipMemcpyHostToDevice, stream_));
  fwd_device(d_x_, B);
  MEK_HIP_LAUNCHED();
  MEK_HIP(hipMemcpyAsync(Y, d_y_, sizeof(float) * B * kNumOutputs, hipMemcpyDeviceToHost, stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
}

float Mlp::loss_grad(const float* X, const float* T, int B, float* grads) {
  MEK_HIP(hipSetDevice(device_));
  ensure_batch(B);
  MEK_HIP(hipMemcpyAsync(d_x_, X, sizeof(float) * B * kNumFeatures, hipMemcpyHostToDevice, stream_));
  MEK_HIP(hipMemcpyAsync(d_t_, T, sizeof(float) * B * kNumOutputs, hipMemcpyHostToDevice, stream_));
  fwd_device(d_x_, B);
  bwd_device(d_x_, d_t_, B);
  MEK_HIP_LAUNCHED();
  float loss = 0.0f;
  MEK_HIP(hipMemcpyAsync(&loss, d_loss_, sizeof(float), hipMemcpyDeviceToHost, stream_));
  if (grads) MEK_HIP(hipMemcpyAsync(grads, d_g_, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
  return loss;
}

float Mlp::train(const float* X, const float* T, int N, int batch, int epochs, float lr, unsigned seed) {
  MEK_HIP(hipSetDevice(device_));
  if (N < 1 || batch < 1) throw Error("ERROR: training needs N >= 1 and batch >= 1");
  batch = std::min(batch, N);
  ensure_batch(std::max(batch, N));
  if (d_data_x_) (void)hipFree(d_data_x_);
  if (d_data_t_) (void)hipFree(d_data_t_);
  d_data_x_ = d_data_t_ = nullptr;
  MEK_HIP(hipMalloc(&d_data_x_, sizeof(float) * N * kNumFeatures));
  MEK_HIP(hipMalloc(&d_data_t_, sizeof(float) * N * kNumOutputs));
  MEK_HIP(hipMemcpy(d_data_x_, X, sizeof(float) * N * kNumFeatures, hipMemcpyHostToDevice));
  MEK_HIP(hipMemcpy(d_data_t_, T, sizeof(float) * N * kNumOutputs, hipMemcpyHostToDevice));
  int* d_perm = nullptr;
  MEK_HIP(hipMalloc(&d_perm, sizeof(int) * N));
  std::vector<int> perm(N);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937 rng(seed);
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(perm.begin(), perm.end(), rng);
    MEK_HIP(hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, stream_));
    for (int s0 = 0; s0 + batch <= N; s0 += batch) {
      mlp_gather(d_data_x_, d_data_t_, d_perm + s0, d_x_, d_t_, batch, kNumFeatures, kNumOutputs, stream_);
      fwd_device(d_x_, batch);
      bwd_device(d_x_, d_t_, batch);
      mlp_adam(d_p_, d_g_, d_m_, d_v_, kMlpParams, lr, 0.9f, 0.999f, 1e-8f, ++step_, stream_);
    }
    MEK_HIP(hipStreamSynchronize(stream_));  // keeps perm alive for the async copy
  }
  MEK_HIP_LAUNCHED();
  // loss over the whole training set
  std::iota(perm.begin(), perm.end(), 0);
  MEK_HIP(hipMemcpy(d_perm, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice));
  mlp_gather(d_data_x_, d_data_t_, d_perm, d_x_, d_t_, N, kNumFeatures, kNumOutputs, stream_);
  fwd_device(d_x_, N);
  mlp_mse(d_y_, d_t_, d_dy_, d_loss_, N * kNumOutputs, stream_);
  float loss = 0.0f;
  MEK_HIP(hipMemcpyAsync(&loss, d_loss_, sizeof(float), hipMemcpyDeviceToHost, stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
  (void)hipFree(d_perm);
  return loss;
}

void features_csr(const CSRHost& A, double t, double tol, float* out, int device) {
  MEK_HIP(hipSetDevice(device));
  const int n = A.rows;
  // slabs of kFeatSlabRows rows from row 0 (a batch cuts each problem the same way)
  std::vector<int> lay;  // slab_row0 (nslab + 1) | slab_prob (nslab) | p_slab0 (2) | p_row0 | p_row1
  for (int r = 0; r < std::max(n, 1); r += kFeatSlabRows) lay.push_back(r);
  const int nslab = static_cast<int>(lay.size());
  lay.push_back(n);
  lay.insert(lay.end(), nslab, 0);
  lay.insert(lay.end(), {0, nslab, 0, n});
  const double tt[3] = {t, tol, A.norm_1()};
  int *d_rp = nullptr, *d_ci = nullptr, *d_lay = nullptr;
  double *d_v = nullptr, *d_stats = nullptr, *d_tt = nullptr;
  float* d_f = nullptr;
  auto cleanup = [&] {
    for (void* p : {static_cast<void*>(d_rp), static_cast<void*>(d_ci), static_cast<void*>(d_lay),
                    static_cast<void*>(d_v), static_cast<void*>(d_stats),
                    static_cast<void*>(d_tt), static_cast<void*>(d_f)})
      if (p) (void)hipFree(p);
  };
  try {
    MEK_HIP(hipMalloc(&d_rp, sizeof(int) * (n + 1)));
    MEK_HIP(hipMalloc(&d_ci, sizeof(int) * std::max<long long>(A.nnz, 1)));
    MEK_HIP(hipMalloc(&d_v, sizeof(double) * std::max<long long>(A.nnz, 1)));
    MEK_HIP(hipMalloc(&d_lay, sizeof(int) * lay.size()));
    MEK_HIP(hipMalloc(&d_stats, sizeof(double) * 8 * std::max(nslab, 1)));
    MEK_HIP(hipMalloc(&d_tt, sizeof(double) * 3));
    MEK_HIP(hipMalloc(&d_f, sizeof(float) * kNumFeatures));
    MEK_HIP(hipMemcpy(d_rp, A.row_ptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (A.nnz) {
      MEK_HIP(hipMemcpy(d_ci, A.col_idx.data(), sizeof(int) * A.nnz, hipMemcpyHostToDevice));
      MEK_HIP(hipMemcpy(d_v, A.values.data(), sizeof(double) * A.nnz, hipMemcpyHostToDevice));
    }
    MEK_HIP(hipMemcpy(d_lay, lay.data(), sizeof(int) * lay.size(), hipMemcpyHostToDevice));
    MEK_HIP(hipMemcpy(d_tt, tt, sizeof(tt), hipMemcpyHostToDevice));
    const int* sl = d_lay + nslab + 1 + nslab;  // p_slab0, then p_row0, p_row1
    FeatureLayout L{d_lay, d_lay + nslab + 1, sl, sl + 2, sl + 3, nslab, 1};
    feature_launch(d_rp, d_ci, d_v, 0, L, d_tt + 2, d_stats, d_tt, d_tt + 1, d_f, nullptr);
    MEK_HIP_LAUNCHED();
    MEK_HIP(hipMemcpy(out, d_f, sizeof(float) * kNumFeatures, hipMemcpyDeviceToHost));
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace mek
