// Learned step selector: MLP and feature kernels (learn.hip).
#pragma once

#include "common.hpp"

namespace mek {

constexpr int kNumFeatures = 16;
constexpr int kHidden = 64;
constexpr int kNumOutputs = 2;  // (m target, log2 substeps target)
// parameter counts of W1 | b1 | W2 | b2 | W3 | b3
constexpr int kMlpParams = kNumFeatures * kHidden + kHidden + kHidden * kHidden + kHidden +
                           kHidden * kNumOutputs + kNumOutputs;

void mlp_dense_fwd(const float* in, const float* W, const float* b, float* Z, float* A, int B, int din,
                   int dout, int act, hipStream_t s);
void mlp_mse(const float* Y, const float* T, float* dY, float* loss, int n, hipStream_t s);
void mlp_dense_bwd_in(const float* dOut, const float* W, const float* Zin, float* dIn, int B, int din,
                      int dout, int act_in, hipStream_t s);
void mlp_dense_grad(const float* In, const float* dOut, float* dW, float* db, int B, int din, int dout,
                    hipStream_t s);
void mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
              int step, hipStream_t s);
void mlp_gather(const float* X, const float* T, const int* idx, float* Xb, float* Tb, int B, int dx, int dt,
                hipStream_t s);
// 16 features per problem for rows [p_row0[p], p_row1[p]) of a
// block-diagonal CSR; colsum (n_total doubles) must be zero on entry.
void feature_launch(const int* row_ptr, const int* col_idx, const void* vals, int vals_f32,
                    const int* p_row0, const int* p_row1, int P, double* colsum, double* stats,
                    const double* t, const double* tol, float* feat, hipStream_t s);

}  // namespace mek
