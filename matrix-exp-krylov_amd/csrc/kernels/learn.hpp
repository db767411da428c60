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
// Where the feature kernels find the problems of a block-diagonal CSR:
// problem p has rows [p_row0[p], p_row1[p]) and is cut into row slabs
// [p_slab0[p], p_slab0[p+1]); slab s starts at row slab_row0[s] and ends at
// slab_row0[s+1] or p_row1[slab_prob[s]], whichever comes first.  Slabs are
// kFeatSlabRows rows from the problem's first row, so a problem gets the same
// slabs alone or in a batch (the batch's streaming blocks are these slabs).
constexpr int kFeatSlabRows = 4096;
struct FeatureLayout {
  const int* slab_row0;  // nslab + 1
  const int* slab_prob;  // nslab
  const int* p_slab0;    // P + 1
  const int* p_row0;     // P
  const int* p_row1;     // P
  int nslab;
  int P;
};
// 16 features per problem from the rows, ||A_p||_1 (norm1, computed by the
// caller), t and tol; stats holds 8 doubles per slab.
void feature_launch(const int* row_ptr, const int* col_idx, const void* vals, int vals_f32,
                    const FeatureLayout& L, const double* norm1, double* stats, const double* t,
                    const double* tol, float* feat, hipStream_t s);

}  // namespace mek
