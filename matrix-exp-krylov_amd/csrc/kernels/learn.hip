// Learned step selector: device feature extraction from CSR matrices and a
// small MLP (16 -> 64 -> 64 -> 2, FP32 weights, GELU) with its own forward,
// backward and Adam kernels.  No reference counterpart (the reference picks
// m and the first sub-step t/max_restarts by hand, main.cu:1292-1305).
//
// Layout: activations row-major [batch][features]; weight W of a dense layer
// row-major [in][out]; all parameters in one FP32 buffer
// W1 | b1 | W2 | b2 | W3 | b3.  Every reduction runs in a fixed order, so
// gradients are deterministic.
#include "learn.hpp"

namespace mek {

namespace {

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

}  // namespace

// out[b][j] = act(bias[j] + sum_i in[b][i] W[i][j]); Z keeps the pre-activation
__global__ void k_dense_fwd(const float* in, const float* W, const float* bias, float* Z, float* A,
                            int B, int din, int dout, int act) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * dout) return;
  const int b = idx / dout, j = idx % dout;
  float z = bias[j];
  for (int i = 0; i < din; ++i) z += in[b * din + i] * W[i * dout + j];
  Z[idx] = z;
  A[idx] = act ? gelu(z) : z;
}

// Mean-squared error over B x dout outputs: dY = 2 (Y - T) / (B dout), and
// the loss (one workgroup, fixed-order reduction).
__global__ void k_mse(const float* Y, const float* T, float* dY, float* loss, int n) {
  __shared__ float red[1024 / 64];
  float acc = 0.0f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d / n;
    acc += d * d;
  }
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) s += red[w];
    *loss = s / n;
  }
}

// dIn[b][i] = (sum_j dOut[b][j] W[i][j]) * gelu'(Zin[b][i])   (act_in = 1)
__global__ void k_dense_bwd_in(const float* dOut, const float* W, const float* Zin, float* dIn, int B,
                               int din, int dout, int act_in) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * din) return;
  const int b = idx / din, i = idx % din;
  float s = 0.0f;
  for (int j = 0; j < dout; ++j) s += dOut[b * dout + j] * W[i * dout + j];
  dIn[idx] = act_in ? s * gelu_grad(Zin[idx]) : s;
}

// dW[i][j] = sum_b In[b][i] dOut[b][j];  db[j] = sum_b dOut[b][j] (row i = din)
__global__ void k_dense_grad(const float* In, const float* dOut, float* dW, float* db, int B, int din,
                             int dout) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (din + 1) * dout) return;
  const int i = idx / dout, j = idx % dout;
  float s = 0.0f;
  if (i < din) {
    for (int b = 0; b < B; ++b) s += In[b * din + i] * dOut[b * dout + j];
    dW[i * dout + j] = s;
  } else {
    for (int b = 0; b < B; ++b) s += dOut[b * dout + j];
    db[j] = s;
  }
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                       float eps, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float mi = b1 * m[i] + (1.0f - b1) * g[i];
  const float vi = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(b1, static_cast<float>(step)));
  const float vh = vi / (1.0f - powf(b2, static_cast<float>(step)));
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// rows of X and T selected by idx into contiguous batch buffers
__global__ void k_gather(const float* X, const float* T, const int* idx, float* Xb, float* Tb, int B,
                         int dx, int dt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B * (dx + dt)) return;
  const int b = k / (dx + dt), c = k % (dx + dt);
  const int r = idx[b];
  if (c < dx) Xb[b * dx + c] = X[static_cast<long long>(r) * dx + c];
  else Tb[b * dt + (c - dx)] = T[static_cast<long long>(r) * dt + (c - dx)];
}

// ------------------------------------------------------------ features

// The eight row statistics of one row slab of a (block-diagonal) CSR whose
// columns are global; one workgroup per slab, one row per thread at a time.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_feat_rows(const int* row_ptr, const int* col_idx,
                                                      const void* vals, FeatureLayout L, double* stats) {
  __shared__ double red[kWavesPerBlock][8];
  const int sl = blockIdx.x;
  const int r0 = L.slab_row0[sl], r1 = min(L.slab_row0[sl + 1], L.p_row1[L.slab_prob[sl]]);
  double s_nnz = 0, s_nnz2 = 0, mx_nnz = 0, mx_rowabs = 0, mx_right = -1e308, mn_left = 1e308,
         mx_diag = 0, s_dom = 0;
  const VT* v = static_cast<const VT*>(vals);
  for (int r = r0 + threadIdx.x; r < r1; r += kBlock) {
    const int k0 = row_ptr[r], k1 = row_ptr[r + 1];
    double diag = 0.0, off = 0.0;
    for (int k = k0; k < k1; ++k) {
      const double a = static_cast<double>(v[k]);
      const int c = col_idx[k];
      if (c == r) diag += a;
      else off += fabs(a);
    }
    const double nz = k1 - k0;
    s_nnz += nz;
    s_nnz2 += nz * nz;
    mx_nnz = fmax(mx_nnz, nz);
    mx_rowabs = fmax(mx_rowabs, fabs(diag) + off);
    mx_right = fmax(mx_right, diag + off);
    mn_left = fmin(mn_left, diag - off);
    mx_diag = fmax(mx_diag, fabs(diag));
    s_dom += fabs(diag) / (off + 1e-300);
  }
  const double vals8[8] = {s_nnz, s_nnz2, mx_nnz, mx_rowabs, mx_right, -mn_left, mx_diag, s_dom};
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bool is_max = (q == 2 || q == 3 || q == 4 || q == 5 || q == 6);
    const double s = is_max ? wave_max(vals8[q]) : wave_sum(vals8[q]);
    if (lane == 0) red[wave][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    const bool is_max = (q == 2 || q == 3 || q == 4 || q == 5 || q == 6);
    double s = red[0][q];
    for (int w = 1; w < kWavesPerBlock; ++w) s = is_max ? fmax(s, red[w][q]) : s + red[w][q];
    stats[sl * 8 + q] = s;
  }
}

// The 16 features of problem p (see learn.hpp) from its slabs' row
// statistics, combined in slab order, and ||A_p||_1.
__global__ void k_feat_final(FeatureLayout L, const double* norm1, const double* stats, const double* t,
                             const double* tol, float* feat) {
  __shared__ double s[8];
  const int p = blockIdx.x;
  const int r0 = L.p_row0[p], r1 = L.p_row1[p];
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    const bool is_max = (q == 2 || q == 3 || q == 4 || q == 5 || q == 6);
    const int s0 = L.p_slab0[p], s1 = L.p_slab0[p + 1];
    double v = stats[s0 * 8 + q];
    for (int sl = s0 + 1; sl < s1; ++sl) v = is_max ? fmax(v, stats[sl * 8 + q]) : v + stats[sl * 8 + q];
    s[q] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double n = fmax(1.0, static_cast<double>(r1 - r0));
  const double mean_nnz = s[0] / n;
  const double var_nnz = fmax(0.0, s[1] / n - mean_nnz * mean_nnz);
  const double tt = t[p];
  const double ninf = s[3];
  auto lg = [](double x) { return log10(fmax(x, 1e-300)); };
  float* f = feat + p * kNumFeatures;
  f[0] = static_cast<float>(lg(n));
  f[1] = static_cast<float>(lg(s[0]));
  f[2] = static_cast<float>(mean_nnz);
  f[3] = static_cast<float>(s[2] / fmax(mean_nnz, 1e-300));
  f[4] = static_cast<float>(sqrt(var_nnz) / fmax(mean_nnz, 1e-300));
  f[5] = static_cast<float>(lg(ninf));
  f[6] = static_cast<float>(lg(norm1[p]));
  f[7] = static_cast<float>(lg(tt));
  f[8] = static_cast<float>(lg(tt * ninf));
  f[9] = static_cast<float>(lg(tt * norm1[p]));
  f[10] = static_cast<float>(tt * s[4]);                    // t * rightmost Gershgorin bound
  f[11] = static_cast<float>(lg(tt * s[5]));                // log t |leftmost Gershgorin bound|
  f[12] = static_cast<float>(lg(tt * s[6]));                // log t max |a_ii|
  f[13] = static_cast<float>(lg(s[7] / n));                 // mean diagonal dominance
  f[14] = static_cast<float>(lg(tol[p]));
  f[15] = static_cast<float>(lg(0.5 * tt * (s[4] + s[5])));  // log half Gershgorin width * t
}

// ------------------------------------------------------------ launchers

void mlp_dense_fwd(const float* in, const float* W, const float* b, float* Z, float* A, int B, int din,
                   int dout, int act, hipStream_t s) {
  const int n = B * dout;
  hipLaunchKernelGGL(k_dense_fwd, dim3((n + 255) / 256), dim3(256), 0, s, in, W, b, Z, A, B, din, dout, act);
}
void mlp_mse(const float* Y, const float* T, float* dY, float* loss, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(1024), 0, s, Y, T, dY, loss, n);
}
void mlp_dense_bwd_in(const float* dOut, const float* W, const float* Zin, float* dIn, int B, int din,
                      int dout, int act_in, hipStream_t s) {
  const int n = B * din;
  hipLaunchKernelGGL(k_dense_bwd_in, dim3((n + 255) / 256), dim3(256), 0, s, dOut, W, Zin, dIn, B, din,
                     dout, act_in);
}
void mlp_dense_grad(const float* In, const float* dOut, float* dW, float* db, int B, int din, int dout,
                    hipStream_t s) {
  const int n = (din + 1) * dout;
  hipLaunchKernelGGL(k_dense_grad, dim3((n + 255) / 256), dim3(256), 0, s, In, dOut, dW, db, B, din, dout);
}
void mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
              int step, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, step);
}
void mlp_gather(const float* X, const float* T, const int* idx, float* Xb, float* Tb, int B, int dx, int dt,
                hipStream_t s) {
  const int n = B * (dx + dt);
  hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, X, T, idx, Xb, Tb, B, dx, dt);
}
void feature_launch(const int* row_ptr, const int* col_idx, const void* vals, int vals_f32,
                    const FeatureLayout& L, const double* norm1, double* stats, const double* t,
                    const double* tol, float* feat, hipStream_t s) {
  if (vals_f32)
    hipLaunchKernelGGL(k_feat_rows<float>, dim3(L.nslab), dim3(kBlock), 0, s, row_ptr, col_idx, vals, L,
                       stats);
  else
    hipLaunchKernelGGL(k_feat_rows<double>, dim3(L.nslab), dim3(kBlock), 0, s, row_ptr, col_idx, vals, L,
                       stats);
  hipLaunchKernelGGL(k_feat_final, dim3(L.P), dim3(kWave), 0, s, L, norm1, stats, t, tol, feat);
}

}  // namespace mek
