// exp(dt * H_k) e1 on the device: one workgroup, FP64 Pade scaling and
// squaring (Higham 2005), replacing the reference's host-side Eigen
// `A.exp()` (main.cu:33-36, used at main.cu:1966-1978 and :2007-2014).  The
// algorithm lives in expm_core.hpp.
#include "expm_core.hpp"
#include "../host/hip_check.hpp"

namespace mek {

template <bool BIG>
__global__ __launch_bounds__(kBlock) void k_expm(ExpmArgs a) {
  extern __shared__ double smem[];
  __shared__ double s_b[14];
  __shared__ int s_int[4];
  __shared__ double s_red[kWavesPerBlock];
  __shared__ int s_dims[2];

  if (threadIdx.x == 0) {
    int k = (a.k_override > 0) ? a.k_override : a.iscal[kKUsed];
    const int brk = (a.k_override > 0) ? a.breakdown_override : a.iscal[kBreakdown];
    int kk = k;
    if (a.mode == 1 && !brk) kk = k + 2;
    s_dims[0] = k;
    s_dims[1] = kk;
  }
  __syncthreads();
  const int k = s_dims[0];
  const int kk = s_dims[1];
  const int brk = (a.k_override > 0) ? a.breakdown_override : a.iscal[kBreakdown];
  if (k <= 0) {
    for (int i = threadIdx.x; i < a.out_len; i += kBlock) a.out[i] = 0.0;
    if (threadIdx.x == 0 && a.mode == 0) a.scal[kErr] = 0.0;
    return;
  }
  double* base = BIG ? a.ws : smem;
  double* S[6];
  for (int i = 0; i < 6; ++i) S[i] = base + static_cast<long long>(i) * kk * kk;

  // M = dt * H_k (Expokit: augmented with H(k+1, k) = 1 before scaling)
  for (int e = threadIdx.x; e < kk * kk; e += kBlock) {
    const int r = e % kk, c = e / kk;
    double v = 0.0;
    if (r < k + 1 && c < k && r < a.ldh) v = a.H[static_cast<long long>(c) * a.ldh + r];
    if (r >= k && kk == k) v = 0.0;  // k x k block only
    if (kk == k + 2 && r == k + 1 && c == k) v = 1.0;
    S[0][e] = a.dt * v;
  }
  __syncthreads();
  double* E = expm_slots<BIG>(S, kk, s_b, s_int, s_red);
  for (int i = threadIdx.x; i < a.out_len; i += kBlock) a.out[i] = (i < kk) ? E[i] : 0.0;
  if (a.full_out)
    for (int e = threadIdx.x; e < kk * kk; e += kBlock) a.full_out[e] = E[e];
  if (threadIdx.x == 0) {
    a.iscal_out[kExpmSquarings] = s_int[1];
    if (a.mode == 0) {
      // reference residual_estimate (main.cu:2004-2020): absolute,
      // |H(k, k-1)| * |[exp(dt H_k) e1]_{k-1}| * ||v||; 0 on breakdown
      const double hk = a.H[static_cast<long long>(k - 1) * a.ldh + k];
      const double err = fabs(hk * E[k - 1]) * a.scal[kVNorm];
      a.scal[kErr] = brk ? 0.0 : err;
    }
  }
}

size_t expm_lds_bytes(int kk) { return static_cast<size_t>(5) * kk * kk * sizeof(double); }

void launch_expm(const ExpmArgs& a, int kk_max, hipStream_t s) {
  if (kk_max <= kMaxKLds) {
    // allow the dynamic LDS of the largest in-LDS problem on the current
    // device; set before every launch (a cheap call) so that each device and
    // each calling thread has it, and report a failure as a launch error
    MEK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_expm<false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(expm_lds_bytes(kMaxKLds))));
    hipLaunchKernelGGL(k_expm<false>, dim3(1), dim3(kBlock), expm_lds_bytes(kk_max), s, a);
  } else {
    hipLaunchKernelGGL(k_expm<true>, dim3(1), dim3(kBlock), 0, s, a);
  }
}

}  // namespace mek
