// exp(dt * H_k) e1 on the device: one workgroup, FP64 Pade scaling and
// squaring (Higham 2005), replacing the reference's host-side Eigen
// `A.exp()` (main.cu:33-36, used at main.cu:1966-1978 and :2007-2014).  The
// algorithm lives in expm_core.hpp.
#include <mutex>
#include <vector>

#include "expm_core.hpp"
#include "launch.hpp"
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

// ---------------------------------------------- allocation-free exponential
//
// Working memory of expm_dense_noalloc, one copy per device for the life of
// the process: 16 scalar slots and 8 int slots (as 4 doubles), the first
// column (k + 4 doubles), then at offset kScrWs the global workspace of six
// k x k matrices that the kernel uses for k > kMaxKLds.
constexpr int kScrWs = 256;
constexpr size_t kScratchDoubles = kScrWs + 6 * static_cast<size_t>(kExpmNoAllocMaxK) * kExpmNoAllocMaxK;
static_assert(24 + kExpmNoAllocMaxK + 4 <= kScrWs, "first column overlaps the workspace");
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // device address of `scratch`, per device

// The device address of this device's `scratch`, looked up once per device
// with that device current (host code never passes the symbol itself).
double* scratch_on_current_device() {
  int dev = 0;
  MEK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (static_cast<int>(g_scratch_ptr.size()) <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MEK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

void expm_dense_noalloc(int k, const double* d_A, double* d_E, hipStream_t s) {
  if (k < 1 || k > kExpmNoAllocMaxK)
    throw Error("ERROR: expm_dense_noalloc: k=" + std::to_string(k) + " outside [1, " +
                std::to_string(kExpmNoAllocMaxK) + "] (the static scratch holds six " +
                std::to_string(kExpmNoAllocMaxK) + "^2 matrices)");
  double* base = scratch_on_current_device();
  ExpmArgs a{};
  a.H = d_A;
  a.ldh = k;
  // mode 2 with k_override reads no slot before writing it: nothing in the
  // scratch needs clearing between calls
  a.iscal = reinterpret_cast<int*>(base + 16);
  a.iscal_out = reinterpret_cast<int*>(base + 16);
  a.scal = base;
  a.out = base + 24;
  a.out_len = k;
  a.dt = 1.0;
  a.mode = 2;  // plain exponential: no residual
  a.ws = base + kScrWs;
  a.k_override = k;
  a.breakdown_override = 0;
  a.full_out = d_E;
  launch_expm(a, k, s);
  MEK_HIP(hipGetLastError());
}

}  // namespace mek
