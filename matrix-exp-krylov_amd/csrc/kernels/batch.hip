// Batched Arnoldi exp(t_p A_p) v_p for many independent problems of uneven
// size, advanced in lockstep.
//
// The matrices are stored as one block-diagonal CSR (columns offset to global
// rows), so one SpMV launch multiplies every active problem at once.  The
// basis V is one global column-major array; column j of problem p is rows
// [row0_p, row0_p + n_p) of column j.  Streaming kernels run over "streaming
// blocks", each a row range of exactly one problem, so every block's partial
// sums belong to one problem; per-problem finalisers (one workgroup each)
// reduce them in block order.  The reference controller
// (main.cu:1425-1518) runs on the device, one thread per problem, so a whole
// cycle (restart, m columns, exp(dt H) e1, lift, accept/reject) is a fixed
// launch sequence that the host captures once in a hipGraph and replays until
// no problem is active.  A rejected problem keeps its basis and only redoes
// the exponential and the lift, like the single-problem runner.
#include "batch.hpp"
#include "expm_core.hpp"

namespace mek {

namespace {

__device__ __forceinline__ double* ps(const BatchBufs& b, int p) { return b.pscal + static_cast<long long>(p) * kPScal; }
__device__ __forceinline__ int* pi(const BatchBufs& b, int p) { return b.pint + static_cast<long long>(p) * kPInt; }
__device__ __forceinline__ double* pcs(const BatchBufs& b, int p) { return b.colscale + static_cast<long long>(p) * (b.M + 2); }

// Problem p takes part in Arnoldi column j of this cycle.
__device__ __forceinline__ bool col_active(const BatchBufs& b, int p, int j) {
  const int* s = pi(b, p);
  return s[kBActive] && s[kBNeedBasis] && !s[kBBreakdown] && j < s[kBM];
}

template <int NQ>
__device__ __forceinline__ void block_partials_b(const double (&v)[NQ], int nq, double* red,
                                                 double* partials, int nsb, int blk) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    if (q < nq) {
      const double s = wave_sum(v[q]);
      if (lane == 0) red[wave * NQ + q] = s;
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nq; q += kBlock) {
    double s = red[q];
    for (int w = 1; w < kWavesPerBlock; ++w) s += red[w * NQ + q];
    partials[static_cast<long long>(q) * nsb + blk] = s;
  }
}

// Sum quantity q over problem p's streaming blocks into dst[q] (one WG).
__device__ void reduce_problem(const BatchBufs& b, int p, int nq, double* dst) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int s0 = b.p_sb0[p], s1 = b.p_sb0[p + 1];
  for (int q = wave; q < nq; q += kWavesPerBlock) {
    double s = 0.0;
    for (int g = s0 + lane; g < s1; g += kWave) s += b.partials[static_cast<long long>(q) * b.nsb + g];
    s = wave_sum(s);
    if (lane == 0) dst[q] = s;
  }
}

template <typename VT>
__device__ __forceinline__ double ldv(const void* vals, long long k) {
  return static_cast<double>(static_cast<const VT*>(vals)[k]);
}

}  // namespace

// --------------------------------------------------------------- start

// partials of ||V[:,0]||^2 per streaming block
__global__ __launch_bounds__(kBlock) void kb_norm0(BatchBufs b) {
  __shared__ double red[kWavesPerBlock];
  const int sb = blockIdx.x;
  double acc = 0.0;
  for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) acc += b.V[r] * b.V[r];
  double o[1] = {acc};
  block_partials_b<1>(o, 1, red, b.partials, b.nsb, sb);
}

// Per problem: beta, colscale[0] (guarded as on the reference's CPU path,
// main_cpu.cpp:473-484) and the controller's initial state
// (TimeSteppingParams, main.cu:1292-1305; first dt, main.cu:1425-1433).
__global__ __launch_bounds__(kBlock) void kb_init(BatchBufs b) {
  __shared__ double tmp[kWavesPerBlock];
  const int p = blockIdx.x;
  reduce_problem(b, p, 1, tmp);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double* s = ps(b, p);
  int* it = pi(b, p);
  const double beta = sqrt(tmp[0]);
  s[kBVNorm] = beta;
  pcs(b, p)[0] = (beta > 1e-14) ? 1.0 / beta : 1.0;
  const double t = s[kBT];
  const int R = it[kBMaxRestarts];
  const double dt_max = t, dt_min = fmax(1e-12, t * 1e-6);
  const double dt_init = (it[kBSubsteps] > 0) ? t / it[kBSubsteps] : t / (R > 0 ? R : 1);
  s[kBDtMax] = dt_max;
  s[kBDtMin] = dt_min;
  s[kBDt] = fmin(fmin(fmax(dt_init, dt_min), dt_max), t);
  s[kBTDone] = 0.0;
  s[kBErr] = 0.0;
  s[kBLastRes] = 0.0;
  it[kBMaxSteps] = (it[kBSubsteps] > 0) ? 1000000 : (R > 0 ? R : 1000);
  it[kBActive] = (t > 0.0) ? 1 : 0;
  it[kBNeedBasis] = 1;
  it[kBRestart] = 0;
  it[kBBreakdown] = 0;
  it[kBKUsed] = 0;
  it[kBAccepted] = 0;
  it[kBRejects] = 0;
  it[kBConverged] = (t > 0.0) ? 0 : 1;
  it[kBAborted] = 0;
  it[kBAttempts] = 0;
  it[kBColumns] = 0;
  it[kBReorthCount] = 0;
}

// ----------------------------------------------------------- cycle start

// Accepted problems restart from y: V[:,0] = y.
__global__ __launch_bounds__(kBlock) void kb_restart_copy(BatchBufs b) {
  const int sb = blockIdx.x;
  const int p = b.sb_prob[sb];
  if (!pi(b, p)[kBRestart]) return;
  for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) b.V[r] = b.y[r];
}

// One thread per problem: finish the restart (main.cu:2023-2073, guarded as
// main_cpu.cpp:589-600) and clip dt to the remaining time (main.cu:1432).
__global__ __launch_bounds__(kBlock) void kb_cycle_begin(BatchBufs b) {
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < b.P; p += gridDim.x * kBlock) {
    double* s = ps(b, p);
    int* it = pi(b, p);
    if (!it[kBActive]) continue;
    if (it[kBRestart]) {
      const double beta = s[kBYNorm];
      s[kBVNorm] = beta;
      pcs(b, p)[0] = (beta > 1e-14) ? 1.0 / beta : 1.0;
      it[kBRestart] = 0;
      it[kBNeedBasis] = 1;
      it[kBBreakdown] = 0;
      it[kBKUsed] = 0;
      it[kBRejects] = 0;
      s[kBDt] = fmin(s[kBDt], s[kBT] - s[kBTDone]);
    }
  }
}

// ------------------------------------------------------------- columns

// w = A q_j for every row block of a problem in column j (block-diagonal CSR).
template <typename VT>
__global__ __launch_bounds__(kBlock) void kb_spmv(BatchBufs b, int j) {
  __shared__ double prod[kBatchNnzCap];
  const long long nrb = b.nrb;
  const int rb0 = static_cast<int>(nrb * blockIdx.x / gridDim.x);
  const int rb1 = static_cast<int>(nrb * (blockIdx.x + 1) / gridDim.x);
  const double* __restrict__ x = b.V + static_cast<long long>(j) * b.ld;
  for (int rb = rb0; rb < rb1; ++rb) {
    const int p = b.rb_prob[rb];
    if (!col_active(b, p, j)) continue;  // uniform over the block
    const double xs = pcs(b, p)[j];
    const int r0 = b.rb_start[rb], r1 = b.rb_start[rb + 1];
    const int k0 = b.row_ptr[r0], k1 = b.row_ptr[r1];
    if (k1 - k0 <= kBatchNnzCap) {
      for (int k = k0 + threadIdx.x; k < k1; k += kBlock)
        prod[k - k0] = ldv<VT>(b.vals, k) * (xs * x[b.col_idx[k]]);
      __syncthreads();
      for (int r = r0 + threadIdx.x; r < r1; r += kBlock) {
        double a = 0.0;
        for (int e = b.row_ptr[r] - k0; e < b.row_ptr[r + 1] - k0; ++e) a += prod[e];
        b.w[r] = a;
      }
      __syncthreads();
    } else {
      double a = 0.0;
      for (int k = k0 + threadIdx.x; k < k1; k += kBlock) a += ldv<VT>(b.vals, k) * (xs * x[b.col_idx[k]]);
      const double s = wave_sum(a);
      if (threadIdx.x % kWave == 0) prod[threadIdx.x / kWave] = s;
      __syncthreads();
      if (threadIdx.x == 0) b.w[r0] = (prod[0] + prod[1]) + (prod[2] + prod[3]);
      __syncthreads();
    }
  }
}

// partials of q_i . w (i <= j) and ||w||^2 (quantity j+1) per streaming block
template <int MAXJ>
__global__ __launch_bounds__(kBlock) void kb_dots1(BatchBufs b, int j) {
  __shared__ double red[kWavesPerBlock * (MAXJ + 1)];
  const int sb = blockIdx.x;
  const int p = b.sb_prob[sb];
  if (!col_active(b, p, j)) return;
  const int nq = j + 1;
  const double* __restrict__ cs = pcs(b, p);
  double acc[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) acc[q] = 0.0;
  for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) {
    const double a = b.w[r];
#pragma unroll
    for (int i = 0; i < MAXJ; ++i)
      if (i < nq) acc[i] += (cs[i] * b.V[static_cast<long long>(i) * b.ld + r]) * a;
    acc[MAXJ] += a * a;
  }
  double o[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) o[q] = (q < nq) ? acc[q] : ((q == nq) ? acc[MAXJ] : 0.0);
  block_partials_b<MAXJ + 1>(o, nq + 1, red, b.partials, b.nsb, sb);
}

// w' = w - sum h1_i q_i into V[:, j+1]; partials of q_i . w' and ||w'||^2
template <int MAXJ>
__global__ __launch_bounds__(kBlock) void kb_update_dot(BatchBufs b, int j) {
  __shared__ double red[kWavesPerBlock * (MAXJ + 1)];
  const int sb = blockIdx.x;
  const int p = b.sb_prob[sb];
  if (!col_active(b, p, j)) return;
  const int nq = j + 1;
  const double* __restrict__ cs = pcs(b, p);
  const double* __restrict__ h = b.h1 + static_cast<long long>(p) * (b.M + 2);
  double* __restrict__ out = b.V + static_cast<long long>(j + 1) * b.ld;
  double acc[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) acc[q] = 0.0;
  for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) {
    double q[MAXJ];
    double a = b.w[r];
#pragma unroll
    for (int i = 0; i < MAXJ; ++i) {
      if (i < nq) {
        q[i] = cs[i] * b.V[static_cast<long long>(i) * b.ld + r];
        a -= h[i] * q[i];
      }
    }
    out[r] = a;
#pragma unroll
    for (int i = 0; i < MAXJ; ++i)
      if (i < nq) acc[i] += q[i] * a;
    acc[MAXJ] += a * a;
  }
  double o[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) o[q] = (q < nq) ? acc[q] : ((q == nq) ? acc[MAXJ] : 0.0);
  block_partials_b<MAXJ + 1>(o, nq + 1, red, b.partials, b.nsb, sb);
}

// w'' = w' - sum h2_i q_i in place; partial ||w''||^2 (only if re-orthogonalising)
__global__ __launch_bounds__(kBlock) void kb_update_norm(BatchBufs b, int j) {
  __shared__ double red[kWavesPerBlock];
  const int sb = blockIdx.x;
  const int p = b.sb_prob[sb];
  if (!col_active(b, p, j) || !pi(b, p)[kBReorth]) return;
  const int nq = j + 1;
  const double* __restrict__ cs = pcs(b, p);
  const double* __restrict__ h = b.h2 + static_cast<long long>(p) * (b.M + 2);
  double* __restrict__ io = b.V + static_cast<long long>(j + 1) * b.ld;
  double acc = 0.0;
  for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) {
    double a = io[r];
    for (int i = 0; i < nq; ++i) a -= h[i] * (cs[i] * b.V[static_cast<long long>(i) * b.ld + r]);
    io[r] = a;
    acc += a * a;
  }
  double o[1] = {acc};
  block_partials_b<1>(o, 1, red, b.partials, b.nsb, sb);
}

__global__ __launch_bounds__(kBlock) void kb_fin_h1(BatchBufs b, int j) {
  __shared__ double tmp[kBatchMaxM + 4];
  const int p = blockIdx.x;
  if (!col_active(b, p, j)) return;
  reduce_problem(b, p, j + 2, tmp);
  __syncthreads();
  double* h1 = b.h1 + static_cast<long long>(p) * (b.M + 2);
  for (int i = threadIdx.x; i <= j; i += kBlock) h1[i] = tmp[i];
  if (threadIdx.x == 0) ps(b, p)[kBWNorm2] = tmp[j + 1];
}

// H(:, j), breakdown (norm < 1e-14 or not finite, main.cu:1911), next scale
__device__ void finish_column_b(const BatchBufs& b, int p, int j, double norm2, int reorth) {
  const int ldh = b.M + 2;
  double* H = b.H + static_cast<long long>(p) * ldh * ldh;
  const double* h1 = b.h1 + static_cast<long long>(p) * ldh;
  const double* h2 = b.h2 + static_cast<long long>(p) * ldh;
  int* it = pi(b, p);
  for (int i = 0; i <= j; ++i) H[static_cast<long long>(j) * ldh + i] = h1[i] + (reorth ? h2[i] : 0.0);
  for (int i = j + 2; i < ldh; ++i) H[static_cast<long long>(j) * ldh + i] = 0.0;
  const double nrm = sqrt(norm2);
  it[kBKUsed] = j + 1;
  it[kBColumns] += 1;
  if (nrm < 1e-14 || isnan(nrm) || isinf(nrm)) {
    H[static_cast<long long>(j) * ldh + j + 1] = 0.0;
    it[kBBreakdown] = 1;
  } else {
    H[static_cast<long long>(j) * ldh + j + 1] = nrm;
    pcs(b, p)[j + 1] = 1.0 / nrm;
  }
  if (reorth) it[kBReorthCount] += 1;
}

__global__ __launch_bounds__(kBlock) void kb_fin_h2(BatchBufs b, int j, int gs_mode, double eta) {
  __shared__ double tmp[kBatchMaxM + 4];
  const int p = blockIdx.x;
  if (!col_active(b, p, j)) return;
  reduce_problem(b, p, j + 2, tmp);
  __syncthreads();
  double* h2 = b.h2 + static_cast<long long>(p) * (b.M + 2);
  for (int i = threadIdx.x; i <= j; i += kBlock) h2[i] = tmp[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    const double wp2 = tmp[j + 1];
    double* s = ps(b, p);
    s[kBWpNorm2] = wp2;
    int reorth = 1;
    if (gs_mode == 0) reorth = 0;
    else if (gs_mode == 2) reorth = (sqrt(wp2) < eta * sqrt(s[kBWNorm2])) ? 1 : 0;
    if (!(wp2 == wp2)) reorth = 0;
    pi(b, p)[kBReorth] = reorth;
    if (!reorth) finish_column_b(b, p, j, wp2, 0);
  }
}

__global__ __launch_bounds__(kBlock) void kb_fin_col(BatchBufs b, int j) {
  __shared__ double tmp[kWavesPerBlock];
  const int p = blockIdx.x;
  if (!col_active(b, p, j) || !pi(b, p)[kBReorth]) return;
  reduce_problem(b, p, 1, tmp);
  __syncthreads();
  if (threadIdx.x == 0) finish_column_b(b, p, j, tmp[0], 1);
}

// The basis of every problem that built one this cycle is now complete.
__global__ __launch_bounds__(kBlock) void kb_after_basis(BatchBufs b) {
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < b.P; p += gridDim.x * kBlock) {
    int* it = pi(b, p);
    if (it[kBActive]) it[kBNeedBasis] = 0;
  }
}

// ------------------------------------------------ exponential, lift, control

// One workgroup per active problem: wH = exp(dt H_k) e1 and the reference's
// residual estimate |H(k,k-1) [exp(dt H_k) e1]_{k-1}| beta (0 on breakdown).
template <bool BIG>
__global__ __launch_bounds__(kBlock) void kb_expm(BatchBufs b) {
  extern __shared__ double smem[];
  __shared__ double s_b[14];
  __shared__ int s_int[4];
  __shared__ double s_red[kWavesPerBlock];
  const int p = blockIdx.x;
  int* it = pi(b, p);
  if (!it[kBActive]) return;
  const int k = it[kBKUsed];
  const int brk = it[kBBreakdown];
  const int ldh = b.M + 2;
  const double* H = b.H + static_cast<long long>(p) * ldh * ldh;
  double* wH = b.wH + static_cast<long long>(p) * (b.M + 4);
  double* s = ps(b, p);
  const double dt = s[kBDt];
  if (k <= 0) {
    for (int i = threadIdx.x; i < b.M + 4; i += kBlock) wH[i] = 0.0;
    if (threadIdx.x == 0) s[kBErr] = 0.0;
    return;
  }
  double* base = BIG ? b.expm_ws + static_cast<long long>(p) * 6 * b.M * b.M : smem;
  double* S[6];
  for (int i = 0; i < 6; ++i) S[i] = base + static_cast<long long>(i) * k * k;
  for (int e = threadIdx.x; e < k * k; e += kBlock) {
    const int r = e % k, c = e / k;
    S[0][e] = dt * H[static_cast<long long>(c) * ldh + r];
  }
  __syncthreads();
  double* E = expm_slots<BIG>(S, k, s_b, s_int, s_red);
  for (int i = threadIdx.x; i < b.M + 4; i += kBlock) wH[i] = (i < k) ? E[i] : 0.0;
  if (threadIdx.x == 0) {
    const double hk = H[static_cast<long long>(k - 1) * ldh + k];
    s[kBErr] = brk ? 0.0 : fabs(hk * E[k - 1]) * s[kBVNorm];
  }
}

// y = sum_{i<=k, |wH_i| > 1e-15} beta wH_i q_i (main.cu:1992-1999); partials ||y||^2
__global__ __launch_bounds__(kBlock) void kb_lift(BatchBufs b) {
  __shared__ double red[kWavesPerBlock];
  const int sb = blockIdx.x;
  const int p = b.sb_prob[sb];
  const int* it = pi(b, p);
  if (!it[kBActive]) return;
  const int k = it[kBKUsed];
  const double beta = ps(b, p)[kBVNorm];
  const double* __restrict__ cs = pcs(b, p);
  const double* __restrict__ wh = b.wH + static_cast<long long>(p) * (b.M + 4);
  double acc = 0.0;
  for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) {
    double a = 0.0;
    for (int i = 0; i <= k; ++i) {
      const double c = wh[i];
      if (fabs(c) > 1e-15) a += (beta * c) * (cs[i] * b.V[static_cast<long long>(i) * b.ld + r]);
    }
    b.y[r] = a;
    acc += a * a;
  }
  double o[1] = {acc};
  block_partials_b<1>(o, 1, red, b.partials, b.nsb, sb);
}

__global__ __launch_bounds__(kBlock) void kb_fin_lift(BatchBufs b) {
  __shared__ double tmp[kWavesPerBlock];
  const int p = blockIdx.x;
  if (!pi(b, p)[kBActive]) return;
  reduce_problem(b, p, 1, tmp);
  __syncthreads();
  if (threadIdx.x == 0) ps(b, p)[kBYNorm] = sqrt(tmp[0]);
}

// The reference controller's accept / reject step (main.cu:1476-1516), one
// thread per problem; writes the number of still-active problems.
__global__ __launch_bounds__(1024) void kb_control(BatchBufs b) {
  __shared__ int cnt[1024 / kWave];
  int active = 0;
  for (int p = threadIdx.x; p < b.P; p += blockDim.x) {
    double* s = ps(b, p);
    int* it = pi(b, p);
    if (!it[kBActive]) continue;
    const double err = s[kBErr];
    const double tol = s[kBTol];
    it[kBAttempts] += 1;
    if (err <= tol || it[kBBreakdown]) {
      s[kBTDone] += s[kBDt];
      it[kBAccepted] += 1;
      s[kBLastRes] = err;
      if (s[kBTDone] >= s[kBT]) {
        it[kBConverged] = 1;
        it[kBActive] = 0;
      } else if (it[kBAccepted] >= it[kBMaxSteps]) {
        it[kBActive] = 0;  // step cap reached with t_done < t
      } else {
        it[kBRestart] = 1;
        if (err < 0.1 * tol) s[kBDt] = fmin(s[kBDt] * 1.5 * 0.8, s[kBDtMax]);
      }
    } else {
      it[kBRejects] += 1;
      if (it[kBRejects] >= 20) {
        it[kBAborted] = 1;
        it[kBActive] = 0;
        s[kBLastRes] = err;
        s[kBTDone] = s[kBT];
      } else {
        s[kBDt] = fmax(s[kBDt] * 0.5 * 0.8, s[kBDtMin]);
      }
    }
    active += it[kBActive];
  }
  // count of active problems (fixed-order reduction, no atomics)
  for (int off = kWave / 2; off >= 1; off >>= 1) active += __shfl_xor(active, off);
  if (threadIdx.x % kWave == 0) cnt[threadIdx.x / kWave] = active;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) tot += cnt[w];
    b.counters[0] = tot;
    b.counters[1] += 1;  // cycles
  }
}

// ------------------------------------------------------------ launchers

void batch_prepare() {
  // the in-LDS exponential of the largest problems needs 154 KB of LDS
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&kb_expm<false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(5 * kMaxKLds * kMaxKLds * sizeof(double)));
}

void batch_launch_init(const BatchBufs& b, hipStream_t s) {
  hipLaunchKernelGGL(kb_norm0, dim3(b.nsb), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(kb_init, dim3(b.P), dim3(kBlock), 0, s, b);
}

static void launch_dots1(const BatchBufs& b, int j, hipStream_t s) {
  const int nq = j + 1;
  if (nq <= 8) hipLaunchKernelGGL(kb_dots1<8>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else if (nq <= 16) hipLaunchKernelGGL(kb_dots1<16>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else if (nq <= 32) hipLaunchKernelGGL(kb_dots1<32>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else if (nq <= 48) hipLaunchKernelGGL(kb_dots1<48>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else hipLaunchKernelGGL(kb_dots1<64>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
}

static void launch_update_dot_b(const BatchBufs& b, int j, hipStream_t s) {
  const int nq = j + 1;
  if (nq <= 8) hipLaunchKernelGGL(kb_update_dot<8>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else if (nq <= 16) hipLaunchKernelGGL(kb_update_dot<16>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else if (nq <= 32) hipLaunchKernelGGL(kb_update_dot<32>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else if (nq <= 48) hipLaunchKernelGGL(kb_update_dot<48>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
  else hipLaunchKernelGGL(kb_update_dot<64>, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
}

void batch_launch_cycle(const BatchBufs& b, int gs_mode, double eta, int spmv_grid, hipStream_t s) {
  const int pg = (b.P + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(kb_restart_copy, dim3(b.nsb), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(kb_cycle_begin, dim3(pg), dim3(kBlock), 0, s, b);
  for (int j = 0; j < b.M; ++j) {
    if (b.vals_f32) hipLaunchKernelGGL(kb_spmv<float>, dim3(spmv_grid), dim3(kBlock), 0, s, b, j);
    else hipLaunchKernelGGL(kb_spmv<double>, dim3(spmv_grid), dim3(kBlock), 0, s, b, j);
    launch_dots1(b, j, s);
    hipLaunchKernelGGL(kb_fin_h1, dim3(b.P), dim3(kBlock), 0, s, b, j);
    launch_update_dot_b(b, j, s);
    hipLaunchKernelGGL(kb_fin_h2, dim3(b.P), dim3(kBlock), 0, s, b, j, gs_mode, eta);
    hipLaunchKernelGGL(kb_update_norm, dim3(b.nsb), dim3(kBlock), 0, s, b, j);
    hipLaunchKernelGGL(kb_fin_col, dim3(b.P), dim3(kBlock), 0, s, b, j);
  }
  hipLaunchKernelGGL(kb_after_basis, dim3(pg), dim3(kBlock), 0, s, b);
  const int kk = b.M;
  if (kk <= kMaxKLds) {
    hipLaunchKernelGGL(kb_expm<false>, dim3(b.P), dim3(kBlock), 5 * kk * kk * sizeof(double), s, b);
  } else {
    hipLaunchKernelGGL(kb_expm<true>, dim3(b.P), dim3(kBlock), 0, s, b);
  }
  hipLaunchKernelGGL(kb_lift, dim3(b.nsb), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(kb_fin_lift, dim3(b.P), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(kb_control, dim3(1), dim3(1024), 0, s, b);
}

}  // namespace mek
