// Batched Arnoldi exp(t_p A_p) v_p for many independent problems of uneven
// size, advanced in lockstep.
//
// The matrices are stored as one block-diagonal CSR (columns offset to global
// rows), so one SpMV launch multiplies every active problem at once.  The
// basis V is one global column-major array; column j of problem p is rows
// [row0_p, row0_p + n_p) of column j; every row0_p is even (an odd n_p is
// followed by one empty row, zero in every vector).  Streaming kernels run over "streaming
// blocks", each a row range of exactly one problem, so every block's partial
// sums belong to one problem; per-problem finalisers (one workgroup each)
// reduce them in block order.  The reference controller
// (main.cu:1425-1518) runs on the device, one thread per problem, so a whole
// cycle (restart, m columns, exp(dt H) e1, lift, accept/reject) is a fixed
// launch sequence that the host captures once in a hipGraph and replays until
// no problem is active.  A rejected problem keeps its basis and only redoes
// the exponential and the lift, like the single-problem runner.
#include "batch.hpp"

#include <algorithm>

#include "expm_core.hpp"
#include "sweeps.hpp"
#include "../host/hip_check.hpp"

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

// Expokit step rounding: up to two significant digits (Sidje 1998).
__device__ __forceinline__ double round_step(double x) {
  const double sc = pow(10.0, floor(log10(x)) - 1.0);
  return ceil(x / sc) * sc;
}

// Expokit constants (Sidje 1998, expv): safety factors and limits.
constexpr double kExpokitGamma = 0.9;
constexpr double kExpokitDelta = 1.2;
constexpr double kExpokitBtol = 1.0e-7;
constexpr int kExpokitMxrej = 10;
constexpr int kExpokitMaxSteps = 100000;

// Problem p takes part in the Expokit ||A v_{m+1}|| product of this cycle.
__device__ __forceinline__ bool av_active(const BatchBufs& b, int p) {
  const int* s = pi(b, p);
  return s[kBActive] && s[kBNeedBasis] && s[kBCtrl] >= 1 && !s[kBBreakdown];
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
  s[kBDtMax] = dt_max;
  s[kBDtMin] = dt_min;
  if (it[kBCtrl] >= 1) {
    // Expokit's first step from ||A||_inf, beta and tol (expv), or t/substeps
    // when the caller predicted the sub-step count
    const int m = it[kBM];
    const double anorm = s[kBAnorm] > 0.0 ? s[kBAnorm] : 1.0;
    const double fact = pow((m + 1) / exp(1.0), m + 1) * sqrt(2.0 * M_PI * (m + 1));
    const double bet = beta > 0.0 ? beta : 1.0;
    const double tol0 = (it[kBCtrl] == 2) ? s[kBTol] * bet : s[kBTol];
    const double t_heur = round_step((1.0 / anorm) * pow((fact * tol0) / (4.0 * bet * anorm), 1.0 / m));
    const double t_new = (it[kBSubsteps] > 0) ? t / it[kBSubsteps] : t_heur;
    s[kBDtMin] = t_heur;  // Expokit's own first step: the fallback after a rejection
    s[kBTNew] = t_new;
    s[kBDt] = fmin(t_new, t);
    s[kBSErr] = 0.0;
  } else {
    const double dt_init = (it[kBSubsteps] > 0) ? t / it[kBSubsteps] : t / (R > 0 ? R : 1);
    s[kBDt] = fmin(fmin(fmax(dt_init, dt_min), dt_max), t);
  }
  s[kBTDone] = 0.0;
  s[kBErr] = 0.0;
  s[kBLastRes] = 0.0;
  if (it[kBCtrl] >= 1) it[kBMaxSteps] = (R > 0) ? max(R, kExpokitMaxSteps) : kExpokitMaxSteps;
  else it[kBMaxSteps] = (it[kBSubsteps] > 0) ? 1000000 : (R > 0 ? R : 1000);
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
  it[kBIReject] = 0;
  it[kBLift] = 0;
  it[kBKLift] = 0;
  it[kBRejTotal] = 0;
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
    it[kBLift] = 0;
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
      it[kBIReject] = 0;
      if (it[kBCtrl] >= 1) s[kBDt] = fmin(s[kBTNew], s[kBT] - s[kBTDone]);
      else s[kBDt] = fmin(s[kBDt], s[kBT] - s[kBTDone]);
    }
  }
}

// ------------------------------------------------------------- columns

// w = A q_j for every row block of a problem in column j (block-diagonal CSR).
// With av = 1 (Expokit), w = A q_{m_p} for the problems in av_active.  The
// blocks are list[0, *cnt) (the compacted blocks of this cycle's problems)
// or, with list == nullptr, row blocks [0, nrb_lim).  They are split into one
// contiguous range per XCD, and the workgroups of an XCD stride through its
// range together, so the rows in flight on an XCD (and the x entries they
// gather) stay close and in its L2.  Short blocks stage val*x in LDS and sum
// each row in storage order.  All nonzero loads and x gathers of a block are
// issued together, and the next block's problem and bounds are loaded one
// iteration ahead.
template <typename VT>
__global__ __launch_bounds__(kBlock) void kb_spmv(BatchBufs b, int j, int av, int nrb_lim, const int* list,
                                                  const int* cnt) {
  constexpr int U = kBatchNnzCap / kBlock;
  __shared__ double prod[kBatchNnzCap];
  int xcd = 0, li = blockIdx.x, gx = gridDim.x, nx = 1;
  if (gridDim.x % kNumXcd == 0) {
    xcd = blockIdx.x % kNumXcd;
    li = blockIdx.x / kNumXcd;
    gx = gridDim.x / kNumXcd;
    nx = kNumXcd;
  }
  const int tid = threadIdx.x;
  const int* __restrict__ rp = b.row_ptr;
  const int* __restrict__ ci = b.col_idx;
  const long long nrb = list ? *cnt : nrb_lim;
  const int rb0 = static_cast<int>(nrb * xcd / nx);
  const int rb1 = static_cast<int>(nrb * (xcd + 1) / nx);
  int rb = rb0 + li;
  auto blk = [&](int i) { return list ? list[i] : i; };
  // current block: problem p, rows [r0, r1), nonzeros [k0, k1); next block:
  // problem np, rows [nr0, nr1)
  int p = 0, r0 = 0, r1 = 0, k0 = 0, k1 = 0, np = 0, nr0 = 0, nr1 = 0;
  if (rb < rb1) {
    const int g = blk(rb);
    p = b.rb_prob[g];
    r0 = b.rb_start[g];
    r1 = b.rb_start[g + 1];
    k0 = rp[r0];
    k1 = rp[r1];
  }
  if (rb + gx < rb1) {
    const int g = blk(rb + gx);
    np = b.rb_prob[g];
    nr0 = b.rb_start[g];
    nr1 = b.rb_start[g + 1];
  }
  for (; rb < rb1; rb += gx) {
    int nk0 = 0, nk1 = 0, mp = 0, mr0 = 0, mr1 = 0;
    if (rb + gx < rb1) {
      nk0 = rp[nr0];
      nk1 = rp[nr1];
    }
    if (rb + 2 * gx < rb1) {
      const int g = blk(rb + 2 * gx);
      mp = b.rb_prob[g];
      mr0 = b.rb_start[g];
      mr1 = b.rb_start[g + 1];
    }
    if (av ? av_active(b, p) : col_active(b, p, j)) {  // uniform over the block
      const int jc = av ? pi(b, p)[kBM] : j;
      const double* __restrict__ x = b.V + static_cast<long long>(jc) * b.ld;
      const double xs = pcs(b, p)[jc];
      if (k1 - k0 <= kBatchNnzCap) {
        int idx[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + tid + u * kBlock;
          idx[u] = 0;
          v[u] = 0.0;
          if (k < k1) {
            idx[u] = ci[k];
            v[u] = ldv<VT>(b.vals, k);
          }
        }
        const int rr = r0 + tid;
        int e0 = 0, e1 = 0;
        if (rr < r1) {
          e0 = rp[rr] - k0;
          e1 = rp[rr + 1] - k0;
        }
        double pr[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pr[u] = v[u] * (xs * x[idx[u]]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + tid + u * kBlock;
          if (k < k1) prod[k - k0] = pr[u];
        }
        __syncthreads();
        for (int r = rr; r < r1; r += kBlock) {
          if (r != rr) {
            e0 = rp[r] - k0;
            e1 = rp[r + 1] - k0;
          }
          double a = 0.0;
          for (int e = e0; e < e1; ++e) a += prod[e];
          __builtin_nontemporal_store(a, b.w + r);
        }
        __syncthreads();
      } else {
        double a = 0.0;
        for (int k = k0 + tid; k < k1; k += kBlock) a += ldv<VT>(b.vals, k) * (xs * x[ci[k]]);
        const double s = wave_sum(a);
        if (tid % kWave == 0) prod[tid / kWave] = s;
        __syncthreads();
        if (tid == 0) b.w[r0] = (prod[0] + prod[1]) + (prod[2] + prod[3]);
        __syncthreads();
      }
    }
    p = np;
    r0 = nr0;
    r1 = nr1;
    k0 = nk0;
    k1 = nk1;
    np = mp;
    nr0 = mr0;
    nr1 = mr1;
  }
}

// Streaming blocks run the column-split sweeps over their own row range, two
// rows per lane (the range starts and ends on an even row).
__device__ __forceinline__ TileRange sb_tiles(const BatchBufs& b, int sb) {
  return block_tiles<2>(b.sb_row0[sb], b.sb_row0[sb + 1]);
}

// The per-column sweeps stride over the streaming blocks of column j in this
// cycle's compacted list (sb_list[0, list_cnt[M + 1 + j])); every block is
// one problem's row range, so the test of whether that problem still takes
// part (a breakdown earlier in the cycle ends it) is uniform per block.
__device__ __forceinline__ int sweep_count(const BatchBufs& b, int j) { return b.list_cnt[b.M + 1 + j]; }

// partials of q_i . w (i <= j) and ||w||^2 (quantity j+1) per streaming block
template <int MAXK>
__global__ __launch_bounds__(kBlock) void kb_dots1(BatchBufs b, int j) {
  const int n = sweep_count(b, j);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int sb = b.sb_list[i];
    const int p = b.sb_prob[sb];
    if (col_active(b, p, j))
      sweep_dots<MAXK, 2>(b.V, b.ld, 0, j + 1, pcs(b, p), b.w, sb_tiles(b, sb), true, b.partials, b.nsb, sb);
    __syncthreads();  // LDS of this block's reduction before the next block's
  }
}

// w' = w - sum h1_i q_i into V[:, j+1]; partials of q_i . w' and ||w'||^2
template <int MAXK>
__global__ __launch_bounds__(kBlock) void kb_update_dot(BatchBufs b, int j) {
  const int n = sweep_count(b, j);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int sb = b.sb_list[i];
    const int p = b.sb_prob[sb];
    if (col_active(b, p, j))
      sweep_update_dot<MAXK, 2>(b.V, b.ld, j + 1, b.h1 + static_cast<long long>(p) * (b.M + 2), pcs(b, p),
                                b.w, b.V + static_cast<long long>(j + 1) * b.ld, sb_tiles(b, sb),
                                b.partials, b.nsb, sb);
    __syncthreads();
  }
}

// w'' = w' - sum h2_i q_i in place; partial ||w''||^2 (only if re-orthogonalising)
template <int MAXK>
__global__ __launch_bounds__(kBlock) void kb_update_norm(BatchBufs b, int j) {
  const int n = sweep_count(b, j);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int sb = b.sb_list[i];
    const int p = b.sb_prob[sb];
    if (col_active(b, p, j) && pi(b, p)[kBReorth])
      sweep_update_norm<MAXK, 2>(b.V, b.ld, j + 1, b.h2 + static_cast<long long>(p) * (b.M + 2), pcs(b, p),
                                 b.V + static_cast<long long>(j + 1) * b.ld, sb_tiles(b, sb), b.partials, sb);
    __syncthreads();
  }
}

// partials of ||w||^2 per streaming block of the av_active problems
__global__ __launch_bounds__(kBlock) void kb_av_norm(BatchBufs b) {
  __shared__ double red[kWavesPerBlock];
  const int sb = blockIdx.x;
  if (!av_active(b, b.sb_prob[sb])) return;
  double acc = 0.0;
  for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) acc += b.w[r] * b.w[r];
  double o[1] = {acc};
  block_partials_b<1>(o, 1, red, b.partials, b.nsb, sb);
}

__global__ __launch_bounds__(kBlock) void kb_fin_av(BatchBufs b) {
  __shared__ double tmp[kWavesPerBlock];
  const int p = blockIdx.x;
  if (!av_active(b, p)) return;
  reduce_problem(b, p, 1, tmp);
  __syncthreads();
  if (threadIdx.x == 0) ps(b, p)[kBAvNorm] = sqrt(tmp[0]);
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

// One workgroup: the row blocks and streaming blocks of the problems that
// build a basis in this cycle (active and needing one), in storage order,
// and the per-column counts (problems are in order of decreasing m, so the
// blocks of column j are a prefix of the lists).  Problems are taken 1024 at
// a time: a scan of their block counts gives each one's offset, then one
// wave per problem writes its block indices.
__global__ __launch_bounds__(1024) void kb_compact(BatchBufs b) {
  constexpr int NT = 1024, NW = NT / kWave;
  __shared__ int s_nr[NT], s_ns[NT], s_or[NT], s_os[NT];
  __shared__ int wsum[2][NW];
  __shared__ int carry[2];
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  if (tid == 0) {
    carry[0] = 0;
    carry[1] = 0;
  }
  __syncthreads();
  for (int base = 0; base < b.P; base += NT) {
    const int p = base + tid;
    int nr = 0, ns = 0;
    if (p < b.P) {
      const int* it = pi(b, p);
      if (it[kBActive] && it[kBNeedBasis]) {
        nr = b.p_rb0[p + 1] - b.p_rb0[p];
        ns = b.p_sb0[p + 1] - b.p_sb0[p];
      }
    }
    int sr = nr, ss = ns;  // inclusive scan within the wave
    for (int off = 1; off < kWave; off <<= 1) {
      const int ur = __shfl_up(sr, off), us = __shfl_up(ss, off);
      if (lane >= off) {
        sr += ur;
        ss += us;
      }
    }
    if (lane == kWave - 1) {
      wsum[0][wave] = sr;
      wsum[1][wave] = ss;
    }
    __syncthreads();
    int pr = carry[0], pq = carry[1];
    for (int w = 0; w < wave; ++w) {
      pr += wsum[0][w];
      pq += wsum[1][w];
    }
    s_nr[tid] = nr;
    s_ns[tid] = ns;
    s_or[tid] = pr + sr - nr;
    s_os[tid] = pq + ss - ns;
    __syncthreads();
    if (tid < b.M) {
      const int q = b.np_col[tid] - base;
      if (b.np_col[tid] < b.P && q >= 0 && q < NT) {
        b.list_cnt[tid] = s_or[q];
        b.list_cnt[b.M + 1 + tid] = s_os[q];
      }
    }
    const int pend = min(NT, b.P - base);
    for (int q = wave; q < pend; q += NW) {
      const int r0 = b.p_rb0[base + q], s0 = b.p_sb0[base + q];
      for (int k = lane; k < s_nr[q]; k += kWave) b.rb_list[s_or[q] + k] = r0 + k;
      for (int k = lane; k < s_ns[q]; k += kWave) b.sb_list[s_os[q] + k] = s0 + k;
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 0; w < NW; ++w) {
        carry[0] += wsum[0][w];
        carry[1] += wsum[1][w];
      }
    }
    __syncthreads();
  }
  if (tid < b.M && b.np_col[tid] >= b.P) {
    b.list_cnt[tid] = carry[0];
    b.list_cnt[b.M + 1 + tid] = carry[1];
  }
  if (tid == 0) {
    b.list_cnt[b.M] = carry[0];
    b.list_cnt[2 * b.M + 1] = carry[1];
  }
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
  const int expokit = it[kBCtrl] >= 1;
  const int ldh = b.M + 2;
  const double* H = b.H + static_cast<long long>(p) * ldh * ldh;
  double* wH = b.wH + static_cast<long long>(p) * (b.M + 4);
  double* s = ps(b, p);
  // Expokit after a breakdown: the step covers the rest of [0, t]
  const double dt = (expokit && brk) ? s[kBT] - s[kBTDone] : s[kBDt];
  if (k <= 0) {
    for (int i = threadIdx.x; i < b.M + 4; i += kBlock) wH[i] = 0.0;
    if (threadIdx.x == 0) s[kBErr] = 0.0;
    return;
  }
  // Expokit without breakdown: augmented (k+2) matrix with H(k+1, k) = 1
  const int kk = (expokit && !brk) ? k + 2 : k;
  const int ord = batch_expm_order(b);
  double* base = BIG ? b.expm_ws + static_cast<long long>(p) * 6 * ord * ord : smem;
  double* S[6];
  for (int i = 0; i < 6; ++i) S[i] = base + static_cast<long long>(i) * kk * kk;
  for (int e = threadIdx.x; e < kk * kk; e += kBlock) {
    const int r = e % kk, c = e / kk;
    double v = 0.0;
    if (c < k && (r < k || (kk > k && r == k))) v = H[static_cast<long long>(c) * ldh + r];
    if (kk > k && r == k + 1 && c == k) v = 1.0;
    S[0][e] = dt * v;
  }
  __syncthreads();
  double* E = expm_slots<BIG>(S, kk, s_b, s_int, s_red);
  for (int i = threadIdx.x; i < b.M + 4; i += kBlock) wH[i] = (i < kk) ? E[i] : 0.0;
  if (threadIdx.x == 0 && expokit && brk) s[kBDt] = dt;
  if (threadIdx.x == 0) {
    const double hk = H[static_cast<long long>(k - 1) * ldh + k];
    s[kBErr] = brk ? 0.0 : fabs(hk * E[k - 1]) * s[kBVNorm];
  }
}

// y = sum_{i < KLift, |wH_i| > 1e-15} beta wH_i q_i (main.cu:1992-1999) for
// the problems the controller marked; partials of ||y||^2.
template <int MAXK>
__global__ __launch_bounds__(kBlock) void kb_lift(BatchBufs b) {
  __shared__ double coef[kSweepMaxCols];
  __shared__ int used[kSweepMaxCols];
  __shared__ int nused;
  const int sb = blockIdx.x;
  const int p = b.sb_prob[sb];
  const int* it = pi(b, p);
  if (!it[kBLift]) return;
  if (threadIdx.x == 0) {
    const int kl = it[kBKLift];
    const double beta = ps(b, p)[kBVNorm];
    const double* cs = pcs(b, p);
    const double* wh = b.wH + static_cast<long long>(p) * (b.M + 4);
    int u = 0;
    for (int i = 0; i < kl && i < kSweepMaxCols; ++i) {
      const double c = wh[i];
      if (fabs(c) > 1e-15) {
        used[u] = i;
        coef[u] = (beta * c) * cs[i];
        ++u;
      }
    }
    nused = u;
  }
  __syncthreads();
  const int nu = nused;
  if (nu == 0) {
    for (int r = b.sb_row0[sb] + threadIdx.x; r < b.sb_row0[sb + 1]; r += kBlock) b.y[r] = 0.0;
    if (threadIdx.x == 0) b.partials[sb] = 0.0;
    return;
  }
  sweep_lift<MAXK, 2>(b.V, b.ld, nu, used, coef, b.y, sb_tiles(b, sb), b.partials, sb);
}

__global__ __launch_bounds__(kBlock) void kb_fin_lift(BatchBufs b) {
  __shared__ double tmp[kWavesPerBlock];
  const int p = blockIdx.x;
  if (!pi(b, p)[kBLift]) return;
  reduce_problem(b, p, 1, tmp);
  __syncthreads();
  if (threadIdx.x == 0) ps(b, p)[kBYNorm] = sqrt(tmp[0]);
}

// The reference controller's accept / reject step (main.cu:1476-1516).  The
// reference lifts y on every attempt, so every attempted problem is lifted.
__device__ void control_reference(const BatchBufs& b, int p) {
  double* s = ps(b, p);
  int* it = pi(b, p);
  const double err = s[kBErr];
  const double tol = s[kBTol];
  it[kBAttempts] += 1;
  it[kBLift] = 1;
  it[kBKLift] = it[kBKUsed] + 1;
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
    it[kBRejTotal] += 1;
    if (it[kBRejects] >= 20) {
      it[kBAborted] = 1;
      it[kBActive] = 0;
      s[kBLastRes] = err;
      s[kBTDone] = s[kBT];
    } else {
      s[kBDt] = fmax(s[kBDt] * 0.5 * 0.8, s[kBDtMin]);
    }
  }
}

// Expokit step control (Sidje 1998, expv): local error from the augmented
// exponential, err_loc <= delta * t_step * tol to accept, the next step from
// gamma * t_step * (t_step * tol / err_loc)^(1/m) rounded to two digits; a
// rejected step keeps its basis and only redoes the exponential.  The lift
// (m + 1 basis vectors without breakdown) runs only for accepted steps.
__device__ void control_expokit(const BatchBufs& b, int p) {
  double* s = ps(b, p);
  int* it = pi(b, p);
  const int m = it[kBM];
  const int brk = it[kBBreakdown];
  const double beta = s[kBVNorm];
  const double* F = b.wH + static_cast<long long>(p) * (b.M + 4);
  // relative mode (2): local error against the smaller of ||w|| at the
  // step's start and at its end, ||w_end|| = beta ||F(0:m)|| (orthonormal basis)
  double tol = s[kBTol];
  if (it[kBCtrl] == 2) {
    double f2 = 0.0;
    const int nf = brk ? it[kBKUsed] : m + 1;
    for (int i = 0; i < nf; ++i) f2 += F[i] * F[i];
    tol *= beta * fmin(1.0, sqrt(f2));
  }
  const double t_step = s[kBDt];
  it[kBAttempts] += 1;
  double err_loc, xm = 1.0 / m;
  bool accept;
  if (brk) {
    err_loc = kExpokitBtol;
    accept = true;
  } else {
    const double phi1 = fabs(beta * F[m]);
    const double phi2 = fabs(beta * F[m + 1] * s[kBAvNorm]);
    if (phi1 > 10.0 * phi2) {
      err_loc = phi2;
    } else if (phi1 > phi2) {
      err_loc = (phi1 * phi2) / (phi1 - phi2);
    } else {
      err_loc = phi1;
      xm = 1.0 / (m - 1);
    }
    accept = err_loc <= kExpokitDelta * t_step * tol;
  }
  if (!accept) {
    it[kBRejTotal] += 1;
    if (it[kBIReject] == kExpokitMxrej) {
      // the requested tolerance is too high: stop with this attempt's y
      it[kBAborted] = 1;
      it[kBActive] = 0;
      it[kBLift] = 1;
      it[kBKLift] = m + 1;
      s[kBLastRes] = err_loc;
      return;
    }
    double t_next = round_step(kExpokitGamma * t_step * pow(t_step * tol / err_loc, xm));
    // a predicted step that fails falls back to Expokit's own (safe) first step
    if (it[kBSubsteps] > 0 && !(t_next <= s[kBDtMin])) t_next = s[kBDtMin];
    s[kBDt] = t_next;
    it[kBIReject] += 1;
    return;
  }
  it[kBLift] = 1;
  it[kBKLift] = brk ? it[kBKUsed] : m + 1;
  s[kBTDone] += t_step;
  it[kBAccepted] += 1;
  s[kBTNew] = round_step(kExpokitGamma * t_step * pow(t_step * tol / err_loc, xm));
  const double rndoff = s[kBAnorm] * 2.220446049250313e-16;
  err_loc = fmax(err_loc, rndoff);
  s[kBSErr] += err_loc;
  s[kBLastRes] = err_loc;
  if (s[kBTDone] >= s[kBT]) {
    it[kBConverged] = 1;
    it[kBActive] = 0;
  } else if (it[kBAccepted] >= it[kBMaxSteps]) {
    it[kBActive] = 0;
  } else {
    it[kBRestart] = 1;
  }
}

// One thread per problem; writes the number of still-active problems.
__global__ __launch_bounds__(1024) void kb_control(BatchBufs b) {
  __shared__ int cnt[1024 / kWave];
  __shared__ int mx[1024 / kWave];
  int active = 0, mmax = 0;
  for (int p = threadIdx.x; p < b.P; p += blockDim.x) {
    int* it = pi(b, p);
    if (!it[kBActive]) continue;
    if (it[kBCtrl] >= 1) control_expokit(b, p);
    else control_reference(b, p);
    active += it[kBActive];
    if (it[kBActive]) mmax = max(mmax, it[kBM]);
  }
  for (int off = kWave / 2; off >= 1; off >>= 1) mmax = max(mmax, __shfl_xor(mmax, off));
  if (threadIdx.x % kWave == 0) mx[threadIdx.x / kWave] = mmax;
  // count of active problems (fixed-order reduction, no atomics)
  for (int off = kWave / 2; off >= 1; off >>= 1) active += __shfl_xor(active, off);
  if (threadIdx.x % kWave == 0) cnt[threadIdx.x / kWave] = active;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) tot += cnt[w];
    b.counters[0] = tot;
    b.counters[1] += 1;  // cycles
    int m_act = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) m_act = max(m_act, mx[w]);
    b.counters[2] = m_act;
  }
}

// ------------------------------------------------------------ launchers

void batch_prepare() {
  // the in-LDS exponential of the largest problems needs 154 KB of LDS (set
  // on the current device before every solve)
  MEK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&kb_expm<false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(5 * kMaxKLds * kMaxKLds * sizeof(double))));
}

void batch_launch_init(const BatchBufs& b, hipStream_t s) {
  hipLaunchKernelGGL(kb_norm0, dim3(b.nsb), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(kb_init, dim3(b.P), dim3(kBlock), 0, s, b);
}

void batch_launch_spmv(const BatchBufs& b, int j, int nrb, int grid, hipStream_t s) {
  const int* none = nullptr;
  if (b.vals_f32) hipLaunchKernelGGL(kb_spmv<float>, dim3(grid), dim3(kBlock), 0, s, b, j, 0, nrb, none, none);
  else hipLaunchKernelGGL(kb_spmv<double>, dim3(grid), dim3(kBlock), 0, s, b, j, 0, nrb, none, none);
}

void batch_launch_cycle(const BatchBufs& b, const BatchPlan& plan, int mcols, int gs_mode, double eta,
                        hipStream_t s) {
  const int pg = (b.P + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(kb_restart_copy, dim3(b.nsb), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(kb_cycle_begin, dim3(pg), dim3(kBlock), 0, s, b);
  hipLaunchKernelGGL(kb_compact, dim3(1), dim3(1024), 0, s, b);
  for (int j = 0; j < mcols; ++j) {
    // problems are ordered by decreasing m: column j concerns a prefix of
    // them; the kernels walk this cycle's compacted blocks of that prefix
    const int np = plan.np[j];
    const int nsb = std::min(plan.nsb[j], kBatchSweepGrid);
    const int sg = plan.spmv_grid[j];
    const int* rl = b.rb_list;
    const int* rc = b.list_cnt + j;
    if (b.vals_f32) hipLaunchKernelGGL(kb_spmv<float>, dim3(sg), dim3(kBlock), 0, s, b, j, 0, 0, rl, rc);
    else hipLaunchKernelGGL(kb_spmv<double>, dim3(sg), dim3(kBlock), 0, s, b, j, 0, 0, rl, rc);
#define MEK_L(K) hipLaunchKernelGGL(kb_dots1<K>, dim3(nsb), dim3(kBlock), 0, s, b, j)
    MEK_SWEEP_BUCKETS(j + 1, MEK_L);
#undef MEK_L
    hipLaunchKernelGGL(kb_fin_h1, dim3(np), dim3(kBlock), 0, s, b, j);
#define MEK_L(K) hipLaunchKernelGGL(kb_update_dot<K>, dim3(nsb), dim3(kBlock), 0, s, b, j)
    MEK_SWEEP_BUCKETS(j + 1, MEK_L);
#undef MEK_L
    hipLaunchKernelGGL(kb_fin_h2, dim3(np), dim3(kBlock), 0, s, b, j, gs_mode, eta);
    if (gs_mode != 0) {
#define MEK_L(K) hipLaunchKernelGGL(kb_update_norm<K>, dim3(nsb), dim3(kBlock), 0, s, b, j)
      MEK_SWEEP_BUCKETS(j + 1, MEK_L);
#undef MEK_L
      hipLaunchKernelGGL(kb_fin_col, dim3(np), dim3(kBlock), 0, s, b, j);
    }
  }
  if (b.any_expokit) {
    // ||A v_{m+1}|| of every Expokit problem that built a full basis
    const int sg = plan.spmv_grid[0];
    const int* rl = b.rb_list;
    const int* rc = b.list_cnt + b.M;
    if (b.vals_f32) hipLaunchKernelGGL(kb_spmv<float>, dim3(sg), dim3(kBlock), 0, s, b, 0, 1, 0, rl, rc);
    else hipLaunchKernelGGL(kb_spmv<double>, dim3(sg), dim3(kBlock), 0, s, b, 0, 1, 0, rl, rc);
    hipLaunchKernelGGL(kb_av_norm, dim3(b.nsb), dim3(kBlock), 0, s, b);
    hipLaunchKernelGGL(kb_fin_av, dim3(b.P), dim3(kBlock), 0, s, b);
  }
  hipLaunchKernelGGL(kb_after_basis, dim3(pg), dim3(kBlock), 0, s, b);
  const int kk = batch_expm_order(b);
  if (kk <= kMaxKLds) {
    hipLaunchKernelGGL(kb_expm<false>, dim3(b.P), dim3(kBlock), 5 * kk * kk * sizeof(double), s, b);
  } else {
    hipLaunchKernelGGL(kb_expm<true>, dim3(b.P), dim3(kBlock), 0, s, b);
  }
  hipLaunchKernelGGL(kb_control, dim3(1), dim3(1024), 0, s, b);
#define MEK_L(K) hipLaunchKernelGGL(kb_lift<K>, dim3(b.nsb), dim3(kBlock), 0, s, b)
  MEK_SWEEP_BUCKETS(mcols + 2, MEK_L);
#undef MEK_L
  hipLaunchKernelGGL(kb_fin_lift, dim3(b.P), dim3(kBlock), 0, s, b);
}

}  // namespace mek
