// Arnoldi column kernels for gfx950 (wave64).
//
// One Arnoldi column j (reference: main.cu:1722-1942, spmv_on_off +
// orthogonalize_mgs + normalize_new_vector) is three streaming sweeps over
// the basis plus three one-workgroup finalisers:
//
//   K1 spmv_dot     w = A q_j  (row-blocked CSR, LDS-staged products) and the
//                   partial sums of h1 = V^T w and ||w||^2          (sweep 1)
//   F1              h1, ||w||^2
//   K2 update_dot   w' = w - V h1, stored unnormalised in V[:, j+1], and the
//                   partial sums of h2 = V^T w' and ||w'||^2        (sweep 2)
//   F2              h2, ||w'||^2, decision on the second pass
//   K3 update_norm  w'' = w' - V h2 in place, partial ||w''||^2     (sweep 3)
//   F3              H(:, j), breakdown test, colscale[j+1] = 1/||w''||
//
// This is classical Gram-Schmidt with a second pass (CGS2), which for the
// reference's unconditional re-orthogonalisation (main.cu:1829-1866) gives
// H(i,j) = h1_i + h2_i exactly as the reference's h_first + h_corr.  When the
// second pass is conditional (DGKS test ||w'|| < eta ||w||) K3/F3 exit at once
// on columns that do not need it.
//
// Every reduction is deterministic: per-thread partials, fixed-order
// in-block reduction (DPP), per-block partials reduced by one workgroup in
// block order.  After a breakdown every later kernel of the attempt returns
// immediately, so the host can enqueue all m columns without reading back.
#include "common.hpp"

namespace mek {

constexpr int kNnzCap = 2048;  // nnz of one short-row SpMV block (LDS products)
constexpr int kMaxQuantities = 512;  // reduced quantities per finaliser (m <= 500)

template <typename VT>
__device__ __forceinline__ double load_val(const void* vals, long long k) {
  return static_cast<double>(static_cast<const VT*>(vals)[k]);
}

// Block-reduce NQ per-thread values (uniform control flow) and write them as
// partials[q*grid + blk].  `red` is LDS scratch of kWavesPerBlock*NQ doubles.
template <int NQ>
__device__ __forceinline__ void block_partials(const double (&v)[NQ], int nq,
                                               double* red, double* partials,
                                               int grid, int blk) {
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
    partials[static_cast<long long>(q) * grid + blk] = s;
  }
}

// ------------------------------------------------------------------- K1

// w = A * (xs * x) with xs = *xs_ptr (1 when null); partial sums of
// q_i . w for i < ndots (quantity i) and of ||w||^2 (quantity normq).  Each
// workgroup owns a contiguous range of SpMV row blocks (XCD-aware), so x
// windows of neighbouring rows stay in one L2.  Short-row blocks stage the
// products val*x in LDS and sum each row in storage order, the same
// operations as a sequential CSR loop.
template <typename VT, int MAXJ>
__global__ __launch_bounds__(kBlock) void k_spmv_dot(ArnoldiBufs b, const double* x_in,
                                                     const double* xs_ptr, double* wout,
                                                     int ndots, int normq, int check_breakdown) {
  if (check_breakdown && b.iscal[kBreakdown]) return;
  __shared__ double prod[kNnzCap];
  __shared__ double red[kWavesPerBlock * (MAXJ + 1)];
  const int lb = xcd_block_id();
  const int tid = threadIdx.x;
  const double* __restrict__ x = x_in;
  const double xs = xs_ptr ? *xs_ptr : 1.0;
  const double* __restrict__ cs = b.colscale;
  const int* __restrict__ rp = b.row_ptr;
  const int* __restrict__ ci = b.col_idx;
  double* __restrict__ w = wout;

  double acc[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) acc[q] = 0.0;

  const long long nrb = b.n_row_blocks;
  const int rb0 = static_cast<int>(nrb * lb / b.grid);
  const int rb1 = static_cast<int>(nrb * (lb + 1) / b.grid);
  for (int rb = rb0; rb < rb1; ++rb) {
    const int r0 = b.row_blocks[rb];
    const int r1 = b.row_blocks[rb + 1];
    const int k0 = rp[r0];
    const int k1 = rp[r1];
    if (k1 - k0 <= kNnzCap) {
      for (int k = k0 + tid; k < k1; k += kBlock)
        prod[k - k0] = load_val<VT>(b.vals, k) * (xs * x[ci[k]]);
      __syncthreads();
      for (int r = r0 + tid; r < r1; r += kBlock) {
        const int e0 = rp[r] - k0, e1 = rp[r + 1] - k0;
        double a = 0.0;
        for (int e = e0; e < e1; ++e) a += prod[e];
        w[r] = a;
        acc[MAXJ] += a * a;
#pragma unroll
        for (int i = 0; i < MAXJ; ++i)
          if (i < ndots) acc[i] += (cs[i] * b.V[static_cast<long long>(i) * b.ld + r]) * a;
      }
      __syncthreads();
    } else {
      // one long row: the block strides through it and reduces
      double a = 0.0;
      for (int k = k0 + tid; k < k1; k += kBlock)
        a += load_val<VT>(b.vals, k) * (xs * x[ci[k]]);
      const double s = wave_sum(a);
      if (tid % kWave == 0) prod[tid / kWave] = s;
      __syncthreads();
      if (tid == 0) {
        double t = prod[0];
        for (int q = 1; q < kWavesPerBlock; ++q) t += prod[q];
        w[r0] = t;
        acc[MAXJ] += t * t;
#pragma unroll
        for (int i = 0; i < MAXJ; ++i)
          if (i < ndots) acc[i] += (cs[i] * b.V[static_cast<long long>(i) * b.ld + r0]) * t;
      }
      __syncthreads();
    }
  }
  block_partials<MAXJ + 1>(acc, ndots, red, b.partials, b.grid, lb);
  __syncthreads();
  double nv[1] = {acc[MAXJ]};
  block_partials<1>(nv, 1, red, b.partials + static_cast<long long>(normq) * b.grid, b.grid, lb);
}

// ------------------------------------------------------------------- K2

// w' = w - sum_{i<=j} h1_i q_i, written to V[:, j+1]; partials of q_i . w'
// (i <= j) and ||w'||^2 (index j+1).  The q_i of a row are loaded once and
// kept in registers for both uses.
template <int MAXJ>
__global__ __launch_bounds__(kBlock) void k_update_dot(ArnoldiBufs b, int j) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double red[kWavesPerBlock * (MAXJ + 1)];
  const int lb = xcd_block_id();
  const int nq = j + 1;
  const double* __restrict__ cs = b.colscale;
  const double* __restrict__ h = b.h1;
  const double* __restrict__ w = b.w;
  double* __restrict__ out = b.V + static_cast<long long>(j + 1) * b.ld;
  double acc[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) acc[q] = 0.0;

  const long long per = (static_cast<long long>(b.n) + b.grid - 1) / b.grid;
  const long long r0 = per * lb;
  const long long r1 = min(static_cast<long long>(b.n), r0 + per);
  for (long long r = r0 + threadIdx.x; r < r1; r += kBlock) {
    double q[MAXJ];
    double a = w[r];
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
  block_partials<MAXJ + 1>(o, nq + 1, red, b.partials, b.grid, lb);
}

// ------------------------------------------------------------------- K3

// w'' = w' - sum_{i<=j} h2_i q_i in place in V[:, j+1]; partial ||w''||^2.
__global__ __launch_bounds__(kBlock) void k_update_norm(ArnoldiBufs b, int j) {
  if (b.iscal[kBreakdown] || !b.iscal[kReorth]) return;
  __shared__ double red[kWavesPerBlock];
  const int lb = xcd_block_id();
  const int nq = j + 1;
  const double* __restrict__ cs = b.colscale;
  const double* __restrict__ h = b.h2;
  double* __restrict__ io = b.V + static_cast<long long>(j + 1) * b.ld;
  double acc = 0.0;
  const long long per = (static_cast<long long>(b.n) + b.grid - 1) / b.grid;
  const long long r0 = per * lb;
  const long long r1 = min(static_cast<long long>(b.n), r0 + per);
  for (long long r = r0 + threadIdx.x; r < r1; r += kBlock) {
    double a = io[r];
    for (int i = 0; i < nq; ++i) a -= h[i] * (cs[i] * b.V[static_cast<long long>(i) * b.ld + r]);
    io[r] = a;
    acc += a * a;
  }
  double o[1] = {acc};
  block_partials<1>(o, 1, red, b.partials, b.grid, lb);
}

// --------------------------------------------- generic kernels (j+1 > 64)

// y -= sum_{i<nq} h_i q_i  (in place on `io`, or io = src - ... if src != io)
__global__ __launch_bounds__(kBlock) void k_update(ArnoldiBufs b, const double* src,
                                                   double* dst, const double* h, int nq) {
  if (b.iscal[kBreakdown]) return;
  const double* __restrict__ cs = b.colscale;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long r = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; r < b.n; r += stride) {
    double a = src[r];
    for (int i = 0; i < nq; ++i) a -= h[i] * (cs[i] * b.V[static_cast<long long>(i) * b.ld + r]);
    dst[r] = a;
  }
}

// partials of q_i . x for i in [i0, i0+nq) (quantity index i - i0 + qoff) and,
// if with_norm, of ||x||^2 at quantity index qoff + nq.
template <int MAXJ>
__global__ __launch_bounds__(kBlock) void k_dots(ArnoldiBufs b, const double* xv, int i0,
                                                 int nq, int with_norm, int qoff) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double red[kWavesPerBlock * (MAXJ + 1)];
  const int lb = xcd_block_id();
  const double* __restrict__ cs = b.colscale;
  double acc[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) acc[q] = 0.0;
  const long long per = (static_cast<long long>(b.n) + b.grid - 1) / b.grid;
  const long long r0 = per * lb;
  const long long r1 = min(static_cast<long long>(b.n), r0 + per);
  for (long long r = r0 + threadIdx.x; r < r1; r += kBlock) {
    const double a = xv[r];
#pragma unroll
    for (int i = 0; i < MAXJ; ++i)
      if (i < nq) acc[i] += (cs[i0 + i] * b.V[static_cast<long long>(i0 + i) * b.ld + r]) * a;
    acc[MAXJ] += a * a;
  }
  const int ntot = nq + (with_norm ? 1 : 0);
  double o[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) o[q] = (q < nq) ? acc[q] : ((q == nq) ? acc[MAXJ] : 0.0);
  block_partials<MAXJ + 1>(o, ntot, red, b.partials + static_cast<long long>(qoff) * b.grid,
                           b.grid, lb);
}

// --------------------------------------------------------------- finalise

// Sum partials[q*grid + 0..grid) for q < nq into dst[q], in block order per
// lane and a fixed-order wave sum.  Runs as one workgroup.
__device__ void reduce_partials(const double* partials, int grid, int nq, double* dst) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  for (int q = wave; q < nq; q += kWavesPerBlock) {
    double s = 0.0;
    for (int g = lane; g < grid; g += kWave) s += partials[static_cast<long long>(q) * grid + g];
    s = wave_sum(s);
    if (lane == 0) dst[q] = s;
  }
}

// Finaliser of K1 (+ generic dots): h1[0..nq), scal[kWNorm2].
__global__ __launch_bounds__(kBlock) void k_fin_h1(ArnoldiBufs b, int nq) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double tmp[kMaxQuantities];
  reduce_partials(b.partials, b.grid, nq + 1, tmp);
  __syncthreads();
  for (int i = threadIdx.x; i < nq; i += kBlock) b.h1[i] = tmp[i];
  if (threadIdx.x == 0) b.scal[kWNorm2] = tmp[nq];
}

// Writes column j of H (zeros below the subdiagonal), the breakdown flag, the
// next column scale and k_used.  reference: normalize_new_vector
// (main.cu:1878-1942), breakdown when the norm is < 1e-14 or not finite.
__device__ void finish_column(ArnoldiBufs& b, int j, double norm2, int reorth) {
  if (threadIdx.x != 0) return;
  const int ldh = b.ldh;
  for (int i = 0; i <= j; ++i) b.H[static_cast<long long>(j) * ldh + i] = b.h1[i] + (reorth ? b.h2[i] : 0.0);
  for (int i = j + 2; i < ldh; ++i) b.H[static_cast<long long>(j) * ldh + i] = 0.0;
  const double nrm = sqrt(norm2);
  b.iscal[kKUsed] = j + 1;
  if (nrm < 1e-14 || isnan(nrm) || isinf(nrm)) {
    b.H[static_cast<long long>(j) * ldh + j + 1] = 0.0;
    b.iscal[kBreakdown] = 1;
  } else {
    b.H[static_cast<long long>(j) * ldh + j + 1] = nrm;
    b.colscale[j + 1] = 1.0 / nrm;
  }
  if (reorth) b.iscal[kReorthCount] += 1;
}

// Finaliser of K2: h2, ||w'||^2 and the second-pass decision.
// gs_mode 0: never re-orthogonalise, 1: always (reference), 2: DGKS with eta.
__global__ __launch_bounds__(kBlock) void k_fin_h2(ArnoldiBufs b, int j, int gs_mode, double eta) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double tmp[kMaxQuantities];
  const int nq = j + 1;
  reduce_partials(b.partials, b.grid, nq + 1, tmp);
  __syncthreads();
  for (int i = threadIdx.x; i < nq; i += kBlock) b.h2[i] = tmp[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    const double wp2 = tmp[nq];
    b.scal[kWpNorm2] = wp2;
    int reorth = 1;
    if (gs_mode == 0) reorth = 0;
    else if (gs_mode == 2) reorth = (sqrt(wp2) < eta * sqrt(b.scal[kWNorm2])) ? 1 : 0;
    if (!(wp2 == wp2)) reorth = 0;  // NaN: the column breaks down either way
    b.iscal[kReorth] = reorth;
  }
  __syncthreads();
  if (!b.iscal[kReorth]) finish_column(b, j, b.scal[kWpNorm2], 0);
}

// Finaliser of K3.
__global__ __launch_bounds__(kBlock) void k_fin_col(ArnoldiBufs b, int j) {
  if (b.iscal[kBreakdown] || !b.iscal[kReorth]) return;
  __shared__ double tmp[kWavesPerBlock];
  reduce_partials(b.partials, b.grid, 1, tmp);
  __syncthreads();
  if (threadIdx.x == 0) b.scal[kWppNorm2] = tmp[0];
  finish_column(b, j, tmp[0], 1);
}

// ---------------------------------------------------- start vector / lift

// partials of ||x||^2 over x[0..n)
__global__ __launch_bounds__(kBlock) void k_norm2(ArnoldiBufs b, const double* xv) {
  __shared__ double red[kWavesPerBlock];
  const int lb = xcd_block_id();
  const long long per = (static_cast<long long>(b.n) + b.grid - 1) / b.grid;
  const long long r0 = per * lb;
  const long long r1 = min(static_cast<long long>(b.n), r0 + per);
  double acc = 0.0;
  for (long long r = r0 + threadIdx.x; r < r1; r += kBlock) acc += xv[r] * xv[r];
  double o[1] = {acc};
  block_partials<1>(o, 1, red, b.partials, b.grid, lb);
}

// Start of a solve or of a new sub-step: beta = ||V[:,0]|| and
// colscale[0] = 1/beta, guarded as on the reference's CPU path
// (main_cpu.cpp:473-484, 589-600: scale 1 when beta <= 1e-14).
__global__ __launch_bounds__(kBlock) void k_fin_start(ArnoldiBufs b) {
  __shared__ double tmp[kWavesPerBlock];
  reduce_partials(b.partials, b.grid, 1, tmp);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double beta = sqrt(tmp[0]);
    b.scal[kVNorm] = beta;
    b.colscale[0] = (beta > 1e-14) ? (1.0 / beta) : 1.0;
    b.iscal[kKUsed] = 0;
    b.iscal[kBreakdown] = 0;
  }
}

// y = sum_{i<ncoef, used} coef_i q_i where coef_i = beta * wH[i]; terms with
// |wH[i]| <= skip_below are left out (reference: 1e-15, main.cu:1992-1994).
// Partials of ||y||^2.
__global__ __launch_bounds__(kBlock) void k_lift(ArnoldiBufs b, int ncoef, double skip_below) {
  __shared__ double red[kWavesPerBlock];
  const int lb = xcd_block_id();
  const double beta = b.scal[kVNorm];
  const double* __restrict__ cs = b.colscale;
  const double* __restrict__ wh = b.wH;
  double* __restrict__ y = b.y;
  const long long per = (static_cast<long long>(b.n) + b.grid - 1) / b.grid;
  const long long r0 = per * lb;
  const long long r1 = min(static_cast<long long>(b.n), r0 + per);
  double acc = 0.0;
  for (long long r = r0 + threadIdx.x; r < r1; r += kBlock) {
    double a = 0.0;
    for (int i = 0; i < ncoef; ++i) {
      const double c = wh[i];
      if (fabs(c) > skip_below) a += (beta * c) * (cs[i] * b.V[static_cast<long long>(i) * b.ld + r]);
    }
    y[r] = a;
    acc += a * a;
  }
  double o[1] = {acc};
  block_partials<1>(o, 1, red, b.partials, b.grid, lb);
}

// scal[slot] = sqrt(sum of the quantity-0 partials)
__global__ __launch_bounds__(kBlock) void k_fin_norm(ArnoldiBufs b, int slot) {
  __shared__ double tmp[kWavesPerBlock];
  reduce_partials(b.partials, b.grid, 1, tmp);
  __syncthreads();
  if (threadIdx.x == 0) b.scal[slot] = sqrt(tmp[0]);
}

// Next sub-step starts from y: V[:,0] = y, beta = ||y||, colscale[0] guarded.
__global__ __launch_bounds__(kBlock) void k_restart(ArnoldiBufs b) {
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long r = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; r < b.n; r += stride)
    b.V[r] = b.y[r];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double beta = b.scal[kYNorm];
    b.scal[kVNorm] = beta;
    b.colscale[0] = (beta > 1e-14) ? (1.0 / beta) : 1.0;
    b.iscal[kKUsed] = 0;
    b.iscal[kBreakdown] = 0;
  }
}

// Host-side launchers (C++ linkage; the C ABI wraps them).
template <int MAXJ>
static void launch_spmv_dot_t(const ArnoldiBufs& b, const double* x, const double* xs, double* w,
                              int ndots, int normq, int chk, hipStream_t s) {
  if (b.vals_f32)
    hipLaunchKernelGGL((k_spmv_dot<float, MAXJ>), dim3(b.grid), dim3(kBlock), 0, s, b, x, xs, w,
                       ndots, normq, chk);
  else
    hipLaunchKernelGGL((k_spmv_dot<double, MAXJ>), dim3(b.grid), dim3(kBlock), 0, s, b, x, xs, w,
                       ndots, normq, chk);
}

// Arnoldi column j: w = A q_j with dots against q_0..q_{ndots-1}; the norm
// partial goes to quantity index normq.
void launch_spmv_dot(const ArnoldiBufs& b, int j, int ndots, int normq, hipStream_t s) {
  const double* x = b.V + static_cast<long long>(j) * b.ld;
  const double* xs = b.colscale + j;
  if (ndots <= 0) launch_spmv_dot_t<1>(b, x, xs, b.w, 0, normq, 1, s);
  else if (ndots <= 8) launch_spmv_dot_t<8>(b, x, xs, b.w, ndots, normq, 1, s);
  else if (ndots <= 16) launch_spmv_dot_t<16>(b, x, xs, b.w, ndots, normq, 1, s);
  else if (ndots <= 32) launch_spmv_dot_t<32>(b, x, xs, b.w, ndots, normq, 1, s);
  else if (ndots <= 48) launch_spmv_dot_t<48>(b, x, xs, b.w, ndots, normq, 1, s);
  else launch_spmv_dot_t<64>(b, x, xs, b.w, ndots < 64 ? ndots : 64, normq, 1, s);
}

// Plain y = A*(xs*x) (xs may be null) with the ||y||^2 partial at quantity 0.
void launch_spmv_plain(const ArnoldiBufs& b, const double* x, const double* xs, double* y,
                       hipStream_t s) {
  launch_spmv_dot_t<1>(b, x, xs, y, 0, 0, 0, s);
}

void launch_update_dot(const ArnoldiBufs& b, int j, hipStream_t s) {
  const int nq = j + 1;
  if (nq <= 8) hipLaunchKernelGGL(k_update_dot<8>, dim3(b.grid), dim3(kBlock), 0, s, b, j);
  else if (nq <= 16) hipLaunchKernelGGL(k_update_dot<16>, dim3(b.grid), dim3(kBlock), 0, s, b, j);
  else if (nq <= 32) hipLaunchKernelGGL(k_update_dot<32>, dim3(b.grid), dim3(kBlock), 0, s, b, j);
  else if (nq <= 48) hipLaunchKernelGGL(k_update_dot<48>, dim3(b.grid), dim3(kBlock), 0, s, b, j);
  else hipLaunchKernelGGL(k_update_dot<64>, dim3(b.grid), dim3(kBlock), 0, s, b, j);
}

void launch_dots(const ArnoldiBufs& b, const double* x, int i0, int nq, int with_norm, int qoff,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_dots<64>, dim3(b.grid), dim3(kBlock), 0, s, b, x, i0, nq, with_norm, qoff);
}

void launch_update(const ArnoldiBufs& b, const double* src, double* dst, const double* h, int nq,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_update, dim3(b.grid), dim3(kBlock), 0, s, b, src, dst, h, nq);
}

void launch_update_norm(const ArnoldiBufs& b, int j, hipStream_t s) {
  hipLaunchKernelGGL(k_update_norm, dim3(b.grid), dim3(kBlock), 0, s, b, j);
}

void launch_fin_h1(const ArnoldiBufs& b, int nq, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_h1, dim3(1), dim3(kBlock), 0, s, b, nq);
}
void launch_fin_h2(const ArnoldiBufs& b, int j, int gs_mode, double eta, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_h2, dim3(1), dim3(kBlock), 0, s, b, j, gs_mode, eta);
}
void launch_fin_col(const ArnoldiBufs& b, int j, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_col, dim3(1), dim3(kBlock), 0, s, b, j);
}
void launch_norm2(const ArnoldiBufs& b, const double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_norm2, dim3(b.grid), dim3(kBlock), 0, s, b, x);
}
void launch_fin_start(const ArnoldiBufs& b, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_start, dim3(1), dim3(kBlock), 0, s, b);
}
void launch_lift(const ArnoldiBufs& b, int ncoef, double skip_below, hipStream_t s) {
  hipLaunchKernelGGL(k_lift, dim3(b.grid), dim3(kBlock), 0, s, b, ncoef, skip_below);
}
void launch_fin_norm(const ArnoldiBufs& b, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(kBlock), 0, s, b, slot);
}
void launch_restart(const ArnoldiBufs& b, hipStream_t s) {
  hipLaunchKernelGGL(k_restart, dim3(b.grid), dim3(kBlock), 0, s, b);
}

}  // namespace mek
