// Arnoldi column kernels for gfx950 (wave64).
//
// One Arnoldi column j (reference: main.cu:1722-1942, spmv_on_off +
// orthogonalize_mgs + normalize_new_vector) is a SpMV and three streaming
// sweeps over the basis, each followed by a finaliser that reduces the
// per-block partial sums (one workgroup per reduced quantity):
//
//   K1 spmv(+dot)   w = A q_j (row-blocked CSR, LDS-staged products); partial
//                   sums of ||w||^2 and, when fused, of h1 = V^T w
//   K1b dots        h1 = V^T w partial sums (when not fused)     (sweep 1)
//   F1              h1, ||w||^2
//   K2 update_dot   w' = w - V h1, stored unnormalised in V[:, j+1], and the
//                   partial sums of h2 = V^T w' and ||w'||^2      (sweep 2)
//   F2              h2, ||w'||^2
//   K3 update_norm  second pass w'' = w' - V h2 in place, ||w''||^2 (sweep 3)
//   F3              H(:, j), breakdown test, colscale[j+1] = 1/||w''||
//
// This is classical Gram-Schmidt with a second pass (CGS2), which for the
// reference's unconditional re-orthogonalisation (main.cu:1829-1866) gives
// H(i,j) = h1_i + h2_i exactly as the reference's h_first + h_corr.  When the
// second pass is conditional (DGKS test ||w'|| < eta ||w||), K3 and F3 decide
// from ||w||, ||w'|| and skip the sweep on columns that do not need it.
//
// Every reduction is deterministic: per-thread partials over a fixed row set,
// a fixed-order in-block reduction (DPP), and per-block partials summed in
// block order.  After a breakdown every later kernel of the attempt returns at
// once, so the host can enqueue all m columns without reading back.
#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "launch.hpp"
#include "sweeps.hpp"

namespace mek {

constexpr int kNnzCap = 4096;  // nnz of one short-row SpMV block (LDS products)

template <typename VT>
__device__ __forceinline__ double load_val(const void* vals, long long k) {
  return static_cast<double>(static_cast<const VT*>(vals)[k]);
}

// The matrix is streamed once per SpMV: values and column indices are read
// with non-temporal loads so that they do not displace x from the caches.
template <typename VT>
__device__ __forceinline__ double load_val_nt(const void* vals, long long k) {
  return static_cast<double>(__builtin_nontemporal_load(static_cast<const VT*>(vals) + k));
}
__device__ __forceinline__ int load_idx_nt(const int* ci, long long k) { return __builtin_nontemporal_load(ci + k); }

// Block-reduce NQ per-thread values (uniform control flow) and write them as
// partials[q*grid + blk].  `red` is LDS scratch of kWavesPerBlock*NQ doubles.
template <int NQ>
__device__ __forceinline__ void block_partials(const double (&v)[NQ], int nq, double* red,
                                               double* partials, int grid, int blk) {
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

// Sum of partials[q*grid + 0..grid) by one workgroup: thread t sums blocks
// t, t+256, ... in order, then a fixed-order block reduction.  Uniform flow.
__device__ double block_sum_quantity(const double* partials, int grid, int q, double* red) {
  double s = 0.0;
  const double* p = partials + static_cast<long long>(q) * grid;
  for (int g = threadIdx.x; g < grid; g += kBlock) s += p[g];
  s = wave_sum(s);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  const double tot = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return tot;
}

__device__ __forceinline__ long long gtid() {
  return static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
}
__device__ __forceinline__ long long gstride() { return static_cast<long long>(gridDim.x) * kBlock; }

// Copy n coefficients into LDS (uniform per kernel) so the column loops read
// them by broadcast instead of issuing per-iteration global loads.
__device__ __forceinline__ void stage(double* dst, const double* src, int n) {
  for (int i = threadIdx.x; i < n; i += kBlock) dst[i] = src[i];
}

// The second-pass decision, from ||w||^2 and ||w'||^2 of the column.
// gs_mode 0: never, 1: always (reference), 2: DGKS ||w'|| < eta ||w||.
__device__ __forceinline__ int reorth_needed(const ArnoldiBufs& b, int gs_mode, double eta) {
  const double wp2 = b.scal[kWpNorm2];
  if (!(wp2 == wp2)) return 0;  // NaN: the column breaks down either way
  if (gs_mode == 0) return 0;
  if (gs_mode == 1) return 1;
  return (sqrt(wp2) < eta * sqrt(b.scal[kWNorm2])) ? 1 : 0;
}

constexpr int kMaxStaged = 512;  // coefficients staged in LDS (m <= 500)
constexpr int kChunk = 16;       // basis columns loaded back-to-back per row

// sum_{i<nq} c_i * (s_i * V[i*ld + r]), loading kChunk columns at a time so
// their loads are in flight together; c and s are LDS-staged.
__device__ __forceinline__ double sweep_combination(const double* __restrict__ V, long long ld,
                                                    long long r, int nq, const double* c,
                                                    const double* s) {
  double sum = 0.0;
  for (int i0 = 0; i0 < nq; i0 += kChunk) {
    double v[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      v[k] = (i0 + k < nq) ? V[static_cast<long long>(i0 + k) * ld + r] : 0.0;
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      if (i0 + k < nq) sum += c[i0 + k] * (s[i0 + k] * v[k]);
  }
  return sum;
}

// ------------------------------------------------------------------- K1

// The Newton shift of a row sum, a - theta (xs x[r]), with a separately
// rounded product (no fused multiply-add), so that every SpMV form gives the
// same bits.
__device__ __forceinline__ double shift_row(double a, double theta, double xs, double xr) {
#pragma clang fp contract(off)
  const double p = theta * (xs * xr);
  return a - p;
}

// w = A * (xs * x) with xs = *xs_ptr (1 when null); partial sums of
// q_i . w for i < ndots (quantity i) and of ||w||^2 (quantity normq).  Each
// workgroup owns a contiguous range of SpMV row blocks (XCD-aware), so x
// windows of neighbouring rows stay in one L2.  Short-row blocks stage the
// products val*x in LDS and sum each row in storage order, the same
// operations as a sequential CSR loop.  All nonzero loads of a block are
// issued together (kNnzCap / kBlock per thread), and the next block's bounds
// are loaded while the current one is processed, so a block costs about two
// dependent memory latencies (nonzeros, then x).  w is written with
// non-temporal stores: its stream then interferes less with the reads
// (measured 7 % on a 5-point stencil).
//
// With shift_ptr set, w = A x' - theta x' (x' = xs x, theta = *shift_ptr):
// the s-step blocks' Newton recurrence; x[r] is read once per row, mostly
// from the lines its gathers brought in.  check_breakdown 1 returns at once
// after a breakdown, 2 also after a failed s-step block.
template <typename VT, int MAXJ>
__global__ __launch_bounds__(kBlock) void k_spmv_dot(ArnoldiBufs b, const double* x_in,
                                                     const double* xs_ptr, double* wout,
                                                     int ndots, int normq, int check_breakdown,
                                                     int i0, const double* shift_ptr) {
  if (check_breakdown && (b.iscal[kBreakdown] || (check_breakdown == 2 && b.iscal[kSStepFail]))) return;
  const bool has_shift = shift_ptr != nullptr;
  const double sh = has_shift ? *shift_ptr : 0.0;
  constexpr int U = kNnzCap / kBlock;
  __shared__ double prod[kNnzCap];
  __shared__ double red[kWavesPerBlock * (MAXJ + 1)];
  const int lb = xcd_block_id();
  const int tid = threadIdx.x;
  const int grid = gridDim.x;
  const double* __restrict__ x = x_in;
  const double xs = xs_ptr ? *xs_ptr : 1.0;
  // fully unrolled column loops below: the coefficients stay in scalar registers;
  // the dots are against columns i0, i0 + 1, ...
  const double* __restrict__ cs = b.colscale + i0;
  const double* __restrict__ Vd = b.V + static_cast<long long>(i0) * b.ld;
  const int* __restrict__ rp = b.row_ptr;
  const int* __restrict__ ci = b.col_idx;
  const int* __restrict__ rbk = b.row_blocks;
  double* __restrict__ w = wout;

  double acc[MAXJ + 1];
#pragma unroll
  for (int q = 0; q <= MAXJ; ++q) acc[q] = 0.0;

  const long long nrb = b.n_row_blocks;
  const int rb0 = static_cast<int>(nrb * lb / grid);
  const int rb1 = static_cast<int>(nrb * (lb + 1) / grid);
  // block rb spans rows [r0, r1) and nonzeros [k0, k1); r2 ends block rb + 1
  int r0 = 0, r1 = 0, k0 = 0, k1 = 0, r2 = 0;
  if (rb0 < rb1) {
    r0 = rbk[rb0];
    r1 = rbk[rb0 + 1];
    k0 = rp[r0];
    k1 = rp[r1];
    r2 = (rb0 + 1 < rb1) ? rbk[rb0 + 2] : r1;
  }
  for (int rb = rb0; rb < rb1; ++rb) {
    const int k2 = (rb + 1 < rb1) ? rp[r2] : k1;
    const int r3 = (rb + 2 < rb1) ? rbk[rb + 3] : r2;
    if (k1 - k0 <= kNnzCap) {
      int idx[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + tid + u * kBlock;
        idx[u] = 0;
        v[u] = 0.0;
        if (k < k1) {
          idx[u] = load_idx_nt(ci, k);
          v[u] = load_val_nt<VT>(b.vals, k);
        }
      }
      const int rr = r0 + tid;
      int e0 = 0, e1 = 0;
      // the shift's x[rr] is loaded with the block's nonzeros, so the row
      // phase waits on no further memory round trip
      double xrr = 0.0;
      if (rr < r1) {
        e0 = rp[rr] - k0;
        e1 = rp[rr + 1] - k0;
        if (has_shift) xrr = x[rr];
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
          if (has_shift) xrr = x[r];
        }
        double a = 0.0;
        for (int e = e0; e < e1; ++e) a += prod[e];
        if (has_shift) a = shift_row(a, sh, xs, xrr);
        __builtin_nontemporal_store(a, w + r);
        acc[MAXJ] += a * a;
#pragma unroll
        for (int i = 0; i < MAXJ; ++i)
          if (i < ndots) acc[i] += (cs[i] * Vd[static_cast<long long>(i) * b.ld + r]) * a;
      }
      __syncthreads();
    } else {
      // one long row: the block strides through it and reduces
      double a = 0.0;
      for (int k = k0 + tid; k < k1; k += kBlock)
        a += load_val_nt<VT>(b.vals, k) * (xs * x[load_idx_nt(ci, k)]);
      const double s = wave_sum(a);
      if (tid % kWave == 0) prod[tid / kWave] = s;
      __syncthreads();
      if (tid == 0) {
        double t = prod[0];
        for (int q = 1; q < kWavesPerBlock; ++q) t += prod[q];
        if (has_shift) t = shift_row(t, sh, xs, x[r0]);
        __builtin_nontemporal_store(t, w + r0);
        acc[MAXJ] += t * t;
#pragma unroll
        for (int i = 0; i < MAXJ; ++i)
          if (i < ndots) acc[i] += (cs[i] * Vd[static_cast<long long>(i) * b.ld + r0]) * t;
      }
      __syncthreads();
    }
    r0 = r1;
    k0 = k1;
    r1 = r2;
    k1 = k2;
    r2 = r3;
  }
  block_partials<MAXJ + 1>(acc, ndots, red, b.partials, grid, lb);
  __syncthreads();
  double nv[1] = {acc[MAXJ]};
  block_partials<1>(nv, 1, red, b.partials + static_cast<long long>(normq) * grid, grid, lb);
}

// K1 on the sliced-ELL copy (ArnoldiBufs::sell_*): the same operation and
// arguments as k_spmv_dot with at most one dot (ndots <= 1).  One wave per
// slice of 64 rows, one row per lane: each entry of the slice is one
// coalesced 512-byte value load and 256-byte column load for the wave, with
// no LDS and no barrier; the x gathers of neighbouring rows fall on the same
// lines.  A row is summed entry by entry in CSR order from 0.0 with separately
// rounded products, the additions of a sequential CSR loop, so w has the same
// bits as k_spmv_dot's.  The waves stride over the slices together: wave v of
// the grid (XCD-aware block order) takes slices v, v + waves, ..., so the
// waves in flight read one contiguous stretch of the matrix and each XCD a
// contiguous part of it (measured 215 us against 270-280 us when each
// workgroup walks its own contiguous slice range, on the 4096^2 Laplacian).
// Per-thread partials, then the block partials in block order as in
// k_spmv_dot.
constexpr int kSellU = 8;  // entries of a slice whose loads are issued together

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(ArnoldiBufs b, const double* x_in,
                                                      const double* xs_ptr, double* wout, int ndots,
                                                      int normq, int check_breakdown, int i0,
                                                      const double* shift_ptr) {
#pragma clang fp contract(off)
  if (check_breakdown && (b.iscal[kBreakdown] || (check_breakdown == 2 && b.iscal[kSStepFail]))) return;
  __shared__ double red[kWavesPerBlock * 2];
  const bool has_shift = shift_ptr != nullptr;
  const double sh = has_shift ? *shift_ptr : 0.0;
  const int lb = xcd_block_id();
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const double* __restrict__ x = x_in;
  const double xs = xs_ptr ? *xs_ptr : 1.0;
  const double cs0 = ndots ? b.colscale[i0] : 0.0;
  const double* __restrict__ Vd = b.V + static_cast<long long>(i0) * b.ld;
  const long long* __restrict__ sp = b.sell_ptr;
  const int* __restrict__ sc = b.sell_col;
  double* __restrict__ w = wout;
  const long long ns = b.n_slices;
  const long long nw = static_cast<long long>(gridDim.x) * kWavesPerBlock;
  double acc[2] = {0.0, 0.0};  // q_i0 . w, ||w||^2
  for (long long s = static_cast<long long>(lb) * kWavesPerBlock + wave; s < ns; s += nw) {
    const long long off = sp[s];
    const int W = static_cast<int>((sp[s + 1] - off) / kWave);
    const long long r = s * kWave + lane;
    double xrr = 0.0;
    if (has_shift && r < b.n) xrr = x[r];
    double a = 0.0;
    for (int k0 = 0; k0 < W; k0 += kSellU) {
      int c[kSellU];
      double v[kSellU];
#pragma unroll
      for (int u = 0; u < kSellU; ++u) {
        const long long e = off + static_cast<long long>(k0 + u) * kWave + lane;
        c[u] = -1;
        v[u] = 0.0;
        if (k0 + u < W) {
          c[u] = load_idx_nt(sc, e);
          v[u] = load_val_nt<VT>(b.sell_val, e);
        }
      }
      double p[kSellU];
#pragma unroll
      for (int u = 0; u < kSellU; ++u) p[u] = (c[u] >= 0) ? v[u] * (xs * x[c[u]]) : 0.0;
#pragma unroll
      for (int u = 0; u < kSellU; ++u)
        if (c[u] >= 0) a += p[u];
    }
    if (r < b.n) {
      if (has_shift) a = shift_row(a, sh, xs, xrr);
      __builtin_nontemporal_store(a, w + r);
      acc[1] += a * a;
      if (ndots) acc[0] += (cs0 * Vd[r]) * a;
    }
  }
  block_partials<2>(acc, ndots, red, b.partials, gridDim.x, lb);
  __syncthreads();
  double nv[1] = {acc[1]};
  block_partials<1>(nv, 1, red, b.partials + static_cast<long long>(normq) * gridDim.x, gridDim.x, lb);
}

// The sliced-ELL copy of a device CSR matrix: one thread per lane of every
// slice (rows past n included) writes its row's entries at their slice
// positions and pads the row to the slice width with column -1 and value 0.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_sell_build(const int* __restrict__ rp, const int* __restrict__ ci,
                                                       const VT* __restrict__ va, int n, long long nslices,
                                                       const long long* __restrict__ sp, int* __restrict__ sc,
                                                       VT* __restrict__ sv) {
  for (long long r = gtid(); r < nslices * kWave; r += gstride()) {
    const long long s = r / kWave;
    const int lane = static_cast<int>(r % kWave);
    const long long off = sp[s];
    const int W = static_cast<int>((sp[s + 1] - off) / kWave);
    const int k0 = (r < n) ? rp[r] : 0;
    const int len = (r < n) ? rp[r + 1] - k0 : 0;
    for (int e = 0; e < W; ++e) {
      const long long d = off + static_cast<long long>(e) * kWave + lane;
      if (e < len) {
        sc[d] = ci[k0 + e];
        sv[d] = va[k0 + e];
      } else {
        sc[d] = -1;
        sv[d] = VT(0);
      }
    }
  }
}

// K1 on the diagonal copy (ArnoldiBufs::dia_*): the same operation and
// arguments as k_spmv_sell.  No column index is read: row r of diagonal d is
// (r, r + off_d).  Each lane takes P consecutive rows (a wave 64 P rows), so
// for each diagonal a lane loads P consecutive values (one 16-byte load for
// P = 2 FP64 values, non-temporal) and the P consecutive x entries
// x[r + off_d] (one 16-byte load whatever the offset's parity), the rows' x
// reads overlapping across diagonals and served by the caches.  The waves
// stride over the 64 P-row groups together in XCD-aware block order, as in
// k_spmv_sell.  Each row is summed diagonal by diagonal in ascending offset
// order, which is its CSR order (the copy exists only for rows with
// ascending columns), from 0.0 with separately rounded products, absent
// entries skipped: w has the bits of the CSR loop.
// The kernel is bound by the latency of a group's loads more than by bytes
// (on the 4096^2 Laplacian a one-diagonal matrix took as long as the five
// diagonals): so a group's loads are issued before the previous group's
// store (a wait for loads also waits for every older store), the interior
// groups load without a branch, and the 3-, 5- and 7-diagonal matrices get
// kernels of their exact count (EX), whose fewer registers keep more waves
// resident (the constant 5-diagonal form 77 -> 66 us per launch).
// P > 1 stores w in 8P-byte groups and needs wout 16-byte aligned (measured
// on the 4096^2 Laplacian in isolation: 139 us against 156-163 us for one row
// per lane; with FP32 values on the 256^3 convection-diffusion matrix P = 4,
// whose value loads are 16 bytes too, took 167 us in the solve against 159 us
// for P = 2; P = 4 for the constant form: 76 against 77 us).
// CV: the constant-diagonal form (ArnoldiBufs::dia_mask / dia_cval).  No
// value is read: a lane loads its P rows' presence masks (P bytes, or 2P)
// and multiplies by the diagonal's one value, so a row costs its mask, its w
// and its share of x instead of 8 (or 4) bytes per diagonal more; the
// products and their order are those above, so w keeps the CSR loop's bits.
template <typename VT>
__device__ __forceinline__ bool dia_absent(VT v);
template <>
__device__ __forceinline__ bool dia_absent<double>(double v) {
  return static_cast<unsigned long long>(__double_as_longlong(v)) == kDiaAbsent64;
}
template <>
__device__ __forceinline__ bool dia_absent<float>(float v) {
  return __float_as_uint(v) == kDiaAbsent32;
}
template <typename VT>
__device__ __forceinline__ VT dia_marker();
template <>
__device__ __forceinline__ double dia_marker<double>() {
  return __longlong_as_double(static_cast<long long>(kDiaAbsent64));
}
template <>
__device__ __forceinline__ float dia_marker<float>() {
  return __uint_as_float(kDiaAbsent32);
}

template <typename VT, int P>
struct DiaVec {
  typedef VT type __attribute__((ext_vector_type(P)));
};
template <typename VT>
struct DiaVec<VT, 1> {
  typedef VT type;
};
// The presence masks of P consecutive rows in one load.
template <int BYTES>
struct DiaMaskLoad;
template <>
struct DiaMaskLoad<1> {
  typedef unsigned char type;
};
template <>
struct DiaMaskLoad<2> {
  typedef unsigned short type;
};
template <>
struct DiaMaskLoad<4> {
  typedef unsigned int type;
};
template <>
struct DiaMaskLoad<8> {
  typedef unsigned long long type;
};

template <typename VT, int ND, int P, bool CV, bool EX, int FD>
__global__ __launch_bounds__(kBlock) void k_spmv_dia(ArnoldiBufs b, const double* x_in, const double* xs_ptr,
                                                     double* wout, int ndots, int normq, int check_breakdown,
                                                     int i0, const double* shift_ptr) {
#pragma clang fp contract(off)
  if (check_breakdown && (b.iscal[kBreakdown] || (check_breakdown == 2 && b.iscal[kSStepFail]))) return;
  __shared__ double red[kWavesPerBlock * (FD + 1)];
  typedef typename DiaVec<VT, P>::type VTP;
  typedef typename DiaVec<double, P>::type DP;
  typedef typename std::conditional<(ND <= 8), unsigned char, unsigned short>::type MT;
  typedef typename DiaMaskLoad<static_cast<int>(sizeof(MT)) * P>::type MTP;
  const bool has_shift = shift_ptr != nullptr;
  const double sh = has_shift ? *shift_ptr : 0.0;
  const int lb = xcd_block_id();
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const double* __restrict__ x = x_in;
  const double xs = xs_ptr ? *xs_ptr : 1.0;
  const double cs0 = ndots ? b.colscale[i0] : 0.0;
  const double* __restrict__ Vd = b.V + static_cast<long long>(i0) * b.ld;
  const VT* __restrict__ dv = static_cast<const VT*>(b.dia_val);
  const MT* __restrict__ mp = static_cast<const MT*>(b.dia_mask);
  const long long n = b.n;
  const long long ld = b.dia_ld;
  const int nd = EX ? ND : b.n_diags;  // EX: the matrix has exactly ND diagonals
  int off[ND];
  double cv[ND];  // CV: the diagonals' values
  int d0 = -1;    // the main diagonal's slot: its x loads give the shift's x[r]
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    off[d] = (d < nd) ? b.dia_off[d] : 0;
    cv[d] = (CV && d < nd) ? b.dia_cval[d] : 0.0;
    if (d < nd && off[d] == 0) d0 = d;
  }
  double* __restrict__ w = wout;
  const long long ngroups = (n + kWave * P - 1) / (kWave * P);
  const long long nw = static_cast<long long>(gridDim.x) * kWavesPerBlock;
  double acc[2] = {0.0, 0.0};  // q_i0 . w (FD = 1), ||w||^2
  double accd[FD];             // FD > 1: sum of V[r, i0 + q] w[r], scaled by colscale at the end
#pragma unroll
  for (int q = 0; q < FD; ++q) accd[q] = 0.0;

  // The operands of one group of 64 P rows: values (or presence masks) and x.
  struct Tile {
    VT v[ND][P];
    double xv[ND][P];
    double xrr[P];
    double vq[FD][P];  // ndots: the rows of columns i0 .. i0 + ndots - 1, for the fused dots
    MTP mk;  // CV: the P rows' presence masks, unpacked only when summed (so
             // that no wait for the load comes before the x loads are issued)
  };
  // An interior group (every x[r + off_d] of its rows exists), with no
  // branch (a branch around a load makes the compiler wait for every load
  // before the next one into the same registers).
  auto load_full = [&](long long g, Tile& t) {
    const long long r0 = (g * kWave + lane) * P;
    if constexpr (CV)
      t.mk = __builtin_nontemporal_load(reinterpret_cast<const MTP*>(mp + r0));
    else
      t.mk = 0;
    if (has_shift && d0 < 0) {
      DP xx;
      __builtin_memcpy(&xx, x + r0, sizeof(DP));
      if constexpr (P == 1) {
        t.xrr[0] = xx;
      } else {
#pragma unroll
        for (int k = 0; k < P; ++k) t.xrr[k] = xx[k];
      }
    }
    if (FD > 1 || ndots) {
      // FD > 1: all FD columns are loaded (surplus ones repeat the last
      // column, a cache hit), so no load is inside a branch
#pragma unroll
      for (int q = 0; q < FD; ++q) {
        const int qc = q < ndots ? q : ndots - 1;
        DP qq;
        __builtin_memcpy(&qq, Vd + static_cast<long long>(qc) * b.ld + r0, sizeof(DP));
        if constexpr (P == 1) {
          t.vq[q][0] = qq;
        } else {
#pragma unroll
          for (int k = 0; k < P; ++k) t.vq[q][k] = qq[k];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      if (d < nd) {
        if constexpr (!CV) {
          const VTP vv =
              __builtin_nontemporal_load(reinterpret_cast<const VTP*>(dv + static_cast<long long>(d) * ld + r0));
          if constexpr (P == 1) {
            t.v[d][0] = vv;
          } else {
#pragma unroll
            for (int k = 0; k < P; ++k) t.v[d][k] = vv[k];
          }
        }
        DP xx;
        __builtin_memcpy(&xx, x + r0 + off[d], sizeof(DP));
        if constexpr (P == 1) {
          t.xv[d][0] = xx;
        } else {
#pragma unroll
          for (int k = 0; k < P; ++k) t.xv[d][k] = xx[k];
        }
      } else {  // an unused slot of the 8- or 16-slot kernel
#pragma unroll
        for (int k = 0; k < P; ++k) {
          t.v[d][k] = dia_marker<VT>();
          t.xv[d][k] = 0.0;
        }
      }
    }
  };
  // A group near either end of the matrix: one entry at a time.
  auto load_edge = [&](long long g, Tile& t) {
    const long long r0 = (g * kWave + lane) * P;
#pragma unroll
    for (int k = 0; k < P; ++k) t.xrr[k] = 0.0;
    t.mk = 0;
    if constexpr (CV) {
      MTP mm = 0;
#pragma unroll
      for (int k = 0; k < P; ++k)
        if (r0 + k < n) mm |= static_cast<MTP>(static_cast<MTP>(mp[r0 + k]) << (8 * sizeof(MT) * k));
      t.mk = mm;
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      if (has_shift && r0 + k < n) t.xrr[k] = x[r0 + k];
#pragma unroll
      for (int q = 0; q < FD; ++q) {
        t.vq[q][k] = 0.0;
        if (q < ndots && r0 + k < n) t.vq[q][k] = Vd[static_cast<long long>(q) * b.ld + r0 + k];
      }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
      for (int k = 0; k < P; ++k) {
        t.v[d][k] = dia_marker<VT>();
        t.xv[d][k] = 0.0;
        const long long r = r0 + k, c = r + off[d];
        if (d < nd && r < n && c >= 0 && c < n) {
          if constexpr (!CV) t.v[d][k] = dv[static_cast<long long>(d) * ld + r];
          t.xv[d][k] = x[c];
        }
      }
    }
  };
  // Row sums of a loaded group (interior: loaded by load_full), then an
  // empty asm that takes them as operands and clobbers memory: the sums exist
  // before the next group's loads are issued (the compiler would otherwise
  // sink them below those loads and keep two groups' operands in registers).
  // The row sums also go into the partial ||w||^2 and, with ndots, q_i0 . w.
  auto sums = [&](long long r0, const Tile& t, bool interior, double (&a)[P]) {
#pragma unroll
    for (int k = 0; k < P; ++k) {
      a[k] = 0.0;
      const unsigned mk = static_cast<unsigned>(static_cast<unsigned long long>(t.mk) >> (8 * sizeof(MT) * k));
      // x[r] for the shift: the main diagonal's x load when the group is
      // interior and the matrix has one, else the load made for it
      double xr = 0.0;
      if (has_shift && !(interior && d0 >= 0)) xr = t.xrr[k];
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const double xd = t.xv[d][k];
        if (has_shift && interior && d == d0) xr = xd;
        if constexpr (CV) {
          if ((mk >> d) & 1u) a[k] += cv[d] * (xs * xd);
        } else {
          if (!dia_absent<VT>(t.v[d][k])) a[k] += static_cast<double>(t.v[d][k]) * (xs * xd);
        }
      }
      if (has_shift) a[k] = shift_row(a[k], sh, xs, xr);
      if (interior || r0 + k < n) {
        acc[1] += a[k] * a[k];
        if constexpr (FD == 1) {
          if (ndots) acc[0] += (cs0 * t.vq[0][k]) * a[k];
        } else {
#pragma unroll
          for (int q = 0; q < FD; ++q)
            if (q < ndots) accd[q] += t.vq[q][k] * a[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) asm volatile("" : "+v"(a[k])::"memory");
  };
  auto store_full = [&](long long r0, const double (&a)[P]) {
    if constexpr (P == 1) {
      __builtin_nontemporal_store(a[0], w + r0);
    } else {
      DP o;
#pragma unroll
      for (int k = 0; k < P; ++k) o[k] = a[k];
      __builtin_nontemporal_store(o, reinterpret_cast<DP*>(w + r0));
    }
  };

  // Interior groups [glo, ghi) (rows whose every x[r + off_d] exists): each
  // group's sums are formed, the next group's loads issued, then the sums
  // stored, so that waiting for a group's loads does not also wait for the
  // previous group's store (loads and stores retire in issue order).  The
  // first group is peeled so that the loop is entered only with a store
  // issued after the pending loads, and the store is unconditional, so the
  // wait at the top of the loop can leave it in flight.  Groups before glo
  // and from ghi on (the matrix's first and last rows, and a last group that
  // runs past n) take the entry-by-entry path.
  const int omin = nd > 0 ? off[0] : 0;
  int omax = 0;
#pragma unroll
  for (int d = 0; d < ND; ++d)
    if (d < nd) omax = off[d];
  const long long gp = static_cast<long long>(kWave) * P;
  const long long glo = omin < 0 ? (-static_cast<long long>(omin) + gp - 1) / gp : 0;
  const long long rhi = n - (omax > 0 ? omax : 0);  // rows r < rhi have x[r + omax]
  const long long ghi = rhi > 0 ? rhi / gp : 0;
  auto edge_group = [&](long long g) {
    Tile t;
    double a[P];
    const long long r0 = (g * kWave + lane) * P;
    load_edge(g, t);
    sums(r0, t, false, a);
#pragma unroll
    for (int k = 0; k < P; ++k)
      if (r0 + k < n) w[r0 + k] = a[k];
  };
  const long long g0 = static_cast<long long>(lb) * kWavesPerBlock + wave;
  long long g = g0;
  for (; g < ngroups && g < glo; g += nw) edge_group(g);
  if (g < ghi) {
    Tile t;
    double a[P];
    load_full(g, t);
    long long r0 = (g * kWave + lane) * P;
    sums(r0, t, true, a);
    long long gn = g + nw;
    if (gn < ghi) load_full(gn, t);
    store_full(r0, a);
    for (g = gn; g < ghi; g += nw) {
      r0 = (g * kWave + lane) * P;
      sums(r0, t, true, a);
      if (g + nw < ghi) load_full(g + nw, t);
      store_full(r0, a);
    }
  }
  for (; g < ngroups; g += nw) edge_group(g);
  if constexpr (FD == 1) {
    block_partials<2>(acc, ndots, red, b.partials, gridDim.x, lb);
  } else {
#pragma unroll
    for (int q = 0; q < FD; ++q) accd[q] *= (q < ndots) ? b.colscale[i0 + q] : 0.0;
    block_partials<FD>(accd, ndots, red, b.partials, gridDim.x, lb);
  }
  __syncthreads();
  double nv[1] = {acc[1]};
  block_partials<1>(nv, 1, red, b.partials + static_cast<long long>(normq) * gridDim.x, gridDim.x, lb);
}

// The diagonal copy of a device CSR matrix: every entry set to the absent
// marker, then one thread per row writes each of its entries into the
// diagonal of its offset.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_dia_fill(VT* __restrict__ dv, long long count) {
  const VT mk = dia_marker<VT>();
  for (long long i = gtid(); i < count; i += gstride()) dv[i] = mk;
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_dia_build(const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const VT* __restrict__ va, int n,
                                                      const int* __restrict__ offs, int nd, long long ld,
                                                      VT* __restrict__ dv) {
  for (long long r = gtid(); r < n; r += gstride()) {
    for (int e = rp[r]; e < rp[r + 1]; ++e) {
      const long long o = static_cast<long long>(ci[e]) - r;
      for (int d = 0; d < nd; ++d)
        if (offs[d] == o) {
          dv[static_cast<long long>(d) * ld + r] = va[e];
          break;
        }
    }
  }
}

// The presence masks of the constant-diagonal form: one thread per row sets
// bit d for each entry on diagonal d; rows n..ld-1 get 0.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_dia_mask_build(const int* __restrict__ rp, const int* __restrict__ ci,
                                                           int n, const int* __restrict__ offs, int nd,
                                                           long long ld, MT* __restrict__ mask) {
  for (long long r = gtid(); r < ld; r += gstride()) {
    unsigned m = 0;
    if (r < n) {
      for (int e = rp[r]; e < rp[r + 1]; ++e) {
        const long long o = static_cast<long long>(ci[e]) - r;
        for (int d = 0; d < nd; ++d)
          if (offs[d] == o) {
            m |= 1u << d;
            break;
          }
      }
    }
    mask[r] = static_cast<MT>(m);
  }
}

// K1 and K1b in one pass for column j (nq = j + 1 <= 64): w = A q_j as in
// k_spmv_dot, and, while each row block's w is still in LDS, the partials of
// q_i . w for i < nq over that block's rows, the basis columns split over the
// waves as in the sweeps (wave w owns columns w, w + 4, ...).  The dot sweep
// then needs no separate pass over w, and the basis reads stream while the
// SpMV waits on its gathers.  Partials: quantity i < nq and ||w||^2 at nq,
// spmv_grid blocks.
template <typename VT, int MAXK>
__global__ __launch_bounds__(kBlock) void k_spmv_dots_cs(ArnoldiBufs b, int j) {
  if (b.iscal[kBreakdown]) return;
  constexpr int U = kNnzCap / kBlock;
  __shared__ double prod[kNnzCap];
  __shared__ double wt[kSpmvRowCap];  // w of the current block's rows
  __shared__ double red[4 * MAXK + 1];
  __shared__ double redn[kWavesPerBlock];
  const int lb = xcd_block_id();
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  const int grid = gridDim.x;
  const int nq = j + 1;
  const double* __restrict__ x = b.V + static_cast<long long>(j) * b.ld;
  const double xs = b.colscale[j];
  const double* __restrict__ V = b.V;
  const int* __restrict__ rp = b.row_ptr;
  const int* __restrict__ ci = b.col_idx;
  const int* __restrict__ rbk = b.row_blocks;
  double* __restrict__ w = b.w;
  long long col[MAXK];
  own_columns<MAXK>(col, b.ld, 0, nq);
  double acc[MAXK];
#pragma unroll
  for (int k = 0; k < MAXK; ++k) acc[k] = 0.0;
  double nrm = 0.0;

  const long long nrb = b.n_row_blocks;
  const int rb0 = static_cast<int>(nrb * lb / grid);
  const int rb1 = static_cast<int>(nrb * (lb + 1) / grid);
  int r0 = 0, r1 = 0, k0 = 0, k1 = 0, r2 = 0;
  if (rb0 < rb1) {
    r0 = rbk[rb0];
    r1 = rbk[rb0 + 1];
    k0 = rp[r0];
    k1 = rp[r1];
    r2 = (rb0 + 1 < rb1) ? rbk[rb0 + 2] : r1;
  }
  for (int rb = rb0; rb < rb1; ++rb) {
    const int k2 = (rb + 1 < rb1) ? rp[r2] : k1;
    const int r3 = (rb + 2 < rb1) ? rbk[rb + 3] : r2;
    if (k1 - k0 <= kNnzCap) {
      int idx[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + tid + u * kBlock;
        idx[u] = 0;
        v[u] = 0.0;
        if (k < k1) {
          idx[u] = load_idx_nt(ci, k);
          v[u] = load_val_nt<VT>(b.vals, k);
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
        __builtin_nontemporal_store(a, w + r);
        wt[r - r0] = a;
        nrm += a * a;
      }
    } else {
      // one long row: the block strides through it and reduces
      double a = 0.0;
      for (int k = k0 + tid; k < k1; k += kBlock)
        a += load_val_nt<VT>(b.vals, k) * (xs * x[load_idx_nt(ci, k)]);
      const double s = wave_sum(a);
      if (lane == 0) prod[tid / kWave] = s;
      __syncthreads();
      if (tid == 0) {
        double t = prod[0];
        for (int q = 1; q < kWavesPerBlock; ++q) t += prod[q];
        __builtin_nontemporal_store(t, w + r0);
        wt[0] = t;
        nrm += t * t;
      }
    }
    __syncthreads();  // the block's w is in wt
    for (int rc = r0; rc < r1; rc += kWave) {
      const int r = rc + lane;
      if (r < r1) {
        const double wr = wt[r - r0];
        double vq[MAXK];
#pragma unroll
        for (int k = 0; k < MAXK; ++k) vq[k] = __builtin_nontemporal_load(V + col[k] + r);
#pragma unroll
        for (int k = 0; k < MAXK; ++k) acc[k] += vq[k] * wr;
      }
    }
    __syncthreads();  // prod and wt are free for the next block
    r0 = r1;
    k0 = k1;
    r1 = r2;
    k1 = k2;
    r2 = r3;
  }
  write_column_partials<MAXK>(acc, 0.0, nq, false, b.colscale, red, b.partials, grid, lb);
  double nv[1] = {nrm};
  block_partials<1>(nv, 1, redn, b.partials + static_cast<long long>(nq) * grid, grid, lb);
}

// ------------------------------------------- K2 and K1b: column-split sweeps (nq <= 64)

// w' = w - sum_i h1_i q_i into V[:, j+1]; partials of q_i . w' (quantity i)
// and ||w'||^2 (quantity j+1).
template <int MAXK>
__global__ __launch_bounds__(kBlock) void k_update_dot_cs(ArnoldiBufs b, int j) {
  if (b.iscal[kBreakdown]) return;
  sweep_update_dot<MAXK, 2>(b.V, b.ld, j + 1, b.h1, b.colscale, b.w,
                            b.V + static_cast<long long>(j + 1) * b.ld, grid_tiles<2>(b.n), b.partials,
                            gridDim.x, blockIdx.x);
}

// partials of q_i . x for i in [i0, i0+nq) (quantity i - i0 + qoff) and, if
// with_norm, ||x||^2 at quantity qoff + nq.
template <int MAXK>
__global__ __launch_bounds__(kBlock) void k_dots_cs(ArnoldiBufs b, const double* xv, int i0, int nq,
                                                    int with_norm, int qoff) {
  if (b.iscal[kBreakdown]) return;
  sweep_dots<MAXK, 2>(b.V, b.ld, i0, nq, b.colscale + i0, xv, grid_tiles<2>(b.n), with_norm != 0,
                   b.partials + static_cast<long long>(qoff) * gridDim.x, gridDim.x, blockIdx.x);
}

// ------------------------------------------------------------------- K3

// Second pass w'' = w' - sum_{i<=j} h2_i q_i in place in V[:, j+1]; partial
// ||w''||^2.  Exits at once when the column needs no second pass.
__global__ __launch_bounds__(kBlock) void k_update_norm(ArnoldiBufs b, int j, int gs_mode, double eta) {
  if (b.iscal[kBreakdown] || !reorth_needed(b, gs_mode, eta)) return;
  __shared__ double red[kWavesPerBlock];
  __shared__ double cs[kMaxStaged];
  __shared__ double h[kMaxStaged];
  const int nq = j + 1;
  stage(cs, b.colscale, nq);
  stage(h, b.h2, nq);
  __syncthreads();
  double* __restrict__ io = b.V + static_cast<long long>(j + 1) * b.ld;
  double acc = 0.0;
  for (long long r = gtid(); r < b.n; r += gstride()) {
    double a = io[r];
    a -= sweep_combination(b.V, b.ld, r, nq, h, cs);
    io[r] = a;
    acc += a * a;
  }
  double o[1] = {acc};
  block_partials<1>(o, 1, red, b.partials, gridDim.x, blockIdx.x);
}

// Column-split form of K3 for nq <= 64.
template <int MAXK>
__global__ __launch_bounds__(kBlock) void k_update_norm_cs(ArnoldiBufs b, int j, int gs_mode,
                                                           double eta) {
  if (b.iscal[kBreakdown] || !reorth_needed(b, gs_mode, eta)) return;
  sweep_update_norm<MAXK, 2>(b.V, b.ld, j + 1, b.h2, b.colscale,
                             b.V + static_cast<long long>(j + 1) * b.ld, grid_tiles<2>(b.n), b.partials,
                             blockIdx.x);
}

// --------------------------------------------- generic kernels (j+1 > 64)

// dst = src - sum_{i<nq} h_i q_i
__global__ __launch_bounds__(kBlock) void k_update(ArnoldiBufs b, const double* src, double* dst,
                                                   const double* h, int nq) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double cs[kMaxStaged];
  __shared__ double hs[kMaxStaged];
  stage(cs, b.colscale, nq);
  stage(hs, h, nq);
  __syncthreads();
  for (long long r = gtid(); r < b.n; r += gstride()) dst[r] = src[r] - sweep_combination(b.V, b.ld, r, nq, hs, cs);
}

// --------------------------------------------------------------- finalise

// Finaliser of K1 (and the chunked dots): one workgroup per quantity; block q
// (q < nq) writes h1[q], block nq writes ||w||^2.  pgrid: blocks of the
// kernels that wrote the partials.
__global__ __launch_bounds__(kBlock) void k_fin_h1(ArnoldiBufs b, int nq, int pgrid) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double red[kWavesPerBlock];
  const int q = blockIdx.x;
  const double s = block_sum_quantity(b.partials, pgrid, q, red);
  if (threadIdx.x == 0) {
    if (q < nq) b.h1[q] = s;
    else b.scal[kWNorm2] = s;
  }
}

// Finaliser of K2: h2[q] (q < nq) and ||w'||^2.
__global__ __launch_bounds__(kBlock) void k_fin_h2(ArnoldiBufs b, int nq, int pgrid) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double red[kWavesPerBlock];
  const int q = blockIdx.x;
  const double s = block_sum_quantity(b.partials, pgrid, q, red);
  if (threadIdx.x == 0) {
    if (q < nq) b.h2[q] = s;
    else b.scal[kWpNorm2] = s;
  }
}

// Writes column j of H (zeros below the subdiagonal), the breakdown flag, the
// next column scale and k_used.  reference: normalize_new_vector
// (main.cu:1878-1942), breakdown when the norm is < 1e-14 or not finite.
__device__ void finish_column(const ArnoldiBufs& b, int j, double norm2, int reorth) {
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

// Finaliser of the column (one workgroup): the second-pass norm if K3 ran.
__global__ __launch_bounds__(kBlock) void k_fin_col(ArnoldiBufs b, int j, int gs_mode, double eta, int pgrid) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double red[kWavesPerBlock];
  const int reorth = reorth_needed(b, gs_mode, eta);
  double norm2 = b.scal[kWpNorm2];
  if (reorth) norm2 = block_sum_quantity(b.partials, pgrid, 0, red);
  if (threadIdx.x == 0) {
    if (reorth) b.scal[kWppNorm2] = norm2;
    b.iscal[kReorth] = reorth;
    finish_column(b, j, norm2, reorth);
  }
}

// ---------------------------------------------------- start vector / lift

// partials of ||x||^2 over x[0..n)
__global__ __launch_bounds__(kBlock) void k_norm2(ArnoldiBufs b, const double* xv) {
  __shared__ double red[kWavesPerBlock];
  double acc = 0.0;
  for (long long r = gtid(); r < b.n; r += gstride()) acc += xv[r] * xv[r];
  double o[1] = {acc};
  block_partials<1>(o, 1, red, b.partials, gridDim.x, blockIdx.x);
}

// Start of a solve: beta = ||V[:,0]|| and colscale[0] = 1/beta, guarded as on
// the reference's CPU path (main_cpu.cpp:473-484: scale 1 when beta <= 1e-14).
__global__ __launch_bounds__(kBlock) void k_fin_start(ArnoldiBufs b, int pgrid) {
  __shared__ double red[kWavesPerBlock];
  const double s = block_sum_quantity(b.partials, pgrid, 0, red);
  if (threadIdx.x == 0) {
    const double beta = sqrt(s);
    b.scal[kVNorm] = beta;
    b.colscale[0] = (beta > 1e-14) ? (1.0 / beta) : 1.0;
    b.iscal[kKUsed] = 0;
    b.iscal[kBreakdown] = 0;
  }
}

// y = sum_{i<ncoef, used} coef_i q_i where coef_i = beta * wH[i]; terms with
// |wH[i]| <= skip_below are left out (reference: 1e-15, main.cu:1992-1994).
// Partials of ||y||^2.
// Kept terms of the lift in order (|wH_i| > skip_below): column used[u] with
// coefficient beta * wH_i (coef) and its column scale (cs); returns the count.
// One thread compacts; the result is in LDS after the barrier.
__device__ int compact_lift_terms(const ArnoldiBufs& b, int ncoef, double skip_below, int* used,
                                  double* coef, double* cs, int* nused) {
  if (threadIdx.x == 0) {
    const double beta = b.scal[kVNorm];
    int u = 0;
    for (int i = 0; i < ncoef; ++i) {
      const double c = b.wH[i];
      if (fabs(c) > skip_below) {
        used[u] = i;
        coef[u] = beta * c;
        cs[u] = b.colscale[i];
        ++u;
      }
    }
    *nused = u;
  }
  __syncthreads();
  return *nused;
}

// Column-split lift for at most 64 terms: y = sum_u (beta wH_u) (cs_u V[:, u]).
template <int MAXK>
__global__ __launch_bounds__(kBlock) void k_lift_cs(ArnoldiBufs b, int ncoef, double skip_below) {
  __shared__ double cs[kSweepMaxCols];
  __shared__ double coef[kSweepMaxCols];
  __shared__ int used[kSweepMaxCols];
  __shared__ int nused;
  const int nu = compact_lift_terms(b, ncoef, skip_below, used, coef, cs, &nused);
  if (nu == 0) {  // no kept term: y = 0
    for (long long r = gtid(); r < b.n; r += gstride()) b.y[r] = 0.0;
    if (threadIdx.x == 0) b.partials[blockIdx.x] = 0.0;
    return;
  }
  // one combined coefficient per kept column, in kept order
  for (int u = threadIdx.x; u < nu; u += kBlock) coef[u] = coef[u] * cs[u];
  __syncthreads();
  sweep_lift<MAXK, 2>(b.V, b.ld, nu, used, coef, b.y, grid_tiles<2>(b.n), b.partials, blockIdx.x);
}

// Generic lift for more than 64 terms (chunked loads, LDS coefficients).
__global__ __launch_bounds__(kBlock) void k_lift(ArnoldiBufs b, int ncoef, double skip_below) {
  __shared__ double red[kWavesPerBlock];
  __shared__ double cs[kMaxStaged];
  __shared__ double coef[kMaxStaged];
  __shared__ int used[kMaxStaged];
  __shared__ int nused;
  const int nu = compact_lift_terms(b, ncoef, skip_below, used, coef, cs, &nused);
  double* __restrict__ y = b.y;
  double acc = 0.0;
  for (long long r = gtid(); r < b.n; r += gstride()) {
    double a = 0.0;
    for (int u0 = 0; u0 < nu; u0 += kChunk) {
      double v[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; ++k)
        v[k] = (u0 + k < nu) ? b.V[static_cast<long long>(used[u0 + k]) * b.ld + r] : 0.0;
#pragma unroll
      for (int k = 0; k < kChunk; ++k)
        if (u0 + k < nu) a += coef[u0 + k] * (cs[u0 + k] * v[k]);
    }
    y[r] = a;
    acc += a * a;
  }
  double o[1] = {acc};
  block_partials<1>(o, 1, red, b.partials, gridDim.x, blockIdx.x);
}

// scal[slot] = sqrt(sum of the quantity-0 partials)
__global__ __launch_bounds__(kBlock) void k_fin_norm(ArnoldiBufs b, int slot, int pgrid) {
  __shared__ double red[kWavesPerBlock];
  const double s = block_sum_quantity(b.partials, pgrid, 0, red);
  if (threadIdx.x == 0) b.scal[slot] = sqrt(s);
}

// Next sub-step starts from y: V[:,0] = y, beta = ||y||, colscale[0] guarded.
__global__ __launch_bounds__(kBlock) void k_restart(ArnoldiBufs b) {
  for (long long r = gtid(); r < b.n; r += gstride()) b.V[r] = b.y[r];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double beta = b.scal[kYNorm];
    b.scal[kVNorm] = beta;
    b.colscale[0] = (beta > 1e-14) ? (1.0 / beta) : 1.0;
    b.iscal[kKUsed] = 0;
    b.iscal[kBreakdown] = 0;
  }
}

// --------------------------------------------------------------- Lanczos
//
// Symmetric A: the three-term recurrence replaces Gram-Schmidt.  Column j is
// the SpMV with the fused partial alpha_j = q_j . A q_j (launch_spmv_lanczos,
// finalised by k_fin_h1 into h1[0]), then one sweep
//   w' = A q_j - alpha_j q_j - beta_{j-1} q_{j-1}  ->  V[:, j+1] (unnormalised)
// with the partial ||w'||^2, and the finaliser below.  H stays tridiagonal
// and symmetric; the exponential, lift and step control are shared with the
// Arnoldi path.

__global__ __launch_bounds__(kBlock) void k_lanczos_update(ArnoldiBufs b, int j) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double red[kWavesPerBlock];
  const double alpha = b.h1[0];
  const double ca = alpha * b.colscale[j];
  const double cb = (j > 0) ? b.H[static_cast<long long>(j - 1) * b.ldh + j] * b.colscale[j - 1] : 0.0;
  const double* __restrict__ qj = b.V + static_cast<long long>(j) * b.ld;
  const double* __restrict__ qm = b.V + static_cast<long long>(j > 0 ? j - 1 : 0) * b.ld;
  const double* __restrict__ w = b.w;
  double* __restrict__ out = b.V + static_cast<long long>(j + 1) * b.ld;
  double acc = 0.0;
  for (long long r = gtid(); r < b.n; r += gstride()) {
    const double a = (w[r] - ca * qj[r]) - cb * qm[r];
    out[r] = a;
    acc += a * a;
  }
  double o[1] = {acc};
  block_partials<1>(o, 1, red, b.partials, gridDim.x, blockIdx.x);
}

// H(j, j) = alpha_j, H(j+1, j) = beta_j = ||w'|| (with the Arnoldi path's
// breakdown test), H(j-1, j) = beta_{j-1}; colscale[j+1] = 1 / beta_j.
__global__ __launch_bounds__(kBlock) void k_lanczos_fin(ArnoldiBufs b, int j, int pgrid) {
  if (b.iscal[kBreakdown]) return;
  __shared__ double red[kWavesPerBlock];
  const double norm2 = block_sum_quantity(b.partials, pgrid, 0, red);
  if (threadIdx.x == 0) {
    const int ldh = b.ldh;
    double* Hc = b.H + static_cast<long long>(j) * ldh;
    for (int i = 0; i < ldh; ++i) Hc[i] = 0.0;
    if (j > 0) Hc[j - 1] = b.H[static_cast<long long>(j - 1) * ldh + j];
    Hc[j] = b.h1[0];
    const double nrm = sqrt(norm2);
    b.iscal[kKUsed] = j + 1;
    if (nrm < 1e-14 || isnan(nrm) || isinf(nrm)) {
      b.iscal[kBreakdown] = 1;
    } else {
      Hc[j + 1] = nrm;
      b.colscale[j + 1] = 1.0 / nrm;
    }
  }
}

// ------------------------------------------------------------ launchers

template <int MAXJ>
static void launch_spmv_dot_t(const ArnoldiBufs& b, const double* x, const double* xs, double* w,
                              int ndots, int normq, int chk, hipStream_t s, int i0 = 0,
                              const double* shift = nullptr) {
  if (b.n_diags > 0 && ndots <= 1) {
#define MEK_DIA(VT, ND, R, CV, EX)                                                                          \
  hipLaunchKernelGGL((k_spmv_dia<VT, ND, R, CV, EX, 1>), dim3(b.spmv_grid), dim3(kBlock), 0, s, b, x, xs, w, \
                     ndots, normq, chk, i0, shift)
// the 3-, 5- and 7-point stencils get kernels of their exact diagonal count
// (fewer registers), other counts the 8- or 16-slot kernel
#define MEK_DIA_ND(VT, R, CV)                                  \
  do {                                                         \
    switch (b.n_diags) {                                       \
      case 3: MEK_DIA(VT, 3, R, CV, true); break;              \
      case 5: MEK_DIA(VT, 5, R, CV, true); break;              \
      case 7: MEK_DIA(VT, 7, R, CV, true); break;              \
      default:                                                 \
        if (b.n_diags <= 8) MEK_DIA(VT, 8, R, CV, false);      \
        else MEK_DIA(VT, kDiaDiagCap, R, CV, false);           \
    }                                                          \
  } while (0)
    // two rows per lane when w can take 16-byte stores (every buffer of the
    // runner; a caller's output of the plain SpMV may not)
    const bool pair = (reinterpret_cast<uintptr_t>(w) & 15) == 0;
    if (!b.dia_val) {  // constant diagonals
      if (pair) MEK_DIA_ND(double, 2, true);
      else MEK_DIA_ND(double, 1, true);
    } else if (b.vals_f32) {
      if (pair) MEK_DIA_ND(float, 2, false);
      else MEK_DIA_ND(float, 1, false);
    } else {
      if (pair) MEK_DIA_ND(double, 2, false);
      else MEK_DIA_ND(double, 1, false);
    }
#undef MEK_DIA_ND
#undef MEK_DIA
    return;
  }
  if (b.sell_ptr && ndots <= 1) {
    if (b.vals_f32)
      hipLaunchKernelGGL(k_spmv_sell<float>, dim3(b.spmv_grid), dim3(kBlock), 0, s, b, x, xs, w, ndots, normq,
                         chk, i0, shift);
    else
      hipLaunchKernelGGL(k_spmv_sell<double>, dim3(b.spmv_grid), dim3(kBlock), 0, s, b, x, xs, w, ndots,
                         normq, chk, i0, shift);
    return;
  }
  if (b.vals_f32)
    hipLaunchKernelGGL((k_spmv_dot<float, MAXJ>), dim3(b.spmv_grid), dim3(kBlock), 0, s, b, x, xs, w,
                       ndots, normq, chk, i0, shift);
  else
    hipLaunchKernelGGL((k_spmv_dot<double, MAXJ>), dim3(b.spmv_grid), dim3(kBlock), 0, s, b, x, xs, w,
                       ndots, normq, chk, i0, shift);
}

void launch_sell_build(const int* row_ptr, const int* col_idx, const void* vals, int vals_f32, int n,
                       const long long* sell_ptr, int* sell_col, void* sell_val, hipStream_t s) {
  if (n <= 0) return;
  const long long nslices = (static_cast<long long>(n) + kWave - 1) / kWave;
  const int grid = static_cast<int>(std::min<long long>((nslices * kWave + kBlock - 1) / kBlock, 16384));
  if (vals_f32)
    hipLaunchKernelGGL(k_sell_build<float>, dim3(grid), dim3(kBlock), 0, s, row_ptr, col_idx,
                       static_cast<const float*>(vals), n, nslices, sell_ptr, sell_col,
                       static_cast<float*>(sell_val));
  else
    hipLaunchKernelGGL(k_sell_build<double>, dim3(grid), dim3(kBlock), 0, s, row_ptr, col_idx,
                       static_cast<const double*>(vals), n, nslices, sell_ptr, sell_col,
                       static_cast<double*>(sell_val));
}

void launch_dia_build(const int* row_ptr, const int* col_idx, const void* vals, int vals_f32, int n,
                      const int* offsets, int n_diags, long long ld, void* dia_val, hipStream_t s) {
  if (n <= 0 || n_diags <= 0) return;
  const long long count = static_cast<long long>(n_diags) * ld;
  const int fgrid = static_cast<int>(std::min<long long>((count + kBlock - 1) / kBlock, 16384));
  const int grid = static_cast<int>(std::min<long long>((static_cast<long long>(n) + kBlock - 1) / kBlock, 16384));
  if (vals_f32) {
    hipLaunchKernelGGL(k_dia_fill<float>, dim3(fgrid), dim3(kBlock), 0, s, static_cast<float*>(dia_val), count);
    hipLaunchKernelGGL(k_dia_build<float>, dim3(grid), dim3(kBlock), 0, s, row_ptr, col_idx,
                       static_cast<const float*>(vals), n, offsets, n_diags, ld, static_cast<float*>(dia_val));
  } else {
    hipLaunchKernelGGL(k_dia_fill<double>, dim3(fgrid), dim3(kBlock), 0, s, static_cast<double*>(dia_val), count);
    hipLaunchKernelGGL(k_dia_build<double>, dim3(grid), dim3(kBlock), 0, s, row_ptr, col_idx,
                       static_cast<const double*>(vals), n, offsets, n_diags, ld, static_cast<double*>(dia_val));
  }
}

void launch_dia_mask_build(const int* row_ptr, const int* col_idx, int n, const int* offsets, int n_diags,
                           long long ld, void* mask, hipStream_t s) {
  if (ld <= 0 || n_diags <= 0) return;
  const int grid = static_cast<int>(std::min<long long>((ld + kBlock - 1) / kBlock, 16384));
  if (n_diags <= 8)
    hipLaunchKernelGGL(k_dia_mask_build<unsigned char>, dim3(grid), dim3(kBlock), 0, s, row_ptr, col_idx, n,
                       offsets, n_diags, ld, static_cast<unsigned char*>(mask));
  else
    hipLaunchKernelGGL(k_dia_mask_build<unsigned short>, dim3(grid), dim3(kBlock), 0, s, row_ptr, col_idx, n,
                       offsets, n_diags, ld, static_cast<unsigned short*>(mask));
}

// s-step block vector i (columns k + i from k - 1 + i): y = (A - theta_i) p
// into V[:, k+i], ||y||^2 partial at quantity 0 (spmv_grid blocks).
void launch_ss_spmv(const ArnoldiBufs& b, const SsArgs& a, int k, int i, hipStream_t s) {
  const long long in = k - 1 + i;
  launch_spmv_dot_t<1>(b, b.V + in * b.ld, b.colscale + in, b.V + (in + 1) * b.ld, 0, 0, 2, s, 0,
                       a.st + kSsTheta + i);
}

// Arnoldi column j: w = A q_j with the ||w||^2 partial at quantity normq
// (spmv_grid blocks).  The form with the first Gram-Schmidt pass fused in is
// launch_spmv_dots.
void launch_spmv_col(const ArnoldiBufs& b, int j, int normq, hipStream_t s) {
  const double* x = b.V + static_cast<long long>(j) * b.ld;
  const double* xs = b.colscale + j;
  launch_spmv_dot_t<1>(b, x, xs, b.w, 0, normq, 1, s);
}

bool spmv_dia_fuses_dots(const ArnoldiBufs& b, int nq) {
  const int nd = b.n_diags;
  return (nd == 3 || nd == 5 || nd == 7) && nq >= 1 && nq <= kDiaFusedDots &&
         (reinterpret_cast<uintptr_t>(b.w) & 15) == 0;
}

// Column j on a diagonal layout with its Gram-Schmidt dots fused into the
// SpMV (spmv_dia_fuses_dots): partials of q_i . w (i <= j) and ||w||^2
// (quantity j + 1), spmv_grid blocks.
void launch_spmv_dia_col_dots(const ArnoldiBufs& b, int j, hipStream_t s) {
  const double* x = b.V + static_cast<long long>(j) * b.ld;
  const double* xs = b.colscale + j;
  const int nq = j + 1;
#define MEK_DD(VT, ND, CV, FD)                                                                              \
  hipLaunchKernelGGL((k_spmv_dia<VT, ND, 2, CV, true, FD>), dim3(b.spmv_grid), dim3(kBlock), 0, s, b, x, xs, \
                     b.w, nq, nq, 1, 0, static_cast<const double*>(nullptr))
#define MEK_DD_FD(VT, ND, CV) \
  do {                        \
    if (nq <= 2)              \
      MEK_DD(VT, ND, CV, 2);  \
    else if (nq <= 4)         \
      MEK_DD(VT, ND, CV, 4);  \
    else                      \
      MEK_DD(VT, ND, CV, 8);  \
  } while (0)
#define MEK_DD_ND(VT, CV)                       \
  do {                                          \
    if (b.n_diags == 3)                         \
      MEK_DD_FD(VT, 3, CV);                     \
    else if (b.n_diags == 5)                    \
      MEK_DD_FD(VT, 5, CV);                     \
    else                                        \
      MEK_DD_FD(VT, 7, CV);                     \
  } while (0)
  if (!b.dia_val)
    MEK_DD_ND(double, true);
  else if (b.vals_f32)
    MEK_DD_ND(float, false);
  else
    MEK_DD_ND(double, false);
#undef MEK_DD_ND
#undef MEK_DD_FD
#undef MEK_DD
}

// Column j with the first Gram-Schmidt pass fused into the SpMV (nq <= 64):
// partials of q_i . w (i < nq) and ||w||^2 (quantity nq), spmv_grid blocks.
void launch_spmv_dots(const ArnoldiBufs& b, int j, hipStream_t s) {
#define MEK_L(K)                                                                                       \
  if (b.vals_f32)                                                                                      \
    hipLaunchKernelGGL((k_spmv_dots_cs<float, K>), dim3(b.spmv_grid), dim3(kBlock), 0, s, b, j);       \
  else                                                                                                 \
    hipLaunchKernelGGL((k_spmv_dots_cs<double, K>), dim3(b.spmv_grid), dim3(kBlock), 0, s, b, j)
  MEK_SWEEP_BUCKETS(j + 1, MEK_L);
#undef MEK_L
}

// Lanczos column j: w = A q_j with the partial q_j . w at quantity 0 and
// ||w||^2 at quantity 1 (spmv_grid blocks).
void launch_spmv_lanczos(const ArnoldiBufs& b, int j, hipStream_t s) {
  launch_spmv_dot_t<1>(b, b.V + static_cast<long long>(j) * b.ld, b.colscale + j, b.w, 1, 1, 1, s, j);
}
void launch_lanczos_update(const ArnoldiBufs& b, int j, hipStream_t s) {
  hipLaunchKernelGGL(k_lanczos_update, dim3(b.grid), dim3(kBlock), 0, s, b, j);
}
void launch_lanczos_fin(const ArnoldiBufs& b, int j, hipStream_t s) {
  hipLaunchKernelGGL(k_lanczos_fin, dim3(1), dim3(kBlock), 0, s, b, j, b.grid);
}

// Plain y = A*(xs*x) (xs may be null) with the ||y||^2 partial at quantity 0.
void launch_spmv_plain(const ArnoldiBufs& b, const double* x, const double* xs, double* y,
                       hipStream_t s) {
  launch_spmv_dot_t<1>(b, x, xs, y, 0, 0, 0, s);
}

void launch_update_dot(const ArnoldiBufs& b, int j, hipStream_t s) {
#define MEK_L(K) hipLaunchKernelGGL(k_update_dot_cs<K>, dim3(b.grid), dim3(kBlock), 0, s, b, j)
  MEK_SWEEP_BUCKETS(j + 1, MEK_L);
#undef MEK_L
}

void launch_dots(const ArnoldiBufs& b, const double* x, int i0, int nq, int with_norm, int qoff,
                 hipStream_t s) {
#define MEK_L(K) \
  hipLaunchKernelGGL(k_dots_cs<K>, dim3(b.grid), dim3(kBlock), 0, s, b, x, i0, nq, with_norm, qoff)
  MEK_SWEEP_BUCKETS(nq, MEK_L);
#undef MEK_L
}

void launch_update(const ArnoldiBufs& b, const double* src, double* dst, const double* h, int nq,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_update, dim3(b.grid), dim3(kBlock), 0, s, b, src, dst, h, nq);
}

void launch_update_norm(const ArnoldiBufs& b, int j, int gs_mode, double eta, hipStream_t s) {
  if (j + 1 > kSweepMaxCols) {
    hipLaunchKernelGGL(k_update_norm, dim3(b.grid), dim3(kBlock), 0, s, b, j, gs_mode, eta);
    return;
  }
#define MEK_L(K) \
  hipLaunchKernelGGL(k_update_norm_cs<K>, dim3(b.grid), dim3(kBlock), 0, s, b, j, gs_mode, eta)
  MEK_SWEEP_BUCKETS(j + 1, MEK_L);
#undef MEK_L
}
void launch_fin_h1(const ArnoldiBufs& b, int nq, int pgrid, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_h1, dim3(nq + 1), dim3(kBlock), 0, s, b, nq, pgrid);
}
void launch_fin_h2(const ArnoldiBufs& b, int j, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_h2, dim3(j + 2), dim3(kBlock), 0, s, b, j + 1, b.grid);
}
void launch_fin_col(const ArnoldiBufs& b, int j, int gs_mode, double eta, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_col, dim3(1), dim3(kBlock), 0, s, b, j, gs_mode, eta, b.grid);
}
void launch_norm2(const ArnoldiBufs& b, const double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_norm2, dim3(b.grid), dim3(kBlock), 0, s, b, x);
}
void launch_fin_start(const ArnoldiBufs& b, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_start, dim3(1), dim3(kBlock), 0, s, b, b.grid);
}
void launch_lift(const ArnoldiBufs& b, int ncoef, double skip_below, hipStream_t s) {
  if (ncoef > kSweepMaxCols) {
    hipLaunchKernelGGL(k_lift, dim3(b.grid), dim3(kBlock), 0, s, b, ncoef, skip_below);
    return;
  }
#define MEK_L(K) hipLaunchKernelGGL(k_lift_cs<K>, dim3(b.grid), dim3(kBlock), 0, s, b, ncoef, skip_below)
  MEK_SWEEP_BUCKETS(ncoef, MEK_L);
#undef MEK_L
}
void launch_fin_norm(const ArnoldiBufs& b, int slot, int pgrid, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(kBlock), 0, s, b, slot, pgrid);
}
void launch_restart(const ArnoldiBufs& b, hipStream_t s) {
  hipLaunchKernelGGL(k_restart, dim3(b.grid), dim3(kBlock), 0, s, b);
}

}  // namespace mek
