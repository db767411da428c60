// Column-split streaming sweeps over a column-major basis, shared by the
// single-problem kernels (arnoldi.hip) and the batched ones (batch.hip).
//
// Wave w of a 256-thread workgroup owns basis columns w, w+4, w+8, ... (at
// most MAXK of them, so up to 4*MAXK columns); a tile is 64 R consecutive rows,
// R per lane (R = 2: a lane loads a 16-byte pair of rows).  Column indices past
// the count are clamped to the last column with a zero coefficient, so every
// load of a tile is issued back-to-back with no branch, and the per-column
// coefficients are wave-uniform registers computed once.  Basis columns are
// read with non-temporal loads: each is streamed once per sweep and should not
// displace the vectors that are reused (x, w) from the caches.  Sweeps that
// produce a combination (update, lift) meet the four waves' partial sums in
// double-buffered LDS, one barrier per tile.
//
// Per-workgroup results go to partials[q * pstride + pidx] for quantity q, in
// a fixed order (DPP wave sums, then the owning wave's lane 0), so the
// reductions stay deterministic.
//
// R = 2 needs an even tile base and even column offsets (ld even), and reads
// (never writes) one row past an odd range end: the vectors must have that
// padding row (the basis columns have it whenever ld > n).
#pragma once

#include "common.hpp"

namespace mek {

// Rows base + t*64*R + R*lane (+ 0..R-1) for t = t0, t0+ts, ... < ntiles;
// rows >= rend are outside the range (their loads are clamped and their
// results dropped).  Requires rend > base.
struct TileRange {
  long long base;
  long long t0;
  long long ntiles;
  long long ts;
  long long rend;
};

template <int R = 1>
__device__ __forceinline__ TileRange grid_tiles(long long n) {
  return TileRange{0, static_cast<long long>(blockIdx.x), (n + kWave * R - 1) / (kWave * R),
                   static_cast<long long>(gridDim.x), n};
}

template <int R = 1>
__device__ __forceinline__ TileRange block_tiles(long long row0, long long row1) {
  return TileRange{row0, 0, (row1 - row0 + kWave * R - 1) / (kWave * R), 1, row1};
}

// R rows of one column starting at p (non-temporal; p 16-byte aligned for R = 2)
template <int R>
__device__ __forceinline__ void load_rows_nt(const double* p, double (&v)[R]) {
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = __builtin_nontemporal_load(p + q);
}

template <int R>
__device__ __forceinline__ void load_rows(const double* p, double (&v)[R]) {
  if constexpr (R == 2) {
    const double2 t = *reinterpret_cast<const double2*>(p);
    v[0] = t.x;
    v[1] = t.y;
  } else {
    v[0] = p[0];
  }
}

template <int R>
__device__ __forceinline__ void store_rows(double* p, const double (&v)[R], bool all) {
  if constexpr (R == 2) {
    if (all) {
      *reinterpret_cast<double2*>(p) = double2{v[0], v[1]};
    } else {
      p[0] = v[0];
    }
  } else {
    p[0] = v[0];
  }
}

// First row of this lane's share of tile t, and the row its loads use (the
// last valid R-row group when the tile runs past the range end).
template <int R>
__device__ __forceinline__ long long tile_row(const TileRange& tr, long long t, int lane) {
  return tr.base + t * (kWave * R) + R * lane;
}
template <int R>
__device__ __forceinline__ long long clamp_row(const TileRange& tr, long long r) {
  if (r < tr.rend) return r;
  const long long last = tr.rend - 1;
  return (R == 2) ? tr.base + ((last - tr.base) & ~1LL) : last;
}

// Wave-owned columns: col[k] = first row of column (i0 + wave + 4k), clamped.
template <int MAXK>
__device__ __forceinline__ void own_columns(long long (&col)[MAXK], long long ld, int i0, int nq) {
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    const int i = wave + 4 * k;
    col[k] = static_cast<long long>(i0 + (i < nq ? i : nq - 1)) * ld;
  }
}

// Write the per-wave column sums acc[k] (column wave + 4k, scaled by cs) and
// an optional norm to the partials.  `red` holds 4*MAXK + 1 doubles.
template <int MAXK>
__device__ __forceinline__ void write_column_partials(const double (&acc)[MAXK], double nrm, int nq,
                                                      bool with_norm, const double* cs, double* red,
                                                      double* partials, long long pstride,
                                                      long long pidx) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) red[wave + 4 * k] = s;
  }
  const double sn = wave_sum(nrm);
  if (threadIdx.x == 0) red[4 * MAXK] = sn;
  __syncthreads();
  const int ntot = nq + (with_norm ? 1 : 0);
  for (int q = threadIdx.x; q < ntot; q += kBlock) {
    const double s = (q < nq) ? cs[q] * red[q] : red[4 * MAXK];
    partials[static_cast<long long>(q) * pstride + pidx] = s;
  }
}

// Partials of q_i . x = cs_i * sum_r V[r, i0+i] x[r] for i < nq (quantity i)
// and, with_norm, of ||x||^2 (quantity nq).  cs points at column i0's scale.
template <int MAXK, int R = 1>
__device__ __forceinline__ void sweep_dots(const double* __restrict__ V, long long ld, int i0, int nq,
                                           const double* cs, const double* __restrict__ x,
                                           TileRange tr, bool with_norm, double* partials,
                                           long long pstride, long long pidx) {
  __shared__ double red[4 * MAXK + 1];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  long long col[MAXK];
  own_columns<MAXK>(col, ld, i0, nq);
  double acc[MAXK];
#pragma unroll
  for (int k = 0; k < MAXK; ++k) acc[k] = 0.0;
  double nrm = 0.0;
  for (long long t = tr.t0; t < tr.ntiles; t += tr.ts) {
    const long long r = tile_row<R>(tr, t, lane);
    if (r >= tr.rend) continue;
    const bool all = (r + R - 1) < tr.rend;
    double v[MAXK][R];
#pragma unroll
    for (int k = 0; k < MAXK; ++k) load_rows_nt<R>(V + col[k] + r, v[k]);
    double a[R];
    load_rows<R>(x + r, a);
    if (!all) a[R - 1] = 0.0;  // the padding row past the range end
#pragma unroll
    for (int k = 0; k < MAXK; ++k)
#pragma unroll
      for (int q = 0; q < R; ++q) acc[k] += v[k][q] * a[q];
    if (wave == 0)
#pragma unroll
      for (int q = 0; q < R; ++q) nrm += a[q] * a[q];
  }
  write_column_partials<MAXK>(acc, nrm, nq, with_norm, cs, red, partials, pstride, pidx);
}

// out = x - sum_{i<nq} c_i V[:, i] with c_i = h_i cs_i; partials of
// q_i . out (quantity i) and ||out||^2 (quantity nq).  out is not a column
// in [0, nq).  Two tiles are in flight per wave: tile t + 2 ts is loaded as
// soon as tile t is done (register sets A and B alternate, so nothing is
// copied), a load is never inside a branch (a tile past the range reloads
// the last one, whose lines are in L2), and each tile's result is stored
// before the next loads are issued.  On the 4096^2 Laplacian this took the
// sweep from 4.30 to 3.85 ms per solve; the same overlap in sweep_dots made
// it slower (3.33 to 3.90 ms), so that sweep loads one tile at a time.
template <int MAXK, int R = 1>
__device__ __forceinline__ void sweep_update_dot(const double* __restrict__ V, long long ld, int nq,
                                                 const double* h, const double* cs,
                                                 const double* __restrict__ x, double* __restrict__ out,
                                                 TileRange tr, double* partials, long long pstride,
                                                 long long pidx) {
  __shared__ double part[2][kWavesPerBlock][kWave * R];
  __shared__ double red[4 * MAXK + 1];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  long long col[MAXK];
  own_columns<MAXK>(col, ld, 0, nq);
  double c[MAXK];
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    const int i = wave + 4 * k;
    c[k] = (i < nq) ? h[i] * cs[i] : 0.0;
  }
  double acc[MAXK];
#pragma unroll
  for (int k = 0; k < MAXK; ++k) acc[k] = 0.0;
  double nrm = 0.0;
  int buf = 0;
  struct Tile {
    double v[MAXK][R];
    double x[R];
  };
  auto load = [&](long long t, Tile& T) {
    const long long tc = t < tr.ntiles ? t : tr.ntiles - 1;
    const long long rc = clamp_row<R>(tr, tile_row<R>(tr, tc, lane));
#pragma unroll
    for (int k = 0; k < MAXK; ++k) load_rows_nt<R>(V + col[k] + rc, T.v[k]);
    load_rows<R>(x + rc, T.x);
  };
  auto use = [&](long long t, const Tile& T) {
    const long long r = tile_row<R>(tr, t, lane);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double p = 0.0;
#pragma unroll
      for (int k = 0; k < MAXK; ++k) p += c[k] * T.v[k][q];
      part[buf][wave][R * lane + q] = p;
    }
    __syncthreads();
    double a[R];
#pragma unroll
    for (int q = 0; q < R; ++q)
      a[q] = T.x[q] - ((part[buf][0][R * lane + q] + part[buf][1][R * lane + q]) +
                       (part[buf][2][R * lane + q] + part[buf][3][R * lane + q]));
    if (r < tr.rend) {
      const bool all = (r + R - 1) < tr.rend;
      if (!all) a[R - 1] = 0.0;
      if (wave == 0) {
        store_rows<R>(out + r, a, all);
#pragma unroll
        for (int q = 0; q < R; ++q) nrm += a[q] * a[q];
      }
#pragma unroll
      for (int k = 0; k < MAXK; ++k)
#pragma unroll
        for (int q = 0; q < R; ++q) acc[k] += T.v[k][q] * a[q];
    }
    buf ^= 1;
  };
  if (tr.t0 < tr.ntiles) {
    Tile A, B;
    long long t = tr.t0;
    load(t, A);
    load(t + tr.ts, B);
    for (;;) {
      use(t, A);
      load(t + 2 * tr.ts, A);
      t += tr.ts;
      if (t >= tr.ntiles) break;
      use(t, B);
      load(t + 2 * tr.ts, B);
      t += tr.ts;
      if (t >= tr.ntiles) break;
    }
  }
  write_column_partials<MAXK>(acc, nrm, nq, true, cs, red, partials, pstride, pidx);
}

// io = io - sum_{i<nq} h_i cs_i V[:, i] in place; partial ||io||^2 (quantity 0).
template <int MAXK, int R = 1>
__device__ __forceinline__ void sweep_update_norm(const double* __restrict__ V, long long ld, int nq,
                                                  const double* h, const double* cs,
                                                  double* __restrict__ io, TileRange tr,
                                                  double* partials, long long pidx) {
  __shared__ double part[2][kWavesPerBlock][kWave * R];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  long long col[MAXK];
  own_columns<MAXK>(col, ld, 0, nq);
  double c[MAXK];
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    const int i = wave + 4 * k;
    c[k] = (i < nq) ? h[i] * cs[i] : 0.0;
  }
  double nrm = 0.0;
  int buf = 0;
  for (long long t = tr.t0; t < tr.ntiles; t += tr.ts) {
    const long long r = tile_row<R>(tr, t, lane);
    const long long rc = clamp_row<R>(tr, r);
    double v[MAXK][R];
#pragma unroll
    for (int k = 0; k < MAXK; ++k) load_rows_nt<R>(V + col[k] + rc, v[k]);
    double xr[R];
    if (wave == 0) load_rows<R>(io + rc, xr);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double p = 0.0;
#pragma unroll
      for (int k = 0; k < MAXK; ++k) p += c[k] * v[k][q];
      part[buf][wave][R * lane + q] = p;
    }
    __syncthreads();
    if (wave == 0 && r < tr.rend) {
      const bool all = (r + R - 1) < tr.rend;
      double a[R];
#pragma unroll
      for (int q = 0; q < R; ++q)
        a[q] = xr[q] - ((part[buf][0][R * lane + q] + part[buf][1][R * lane + q]) +
                        (part[buf][2][R * lane + q] + part[buf][3][R * lane + q]));
      if (!all) a[R - 1] = 0.0;
      store_rows<R>(io + r, a, all);
#pragma unroll
      for (int q = 0; q < R; ++q) nrm += a[q] * a[q];
    }
    buf ^= 1;
  }
  // only wave 0 holds a norm; lane 0 of wave 0 writes it after the wave sum
  const double sn = wave_sum(nrm);
  if (threadIdx.x == 0) partials[pidx] = sn;
}

// y = sum_{u<nu} coef_u V[:, used_u] (coefficients already include the
// column scales); partial ||y||^2 (quantity 0).  used and coef live in LDS.
template <int MAXK, int R = 1>
__device__ __forceinline__ void sweep_lift(const double* __restrict__ V, long long ld, int nu,
                                           const int* used, const double* coef,
                                           double* __restrict__ y, TileRange tr, double* partials,
                                           long long pidx) {
  __shared__ double part[2][kWavesPerBlock][kWave * R];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  long long col[MAXK];
  double c[MAXK];
#pragma unroll
  for (int k = 0; k < MAXK; ++k) {
    const int u = wave + 4 * k;
    const int uc = u < nu ? u : nu - 1;
    col[k] = static_cast<long long>(used[uc]) * ld;
    c[k] = (u < nu) ? coef[uc] : 0.0;
  }
  double nrm = 0.0;
  int buf = 0;
  // two tiles in flight per wave, as in sweep_update_dot (on the 4096^2
  // Laplacian, m = 40: 3.27 -> 2.92 ms of lifts per solve, although the
  // 12-column kernel then holds 2 waves per SIMD instead of 6)
  struct Tile {
    double v[MAXK][R];
  };
  auto load = [&](long long t, Tile& T) {
    const long long tc = t < tr.ntiles ? t : tr.ntiles - 1;
    const long long rc = clamp_row<R>(tr, tile_row<R>(tr, tc, lane));
#pragma unroll
    for (int k = 0; k < MAXK; ++k) load_rows_nt<R>(V + col[k] + rc, T.v[k]);
  };
  auto use = [&](long long t, const Tile& T) {
    const long long r = tile_row<R>(tr, t, lane);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double p = 0.0;
#pragma unroll
      for (int k = 0; k < MAXK; ++k) p += c[k] * T.v[k][q];
      part[buf][wave][R * lane + q] = p;
    }
    __syncthreads();
    if (wave == 0 && r < tr.rend) {
      const bool all = (r + R - 1) < tr.rend;
      double a[R];
#pragma unroll
      for (int q = 0; q < R; ++q)
        a[q] = (part[buf][0][R * lane + q] + part[buf][1][R * lane + q]) +
               (part[buf][2][R * lane + q] + part[buf][3][R * lane + q]);
      if (!all) a[R - 1] = 0.0;
      store_rows<R>(y + r, a, all);
#pragma unroll
      for (int q = 0; q < R; ++q) nrm += a[q] * a[q];
    }
    buf ^= 1;
  };
  if (tr.t0 < tr.ntiles) {
    Tile A, B;
    long long t = tr.t0;
    load(t, A);
    load(t + tr.ts, B);
    for (;;) {
      use(t, A);
      load(t + 2 * tr.ts, A);
      t += tr.ts;
      if (t >= tr.ntiles) break;
      use(t, B);
      load(t + 2 * tr.ts, B);
      t += tr.ts;
      if (t >= tr.ntiles) break;
    }
  }
  const double sn = wave_sum(nrm);
  if (threadIdx.x == 0) partials[pidx] = sn;
}

// Dispatch a sweep on the column count: MAXK = ceil(count / 4) rounded up to
// an even bucket (2, 4, ..., 16), so up to 64 columns.
#define MEK_SWEEP_BUCKETS(count, LAUNCH) \
  do {                                   \
    const int mek_c_ = (count);          \
    if (mek_c_ <= 8) LAUNCH(2);          \
    else if (mek_c_ <= 16) LAUNCH(4);    \
    else if (mek_c_ <= 24) LAUNCH(6);    \
    else if (mek_c_ <= 32) LAUNCH(8);    \
    else if (mek_c_ <= 40) LAUNCH(10);   \
    else if (mek_c_ <= 48) LAUNCH(12);   \
    else if (mek_c_ <= 56) LAUNCH(14);   \
    else LAUNCH(16);                     \
  } while (0)

constexpr int kSweepMaxCols = 64;

}  // namespace mek
