// s-step (block) Arnoldi kernels for one problem on gfx950 (wave64).
//
// The column-by-column path (arnoldi.hip) reads the basis twice per Arnoldi
// column (dots, then the update).  A block builds s columns at a time and
// reads the basis twice per block:
//
//   p_i = (A - theta_i I) p_{i-1} / sigma_i, p_0 = q_{k-1}      s SpMVs
//   C   = [Q_k P]^T P                                          dots sweep
//   P'  = P - Q_k C_{Q_k}, with [Q_k P']^T P' on the side      update sweep
//   (P'' = P' - Q_k C2, when the side dots show that P' is not yet
//    orthogonal to Q_k to 1e-12 of its norm: the second pass)
//   Q_new = P' R^-1, R^T R = P'^T P' (Cholesky QR), with Q_new^T Q_new on the
//    side (Q_new R2^-1 once more when that Gram matrix is off I by > 1e-12)
//
// When every column of P keeps at least a tenth of its norm in P' (the
// DGKS test), P'^T P' follows from the dots by Pythagoras, P^T P - C1^T C1,
// and the update writes Q_new = (P - Q_k C1) R^-1 at once; its side dots are
// then Q_k^T Q_new and Q_new^T Q_new, and a Q_new that passes both checks
// (1e-12) is accepted without the normalising sweep.  One that does not goes
// on as a P' would: a second update, or a second Cholesky QR sweep.
//
// and recovers the new Hessenberg columns from the change of basis: with
// [p_0 .. p_s] = [Q_k Q_new] T and A [p_0 .. p_{s-1}] = [p_0 .. p_s] B (B
// bidiagonal: theta_i on the diagonal, sigma_i below it), the columns k-1 ..
// k+s-2 of H are (T B - H_old U) W^-1, U and W the upper and lower parts of
// T's first s columns (W upper triangular).
//
// The sweeps use v_mfma_f64_16x16x4f64: operand A[m][k] sits in lane
// m + 16 k, B[k][n] in lane n + 16 k, and result element q of lane l is
// D[l / 16 + 4 q][l % 16].  A wave takes 16-row chunks; in the dots the K
// index runs over the chunk's rows (lane group g holds rows 4g .. 4g+3, so
// each lane loads 32 contiguous bytes of a column), in the update over the
// window's columns, with the rows of the A operand permuted (row 4 (i % 4) +
// i / 4 in lane i) so that lane (g, j) receives rows 4g .. 4g+3 of output
// column j: exactly the B operand layout of the side dots, which therefore
// read the new rows straight from the registers.
//
// Every reduction has a fixed order (MFMA per chunk, chunks in order per
// wave, waves in order per workgroup, workgroups in order in the reducer), so
// a solve gives the same bits every time.
#include "common.hpp"
#include "launch.hpp"

namespace mek {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bool ss_skip(const SsArgs& a, int cond) {
  return a.iscal[kBreakdown] || a.iscal[kSStepFail] || (cond >= 0 && !a.fl[cond]);
}

// Per-workgroup partial of quantity (c, j): the four waves' tiles summed in
// wave order through the LDS scratch red (>= 1024 NT doubles); lane (g, j)
// of tile ct holds D[16 ct + g + 4 q][j] in element q.
template <int NT>
__device__ __forceinline__ void write_tile_partials(const dbl4 (&acc)[NT], int ncol, int s, double* red,
                                                    double* partials, int grid) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int ct = 0; ct < NT; ++ct)
#pragma unroll
    for (int q = 0; q < 4; ++q) red[(wave * NT * 4 + ct * 4 + q) * kWave + lane] = acc[ct][q];
  __syncthreads();
  for (int e = threadIdx.x; e < NT * 4 * kWave; e += kBlock) {
    const int l = e % kWave, tq = e / kWave;
    const int ct = tq / 4, q = tq % 4;
    const int c = 16 * ct + l / 16 + 4 * q, j = l % 16;
    constexpr int W = NT * 4 * kWave;  // one wave's share
    if (c < ncol && j < s)
      partials[static_cast<long long>(c * kSsMaxS + j) * grid + blockIdx.x] =
          (red[e] + red[W + e]) + (red[2 * W + e] + red[3 * W + e]);
  }
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

constexpr int kSsRT = 128;        // rows of a block-sweep tile (two per lane)
constexpr int kSsLds = kSsRT + 2;  // LDS stride of one staged column (doubles)

// This wave's columns (wave, wave + 4, ...) of the window [c0, c0 + ncol),
// rows [r0, r0 + 128) of the tile, two per lane with non-temporal 16-byte
// loads (1 KB of one column per instruction); zeros past the window and
// past the swept rows.
template <int NT>
__device__ __forceinline__ void ss_load_tile(const SsArgs& a, int c0, int ncol, long long r0,
                                             dbl2 (&reg)[4 * NT]) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const long long r = r0 + 2 * lane;
  const bool ok = r < a.rows;
#pragma unroll
  for (int u = 0; u < 4 * NT; ++u) {
    const int c = wave + 4 * u;
    reg[u] = (c < ncol && ok)
                 ? __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(a.V + static_cast<long long>(c0 + c) * a.ld + r))
                 : dbl2{0.0, 0.0};
  }
}

template <int NT>
__device__ __forceinline__ void ss_store_tile(double (*T)[kSsLds], const dbl2 (&reg)[4 * NT]) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int u = 0; u < 4 * NT; ++u) *reinterpret_cast<dbl2*>(&T[wave + 4 * u][2 * lane]) = reg[u];
}

// Sum of a quantity's per-workgroup partials in workgroup order.
__device__ __forceinline__ double ss_block_sum(const double* p, int pgrid, double* red) {
  double acc = 0.0;
  for (int g = threadIdx.x; g < pgrid; g += kBlock) acc += p[g];
  acc = wave_sum(acc);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = acc;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace

// ---------------------------------------------------------------- kernels

// Raw dots D[c][j] = V[:, c] . V[:, pcol + j] for c < ncol, j < s.  Each
// tile of 128 rows is staged in LDS with coalesced loads (the next tile's
// loads are in flight during this tile's MFMAs); wave w then takes rows
// 32w .. 32w+31 of the tile, lane group g the rows 4u + g of each 4-row step.
template <int NT>
__global__ __launch_bounds__(kBlock) void k_ss_dots(SsArgs a, int ncol, int pcol, int s) {
  if (ss_skip(a, -1)) return;
  __shared__ double T[16 * NT][kSsLds];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int g = lane / 16, i = lane % 16;
  const int pi = pcol + (i < s ? i : 0);
  const double bmask = i < s ? 1.0 : 0.0;
  dbl4 acc[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
  const long long ntiles = (a.rows + kSsRT - 1) / kSsRT;
  dbl2 reg[4 * NT];
  long long t = blockIdx.x;
  if (t < ntiles) ss_load_tile<NT>(a, 0, ncol, t * kSsRT, reg);
  for (; t < ntiles; t += gridDim.x) {
    ss_store_tile<NT>(T, reg);
    __syncthreads();
    if (t + gridDim.x < ntiles) ss_load_tile<NT>(a, 0, ncol, (t + gridDim.x) * kSsRT, reg);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int row = 32 * wave + 4 * u + g;
      const double bv = T[pi][row] * bmask;
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[ct] = mfma4(T[16 * ct + i][row], bv, acc[ct]);
    }
    __syncthreads();
  }
  write_tile_partials<NT>(acc, ncol, s, &T[0][0], a.partials, gridDim.x);
}

// Out = V[:, c0 .. c0+ncol) M (M: ncol x s raw-column coefficients at
// st[kSsM]) written over V[:, out0 .. out0+s) (a range inside the window),
// and the side dots D[c][j] = Vnew[:, c0+c] . Out[:, j] over the window with
// the new values in the output columns.  Per 128-row tile staged in LDS:
// the update MFMAs (each wave two 16-row blocks, K over the window's
// columns), the new rows written into the staged output columns, stored
// from there with coalesced 16-byte stores, and the side dots read from the
// same tile.
template <int NT>
__global__ __launch_bounds__(kBlock) void k_ss_update(SsArgs a, int c0, int ncol, int out0, int s,
                                                      int cond) {
  if (ss_skip(a, cond)) return;
  __shared__ double T[16 * NT][kSsLds];
  constexpr int KT = 4 * NT;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int g = lane / 16, i = lane % 16;
  const int ol = out0 - c0;
  const int oi = ol + (i < s ? i : 0);
  const double bmask = i < s ? 1.0 : 0.0;
  double bm[KT];
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    const int c = 4 * t + g;
    bm[t] = (c < ncol && i < s) ? a.st[kSsM + c * kSsMaxS + i] : 0.0;
  }
  dbl4 acc[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
  const long long ntiles = (a.rows + kSsRT - 1) / kSsRT;
  dbl2 reg[4 * NT];
  long long t = blockIdx.x;
  if (t < ntiles) ss_load_tile<NT>(a, c0, ncol, t * kSsRT, reg);
  for (; t < ntiles; t += gridDim.x) {
    ss_store_tile<NT>(T, reg);
    __syncthreads();
    if (t + gridDim.x < ntiles) ss_load_tile<NT>(a, c0, ncol, (t + gridDim.x) * kSsRT, reg);
    dbl4 o[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int rb = 32 * wave + 16 * h;
      o[h] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kt = 0; kt < KT; ++kt) o[h] = mfma4(T[4 * kt + g][rb + i], bm[kt], o[h]);
    }
    __syncthreads();  // every wave is done reading the old tile
    if (i < s) {
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int q = 0; q < 4; ++q) T[ol + i][32 * wave + 16 * h + g + 4 * q] = o[h][q];
    }
    __syncthreads();
    const long long r = t * kSsRT + 2 * lane;
    if (r < a.rows)
      for (int j = wave; j < s; j += kWavesPerBlock)
        __builtin_nontemporal_store(*reinterpret_cast<const dbl2*>(&T[ol + j][2 * lane]),
                                    reinterpret_cast<dbl2*>(a.V + static_cast<long long>(out0 + j) * a.ld + r));
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int row = 32 * wave + 4 * u + g;
      const double bv = T[oi][row] * bmask;
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[ct] = mfma4(T[16 * ct + i][row], bv, acc[ct]);
    }
    __syncthreads();
  }
  write_tile_partials<NT>(acc, ncol, s, &T[0][0], a.partials, gridDim.x);
}

// D[c][j] = sum of the block partials of quantity (c, j), one workgroup each
// (blockIdx.x = c * s + j).
__global__ __launch_bounds__(kBlock) void k_ss_reduce(SsArgs a, int s, int pgrid, int cond) {
  if (ss_skip(a, cond)) return;
  __shared__ double red[kWavesPerBlock];
  const int c = blockIdx.x / s, j = blockIdx.x % s;
  const double v = ss_block_sum(a.partials + static_cast<long long>(c * kSsMaxS + j) * pgrid, pgrid, red);
  if (threadIdx.x == 0) a.st[kSsDots + c * kSsMaxS + j] = v;
}

// sigma_i = ||y_i|| from the SpMV's partials: colscale of column k + i, and
// at i = 0 the block's flags cleared.
__global__ __launch_bounds__(kBlock) void k_ss_sigma(SsArgs a, int k, int i, int pgrid) {
  if (a.iscal[kBreakdown] || a.iscal[kSStepFail]) return;
  __shared__ double red[kWavesPerBlock];
  const double s2 = ss_block_sum(a.spmv_partials, pgrid, red);
  if (threadIdx.x == 0) {
    const double sig = sqrt(s2);
    if (!(sig > 1e-280) || isinf(sig)) {
      a.iscal[kSStepFail] = 1;
      a.fl[kSsFailK] = k;
    } else {
      a.colscale[k + i] = 1.0 / sig;
      a.st[kSsSigma + i] = sig;
    }
    if (i == 0) {
      a.fl[kSsPass2] = 0;
      a.fl[kSsNormReady] = 0;
      a.fl[kSsNorm2] = 0;
      a.fl[kSsPasses] = 1;
      a.fl[kSsFused] = 0;
    }
  }
}

namespace {

// Upper Cholesky factor R (R^T R = G) of an s x s matrix and its inverse;
// false when a pivot falls below 1e-28 (the block has no new direction left
// at the column-by-column path's breakdown scale) or is not finite.
__device__ bool chol_inv(const double (*G)[kSsMaxS], int s, double (*R)[kSsMaxS], double (*Ri)[kSsMaxS]) {
  for (int r = 0; r < s; ++r)
    for (int c = 0; c < s; ++c) R[r][c] = Ri[r][c] = 0.0;
  for (int j = 0; j < s; ++j) {
    double d = G[j][j];
    for (int l = 0; l < j; ++l) d -= R[l][j] * R[l][j];
    if (!(d > 1e-28) || isinf(d)) return false;
    const double rjj = sqrt(d);
    R[j][j] = rjj;
    for (int c = j + 1; c < s; ++c) {
      double x = G[j][c];
      for (int l = 0; l < j; ++l) x -= R[l][j] * R[l][c];
      R[j][c] = x / rjj;
    }
  }
  for (int c = 0; c < s; ++c) {  // R Ri = I, column c by back substitution
    for (int r = c; r >= 0; --r) {
      double x = (r == c) ? 1.0 : 0.0;
      for (int l = r + 1; l <= c; ++l) x -= R[r][l] * Ri[l][c];
      Ri[r][c] = x / R[r][r];
    }
  }
  return true;
}

}  // namespace

// The block's small algebra (one workgroup).  Phase A: from the dots, C1 and
// the coefficients of the update.  Phase B: from the update's side dots,
// either the second update (call 0 only) or the Cholesky factor and the
// coefficients of the normalising sweep.  Phase D: from the normalising
// sweep's Gram matrix, either a second Cholesky QR sweep (call 0 only) or the
// Hessenberg columns.  Any failure sets iscal[kSStepFail] (and records k).
// The operands are staged in LDS first; every loop over the block's entries
// is spread over the workgroup.
__global__ __launch_bounds__(kBlock) void k_ss_plan(SsArgs a, int phase, int call, int k, int s,
                                                    int force2, int cond) {
  if (ss_skip(a, cond)) return;
  __shared__ double sD[kSsMaxCols][kSsMaxS];  // reduced dots
  __shared__ double sC[kSsMaxCols][kSsMaxS];  // C_tot
  __shared__ double sX[kSsMaxCols][kSsMaxS];  // T B - H_old U
  __shared__ double sH[kSsMaxCols][kSsMaxCols - 1];  // H_old rows < k, columns < k - 1
  __shared__ double sCs[kSsMaxCols];          // column scales of the window
  __shared__ double G[kSsMaxS][kSsMaxS];
  __shared__ double R[kSsMaxS][kSsMaxS];
  __shared__ double Ri[kSsMaxS][kSsMaxS];
  __shared__ int need_col[kSsMaxS];
  __shared__ int state;  // 0 go on, 1 second pass, 2 failed
  const int tid = threadIdx.x;
  double* Ct = a.st + kSsCtot;
  double* M = a.st + kSsM;
  double* Rt = a.st + kSsR;
  const int nd = (phase == kSsPhaseD) ? s : k + s;  // rows of the reduced dots
  for (int e = tid; e < kSsMaxCols * kSsMaxS; e += kBlock) {
    const int c = e / kSsMaxS, j = e % kSsMaxS;
    sD[c][j] = (c < nd && j < s) ? a.st[kSsDots + e] : 0.0;
  }
  for (int c = tid; c < kSsMaxCols; c += kBlock) sCs[c] = (c < k + s) ? a.colscale[c] : 0.0;
  __syncthreads();
  if (phase == kSsPhaseA) {
    // true C1 = Q_k^T P; update coefficients over the window [0, k+s)
    for (int e = tid; e < kSsMaxCols * kSsMaxS; e += kBlock) {
      const int c = e / kSsMaxS, j = e % kSsMaxS;
      double mv = 0.0;
      if (j < s && c < k + s) {
        if (c < k) {
          const double c1 = sD[c][j] * sCs[c] * sCs[k + j];
          Ct[e] = c1;
          sC[c][j] = c1;
          mv = -c1 * sCs[c];
          if (!isfinite(c1)) {
            a.iscal[kSStepFail] = 1;
            a.fl[kSsFailK] = k;
          }
        } else {
          mv = (c - k == j) ? sCs[k + j] : 0.0;
        }
      }
      M[e] = mv;
    }
    if (force2) return;
    // P'^T P' = P^T P - C1^T C1 (true columns) and the DGKS test per column
    __syncthreads();
    if (tid < kSsMaxS * kSsMaxS) {
      const int l = tid / kSsMaxS, j = tid % kSsMaxS;
      double g = 0.0;
      if (l < s && j < s) {
        g = sD[k + l][j] * sCs[k + l] * sCs[k + j];
        for (int c = 0; c < k; ++c) g -= sC[c][l] * sC[c][j];
      }
      G[l][j] = g;
    }
    __syncthreads();
    if (tid == 0) {
      bool ok = true;
      for (int j = 0; j < s && ok; ++j) {
        const double p2 = sD[k + j][j] * sCs[k + j] * sCs[k + j];
        ok = G[j][j] >= 0.01 * p2;
      }
      ok = ok && chol_inv(G, s, R, Ri);
      if (ok) {
        double lo = R[0][0], hi = R[0][0];
        for (int j = 1; j < s; ++j) {
          lo = fmin(lo, R[j][j]);
          hi = fmax(hi, R[j][j]);
        }
        ok = hi <= 1e7 * lo;
      }
      state = ok ? 1 : 0;
    }
    __syncthreads();
    if (state != 1) return;
    // Q_new = (P - Q_k C1) R^-1: coefficients of the raw window columns
    for (int e = tid; e < kSsMaxCols * kSsMaxS; e += kBlock) {
      const int c = e / kSsMaxS, j = e % kSsMaxS;
      double mv = 0.0;
      if (j < s && c < k + s) {
        if (c < k) {
          double x = 0.0;
          for (int l = 0; l <= j; ++l) x += sC[c][l] * Ri[l][j];
          mv = -x * sCs[c];
        } else if (c - k <= j) {
          mv = Ri[c - k][j] * sCs[c];
        }
      }
      M[e] = mv;
    }
    if (tid < kSsMaxS * kSsMaxS) {
      const int l = tid / kSsMaxS, j = tid % kSsMaxS;
      Rt[tid] = R[l][j];
    }
    if (tid == 0) a.fl[kSsFused] = 1;
    return;
  }
  if (phase == kSsPhaseB) {
    // side dots of P' (stored true, column scale 1 from here on); after a
    // fused first update (call 0) they are those of Q_new = P' R^-1, and R
    // (R_tot so far) is staged in sX
    const bool fused = a.fl[kSsFused] != 0;
    if (tid < kSsMaxS * kSsMaxS) {
      const int l = tid / kSsMaxS, j = tid % kSsMaxS;
      G[l][j] = (l < s && j < s) ? sD[k + l][j] : 0.0;
      sX[l][j] = Rt[tid];
    }
    if (tid < s) {
      // second update due when ||Q_k^T p'_j|| > 1e-12 ||p'_j|| (also on NaN)
      double n2 = 0.0;
      for (int c = 0; c < k; ++c) {
        const double x = sD[c][tid] * sCs[c];
        n2 += x * x;
      }
      const double g2 = sD[k + tid][tid];
      need_col[tid] = !(n2 <= 1e-24 * g2);
      a.colscale[k + tid] = 1.0;
    }
    __syncthreads();
    if (tid == 0) {
      bool need = force2 && call == 0;
      for (int j = 0; j < s; ++j) need = need || need_col[j];
      if (need) {
        state = (call == 0) ? 1 : 2;
      } else if (fused && call == 0) {
        // Q_new orthogonal to Q_k: accept it when Q_new^T Q_new is I to
        // 1e-12, else one Cholesky QR sweep more
        double err = 0.0;
        for (int l = 0; l < s; ++l)
          for (int j = 0; j < s; ++j) {
            const double e = fabs(G[l][j] - (l == j ? 1.0 : 0.0));
            err = (e > err || !(e == e)) ? e : err;
          }
        if (err <= 1e-12) state = 3;
        else state = (err <= 1e-2 && chol_inv(G, s, R, Ri)) ? 0 : 2;
      } else {
        state = chol_inv(G, s, R, Ri) ? 0 : 2;
        if (state == 0) {
          // a block whose new directions are nearly dependent (pivot ratio
          // above 1e7) would carry that conditioning into H: refuse it
          double lo = R[0][0], hi = R[0][0];
          for (int j = 1; j < s; ++j) {
            lo = fmin(lo, R[j][j]);
            hi = fmax(hi, R[j][j]);
          }
          if (!(hi <= 1e7 * lo)) state = 2;
        }
      }
    }
    __syncthreads();
    if (state == 2) {
      if (tid == 0) {
        a.iscal[kSStepFail] = 1;
        a.fl[kSsFailK] = k;
      }
      return;
    }
    if (state == 3) goto accepted;
    if (state == 1) {
      // second update: P'' = P' - Q_k C2; C_tot += C2 (after a fused
      // update the stored columns are Q_new = P' R^-1, so C_tot += C2 R)
      for (int e = tid; e < kSsMaxCols * kSsMaxS; e += kBlock) {
        const int c = e / kSsMaxS, j = e % kSsMaxS;
        double mv = 0.0;
        if (j < s && c < k + s) {
          if (c < k) {
            const double c2 = sD[c][j] * sCs[c];
            if (fused) {
              double x = 0.0;
              for (int l = 0; l <= j; ++l) x += sD[c][l] * sCs[c] * sX[l][j];
              Ct[e] += x;
            } else {
              Ct[e] += c2;
            }
            mv = -c2 * sCs[c];
          } else {
            mv = (c - k == j) ? 1.0 : 0.0;
          }
        }
        M[e] = mv;
      }
      if (tid == 0) {
        a.fl[kSsPass2] = 1;
        a.fl[kSsPasses] = 2;
      }
      return;
    }
    // normalising sweep over the window [k, k+s): Q_new = P' R^-1 (after a
    // fused update, a second Cholesky QR of its Q_new); R_tot = R R_tot then
    for (int e = tid; e < kSsMaxCols * kSsMaxS; e += kBlock) {
      const int l = e / kSsMaxS, j = e % kSsMaxS;
      M[e] = (l < s && j < s) ? Ri[l][j] : 0.0;
    }
    if (tid < kSsMaxS * kSsMaxS) {
      const int l = tid / kSsMaxS, j = tid % kSsMaxS;
      double x = R[l][j];
      if (fused) {
        x = 0.0;
        for (int q = l; q <= j && q < s; ++q) x += R[l][q] * sX[q][j];
      }
      Rt[tid] = x;
    }
    if (tid == 0) {
      a.fl[kSsNormReady] = 1;
      a.fl[kSsPass2] = 0;
      if (fused && call == 0) a.fl[kSsPasses] = 2;
    }
    return;
  }
  // phase D: Gram matrix of Q_new
  if (tid < kSsMaxS * kSsMaxS) {
    const int l = tid / kSsMaxS, j = tid % kSsMaxS;
    G[l][j] = (l < s && j < s) ? sD[l][j] : 0.0;
  }
  __syncthreads();
  if (tid == 0) {
    double err = 0.0;
    for (int l = 0; l < s; ++l)
      for (int j = 0; j < s; ++j) {
        const double e = fabs(G[l][j] - (l == j ? 1.0 : 0.0));
        err = (e > err || !(e == e)) ? e : err;
      }
    if (!(err <= 1e-12)) {
      if (call == 0 && err <= 1e-2 && chol_inv(G, s, R, Ri)) state = 1;
      else state = (call == 1 && err <= 1e-10) ? 0 : 2;
    } else {
      state = 0;
    }
  }
  __syncthreads();
  if (state == 2) {
    if (tid == 0) {
      a.iscal[kSStepFail] = 1;
      a.fl[kSsFailK] = k;
    }
    return;
  }
  if (state == 1) {
    // second Cholesky QR: Q_new <- Q_new R2^-1, R_tot <- R2 R_tot
    for (int e = tid; e < kSsMaxCols * kSsMaxS; e += kBlock) {
      const int l = e / kSsMaxS, j = e % kSsMaxS;
      M[e] = (l < s && j < s) ? Ri[l][j] : 0.0;
    }
    if (tid < kSsMaxS * kSsMaxS) {
      const int l = tid / kSsMaxS, j = tid % kSsMaxS;
      double x = 0.0;
      for (int q = l; q <= j && q < s; ++q) x += R[l][q] * Rt[q * kSsMaxS + j];
      G[l][j] = x;  // G is free again: stage the product
    }
    __syncthreads();
    if (tid < kSsMaxS * kSsMaxS) Rt[tid] = G[tid / kSsMaxS][tid % kSsMaxS];
    if (tid == 0) {
      a.fl[kSsNorm2] = 1;
      a.fl[kSsNormReady] = 0;
    }
    return;
  }
  // accepted: Hessenberg columns k-1 .. k+s-2 = (T B - H_old U) W^-1, with
  // [p_0 .. p_s] = [Q_k Q_new] T
accepted:
  const double* th = a.st + kSsTheta;
  const double* sg = a.st + kSsSigma;
  const int ldh = a.ldh;
  for (int e = tid; e < kSsMaxCols * kSsMaxS; e += kBlock) sC[e / kSsMaxS][e % kSsMaxS] = Ct[e];
  if (tid < kSsMaxS * kSsMaxS) R[tid / kSsMaxS][tid % kSsMaxS] = Rt[tid];
  for (int e = tid; e < k * (k - 1); e += kBlock) {
    const int r = e % k, q = e / k;
    sH[r][q] = a.H[static_cast<long long>(q) * ldh + r];
  }
  __syncthreads();
  auto T = [&](int r, int c) -> double {
    if (c == 0) return r == k - 1 ? 1.0 : 0.0;
    if (r < k) return sC[r][c - 1];
    return (r - k <= c - 1) ? R[r - k][c - 1] : 0.0;
  };
  for (int e = tid; e < (k + s) * s; e += kBlock) {
    const int r = e / s, c = e % s;
    double x = T(r, c) * th[c] + T(r, c + 1) * sg[c];
    if (r < k && c >= 1)
      for (int q = 0; q < k - 1; ++q) x -= sH[r][q] * sC[q][c - 1];
    sX[r][c] = x;
  }
  __syncthreads();
  for (int r = tid; r < k + s; r += kBlock) {
    double hn[kSsMaxS];
    for (int c = 0; c < s; ++c) {
      double x = sX[r][c];
      for (int l = 0; l < c; ++l) x -= hn[l] * T(k - 1 + l, c);
      hn[c] = x / T(k - 1 + c, c);
    }
    for (int c = 0; c < s; ++c) a.H[static_cast<long long>(k - 1 + c) * ldh + r] = (r <= k + c) ? hn[c] : 0.0;
  }
  for (int e = tid; e < s * ldh; e += kBlock) {
    const int c = e / ldh, r = e % ldh;
    if (r >= k + s) a.H[static_cast<long long>(k - 1 + c) * ldh + r] = 0.0;
  }
  if (tid == 0) {
    a.iscal[kKUsed] = k + s - 1;
    a.fl[kSsNormReady] = 0;
    a.fl[kSsNorm2] = 0;
    a.fl[kSsBlocks] += 1;
    if (a.fl[kSsPasses] > 1 || call == 1) a.fl[kSsSecond] += 1;
    if (phase == kSsPhaseB) a.fl[kSsFusedBlocks] += 1;
  }
}

// ------------------------------------------------------------ launchers

#define MEK_SS_TILES(ncol, LAUNCH) \
  do {                             \
    const int mek_n_ = (ncol);     \
    if (mek_n_ <= 16) LAUNCH(1);   \
    else if (mek_n_ <= 32) LAUNCH(2); \
    else if (mek_n_ <= 48) LAUNCH(3); \
    else LAUNCH(4);                \
  } while (0)

void launch_ss_sigma(const ArnoldiBufs& b, const SsArgs& a, int k, int i, hipStream_t s) {
  hipLaunchKernelGGL(k_ss_sigma, dim3(1), dim3(kBlock), 0, s, a, k, i, b.spmv_grid);
}

void launch_ss_dots(const SsArgs& a, int k, int s, hipStream_t st) {
#define MEK_L(NT) hipLaunchKernelGGL(k_ss_dots<NT>, dim3(a.grid), dim3(kBlock), 0, st, a, k + s, k, s)
  MEK_SS_TILES(k + s, MEK_L);
#undef MEK_L
}

void launch_ss_update(const SsArgs& a, int c0, int ncol, int out0, int s, int cond, hipStream_t st) {
#define MEK_L(NT) \
  hipLaunchKernelGGL(k_ss_update<NT>, dim3(a.grid), dim3(kBlock), 0, st, a, c0, ncol, out0, s, cond)
  MEK_SS_TILES(ncol, MEK_L);
#undef MEK_L
}

void launch_ss_reduce(const SsArgs& a, int ncol, int s, int cond, hipStream_t st) {
  hipLaunchKernelGGL(k_ss_reduce, dim3(ncol * s), dim3(kBlock), 0, st, a, s, a.grid, cond);
}

void launch_ss_plan(const SsArgs& a, int phase, int call, int k, int s, int force_pass2, int cond,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_ss_plan, dim3(1), dim3(kBlock), 0, st, a, phase, call, k, s, force_pass2, cond);
}

}  // namespace mek
