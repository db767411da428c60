// Block Krylov kernels for 32 right-hand sides on one shared A (gfx950).
//
// Block Arnoldi with classical block Gram-Schmidt and a second pass (BCGS2),
// the intra-block orthonormalisation by Cholesky QR, and the block Hessenberg
// exponential by Taylor / Paterson-Stockmeyer with scaling and squaring on a
// dense FP64 MFMA GEMM.  The basis blocks are stored raw (P_i, row-major
// n_pad x 32) with their triangular scales S_i = R_i^{-1} applied inside the
// small 32 x 32 coefficient algebra, so no sweep rescales a block.
//
// MFMA: v_mfma_f64_16x16x4f64.  Operand A[m][k] sits in lane m + 16 k, B[k][n]
// in lane n + 16 k; result element q of lane l is D[l / 16 + 4 q][l % 16].
#include <algorithm>

#include "block.hpp"

namespace mek {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ dbl4 zero4() { return dbl4{0.0, 0.0, 0.0, 0.0}; }

// Orders a wave's LDS stores before its later LDS loads (and the loads before
// the next stores) when the LDS region is private to the wave: no workgroup
// barrier is needed then.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace

// ------------------------------------------------------------------ SpMM

// Out[r, c] = sum_k A[r, k] X[k, c] for 32 columns c: a quarter-wave per
// row, two right-hand sides per lane, so every gathered row of X is one
// 256-byte segment read with 16-byte loads and one wave instruction carries
// four rows (the gather instruction rate bounds this kernel: one column per
// lane took 1.6x as long).  Each column is summed in storage order with
// separately rounded products (the same operations as a sequential CSR loop
// per column).  Workgroups own contiguous row ranges; rows in [n, n_pad) are
// written as zeros.
__global__ __launch_bounds__(kBlock) void k_spmm(BlockBufs b, const double* __restrict__ X,
                                                 double* __restrict__ Out) {
#pragma clang fp contract(off)
  if (b.flags[0]) return;
  constexpr int kLanesPerRow = kRhs / 2;
  constexpr int kRowsPerPass = kBlock / kLanesPerRow;
  const int c2 = 2 * (threadIdx.x % kLanesPerRow);  // this lane's columns c2, c2 + 1
  const long long rows_per_blk =
      ((b.n_pad + gridDim.x - 1) / gridDim.x + kRowsPerPass - 1) / kRowsPerPass * kRowsPerPass;
  const long long r0 = static_cast<long long>(xcd_block_id()) * rows_per_blk;
  const long long r1 = r0 + rows_per_blk < b.n_pad ? r0 + rows_per_blk : b.n_pad;
  for (long long r = r0 + threadIdx.x / kLanesPerRow; r < r1; r += kRowsPerPass) {
    double a0 = 0.0, a1 = 0.0;
    if (r < b.n) {
      const int k0 = b.row_ptr[r], k1 = b.row_ptr[r + 1];
      int k = k0;
      for (; k + 4 <= k1; k += 4) {  // four gathers in flight, summed in order
        double v[4];
        double2 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v[u] = b.vals[k + u];
          x[u] = *reinterpret_cast<const double2*>(X + static_cast<long long>(b.col_idx[k + u]) * kRhs + c2);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double p0 = v[u] * x[u].x;
          const double p1 = v[u] * x[u].y;
          a0 = a0 + p0;
          a1 = a1 + p1;
        }
      }
      for (; k < k1; ++k) {
        const double vv = b.vals[k];
        const double2 xx = *reinterpret_cast<const double2*>(X + static_cast<long long>(b.col_idx[k]) * kRhs + c2);
        const double p0 = vv * xx.x;
        const double p1 = vv * xx.y;
        a0 = a0 + p0;
        a1 = a1 + p1;
      }
    }
    __builtin_nontemporal_store(a0, Out + r * kRhs + c2);
    __builtin_nontemporal_store(a1, Out + r * kRhs + c2 + 1);
  }
}

void block_launch_spmm(const BlockBufs& b, const double* X, double* Out, hipStream_t s) {
  hipLaunchKernelGGL(k_spmm, dim3(b.spmm_grid > 0 ? b.spmm_grid : 4096), dim3(kBlock), 0, s, b, X, Out);
}

// ---------------------------------------------------------- block dots

// Partial C_i = X1_i^T X2 over one chunk of kBlkRows rows for the blocks
// i0 <= i < i0 + cnt: block i < nb of X1, or (i == nb) X2 itself for its
// Gram matrix.  Partials are laid out [block i][element e][chunk g], so the
// finaliser reads each element's chunks contiguously.

// Split-K form for a few blocks: logical workgroup L = g * cnt + (i - i0)
// (chunk g); consecutive logical ids run on one XCD, so the workgroups of a
// chunk share X2 through L2.  Each wave takes every 4th k-step of 4 rows; the
// four waves' 32 x 32 tiles are summed in LDS in wave order.
__global__ __launch_bounds__(kBlock) void k_bdots(BlockBufs b, const double* __restrict__ X1, int nb,
                                                  const double* __restrict__ X2, int i0, int cnt, int cond) {
  if (b.flags[0] || (cond && !b.flags[2])) return;
  __shared__ double red[kWavesPerBlock][kRhs * kRhs];
  const int L = xcd_block_id();
  const int i = i0 + L % cnt;
  const long long g = L / cnt;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int lr = lane / 16, lc = lane % 16;
  const double* __restrict__ A = (i < nb) ? X1 + static_cast<long long>(i) * b.n_pad * kRhs : X2;
  const long long row0 = g * kBlkRows;
  dbl4 c00 = zero4(), c01 = zero4(), c10 = zero4(), c11 = zero4();
  constexpr int kSteps = kBlkRows / 4;
  for (int st = wave; st < kSteps; st += 4 * kWavesPerBlock) {
    double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long r = row0 + 4LL * (st + u * kWavesPerBlock) + lr;
      a0[u] = A[r * kRhs + lc];
      a1[u] = A[r * kRhs + 16 + lc];
      b0[u] = X2[r * kRhs + lc];
      b1[u] = X2[r * kRhs + 16 + lc];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c00 = mfma(a0[u], b0[u], c00);
      c01 = mfma(a0[u], b1[u], c01);
      c10 = mfma(a1[u], b0[u], c10);
      c11 = mfma(a1[u], b1[u], c11);
    }
  }
  // tile (mh, nh), element q: C[16 mh + lr + 4 q][16 nh + lc]
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int m = lr + 4 * q;
    red[wave][m * kRhs + lc] = c00[q];
    red[wave][m * kRhs + 16 + lc] = c01[q];
    red[wave][(16 + m) * kRhs + lc] = c10[q];
    red[wave][(16 + m) * kRhs + 16 + lc] = c11[q];
  }
  __syncthreads();
  double* out = b.partials + static_cast<long long>(i) * (kRhs * kRhs) * b.nchunks + g;
  for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock)
    out[static_cast<long long>(e) * b.nchunks] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
}

// Wave-per-block form for kDotsWavesMin .. kDotsWavesMax blocks: one
// workgroup per chunk, wave w takes block i0 + w over all the chunk's rows
// (no split-K, no imbalance), and each 64-row slab of X2 is loaded from HBM
// once, into LDS (double buffered), and read from there by every wave.
// Relying on L2 instead (the split-K form) reads X2 about 1.2 times per
// block step; here it is read once.
constexpr int kDotsWavesMin = 8;
constexpr int kDotsWavesMax = 16;
constexpr int kZRows = 64;       // rows of an X2 slab
constexpr int kZLd = kRhs + 4;   // padded LDS row of the slab

__global__ __launch_bounds__(kWave * kDotsWavesMax) void k_bdots_w(BlockBufs b, const double* __restrict__ X1,
                                                                    int nb, const double* __restrict__ X2,
                                                                    int i0, int cond) {
  if (b.flags[0] || (cond && !b.flags[2])) return;
  __shared__ __align__(16) double zs[2][kZRows * kZLd];
  const long long g = xcd_block_id();
  const int tid = threadIdx.x;
  const int nt = blockDim.x;  // >= 64 kDotsWavesMin
  const int lane = tid % kWave;
  const int wave = tid / kWave;
  const int lr = lane / 16, lc = lane % 16;
  const long long row0 = g * kBlkRows;
  const int i = i0 + wave;
  const double* __restrict__ A = (i < nb) ? X1 + static_cast<long long>(i) * b.n_pad * kRhs : X2;
  // the four MFMA tiles take the even and odd columns of X1_i and of X2, so
  // one 16-byte load feeds two of them
  dbl4 cee = zero4(), ceo = zero4(), coe = zero4(), coo = zero4();
  // a slab is 1024 pairs of doubles; each thread moves at most two
  constexpr int kPairs = kZRows * kRhs / 2;
  static_assert(kPairs <= 2 * kWave * kDotsWavesMin, "two pairs per thread");
  double2 zr[2];
  auto load_slab = [&](int s) {
    const double* src = X2 + (row0 + static_cast<long long>(s) * kZRows) * kRhs;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int q = tid + nt * p;
      if (q < kPairs) zr[p] = *reinterpret_cast<const double2*>(src + 2 * q);
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int q = tid + nt * p;
      if (q < kPairs) {
        const int r = (2 * q) / kRhs, c = (2 * q) % kRhs;
        zs[buf][r * kZLd + c] = zr[p].x;
        zs[buf][r * kZLd + c + 1] = zr[p].y;
      }
    }
  };
  constexpr int kSlabs = kBlkRows / kZRows;
  load_slab(0);
  store_slab(0);
  __syncthreads();
  for (int s = 0; s < kSlabs; ++s) {
    const int buf = s & 1;
    if (s + 1 < kSlabs) load_slab(s + 1);  // in flight during this slab
#pragma unroll
    for (int kq = 0; kq < kZRows / 16; ++kq) {
      double2 av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long long r = row0 + static_cast<long long>(s) * kZRows + 16 * kq + 4 * u + lr;
        av[u] = *reinterpret_cast<const double2*>(A + r * kRhs + 2 * lc);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int rz = 16 * kq + 4 * u + lr;
        bv[u] = *reinterpret_cast<const double2*>(&zs[buf][rz * kZLd + 2 * lc]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cee = mfma(av[u].x, bv[u].x, cee);
        ceo = mfma(av[u].x, bv[u].y, ceo);
        coe = mfma(av[u].y, bv[u].x, coe);
        coo = mfma(av[u].y, bv[u].y, coo);
      }
    }
    // every wave is done with buffer buf ^ 1 (read during slab s - 1) since
    // the barrier that ended that slab
    if (s + 1 < kSlabs) store_slab(buf ^ 1);
    __syncthreads();
  }
  // tile element q of lane l is (m, n) = (l / 16 + 4 q, l % 16): column 2m
  // (+1) of X1_i against column 2n (+1) of X2
  double* out = b.partials + static_cast<long long>(i) * (kRhs * kRhs) * b.nchunks + g;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int m2 = 2 * (lr + 4 * q), n2 = 2 * lc;
    out[static_cast<long long>(m2 * kRhs + n2) * b.nchunks] = cee[q];
    out[static_cast<long long>(m2 * kRhs + n2 + 1) * b.nchunks] = ceo[q];
    out[static_cast<long long>((m2 + 1) * kRhs + n2) * b.nchunks] = coe[q];
    out[static_cast<long long>((m2 + 1) * kRhs + n2 + 1) * b.nchunks] = coo[q];
  }
}

// C[i][e] = the sum of the chunk partials: one wave per element, lane l
// summing chunks l, l + 64, ... in order, then the fixed-order wave sum.
__global__ __launch_bounds__(kBlock) void k_bdots_fin(BlockBufs b, int cond) {
  if (b.flags[0] || (cond && !b.flags[2])) return;
  constexpr int kPerBlk = kBlock / kWave;  // elements per workgroup
  const int i = blockIdx.x / (kRhs * kRhs / kPerBlk);
  const int e = (blockIdx.x % (kRhs * kRhs / kPerBlk)) * kPerBlk + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const double* p = b.partials + (static_cast<long long>(i) * (kRhs * kRhs) + e) * b.nchunks;
  double s = 0.0;
  for (int g = lane; g < b.nchunks; g += kWave) s += p[g];
  s = wave_sum(s);
  if (lane == 0) b.C[static_cast<long long>(i) * kRhs * kRhs + e] = s;
}

// Groups of kDotsWavesMin .. kDotsWavesMax blocks take the wave-per-block
// form (one launch each, X2 read once per launch); fewer remaining blocks take
// the split-K form.
void block_launch_dots_partial(const BlockBufs& b, const double* X1, int nb, const double* X2, int gram,
                               int cond, hipStream_t s) {
  const int ntot = nb + (gram ? 1 : 0);
  int i0 = 0;
  while (ntot - i0 >= kDotsWavesMin) {
    const int w = std::min(ntot - i0, kDotsWavesMax);
    hipLaunchKernelGGL(k_bdots_w, dim3(b.nchunks), dim3(kWave * w), 0, s, b, X1, nb, X2, i0, cond);
    i0 += w;
  }
  if (i0 < ntot) {
    const int cnt = ntot - i0;
    hipLaunchKernelGGL(k_bdots, dim3(cnt * b.nchunks), dim3(kBlock), 0, s, b, X1, nb, X2, i0, cnt, cond);
  }
}

void block_launch_dots_fin(const BlockBufs& b, int nb, int gram, int cond, hipStream_t s) {
  const int ntot = nb + (gram ? 1 : 0);
  hipLaunchKernelGGL(k_bdots_fin, dim3(ntot * (kRhs * kRhs / (kBlock / kWave))), dim3(kBlock), 0, s, b, cond);
}

void block_launch_dots(const BlockBufs& b, const double* X1, int nb, const double* X2, int gram, int cond,
                       hipStream_t s) {
  block_launch_dots_partial(b, X1, nb, X2, gram, cond, s);
  block_launch_dots_fin(b, nb, gram, cond, s);
}

// -------------------------------------------------------- block update

// Out = X - sum_{i<nb} P_i G_i (X null: Out = -sum ...).  A workgroup takes
// 256 rows at a time, a wave 64 of them (four 16-row MFMA tiles).  Per basis
// block the wave's 64 x 32 slab of P_i (16 KB, contiguous) is loaded with
// fully coalesced 16-byte loads into LDS (rows padded to 33 doubles against
// bank conflicts), and the MFMA A operands are read from there; G_i is loaded
// once as eight 4 x 32 B-operand slabs.  The loads of block i + 1's slab are
// issued as soon as block i's slab is in LDS, so they are in flight during
// block i's MFMAs (the MFMA time is about half the load time at this
// arithmetic intensity, 8 flop per byte).  Two waves per SIMD: the registers
// (slab, accumulators, B operands) fit 256 VGPRs, and LDS allows two
// workgroups per CU anyway.  Each wave's slab is private, so the LDS stores
// and loads are ordered within the wave and the waves of a workgroup never
// wait for each other.  In place (Out == X) is allowed: every element is read
// and written by one lane.
constexpr int kUpdPad = kRhs + 1;

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_bupdate(
    BlockBufs b, int nb, const double* X, double* Out, int cond) {
  if (b.flags[0] || (cond && !b.flags[2])) return;
  __shared__ double tile[kWavesPerBlock][64 * kUpdPad];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int lr = lane / 16, lc = lane % 16;
  const long long ntile256 = b.n_pad / 256;
  const long long bstride = b.n_pad * kRhs;  // doubles of one basis block
  double* my = tile[wave];
  for (long long T = blockIdx.x; T < ntile256; T += gridDim.x) {
    const long long rbase = T * 256 + 64LL * wave;
    dbl4 acc[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = zero4();
    const double* __restrict__ src = b.P + rbase * kRhs + 2 * lane;
    double2 v[16];  // the slab of block i, lane-interleaved
    if (nb > 0) {
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = *reinterpret_cast<const double2*>(src + u * 128);
    }
    for (int i = 0; i < nb; ++i) {
      const double* __restrict__ Gi = b.G + static_cast<long long>(i) * kRhs * kRhs;
      double g0[8], g1[8];
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        g0[ks] = Gi[(4 * ks + lr) * kRhs + lc];
        g1[ks] = Gi[(4 * ks + lr) * kRhs + 16 + lc];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = u * 128 + 2 * lane;  // element of the 64 x 32 slab
        const int row = e / kRhs, col = e % kRhs;
        my[row * kUpdPad + col] = v[u].x;
        my[row * kUpdPad + col + 1] = v[u].y;
      }
      if (i + 1 < nb) {
        const double* __restrict__ nsrc = src + (i + 1) * bstride;
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = *reinterpret_cast<const double2*>(nsrc + u * 128);
      }
      wave_lds_sync();
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double* prow = my + (16 * t + lc) * kUpdPad + lr;
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
          const double a = prow[4 * ks];
          acc[t][0] = mfma(a, g0[ks], acc[t][0]);
          acc[t][1] = mfma(a, g1[ks], acc[t][1]);
        }
      }
      wave_lds_sync();
    }
    // acc[t][nh] element q: row 16 t + lr + 4 q, column 16 nh + lc
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long long r = rbase + 16 * t + lr + 4 * q;
        const double x0 = X ? X[r * kRhs + lc] : 0.0;
        const double x1 = X ? X[r * kRhs + 16 + lc] : 0.0;
        Out[r * kRhs + lc] = x0 - acc[t][0][q];
        Out[r * kRhs + 16 + lc] = x1 - acc[t][1][q];
      }
    }
  }
}

void block_launch_update(const BlockBufs& b, int nb, const double* X, double* Out, int cond, hipStream_t s) {
  hipLaunchKernelGGL(k_bupdate, dim3(b.update_grid > 0 ? b.update_grid : 4096), dim3(kBlock), 0, s, b, nb, X,
                     Out, cond);
}

// ------------------------------------------------- small block algebra

// 32 x 32 row-major helpers for one workgroup (256 threads, 4 elements each).
__device__ __forceinline__ void load32(double* dst, const double* src) {
  for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) dst[e] = src[e];
}

// Pass `pass` (1 or 2) of column j: workgroup i < nb forms Hh = S_i^T C_i
// (= Q_i^T Z), stores it into block (i, j) of H (pass 1) or adds it (pass 2),
// and the update coefficients G_i = S_i Hh.
__global__ __launch_bounds__(kBlock) void k_bcoeffs(BlockBufs b, int j, int pass) {
  if (b.flags[0] || (pass == 2 && !b.flags[2])) return;
  __shared__ double S[kRhs * kRhs], C[kRhs * kRhs], Hh[kRhs * kRhs];
  const int i = blockIdx.x;
  load32(S, b.S + static_cast<long long>(i) * kRhs * kRhs);
  load32(C, b.C + static_cast<long long>(i) * kRhs * kRhs);
  __syncthreads();
  for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
    const int r = e / kRhs, c = e % kRhs;
    double s = 0.0;
    for (int k = 0; k <= r; ++k) s += S[k * kRhs + r] * C[k * kRhs + c];  // S upper: S[k][r], k <= r
    Hh[e] = s;
  }
  __syncthreads();
  double* G = b.G + static_cast<long long>(i) * kRhs * kRhs;
  for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
    const int r = e / kRhs, c = e % kRhs;
    double s = 0.0;
    for (int k = r; k < kRhs; ++k) s += S[r * kRhs + k] * Hh[k * kRhs + c];
    G[e] = s;
    double& h = b.Hb[static_cast<long long>(kRhs * j + c) * b.ldh + kRhs * i + r];
    h = (pass == 1) ? Hh[e] : h + Hh[e];
  }
}

void block_launch_coeffs(const BlockBufs& b, int j, int nb, int pass, hipStream_t s) {
  hipLaunchKernelGGL(k_bcoeffs, dim3(nb), dim3(kBlock), 0, s, b, j, pass);
}

// After the first pass of column j (C[0..j] = P_i^T Z, C[j+1] = Z^T Z, and
// block (i, j) of H = Hh_i = Q_i^T Z): Gm = Z'^T Z' = Z^T Z - sum_i Hh_i^T
// Hh_i, exact for an orthonormal basis, and the second-pass decision into
// flags[2] (counted in flags[3]).  DGKS per column: a second pass when
// ||z'_c||^2 < eta^2 ||z_c||^2 for some column c.
__global__ __launch_bounds__(kBlock) void k_bgram(BlockBufs b, int j, int gs_mode, double eta) {
  if (b.flags[0]) return;
  __shared__ double Gz[kRhs * kRhs], Gp[kRhs * kRhs];
  __shared__ int need;
  load32(Gz, b.C + static_cast<long long>(j + 1) * kRhs * kRhs);
  if (threadIdx.x == 0) need = 0;
  __syncthreads();
  const int rows = kRhs * (j + 1);
  const double* Hc = b.Hb + static_cast<long long>(kRhs * j) * b.ldh;
  double acc[kRhs * kRhs / kBlock] = {};
  for (int r0 = 0; r0 < rows; r0 += kRhs) {  // 32 rows of block column j at a time
    for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
      const int r = e % kRhs, c = e / kRhs;
      Gp[r * kRhs + c] = Hc[static_cast<long long>(c) * b.ldh + r0 + r];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kRhs * kRhs / kBlock; ++u) {
      const int e = threadIdx.x + u * kBlock;
      const int a = e / kRhs, c = e % kRhs;
      for (int r = 0; r < kRhs; ++r) acc[u] += Gp[r * kRhs + a] * Gp[r * kRhs + c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kRhs * kRhs / kBlock; ++u) {
    const int e = threadIdx.x + u * kBlock;
    Gp[e] = Gz[e] - acc[u];
  }
  __syncthreads();
  if (threadIdx.x < kRhs) {
    const int c = threadIdx.x;
    const double zz = Gz[c * kRhs + c], zp = Gp[c * kRhs + c];
    int bad = 1;
    if (gs_mode == 0) bad = 0;
    else if (gs_mode == 2) bad = !(zp >= eta * eta * zz);  // also true for NaN
    if (bad) atomicOr(&need, 1);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) b.Gm[e] = Gp[e];
  if (threadIdx.x == 0) {
    b.flags[2] = need;
    b.flags[3] += need;
  }
}

void block_launch_gram(const BlockBufs& b, int j, int gs_mode, double eta, hipStream_t s) {
  hipLaunchKernelGGL(k_bgram, dim3(1), dim3(kBlock), 0, s, b, j, gs_mode, eta);
}

// Cholesky of the Gram block C[0] = R^T R (R upper) and S = R^{-1}.  j >= 0:
// the new block is j + 1; block column j of H is right-multiplied by S_j and
// H(j+1, j) = R S_j.  j = -1: the start block, R0 = R.  A pivot below 1e-14
// (the reference's breakdown threshold for a vector norm) or a non-finite one
// marks a breakdown: the basis stops at j + 1 blocks and H(j+1, j) = 0.
// For j >= 0 without a second pass the Gram block is Gm (from k_bgram).
__global__ __launch_bounds__(kBlock) void k_bchol(BlockBufs b, int j) {
  if (b.flags[0]) return;
  __shared__ double R[kRhs * kRhs], Sn[kRhs * kRhs], Sj[kRhs * kRhs];
  __shared__ int brk;
  load32(R, (j >= 0 && !b.flags[2]) ? b.Gm : b.C);
  if (j >= 0) load32(Sj, b.S + static_cast<long long>(j) * kRhs * kRhs);
  __syncthreads();
  // upper Cholesky in place (row-major R[r][c], c >= r), row k at a time:
  // every thread forms the pivot (same operations, same value), thread c > k
  // its entry of row k
  int bad = 0;
  for (int k = 0; k < kRhs; ++k) {
    double d = R[k * kRhs + k];
    for (int p = 0; p < k; ++p) d -= R[p * kRhs + k] * R[p * kRhs + k];
    if (!(d > 1e-28) || !isfinite(d)) {
      bad = 1;
      break;
    }
    const double rkk = sqrt(d);
    double s = 0.0;
    const int c = threadIdx.x;
    if (c > k && c < kRhs) {
      s = R[k * kRhs + c];
      for (int p = 0; p < k; ++p) s -= R[p * kRhs + k] * R[p * kRhs + c];
    }
    __syncthreads();  // all pivots read before row k changes
    if (c == k) R[k * kRhs + k] = rkk;
    if (c > k && c < kRhs) R[k * kRhs + c] = s / rkk;
    __syncthreads();
  }
  if (threadIdx.x == 0) brk = bad;
  __syncthreads();
  if (brk) {
    if (threadIdx.x == 0) {
      b.flags[0] = 1;
      b.flags[1] = j + 1;
    }
    if (j >= 0) {  // H(j+1, j) = 0 and column j scaled by S_j below
      for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
        const int r = e / kRhs, c = e % kRhs;
        b.Hb[static_cast<long long>(kRhs * j + c) * b.ldh + kRhs * (j + 1) + r] = 0.0;
      }
    }
  } else {
    for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
      const int r = e / kRhs, c = e % kRhs;
      if (c < r) R[e] = 0.0;
    }
    __syncthreads();
    // S = R^{-1}, upper triangular: one column per thread (back substitution)
    if (threadIdx.x < kRhs) {
      const int c = threadIdx.x;
      for (int r = kRhs - 1; r >= 0; --r) {
        double s = (r == c) ? 1.0 : 0.0;
        for (int k = r + 1; k <= c; ++k) s -= R[r * kRhs + k] * Sn[k * kRhs + c];
        Sn[r * kRhs + c] = (r <= c) ? s / R[r * kRhs + r] : 0.0;
      }
    }
    __syncthreads();
    double* Snext = b.S + static_cast<long long>(j + 1) * kRhs * kRhs;
    for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) Snext[e] = Sn[e];
    if (j < 0) {
      for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) b.R0[e] = R[e];
      if (threadIdx.x == 0) b.flags[1] = 1;
      return;
    }
    // H(j+1, j) = R S_j
    for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
      const int r = e / kRhs, c = e % kRhs;
      double s = 0.0;
      for (int k = r; k <= c; ++k) s += R[r * kRhs + k] * Sj[k * kRhs + c];
      b.Hb[static_cast<long long>(kRhs * j + c) * b.ldh + kRhs * (j + 1) + r] = s;
    }
    if (threadIdx.x == 0) b.flags[1] = j + 2;
  }
  if (j < 0) return;
  // block column j of H (rows < 32 (j+1)) times S_j, 32 rows at a time
  // through LDS (R is free again)
  const int rows = kRhs * (j + 1);
  double* Hc = b.Hb + static_cast<long long>(kRhs * j) * b.ldh;
  for (int r0 = 0; r0 < rows; r0 += kRhs) {
    __syncthreads();
    for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
      const int r = e % kRhs, c = e / kRhs;
      R[r * kRhs + c] = Hc[static_cast<long long>(c) * b.ldh + r0 + r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
      const int r = e % kRhs, c = e / kRhs;
      double s = 0.0;
      for (int k = 0; k <= c; ++k) s += R[r * kRhs + k] * Sj[k * kRhs + c];
      Hc[static_cast<long long>(c) * b.ldh + r0 + r] = s;
    }
  }
}

void block_launch_chol(const BlockBufs& b, int j, hipStream_t s) {
  hipLaunchKernelGGL(k_bchol, dim3(1), dim3(kBlock), 0, s, b, j);
}

// G_i = -S_i (E_i R0) for the lift Y = sum_i Q_i F_i = -sum_i P_i G_i, with
// E_i the i-th 32 x 32 block of the first block column of exp(dt H).
__global__ __launch_bounds__(kBlock) void k_blift_coeffs(BlockBufs b, const double* F, int ldf) {
  __shared__ double S[kRhs * kRhs], R[kRhs * kRhs], Fi[kRhs * kRhs];
  const int i = blockIdx.x;
  load32(S, b.S + static_cast<long long>(i) * kRhs * kRhs);
  load32(R, b.R0);
  __syncthreads();
  for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
    const int r = e / kRhs, c = e % kRhs;
    double s = 0.0;
    for (int k = 0; k <= c; ++k) s += F[static_cast<long long>(k) * ldf + kRhs * i + r] * R[k * kRhs + c];
    Fi[e] = s;
  }
  __syncthreads();
  double* G = b.G + static_cast<long long>(i) * kRhs * kRhs;
  for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
    const int r = e / kRhs, c = e % kRhs;
    double s = 0.0;
    for (int k = r; k < kRhs; ++k) s += S[r * kRhs + k] * Fi[k * kRhs + c];
    G[e] = -s;
  }
}

void block_launch_lift_coeffs(const BlockBufs& b, int nb, const double* F, int ldf, hipStream_t s) {
  hipLaunchKernelGGL(k_blift_coeffs, dim3(nb), dim3(kBlock), 0, s, b, F, ldf);
}

// scal[0] = ||E R0||_F over the k blocks (= ||Y||_F), scal[1] = ||H(k, k-1)
// E_{k-1} R0||_F (0 after a breakdown): the block form of the reference's
// residual estimate |h_{k+1,k}| |[exp(dt H) e1]_k| beta.
__global__ __launch_bounds__(kBlock) void k_bresidual(BlockBufs b, int k, const double* F, int ldf) {
  __shared__ double R[kRhs * kRhs], FR[kRhs * kRhs], red[kWavesPerBlock];
  load32(R, b.R0);
  __syncthreads();
  // ||F R0||_F over all rows of the first block column
  double s2 = 0.0;
  const int rows = kRhs * k;
  for (int e = threadIdx.x; e < rows * kRhs; e += kBlock) {
    const int r = e / kRhs, c = e % kRhs;
    double s = 0.0;
    for (int q = 0; q <= c; ++q) s += F[static_cast<long long>(q) * ldf + r] * R[q * kRhs + c];
    s2 += s * s;
    if (r >= rows - kRhs) FR[(r - (rows - kRhs)) * kRhs + c] = s;
  }
  s2 = wave_sum(s2);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s2;
  __syncthreads();
  const double fn = sqrt((red[0] + red[1]) + (red[2] + red[3]));
  __syncthreads();
  double e2 = 0.0;
  if (!b.flags[0]) {
    const double* Hk = b.Hb + static_cast<long long>(kRhs * (k - 1)) * b.ldh + kRhs * k;  // H(k, k-1)
    for (int e = threadIdx.x; e < kRhs * kRhs; e += kBlock) {
      const int r = e / kRhs, c = e % kRhs;
      double s = 0.0;
      for (int q = 0; q < kRhs; ++q) s += Hk[static_cast<long long>(q) * b.ldh + r] * FR[q * kRhs + c];
      e2 += s * s;
    }
  }
  e2 = wave_sum(e2);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = e2;
  __syncthreads();
  if (threadIdx.x == 0) {
    b.scal[0] = fn;
    b.scal[1] = sqrt((red[0] + red[1]) + (red[2] + red[3]));
    double r2 = 0.0;  // ||R0||_F = ||V_step||_F
    for (int e = 0; e < kRhs * kRhs; ++e) r2 += R[e] * R[e];
    b.scal[2] = sqrt(r2);
  }
}

void block_launch_residual(const BlockBufs& b, int k, const double* F, int ldf, hipStream_t s) {
  hipLaunchKernelGGL(k_bresidual, dim3(1), dim3(kBlock), 0, s, b, k, F, ldf);
}

// ------------------------------------------------------ dense GEMM

// D = alpha A B + beta E, column-major, N x N, N % 64 == 0.  A 64 x 64 tile
// per workgroup, a 32 x 32 quarter per wave (2 x 2 MFMA tiles); operands come
// straight from global memory (the matrices are small enough for L2).
__global__ __launch_bounds__(kBlock) void k_dgemm(int N, const double* __restrict__ A, int lda,
                                                  const double* __restrict__ B, int ldb, double alpha,
                                                  const double* E, int lde, double beta, double* D,
                                                  int ldd) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int lr = lane / 16, lc = lane % 16;
  const int tiles = N / kGemmTile;
  const int tm = blockIdx.x % tiles, tn = blockIdx.x / tiles;
  const int m0 = tm * kGemmTile + 32 * (wave % 2);
  const int n0 = tn * kGemmTile + 32 * (wave / 2);
  dbl4 c00 = zero4(), c01 = zero4(), c10 = zero4(), c11 = zero4();
  for (int k0 = 0; k0 < N; k0 += 4) {
    // A[m][k] at A[k * lda + m]: lane -> m = lc, k = lr
    const double a0 = A[static_cast<long long>(k0 + lr) * lda + m0 + lc];
    const double a1 = A[static_cast<long long>(k0 + lr) * lda + m0 + 16 + lc];
    // B[k][n] at B[n * ldb + k]: lane -> n = lc, k = lr
    const double b0 = B[static_cast<long long>(n0 + lc) * ldb + k0 + lr];
    const double b1 = B[static_cast<long long>(n0 + 16 + lc) * ldb + k0 + lr];
    c00 = mfma(a0, b0, c00);
    c01 = mfma(a0, b1, c01);
    c10 = mfma(a1, b0, c10);
    c11 = mfma(a1, b1, c11);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int m = lr + 4 * q;
    const int mm[2] = {m0 + m, m0 + 16 + m};
    const int nn[2] = {n0 + lc, n0 + 16 + lc};
    const double v[2][2] = {{c00[q], c01[q]}, {c10[q], c11[q]}};
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const long long oe = static_cast<long long>(nn[y]) * lde + mm[x];
        const long long od = static_cast<long long>(nn[y]) * ldd + mm[x];
        const double e = (E && beta != 0.0) ? beta * E[oe] : 0.0;
        D[od] = alpha * v[x][y] + e;
      }
  }
}

void dense_gemm(int N, const double* A, int lda, const double* B, int ldb, double alpha, const double* E,
                int lde, double beta, double* D, int ldd, hipStream_t s) {
  const int tiles = N / kGemmTile;
  hipLaunchKernelGGL(k_dgemm, dim3(tiles * tiles), dim3(kBlock), 0, s, N, A, lda, B, ldb, alpha, E, lde,
                     beta, D, ldd);
}

__global__ __launch_bounds__(kBlock) void k_dcombine(int N, int ld, double* D, const double* X0, double c0,
                                                     const double* X1, double c1, const double* X2, double c2,
                                                     const double* X3, double c3, double diag) {
  const long long tot = static_cast<long long>(N) * N;
  for (long long e = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; e < tot;
       e += static_cast<long long>(gridDim.x) * kBlock) {
    const int r = static_cast<int>(e % N), c = static_cast<int>(e / N);
    const long long o = static_cast<long long>(c) * ld + r;
    double v = (r == c) ? diag : 0.0;
    if (X0) v += c0 * X0[o];
    if (X1) v += c1 * X1[o];
    if (X2) v += c2 * X2[o];
    if (X3) v += c3 * X3[o];
    D[o] = v;
  }
}

void dense_combine(int N, int ld, double* D, const double* X0, double c0, const double* X1, double c1,
                   const double* X2, double c2, const double* X3, double c3, double diag, hipStream_t s) {
  const long long tot = static_cast<long long>(N) * N;
  const int grid = static_cast<int>((tot + kBlock - 1) / kBlock < 2048 ? (tot + kBlock - 1) / kBlock : 2048);
  hipLaunchKernelGGL(k_dcombine, dim3(grid), dim3(kBlock), 0, s, N, ld, D, X0, c0, X1, c1, X2, c2, X3, c3,
                     diag);
}

// D (N x N, ld N) = scale * the leading kk x kk block of src (ld lds), zero
// elsewhere.
__global__ __launch_bounds__(kBlock) void k_dload(int N, int kk, const double* src, int lds, double scale,
                                                  double* D) {
  const long long tot = static_cast<long long>(N) * N;
  for (long long e = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; e < tot;
       e += static_cast<long long>(gridDim.x) * kBlock) {
    const int r = static_cast<int>(e % N), c = static_cast<int>(e / N);
    D[e] = (r < kk && c < kk) ? scale * src[static_cast<long long>(c) * lds + r] : 0.0;
  }
}

void dense_load(int N, int kk, const double* src, int lds, double scale, double* D, hipStream_t s) {
  const long long tot = static_cast<long long>(N) * N;
  const int grid = static_cast<int>((tot + kBlock - 1) / kBlock < 2048 ? (tot + kBlock - 1) / kBlock : 2048);
  hipLaunchKernelGGL(k_dload, dim3(grid), dim3(kBlock), 0, s, N, kk, src, lds, scale, D);
}

// max_c sum_r |A[r][c]|: one workgroup, thread per column, fixed-order max.
__global__ __launch_bounds__(kBlock) void k_dnorm1(int N, const double* A, int lda, double* out) {
  __shared__ double red[kBlock];
  double mx = 0.0;
  for (int c = threadIdx.x; c < N; c += kBlock) {
    double s = 0.0;
    for (int r = 0; r < N; ++r) s += fabs(A[static_cast<long long>(c) * lda + r]);
    mx = fmax(mx, s);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int t = 0; t < kBlock; ++t) m = fmax(m, red[t]);
    *out = m;
  }
}

void dense_norm1(int N, const double* A, int lda, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_dnorm1, dim3(1), dim3(kBlock), 0, s, N, A, lda, out);
}

}  // namespace mek
