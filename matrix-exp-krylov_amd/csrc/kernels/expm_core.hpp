#pragma once
// Device core of the small dense exponential (shared by the single-problem
// and batched kernels): one workgroup per small dense matrix,
// FP64 Pade scaling and squaring (Higham 2005, SIAM J. Matrix Anal. Appl.
// 26(4), Algorithm 2.3), replacing the reference's host-side Eigen
// `A.exp()` (main.cu:33-36, used at main.cu:1966-1978 and :2007-2014).
//
// Matrices are column-major k x k.  For k <= kMaxKLds five matrices live in
// LDS (5 * 62 * 62 * 8 B = 154 KB of the CU's 160 KB); larger k uses a global
// workspace of six matrices.  Products B <- A*B are done in place on the right
// operand: each wave owns whole columns of B, so a column is read completely
// (by broadcast) before the wave overwrites it.
#include "common.hpp"

namespace mek {

constexpr int kMaxKLds = 62;  // 5 * 62 * 62 * 8 B = 154 KB of LDS

// Higham 2005, Table 2.3 (theta_m for m = 3, 5, 7, 9, 13, unit roundoff 2^-53)
static __constant__ double kTheta[5] = {1.495585217958292e-2, 2.539398330063230e-1,
                                 9.504178996162932e-1, 2.097847961257068e0,
                                 5.371920351148152e0};


// B <- A * B for k <= 64 (each lane one row, each wave whole column groups)
__device__ inline void mm_right_small(const double* A, double* B, int k) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int ngroups = (k + 3) / 4;
  for (int g = wave; g < ngroups; g += kWavesPerBlock) {
    const int c0 = 4 * g;
    const int nc = min(4, k - c0);
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
    if (lane < k) {
      for (int l = 0; l < k; ++l) {
        const double a = A[l * k + lane];
        acc0 += a * B[(c0 + 0) * k + l];
        if (nc > 1) acc1 += a * B[(c0 + 1) * k + l];
        if (nc > 2) acc2 += a * B[(c0 + 2) * k + l];
        if (nc > 3) acc3 += a * B[(c0 + 3) * k + l];
      }
    }
    // all lanes of this wave have finished reading the group's columns
    __builtin_amdgcn_wave_barrier();
    if (lane < k) {
      B[(c0 + 0) * k + lane] = acc0;
      if (nc > 1) B[(c0 + 1) * k + lane] = acc1;
      if (nc > 2) B[(c0 + 2) * k + lane] = acc2;
      if (nc > 3) B[(c0 + 3) * k + lane] = acc3;
    }
  }
  __syncthreads();
}

// B <- A * B for any k through a temporary (global workspace path)
__device__ inline void mm_right_big(const double* A, double* B, double* T, int k) {
  const int kk = k * k;
  for (int e = threadIdx.x; e < kk; e += kBlock) {
    const int i = e % k, c = e / k;
    double acc = 0.0;
    for (int l = 0; l < k; ++l) acc += A[l * k + i] * B[c * k + l];
    T[e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kk; e += kBlock) B[e] = T[e];
  __syncthreads();
}

template <bool BIG>
__device__ __forceinline__ void mm_right(const double* A, double* B, double* T, int k) {
  if (BIG) mm_right_big(A, B, T, k);
  else mm_right_small(A, B, k);
}

__device__ inline void copy_mat(double* dst, const double* src, int k) {
  for (int e = threadIdx.x; e < k * k; e += kBlock) dst[e] = src[e];
  __syncthreads();
}

// Solve Q X = P in place (P <- X), LU with partial pivoting: the pivot is the
// row of largest |entry| in the column, the first such row on ties.
__device__ inline void lu_solve_dev(double* Q, double* P, int k, int* s_piv) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  for (int p = 0; p < k; ++p) {
    if (wave == 0) {
      double best = -1.0;
      int bi = 0x7fffffff;
      for (int r = p + lane; r < k; r += kWave) {
        const double a = fabs(Q[p * k + r]);
        if (a > best) { best = a; bi = r; }
      }
      for (int off = kWave / 2; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) *s_piv = (bi == 0x7fffffff) ? p : bi;
    }
    __syncthreads();
    const int piv = *s_piv;
    if (piv != p) {
      for (int c = threadIdx.x; c < 2 * k; c += kBlock) {
        double* M = (c < k) ? Q : P;
        const int cc = (c < k) ? c : c - k;
        const double t = M[cc * k + p];
        M[cc * k + p] = M[cc * k + piv];
        M[cc * k + piv] = t;
      }
    }
    __syncthreads();
    const double d = Q[p * k + p];
    for (int r = p + 1 + threadIdx.x; r < k; r += kBlock) Q[p * k + r] = Q[p * k + r] / d;
    __syncthreads();
    // trailing update of Q (columns > p) and of every column of P
    const int rows = k - p - 1;
    const int cols = (k - p - 1) + k;
    for (int e = threadIdx.x; e < rows * cols; e += kBlock) {
      const int r = p + 1 + e % rows;
      const int cidx = e / rows;
      const double l = Q[p * k + r];
      if (cidx < k - p - 1) {
        const int c = p + 1 + cidx;
        Q[c * k + r] -= l * Q[c * k + p];
      } else {
        const int c = cidx - (k - p - 1);
        P[c * k + r] -= l * P[c * k + p];
      }
    }
    __syncthreads();
  }
  // back substitution, one column of P per thread
  for (int c = threadIdx.x; c < k; c += kBlock) {
    for (int r = k - 1; r >= 0; --r) {
      double s = P[c * k + r];
      for (int l = r + 1; l < k; ++l) s -= Q[l * k + r] * P[c * k + l];
      P[c * k + r] = s / Q[r * k + r];
    }
  }
  __syncthreads();
}

// ceil(log2(x)) for finite x > 0, exactly (no libm rounding).
__device__ __forceinline__ int ceil_log2(double x) {
  int e;
  const double f = frexp(x, &e);
  return (f == 0.5) ? e - 1 : e;
}

// E = exp(S0) for the k x k matrix in slot S0; returns E in *result (one of
// the slots).  Slots S[0..4] (+ S[5] temporary on the BIG path).
template <bool BIG>
__device__ inline double* expm_slots(double* const* S, int k, double* s_b, int* s_int, double* s_red) {
  const int kk = k * k;
  double* A = S[0];
  double* T = BIG ? S[5] : nullptr;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;

  // ||A||_1 and finiteness
  if (threadIdx.x == 0) s_int[0] = 0;
  __syncthreads();
  double cmax = 0.0;
  for (int c = threadIdx.x; c < k; c += kBlock) {
    double s = 0.0;
    for (int r = 0; r < k; ++r) {
      const double a = A[c * k + r];
      if (!isfinite(a)) s_int[0] = 1;
      s += fabs(a);
    }
    cmax = fmax(cmax, s);
  }
  cmax = wave_max(cmax);
  if (lane == 0) s_red[wave] = cmax;
  __syncthreads();
  if (s_int[0]) {
    for (int e = threadIdx.x; e < kk; e += kBlock) A[e] = __longlong_as_double(0x7ff8000000000000LL);
    if (threadIdx.x == 0) s_int[1] = -1;
    __syncthreads();
    return A;
  }
  double nrm = s_red[0];
  for (int w = 1; w < kWavesPerBlock; ++w) nrm = fmax(nrm, s_red[w]);

  int q = 13, s = 0;
  for (int i = 0; i < 4; ++i) {
    if (nrm <= kTheta[i]) { q = (i == 0) ? 3 : (i == 1) ? 5 : (i == 2) ? 7 : 9; break; }
  }
  if (q == 13 && nrm > kTheta[4]) {
    s = ceil_log2(nrm / kTheta[4]);
    if (s < 0) s = 0;
  }
  if (threadIdx.x == 0) {
    // b_j = (2q-j)!/(j!(q-j)!), exact in 64-bit integers
    unsigned long long bj = 1ULL;
    s_b[q] = 1.0;
    for (int j = q; j >= 1; --j) {
      bj = bj * static_cast<unsigned long long>(2 * q - j + 1) * static_cast<unsigned long long>(j) /
           static_cast<unsigned long long>(q - j + 1);
      s_b[j - 1] = static_cast<double>(bj);
    }
    s_int[1] = s;
  }
  if (s > 0)
    for (int e = threadIdx.x; e < kk; e += kBlock) A[e] = ldexp(A[e], -s);
  __syncthreads();
  const double* b = s_b;

  double* A2 = S[1];
  copy_mat(A2, A, k);
  mm_right<BIG>(A, A2, T, k);  // A2 = A*A
  double *Pm, *Qm;
  if (q == 13) {
    double* A4 = S[2];
    double* A6 = S[3];
    double* W = S[4];
    copy_mat(A4, A2, k);
    mm_right<BIG>(A2, A4, T, k);  // A4
    copy_mat(A6, A2, k);
    mm_right<BIG>(A4, A6, T, k);  // A6 = A4*A2
    for (int e = threadIdx.x; e < kk; e += kBlock) W[e] = b[13] * A6[e] + b[11] * A4[e] + b[9] * A2[e];
    __syncthreads();
    mm_right<BIG>(A6, W, T, k);
    for (int e = threadIdx.x; e < kk; e += kBlock) {
      double v = W[e] + b[7] * A6[e] + b[5] * A4[e] + b[3] * A2[e];
      if (e % (k + 1) == 0) v += b[1];
      W[e] = v;
    }
    __syncthreads();
    mm_right<BIG>(A, W, T, k);  // U = A * W
    // V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I, into S0
    for (int e = threadIdx.x; e < kk; e += kBlock) A[e] = b[12] * A6[e] + b[10] * A4[e] + b[8] * A2[e];
    __syncthreads();
    mm_right<BIG>(A6, A, T, k);
    for (int e = threadIdx.x; e < kk; e += kBlock) {
      double v = A[e] + b[6] * A6[e] + b[4] * A4[e] + b[2] * A2[e];
      if (e % (k + 1) == 0) v += b[0];
      A[e] = v;
    }
    __syncthreads();
    // P = V + U -> S1, Q = V - U -> S2
    Pm = S[1];
    Qm = S[2];
    for (int e = threadIdx.x; e < kk; e += kBlock) {
      const double u = W[e], v = A[e];
      Pm[e] = v + u;
      Qm[e] = v - u;
    }
    __syncthreads();
  } else {
    if (q >= 5) { copy_mat(S[2], A2, k); mm_right<BIG>(A2, S[2], T, k); }       // A4
    if (q >= 7) { copy_mat(S[3], A2, k); mm_right<BIG>(S[2], S[3], T, k); }     // A6
    if (q >= 9) { copy_mat(S[4], A2, k); mm_right<BIG>(S[3], S[4], T, k); }     // A8
    // odd part W = sum b_{2p+1} A^{2p} -> S4, even part V = sum b_{2p} A^{2p} -> S3
    for (int e = threadIdx.x; e < kk; e += kBlock) {
      const bool diag = (e % (k + 1) == 0);
      double wv = diag ? b[1] : 0.0;
      double vv = diag ? b[0] : 0.0;
      const double p1 = A2[e];
      const double p2 = (q >= 5) ? S[2][e] : 0.0;
      const double p3 = (q >= 7) ? S[3][e] : 0.0;
      const double p4 = (q >= 9) ? S[4][e] : 0.0;
      wv += b[3] * p1; vv += b[2] * p1;
      if (q >= 5) { wv += b[5] * p2; vv += b[4] * p2; }
      if (q >= 7) { wv += b[7] * p3; vv += b[6] * p3; }
      if (q >= 9) { wv += b[9] * p4; vv += b[8] * p4; }
      S[4][e] = wv;
      S[3][e] = vv;
    }
    __syncthreads();
    mm_right<BIG>(A, S[4], T, k);  // U = A * W
    Pm = S[1];
    Qm = S[2];
    for (int e = threadIdx.x; e < kk; e += kBlock) {
      const double u = S[4][e], v = S[3][e];
      Pm[e] = v + u;
      Qm[e] = v - u;
    }
    __syncthreads();
  }
  lu_solve_dev(Qm, Pm, k, s_int + 2);
  double* X = Pm;
  double* Y = S[0];
  for (int i = 0; i < s; ++i) {
    copy_mat(Y, X, k);
    mm_right<BIG>(X, Y, T, k);  // Y = X*X
    double* t = X; X = Y; Y = t;
  }
  return X;
}

}  // namespace mek
