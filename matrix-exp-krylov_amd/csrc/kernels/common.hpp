// Shared device helpers for the gfx950 kernels: wave64 DPP reductions,
// block reductions, XCD-aware block remapping and the device buffer bundle
// that the Arnoldi kernels receive by value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mek {

constexpr int kWave = 64;
constexpr int kBlock = 256;            // threads per workgroup (4 waves)
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kNumXcd = 8;             // MI355X: 8 XCDs, round-robin dispatch

// ---------------------------------------------------------------- DPP

// Move a double across lanes with one DPP control (two 32-bit halves).
// bound_ctrl=true makes lanes without a source read 0.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long b = __double_as_longlong(x);
  int lo = static_cast<int>(b & 0xffffffffLL);
  int hi = static_cast<int>(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Sum over the 64 lanes of a wave, returned wave-uniform.  Within each row of
// 16 lanes the xor-butterfly (quad_perm 1,0,3,2 / 2,3,0,1, row_half_mirror,
// row_mirror) leaves every lane with the same row sum (each step adds two
// identical operands in swapped order, so all lanes round identically); the
// four row sums are then added in a fixed order, so the result is
// deterministic.  All 64 lanes must be active.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]: xor 1
  x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]: xor 2
  x += dpp_f64<0x141>(x);  // row_half_mirror: 8-lane sums
  x += dpp_f64<0x140>(x);  // row_mirror: 16-lane sums
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) +
         (readlane_f64(x, 32) + readlane_f64(x, 48));
}

__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp_f64<0xB1>(x));
  x = fmax(x, dpp_f64<0x4E>(x));
  x = fmax(x, dpp_f64<0x141>(x));
  x = fmax(x, dpp_f64<0x140>(x));
  return fmax(fmax(readlane_f64(x, 0), readlane_f64(x, 16)),
              fmax(readlane_f64(x, 32), readlane_f64(x, 48)));
}

// ---------------------------------------------------------------- launch

// XCD-aware logical block id: hardware dispatches block b to XCD b % 8, so
// give each XCD a contiguous range of logical blocks (neighbouring rows share
// that XCD's L2).  Requires gridDim.x % 8 == 0 for an exact permutation;
// otherwise the identity is used.
__device__ __forceinline__ int xcd_block_id() {
  const int b = blockIdx.x;
  const int g = gridDim.x;
  if (g % kNumXcd != 0) return b;
  return (b % kNumXcd) * (g / kNumXcd) + (b / kNumXcd);
}

// ---------------------------------------------------------------- buffers

// Device buffers of one Arnoldi problem.  Column i of the Krylov basis is
// stored unnormalised at V + i*ld; its true value is colscale[i] * V[:, i]
// (the multiply is done on the fly, with the same single rounding as an
// explicit q = w * (1/||w||)).
struct ArnoldiBufs {
  int n;                 // rows
  int m;                 // allocated Krylov dimension
  const int* row_ptr;    // n+1
  const int* col_idx;    // nnz
  const void* vals;      // nnz, double or float
  int vals_f32;          // 1 if vals is float
  const int* row_blocks; // n_row_blocks+1 row starts of SpMV row blocks
  int n_row_blocks;
  double* V;             // ld x (m+2), column-major
  long long ld;
  double* w;             // n
  double* y;             // n
  double* colscale;      // m+2
  double* h1;            // m+2: first-pass coefficients
  double* h2;            // m+2: second-pass coefficients
  double* H;             // ldh x ldh Hessenberg, column-major
  int ldh;               // m+2
  double* wH;            // m+2: exp(dt H) e1 (and Expokit extras)
  double* scal;          // scalar slots, see ScalarSlot
  int* iscal;            // int slots, see IntSlot
  double* partials;      // per-block partial sums: quantity q of block g at q*grid + g
  int grid;              // blocks of the streaming kernels
  int spmv_grid;         // blocks of the SpMV kernel
  // Sliced-ELL copy of the matrix (null when the SpMV uses the CSR row
  // blocks): slices of 64 consecutive rows, one row per lane; slice s holds
  // W_s = (sell_ptr[s+1] - sell_ptr[s]) / 64 entries per row, stored entry
  // by entry (entry e of rows 0..63, then entry e+1, ...), each row's
  // entries in CSR order; a row shorter than W_s is padded with column -1.
  const long long* sell_ptr;  // n_slices + 1 entry offsets
  const int* sell_col;
  const void* sell_val;       // double, or float when vals_f32
  int n_slices;
  // Diagonal copy of the matrix (null when unused): diagonal d holds the
  // entries (r, r + dia_off[d]) at dia_val[d * dia_ld + r], offsets in
  // ascending order; an entry the CSR does not hold is a NaN marker
  // (kDiaAbsent64 / kDiaAbsent32 bits), which the SpMV skips.
  const void* dia_val;        // double, or float when vals_f32
  const int* dia_off;
  int n_diags;                // > 0 when the SpMV uses a diagonal layout
  long long dia_ld;
  // Constant-diagonal form (dia_val null): every entry of diagonal d has the
  // value dia_cval[d] (the stored value, FP32-rounded when vals_f32), and bit
  // d of row r's presence mask dia_mask[r] (uint8 when n_diags <= 8, else
  // uint16; dia_ld entries, 0 past the last row) says whether the CSR holds
  // (r, r + dia_off[d]).
  const void* dia_mask;
  const double* dia_cval;
};

// Marker bits of an absent entry in the diagonal copy (quiet NaNs; the copy
// is only built for matrices without NaN values).
constexpr unsigned long long kDiaAbsent64 = 0x7FF8DEADBEEF0001ULL;
constexpr unsigned int kDiaAbsent32 = 0x7FC0BEEFU;
constexpr int kDiaDiagCap = 16;  // most diagonals the diagonal-copy SpMV takes

enum ScalarSlot : int {
  kVNorm = 0,     // beta = ||start vector||
  kWNorm2,        // ||A q_j||^2
  kWpNorm2,       // ||w'||^2 after the first pass
  kWppNorm2,      // ||w''||^2 after the second pass
  kErr,           // residual estimate of the last attempt
  kYNorm,         // ||y|| of the last lift
  kAvNorm,        // ||A q_m|| (Expokit)
  kNumScalars = 16
};

enum IntSlot : int {
  kKUsed = 0,     // Krylov columns built in this attempt
  kBreakdown,     // 1 after a happy breakdown
  kReorth,        // 1 if the current column takes the second pass
  kReorthCount,   // columns that took the second pass (statistics)
  kExpmSquarings,
  kSStepFail,     // 1 when an s-step block could not be completed
  kNumInts = 8
};

// Arguments of the one-workgroup expm kernel (expm.hip).
struct ExpmArgs {
  const double* H;   // source Hessenberg, column-major, leading dimension ldh
  int ldh;
  const int* iscal;  // kKUsed, kBreakdown
  int* iscal_out;    // kExpmSquarings
  double* scal;      // reads kVNorm, writes kErr
  double* out;       // first column of exp(dt*M): kk entries (+ zero pad)
  int out_len;       // entries of `out` to write (pads with zeros)
  double dt;
  int mode;          // 0: reference residual; 1: Expokit augmented matrix
  double* ws;        // global workspace (6*kk*kk doubles) when kk > kMaxKLds
  int k_override;    // >0: use this k instead of iscal[kKUsed] (standalone use)
  int breakdown_override;  // used with k_override
  double* full_out;  // optional: the whole kk x kk exponential
};

}  // namespace mek
