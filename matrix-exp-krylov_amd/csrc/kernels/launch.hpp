// Host-side launchers of the gfx950 kernels (defined in arnoldi.hip, expm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace mek {

constexpr int kMaxFusedDots = 64;  // dots carried in registers by one sweep
constexpr int kSpmvNnzCap = 4096;  // must equal kNnzCap in arnoldi.hip
constexpr int kSpmvRowCap = 2048;
constexpr int kMaxKrylov = 500;    // largest supported m
constexpr int kMaxGrid = 4096;     // blocks of the streaming and SpMV kernels

// Column j: w = A q_j with the ||w||^2 partial at quantity normq.
void launch_spmv_col(const ArnoldiBufs& b, int j, int normq, hipStream_t s);
// Column j (j + 1 <= kSweepMaxCols) with the first Gram-Schmidt pass fused
// into the SpMV: partials of q_i . w (i <= j) and ||w||^2 (quantity j + 1),
// spmv_grid blocks (finalise with launch_fin_h1(b, j + 1, b.spmv_grid)).
void launch_spmv_dots(const ArnoldiBufs& b, int j, hipStream_t s);
void launch_update_dot(const ArnoldiBufs& b, int j, hipStream_t s);
void launch_update_norm(const ArnoldiBufs& b, int j, int gs_mode, double eta, hipStream_t s);
void launch_dots(const ArnoldiBufs& b, const double* x, int i0, int nq, int with_norm, int qoff,
                 hipStream_t s);
void launch_update(const ArnoldiBufs& b, const double* src, double* dst, const double* h, int nq,
                   hipStream_t s);
// finalisers read partials written by pgrid blocks
void launch_fin_h1(const ArnoldiBufs& b, int nq, int pgrid, hipStream_t s);
void launch_fin_h2(const ArnoldiBufs& b, int j, hipStream_t s);
void launch_fin_col(const ArnoldiBufs& b, int j, int gs_mode, double eta, hipStream_t s);
void launch_norm2(const ArnoldiBufs& b, const double* x, hipStream_t s);
void launch_fin_start(const ArnoldiBufs& b, hipStream_t s);
void launch_lift(const ArnoldiBufs& b, int ncoef, double skip_below, hipStream_t s);
void launch_fin_norm(const ArnoldiBufs& b, int slot, int pgrid, hipStream_t s);
void launch_restart(const ArnoldiBufs& b, hipStream_t s);
// Symmetric Lanczos column j (three-term recurrence): SpMV with the fused
// alpha partial (finalise with launch_fin_h1(b, 1, b.spmv_grid)), the update
// sweep into V[:, j+1] and its finaliser (H entries, breakdown, colscale).
void launch_spmv_lanczos(const ArnoldiBufs& b, int j, hipStream_t s);
void launch_lanczos_update(const ArnoldiBufs& b, int j, hipStream_t s);
void launch_lanczos_fin(const ArnoldiBufs& b, int j, hipStream_t s);
void launch_spmv_plain(const ArnoldiBufs& b, const double* x, const double* xs, double* y,
                       hipStream_t s);
// Sliced-ELL copy (ArnoldiBufs::sell_*) of a device CSR matrix whose slice
// offsets sell_ptr (n_slices + 1, from sell_layout) are on the device.  With
// sell_ptr set in the buffers, every SpMV launcher above except the fused
// spmv_dots uses it.
void launch_sell_build(const int* row_ptr, const int* col_idx, const void* vals, int vals_f32, int n,
                       const long long* sell_ptr, int* sell_col, void* sell_val, hipStream_t s);
// Diagonal copy (ArnoldiBufs::dia_*) of a device CSR matrix: n_diags x ld
// entries (ld >= n), offsets (ascending, on the device) from dia_offsets.
// With dia_val set in the buffers, every SpMV launcher above except the
// fused spmv_dots uses it (before the sliced-ELL copy).
void launch_dia_build(const int* row_ptr, const int* col_idx, const void* vals, int vals_f32, int n,
                      const int* offsets, int n_diags, long long ld, void* dia_val, hipStream_t s);
// Column j on a 3-, 5- or 7-diagonal layout with its first Gram-Schmidt pass
// (q_i . w for i <= j, and ||w||^2 at quantity j + 1, spmv_grid blocks)
// fused into the SpMV, for j + 1 <= kDiaFusedDots (spmv_dia_fuses_dots).
constexpr int kDiaFusedDots = 8;
bool spmv_dia_fuses_dots(const ArnoldiBufs& b, int nq);
void launch_spmv_dia_col_dots(const ArnoldiBufs& b, int j, hipStream_t s);
// Presence masks of the constant-diagonal form (ArnoldiBufs::dia_mask): ld
// entries, uint8 for n_diags <= 8 else uint16; with dia_mask and dia_cval
// set and dia_val null the diagonal SpMV reads them instead of values.
void launch_dia_mask_build(const int* row_ptr, const int* col_idx, int n, const int* offsets, int n_diags,
                           long long ld, void* mask, hipStream_t s);

// ------------------------------------------------------------ s-step blocks
//
// Block (s-step) Arnoldi: s basis vectors at a time from a Newton recurrence
// p_i = (A - theta_i) p_{i-1} / sigma_i, orthogonalised against the basis by
// one block Gram-Schmidt sweep pair (a second pair when the verification dots
// ask for it) and among themselves by Cholesky QR; the Hessenberg columns
// follow from the change of basis (sstep.hip).
constexpr int kSsMaxS = 16;     // block size limit: one 16-wide MFMA tile
constexpr int kSsMaxCols = 64;  // columns one block sweep covers: four tiles
constexpr int kSsMaxGrid = 2048;  // workgroups of a block sweep

// State of the block in progress (doubles), then its flags (ints).
enum SsSlot : int {
  kSsTheta = 0,                       // shifts theta_i
  kSsSigma = kSsTheta + kSsMaxS,      // sigma_i = ||(A - theta_i) p_{i-1}||
  kSsDots = kSsSigma + kSsMaxS,       // reduced sweep dots [64][16]
  kSsCtot = kSsDots + kSsMaxCols * kSsMaxS,  // P on Q_k [64][16]
  kSsM = kSsCtot + kSsMaxCols * kSsMaxS,     // raw-column coefficients of the next update [64][16]
  kSsR = kSsM + kSsMaxCols * kSsMaxS,        // R_tot [16][16]
  kSsDoubles = kSsR + kSsMaxS * kSsMaxS
};
enum SsFlag : int {
  kSsPass2 = 0,    // the second Gram-Schmidt update is due
  kSsNormReady,    // Cholesky factor ready: the normalising sweep is due
  kSsNorm2,        // the second Cholesky QR sweep is due
  kSsPasses,       // Gram-Schmidt updates taken in this block
  kSsBlocks,       // blocks accepted (statistics, kept across solves)
  kSsSecond,       // blocks that took a second update or Cholesky pass (statistics)
  kSsFailK,        // k of the block that failed (iscal[kSStepFail] set)
  kSsFused,        // the first update wrote Q_new = (P - Q_k C1) R^-1 (R from the dots)
  kSsFusedBlocks,  // blocks accepted after the fused update alone (statistics, kept across solves)
  kSsNumFlags = 16
};

struct SsArgs {
  double* V;             // the basis (ArnoldiBufs::V)
  long long ld;
  long long rows;        // rows swept: n rounded up to 16 (rows >= n are zero)
  double* colscale;
  double* H;
  int ldh;
  int* iscal;            // ArnoldiBufs::iscal (breakdown, kKUsed, kSStepFail)
  const double* spmv_partials;  // ArnoldiBufs::partials (the SpMV's ||y||^2)
  int spmv_grid;
  double* st;            // SsSlot state
  int* fl;               // SsFlag flags
  double* partials;      // block sweep partials: quantity c*16 + j of block g at (c*16 + j)*grid + g
  int grid;              // blocks of the block sweeps
};

// Launch sequence of one block, columns k .. k+s-1 from q_{k-1} (k >= 1,
// k + s <= 64, s <= 16); every launch returns at once after a breakdown or a
// failed block (iscal[kSStepFail]):
//   for i < s: launch_ss_spmv(i), launch_ss_sigma(i)      p_i, sigma_i
//   launch_ss_dots                                        Q^T P over [0, k+s)
//   launch_ss_reduce, launch_ss_plan(kSsPhaseA)
//   launch_ss_update(window [0, k+s), cond -1)            P' = P - Q_k C, or
//                                                         Q_new = P' R^-1 at once
//   launch_ss_reduce, launch_ss_plan(kSsPhaseB, call 0)   (accepts a Q_new that
//                                                         passes both checks)
//   the same three with cond kSsPass2 and call 1         second update if due
//   launch_ss_update(window [k, k+s), cond kSsNormReady)  Q_new = P' R^-1
//   launch_ss_reduce, launch_ss_plan(kSsPhaseD, call 0)
//   the same three with cond kSsNorm2 and call 1         second Cholesky QR if due
// A block that cannot be completed sets iscal[kSStepFail]; the caller then
// rebuilds the columns with the column-by-column path.
enum SsPhase : int { kSsPhaseA = 0, kSsPhaseB, kSsPhaseD };
void launch_ss_spmv(const ArnoldiBufs& b, const SsArgs& a, int k, int i, hipStream_t s);
void launch_ss_sigma(const ArnoldiBufs& b, const SsArgs& a, int k, int i, hipStream_t s);
void launch_ss_dots(const SsArgs& a, int k, int s, hipStream_t st);
void launch_ss_update(const SsArgs& a, int c0, int ncol, int out0, int s, int cond, hipStream_t st);
void launch_ss_reduce(const SsArgs& a, int ncol, int s, int cond, hipStream_t st);
void launch_ss_plan(const SsArgs& a, int phase, int call, int k, int s, int force_pass2, int cond,
                    hipStream_t st);
// Doubles of SsArgs::partials for a sweep grid.
inline size_t sstep_partials_doubles(int grid) {
  return static_cast<size_t>(kSsMaxCols) * kSsMaxS * static_cast<size_t>(grid);
}

void launch_expm(const ExpmArgs& a, int kk_max, hipStream_t s);
size_t expm_lds_bytes(int kk);
constexpr int kExpmMaxKLds = 62;  // must equal kMaxKLds in expm_core.hpp

// exp(A) of a dense k x k column-major matrix into d_E, asynchronous on s,
// with every working buffer in the library's static device scratch (no
// allocation); k in [1, kExpmNoAllocMaxK].  Calls that overlap on one device
// share that scratch and race.
constexpr int kExpmNoAllocMaxK = 128;
void expm_dense_noalloc(int k, const double* d_A, double* d_E, hipStream_t s);

}  // namespace mek
