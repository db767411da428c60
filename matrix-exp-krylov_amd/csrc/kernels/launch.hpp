// Host-side launchers of the gfx950 kernels (defined in arnoldi.hip, expm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace mek {

constexpr int kMaxFusedDots = 64;  // dots carried in registers by one sweep
constexpr int kSpmvNnzCap = 2048;  // must equal kNnzCap in arnoldi.hip
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
