// DeviceContext: one GPU's resources for the Arnoldi solve.
// reference: struct DeviceContext, main.cu:747-1125 (handles, streams, device
// CSR, V_m, work vectors, scalar scratch, SpMV buffer, destroy()).
//
// MI355X layout: the basis V is column-major with a leading dimension padded
// to 64 doubles (512 B), holding m+2 columns; values are stored as FP64 or, in
// the mixed form, FP32 (x, y and every accumulation stay FP64).  One HIP
// stream.  There are no vendor BLAS/sparse handles: every kernel is this
// project's own.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mek/csr_host.hpp"

namespace mek {
struct ArnoldiBufs;
struct SsArgs;
}

struct DeviceContext {
  int device_id;
  hipStream_t stream_compute = nullptr;

  // device CSR
  int* d_row_ptr = nullptr;
  int* d_col_idx = nullptr;
  void* d_values = nullptr;       // double, or float when values_f32
  bool values_f32 = false;
  int* d_row_blocks = nullptr;    // SpMV row blocks (the "SpMV buffer")
  int n_row_blocks = 0;

  // Krylov basis and work vectors
  double* d_V_m = nullptr;        // ld x (allocated_m_max + 2)
  long long ld = 0;
  double* d_w = nullptr;
  double* d_y = nullptr;
  double* d_small = nullptr;      // colscale, h1, h2, H, wH, scalars (one block)
  double* d_partials = nullptr;   // per-block partial sums
  double* d_expm_ws = nullptr;    // global expm workspace for large m
  double* h_status = nullptr;     // pinned host copy of the scalar slots
  int grid = 0;                   // blocks of the streaming kernels
  int spmv_grid = 1;              // blocks of the SpMV kernel

  // sliced-ELL copy of the matrix (build_sell): when set, the SpMV reads it
  // instead of the CSR row blocks
  long long* d_sell_ptr = nullptr;  // n_slices + 1 entry offsets
  int* d_sell_col = nullptr;        // padding: -1
  void* d_sell_val = nullptr;       // double, or float when values_f32
  int n_slices = 0;
  long long sell_entries = 0;       // entries with padding
  int sell_width = 0;               // widest slice (entries per row)

  // diagonal copy of the matrix (build_dia): when set, the SpMV reads it
  // instead of the sliced-ELL copy or the CSR row blocks
  void* d_dia_val = nullptr;        // n_diags x dia_ld, diagonal-major; absent entries hold the marker
  int* d_dia_off = nullptr;         // column - row of each diagonal, ascending
  int n_diags = 0;
  long long dia_ld = 0;             // rows rounded up to 64
  // constant-diagonal form (build_dia with values; d_dia_val stays null):
  // one value per diagonal and a presence mask per row
  void* d_dia_mask = nullptr;       // dia_ld entries: uint8 (n_diags <= 8) or uint16, bit d = diagonal d
  double* d_dia_cval = nullptr;     // n_diags values

  // s-step blocks (allocated by alloc_sstep): state, flags, sweep partials
  double* d_ss_state = nullptr;
  int* d_ss_flags = nullptr;
  double* d_ss_partials = nullptr;
  int ss_grid = 0;
  int ss_grid_small = 0;          // for block windows of up to 32 columns (0: ss_grid)

  int local_rows = 0;
  long long local_nnz = 0;
  int allocated_m_max = 0;
  bool log = true;                // print the reference's allocation lines

  explicit DeviceContext(int dev_id) : device_id(dev_id) {}

  void init_handles();
  void create_streams();
  // reference: alloc_from (main.cu:862-897).  Additive last argument: store
  // the values in FP32 (the caller decides whether that is lossless).
  void alloc_from(const CSRHost& part, int global_cols, int m_alloc_max, bool store_f32 = false);
  // reference: create_descriptors (main.cu:900-1088).  On one GPU the matrix
  // has no off-diagonal block; the call uploads the adaptive SpMV row blocks.
  void create_descriptors(const CSRHost& part, int global_cols, const CSRHost::GhostMap& ghost_map,
                          int global_row_offset, int m);
  void destroy();
  // Additive: the sliced-ELL copy of the device CSR (after alloc_from).
  // format 0: built when its padding is at most mek::kSellMaxPad of nnz and
  // no slice is wider than mek::kSellMaxWidth; 1: never (CSR row blocks);
  // 2: always.  Returns whether the SpMV uses it.
  bool build_sell(const CSRHost& part, int format);
  void free_sell();
  // Additive: the diagonal copy of the device CSR (after alloc_from), for
  // matrices whose entries lie on at most mek::kDiaMaxDiags diagonals and
  // whose rows list their columns in ascending order (mek::dia_offsets).
  // With cvals (one value per diagonal, from mek::dia_constants) it builds
  // the constant-diagonal form instead: presence masks and the values, no
  // per-entry copy.  Returns whether it was built.
  bool build_dia(const std::vector<int>& offsets, const std::vector<double>& cvals = {});
  void free_dia();
  // Additive: the SpMV layout.  format 0: the diagonal layout when the
  // matrix has one with at most mek::kDiaMaxPad padding (its constant form
  // when every diagonal holds one value), else the sliced-ELL copy under
  // build_sell's rule, else the CSR row blocks; 1 CSR row blocks; 2 sliced
  // ELL; 3 diagonal copy with per-entry values; 4 constant diagonals (3 and
  // 4 are errors when the matrix has no such form).  Returns the layout in
  // use (1, 2, 3 or 4).
  int build_spmv_layout(const CSRHost& part, int format);
  // Additive: buffers of the s-step block path (sstep.hip), shifts uploaded.
  void alloc_sstep(const std::vector<double>& shifts);
  mek::SsArgs ss_args() const;

  // Additive: device buffer bundle handed to the kernels.
  mek::ArnoldiBufs bufs() const;
  size_t small_doubles() const;  // size of d_small
};

namespace mek {
// Adaptive CSR row blocks: consecutive rows whose nnz sum to at most nnz_cap
// (and at most row_cap rows); a row longer than nnz_cap is a block of its own.
// Returns the n_blocks+1 block start rows.
std::vector<int> build_row_blocks(const int* row_ptr, int n, int nnz_cap, int row_cap);

// Sliced-ELL layout: slices of 64 consecutive rows, each padded to its
// longest row.  Fills the n_slices + 1 entry offsets and returns the padded
// entry count; *max_width receives the widest slice's row length.
long long sell_layout(const int* row_ptr, int n, std::vector<long long>& ptr, int* max_width);
constexpr double kSellMaxPad = 0.15;  // automatic choice: padding <= 15 % of nnz
constexpr int kSellMaxWidth = 256;    // ... and no row longer than this

// Diagonal layout: the distinct column - row offsets of a CSR matrix, in
// ascending order, when there are at most kDiaMaxDiags of them, every row
// lists its columns in strictly ascending order (so summing a row diagonal
// by diagonal adds its products in CSR order) and no value is NaN (a NaN
// marks the absent entries of the copy); empty otherwise.
constexpr int kDiaMaxDiags = 16;
constexpr double kDiaMaxPad = 0.2;    // automatic choice: n * diagonals <= 1.2 nnz
std::vector<int> dia_offsets(const int* row_ptr, const int* col_idx, const double* values, int n);
// Constant-diagonal form: the value of each diagonal of `offsets` (from
// dia_offsets) when all entries on it have the same bits as stored on the
// device (rounded to FP32 first when f32), else empty.  The SpMV multiplies
// by that value, so the products are those of the per-entry copy.
std::vector<double> dia_constants(const int* row_ptr, const int* col_idx, const double* values, int n,
                                  const std::vector<int>& offsets, bool f32);
}  // namespace mek
