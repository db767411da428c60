// BlockRunner: Y = exp(t A) V for a block V of 32 right-hand sides on one
// shared A, by block Arnoldi (additive; the reference solves one vector per
// process, main.cu:1411-1571).  A Krylov step multiplies A by a 32-column
// block (SpMM), orthogonalises it against the basis by block classical
// Gram-Schmidt with a second pass on FP64 MFMA, and orthonormalises it inside
// by Cholesky QR; the block Hessenberg exponential is a Taylor polynomial
// with scaling and squaring on a dense FP64 MFMA GEMM.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mek/csr_host.hpp"

namespace mek {

struct BlockStats {
  int converged = 0;
  int accepted = 0;
  int attempts = 0;
  int rejected = 0;
  int block_steps = 0;     // SpMMs of the basis
  int squarings = 0;       // in the last exponential
  int second_passes = 0;   // block steps that took the second Gram-Schmidt pass
  double t_done = 0.0;
  double last_residual = 0.0;
  double solve_ms = 0.0;
  // HIP-event kernel time (ms) and algorithmic HBM bytes per kind:
  // 0 spmm, 1 dots, 2 update, 3 expm, 4 lift
  double kernel_ms[5] = {0, 0, 0, 0, 0};
  double kernel_bytes[5] = {0, 0, 0, 0, 0};
  long long launches[5] = {0, 0, 0, 0, 0};
};

class BlockRunner {
 public:
  // m: block Arnoldi steps (basis of 32 m vectors); substeps > 0 fixes the
  // first sub-step to t / substeps, 0 takes Expokit's first-step formula.
  // The error test is relative: ||H(k,k-1) E_k R0||_F <= 1.2 (dt/t) tol
  // min(||V_step||_F, ||Y_step||_F) per sub-step.
  // gs_mode: 1 second Gram-Schmidt pass always, 2 only when a column of the
  // new block keeps less than eta of its norm in the first pass (DGKS), 0 never.
  // autotune: time the SpMM and update-sweep grids at the first solve.
  BlockRunner(const CSRHost& A, int m, double t, double tol, int substeps, int device_id,
              bool timing = false, int gs_mode = 2, double eta = 0.1, bool autotune = false);
  ~BlockRunner();
  BlockRunner(const BlockRunner&) = delete;
  BlockRunner& operator=(const BlockRunner&) = delete;

  int rows() const { return n_; }
  // V, Y: n x 32 row-major (host or device).
  void expmv(const double* h_V, double* h_Y);
  void expmv_device(const double* d_V, double* d_Y);
  const BlockStats& stats() const { return stats_; }
  int spmm_grid() const { return tuned_spmm_grid_; }
  int update_grid() const { return tuned_update_grid_; }
  hipStream_t stream() const { return stream_; }
  void destroy();

 private:
  void start_block(const double* d_src, bool device_src, const double* h_src);
  void autotune();
  int build_basis();
  double exponential(double dt, int k);  // returns ||dt H||_1; F in the work matrix
  void timed_begin(int kind);
  void timed_end(int kind, double bytes);
  void flush_timing();

  int n_ = 0;
  long long n_pad_ = 0;
  long long nnz_ = 0;
  int m_ = 0;
  double t_ = 0.0, tol_ = 0.0, anorm_ = 1.0;
  int substeps_ = 0;
  int device_ = 0;
  bool timing_ = false;
  int gs_mode_ = 2;
  double eta_ = 0.1;
  bool autotune_ = false;
  int tuned_spmm_grid_ = 0, tuned_update_grid_ = 0;
  hipStream_t stream_ = nullptr;
  void* bufs_ = nullptr;  // BlockBufs
  // device memory
  int *d_row_ptr_ = nullptr, *d_col_idx_ = nullptr, *d_flags_ = nullptr;
  double *d_vals_ = nullptr, *d_P_ = nullptr, *d_W_ = nullptr, *d_small_ = nullptr;
  double *d_partials_ = nullptr, *d_mat_ = nullptr, *h_pinned_ = nullptr;
  int* h_flags_ = nullptr;
  int npad_max_ = 0;  // largest exponential order (multiple of 64)
  double* F_ = nullptr;  // result of the last exponential (column-major, ld F_ld_)
  int F_ld_ = 0;
  BlockStats stats_;
  struct Ev {
    hipEvent_t a, b;
    int kind;
    double bytes;
  };
  std::vector<Ev> ev_;
  size_t ev_used_ = 0;
};

}  // namespace mek
