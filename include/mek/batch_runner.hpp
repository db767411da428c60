// BatchRunner: many independent y_p = exp(t_p A_p) v_p solves of uneven size
// on one GPU, advanced in lockstep (additive; the reference solves one
// problem per process).  Each problem follows the reference controller
// (main.cu:1425-1518) or Expokit step control with its own t, m, tol and step
// cap; the controller runs on the device and one cycle of the solve is
// replayed from a hipGraph.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mek/csr_host.hpp"

namespace mek {

struct BatchProblem {
  const CSRHost* A;
  double t;
  int m;               // Krylov dimension (1..62)
  double tol;
  int max_restarts;    // as ArnoldiParams::max_restarts
  int substeps;        // >0: first sub-step t/substeps and no step cap (learned selector)
  int control = 0;     // 0 reference controller, 1 Expokit, 2 Expokit relative
};

// Launch sizes of one cycle, per Arnoldi column j: the problems are stored in
// order of decreasing m, so column j involves the first np[j] problems, their
// nsb[j] streaming blocks and nrb[j] SpMV row blocks.
struct BatchPlan {
  std::vector<int> np, nsb, nrb, spmv_grid;
};

struct BatchProblemStats {
  int converged, accepted, attempts, rejected, columns, aborted, reorth_columns;
  double t_done, last_residual;
};

class BatchRunner {
 public:
  // values_f32: 0 FP64 values; 1 FP32 when every value of every problem is
  // exact in FP32, else FP64; 2 FP32 always (the caller accepts the loss).
  // gs_mode: 0 never, 1 always (reference), 2 DGKS with eta.
  // autotune: time the SpMV grid at the first solve and keep the fastest.
  // t_p must be finite and >= 0; a problem with t_p = 0 returns y_p = v_p.
  BatchRunner(int device_id, const std::vector<BatchProblem>& problems, int values_f32,
              int gs_mode, double dgks_eta, bool use_graph, bool autotune = false);
  ~BatchRunner();
  BatchRunner(const BatchRunner&) = delete;
  BatchRunner& operator=(const BatchRunner&) = delete;

  int num_problems() const { return P_; }
  long long total_rows() const { return n_tot_; }
  // rows of problem p (the caller's index) in the batch's global row space;
  // each problem takes an even number of rows there (one empty padding row
  // after an odd count), so that the sweeps can take two rows per lane
  int row_offset(int p) const { return row0_[ext2int_[p]]; }
  int rows(int p) const { return nrows_[ext2int_[p]]; }

  // Start vectors (host, one array per problem) into the device copy of v.
  void upload_v(const std::vector<const double*>& v);
  // Solve all problems from the uploaded v; y stays on the device.
  void solve_device();
  void download_y(const std::vector<double*>& y) const;
  BatchProblemStats stats(int p) const;
  // The 16 learned-selector features of every problem (P x 16, host).
  void features(float* out);
  // Median microseconds of the first column's SpMV over every problem (all
  // of them active, the solver's SpMV grid) over `reps` timed launches; the
  // batch is left at its start state.
  double time_spmv(int reps);
  int cycles() const { return cycles_; }
  int spmv_grid() const { return tuned_spmv_grid_; }
  bool values_f32_used() const { return values_f32_used_; }
  double solve_ms() const { return solve_ms_; }
  hipStream_t stream() const { return stream_; }
  void destroy();

 private:
  void reset_state();
  void run_cycle_eager(int mcols);
  hipGraphExec_t cycle_graph(int mcols);
  int P_ = 0;
  int M_ = 0;
  long long n_tot_ = 0;
  std::vector<int> row0_;   // padded offsets (even), P + 1 entries
  std::vector<int> nrows_;  // rows of each problem
  int device_ = 0;
  bool use_graph_ = true;
  int gs_mode_ = 1;
  double eta_ = 0.1;
  bool autotune_ = false;
  bool values_f32_used_ = false;
  int tuned_spmv_grid_ = 0;
  std::vector<int> ext2int_;   // caller's problem index -> storage order
  BatchPlan plan_;
  hipStream_t stream_ = nullptr;
  // one captured cycle per basis size in use: a cycle builds only as many
  // columns as the largest m among the still-active problems
  std::vector<hipGraphExec_t> graph_exec_;
  std::vector<hipGraph_t> graph_;
  int cycles_ = 0;
  double solve_ms_ = 0.0;
  // device buffers
  int *d_row_ptr_ = nullptr, *d_col_idx_ = nullptr, *d_rb_start_ = nullptr, *d_rb_prob_ = nullptr;
  int *d_sb_prob_ = nullptr, *d_sb_row0_ = nullptr, *d_p_sb0_ = nullptr, *d_pint_ = nullptr;
  // per-cycle compaction: first row block of each problem, problems per
  // column, the compacted block lists and their counts
  int *d_p_rb0_ = nullptr, *d_np_col_ = nullptr, *d_rb_list_ = nullptr, *d_sb_list_ = nullptr;
  int* d_list_cnt_ = nullptr;
  int *d_counters_ = nullptr, *h_counters_ = nullptr;
  void* d_vals_ = nullptr;
  double *d_V_ = nullptr, *d_w_ = nullptr, *d_y_ = nullptr, *d_v0_ = nullptr, *d_small_ = nullptr;
  double *d_partials_ = nullptr, *d_expm_ws_ = nullptr, *d_pscal_ = nullptr;
  // feature extraction (allocated by the first features() call)
  int *d_row0_ = nullptr, *d_feat_int_ = nullptr;
  double* d_feat_dbl_ = nullptr;
  float* d_feat_out_ = nullptr;
  std::vector<int> pint_init_;
  std::vector<double> pscal_init_;
  std::vector<double> norm1_;  // ||A_p||_1 (host), a selector feature
  void* bufs_ = nullptr;  // BatchBufs (opaque here)
};

}  // namespace mek
