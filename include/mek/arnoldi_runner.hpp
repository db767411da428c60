// ArnoldiRunner: y = exp(t*A)*v by adaptive time-stepping Arnoldi.
// reference: ArnoldiParams / ExpmvMode / TimeSteppingParams / ArnoldiRunner,
// main.cu:1220-2081.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "mek/autotune.hpp"
#include "mek/csr_host.hpp"
#include "mek/device_context.hpp"

// reference: main.cu:81-82
constexpr double ARNOLDI_TOL = 1e-8;
constexpr int MAX_RESTARTS = 10;

// reference: main.cu:1221-1230
struct ArnoldiParams {
  int m;
  double t;
  double tol;
  int max_restarts;
  bool reorthogonalize;
  ArnoldiParams(int m_val = 100, double t_val = 1.0, double tol_val = ARNOLDI_TOL,
                int max_restarts_val = MAX_RESTARTS, bool reortho = true)
      : m(m_val), t(t_val), tol(tol_val), max_restarts(max_restarts_val), reorthogonalize(reortho) {}
};

// reference: main.cu:1232-1235
enum class ExpmvMode { AdaptiveTimeStepping, RestartedKrylov };

// reference: main.cu:1237-1250
struct TimeSteppingParams {
  double dt_init{0.0};
  double dt_min{0.0};
  double dt_max{0.0};
  double tol_step{0.0};
  int max_steps{0};
  int max_rejects{20};
  double safety{0.8};
  double grow{1.5};
  double shrink{0.5};
  bool adapt_m{false};
  int m_max{0};
  int m_step{0};
};

namespace mek {

// Additive: how sub-steps are chosen.
enum class StepControl {
  Reference = 0,  // the reference controller (main.cu:1425-1518)
  Expokit = 1,    // Sidje 1998 (ACM TOMS 24(1)) step-size heuristic
  // Expokit with the local error measured relative to the norm of the
  // step's start vector: err_loc <= 1.2 dt tol ||w||
  ExpokitRelative = 2,
};

// Additive solver options (defaults reproduce the reference exactly).
struct SolverOptions {
  StepControl control = StepControl::Reference;
  // Second Gram-Schmidt pass: -1 follows ArnoldiParams::reorthogonalize
  // (always / never); 2 = conditional (DGKS test ||w'|| < eta ||w||).
  int gs_mode = -1;
  double dgks_eta = 0.1;
  // Matrix value storage on the device: 0 FP64; 1 FP32 when every value is
  // exact in FP32, else FP64 (FP32 storage would move y by about t ||dA||,
  // far above a 1e-8 tolerance); 2 FP32 always (the caller accepts the loss).
  int values_f32 = 0;
  bool verbose = true;      // print the reference's per-column and [STEP] lines
  bool timing = false;      // HIP-event timing of every kernel launch
  bool fuse_spmv_dot = false;  // first-pass dots inside the SpMV kernel
  // Symmetric A only (checked at init): the Lanczos three-term recurrence
  // instead of Gram-Schmidt; H is tridiagonal.
  bool lanczos = false;
  // Time candidate SpMV row-block caps and grids and sweep grids on this
  // matrix at init and keep the fastest (mek/autotune.hpp).
  bool autotune = false;
  // s-step blocks (2..16; 0 off): after the first sstep_first columns of
  // each Arnoldi cycle (-1: sstep columns) the basis grows sstep vectors at a
  // time from a Newton recurrence with shifts at Chebyshev points of the
  // Gershgorin interval, each block orthogonalised by one block Gram-Schmidt
  // sweep pair (two when needed) and Cholesky QR, so the basis is read twice
  // per block instead of twice per column.  Needs m + 1 <= 64; a block that
  // cannot be completed falls back to the column-by-column path for the
  // rest of that cycle.
  int sstep = 0;
  int sstep_first = -1;
  // SpMV layout: 0 the diagonal copy when the entries lie on at most 16
  // diagonals with little padding (stencils), else the sliced-ELL copy when
  // its padding is small (banded matrices), else the CSR row blocks; 1 CSR
  // row blocks; 2 sliced ELL; 3 diagonal.  All give the same bits.
  int spmv_format = 0;
};

struct StepRecord {
  double t, dt, err;
  int m_used, accepts, rejects;
};

// Kernel classes for event timing.
// kKUpdateNorm also holds the conditional passes of the s-step blocks (second
// update, normalising sweeps), which count no bytes since they may return at
// once; kKSsDots / kKSsUpdate are the s-step blocks' dots and first update.
enum KernelKind : int {
  kKSpmv = 0, kKUpdateDot, kKUpdateNorm, kKDots, kKFinal, kKExpm, kKLift, kKOther, kKSsDots, kKSsUpdate, kNumKinds
};

struct KernelTiming {
  double ms = 0.0;        // summed event time
  long long launches = 0;
  double bytes = 0.0;     // summed algorithmic bytes
};

struct SolveStats {
  int attempts = 0;
  int accepted = 0;
  int rejected = 0;
  int spmv_count = 0;       // Krylov columns built (SpMV launches of the basis)
  int extra_spmv = 0;       // Expokit ||A v_{m+1}|| products
  int reorth_columns = 0;
  double t_done = 0.0;
  double solve_ms = 0.0;    // the METRIC total_compute_expmv_ms span
  double err_sum = 0.0;     // Expokit: summed local error estimates
  bool aborted = false;
  int sstep_blocks = 0;     // s-step blocks accepted
  int sstep_second = 0;     // of them, blocks that took a second update or Cholesky pass
  int sstep_fallbacks = 0;  // cycles rebuilt column by column after a failed block
  int sstep_fused = 0;      // blocks done by the fused update alone (no normalising sweep)
  std::vector<StepRecord> steps;
  KernelTiming kernels[kNumKinds];
};

}  // namespace mek

struct ArnoldiRunner {
  int num_gpus;
  std::vector<DeviceContext> device_contexts;
  ArnoldiParams params;
  ExpmvMode mode{ExpmvMode::AdaptiveTimeStepping};
  TimeSteppingParams ts_params;
  bool converged = false;
  int restarts_done = 0;
  double last_residual = 0.0;
  double current_segment_t = 0.0;
  std::vector<CSRHost::GhostMap> ghost_maps;
  double v_norm = 0.0;
  int global_rows;
  int global_cols;
  long long global_nnz;

  // Additive
  mek::SolverOptions options;
  mek::SolveStats stats;
  mek::TuneChoice tuning;  // additive: launch shapes chosen when options.autotune
  bool values_f32_used = false;  // additive: the values are stored in FP32
  double anorm_inf = 0.0;   // ||A||_inf, for the Expokit initial step

  ArnoldiRunner(int n_gpus, const ArnoldiParams& arnoldi_params, int total_rows, int total_cols,
                long long total_nnz);

  // reference: main.cu:1308-1409.  One GPU: device_ids.size() must be 1.
  void init_all_devices(const std::vector<CSRHost>& partitions, const std::vector<int>& device_ids,
                        const std::vector<CSRHost::GhostMap>& ghost_maps_in);
  // Additive: the same for the whole matrix on one device, reading A in
  // place (no partition copy).
  void init_device(const CSRHost& A, int device_id);
  // reference: main.cu:1411-1571.  Host v in, host y out (y pre-sized).
  void compute_expmv(const std::vector<double>& v_host, std::vector<double>& y_host);
  void destroy();

  // Additive: device-resident input and output (device pointers of n doubles).
  void compute_expmv_device(const double* d_v, double* d_y);
  // Additive: compute_expmv on host arrays of n doubles (no vector copies).
  // v and y cross PCIe through a pinned staging buffer of n doubles (kept by
  // the runner), filled and emptied by several host threads.
  void compute_expmv_host(const double* v_host, double* y_host);
  // Additive: one Arnoldi decomposition from v (no time stepping): fills
  // H ((m+1) x m, column-major) and returns the columns built.
  int arnoldi_decomposition(const std::vector<double>& v_host, std::vector<double>& H_out);

 private:
  void init_one(const CSRHost& part, int device_id);
  void solve_on_device(const double* d_v_or_null, const double* h_v_or_null);
  void run_columns(int m_cols);
  void run_columns_range(int j0, int j1);
  void run_sstep_block(int k, int ss);
  void run_lanczos_columns(int m_cols);
  int read_status();
  void begin_timed(int kind);
  void end_timed(int kind, double bytes);
  void flush_timing();
  void init_q1_device(const double* d_v, const double* h_v);
  void launch_expm(double dt, int mode);
  void lift(int ncoef, double skip_below);
  void restart_from_y();
  void reference_controller();
  void expokit_controller();

  bool ss_active_ = false;
  std::vector<double> ss_shifts_;
  int ss_blocks_before_ = 0, ss_second_before_ = 0, ss_fused_before_ = 0;
  void stage_copy(double* dst, const double* src, size_t n);
  void stage_pipelined(double* dst, const double* src, size_t n, size_t nchunk,
                       const std::function<void(size_t, size_t)>& after, const hipEvent_t* ready);
  double* h_stage_ = nullptr;  // pinned staging buffer for host v and y
  size_t h_stage_n_ = 0;
  std::vector<hipEvent_t> stage_ev_;  // one per chunk of the staged copy of y
  // host v and y cross in chunks (copy of one chunk overlapping the DMA of
  // the one before) when n holds at least two chunks of kStageChunkMin
  static constexpr size_t kStageChunks = 8;
  static constexpr size_t kStageChunkMin = size_t(1) << 20;  // doubles
  struct EventPair { hipEvent_t a, b; int kind; double bytes; };
  std::vector<EventPair> ev_pool_;
  size_t ev_used_ = 0;
};
