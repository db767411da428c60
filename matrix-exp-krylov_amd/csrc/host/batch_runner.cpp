// BatchRunner: lockstep batched exp(t_p A_p) v_p on one MI355X (batch.hip).
#include "mek/batch_runner.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "../kernels/batch.hpp"
#include "../kernels/launch.hpp"
#include "../kernels/learn.hpp"
#include "mek/autotune.hpp"
#include "hip_check.hpp"
#include "mek/device_context.hpp"

namespace mek {

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

template <typename T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  MEK_HIP(hipMalloc(&d, sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) MEK_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return d;
}

}  // namespace

BatchRunner::BatchRunner(int device_id, const std::vector<BatchProblem>& probs_in, int values_f32,
                         int gs_mode, double dgks_eta, bool use_graph, bool autotune)
    : P_(static_cast<int>(probs_in.size())), device_(device_id), use_graph_(use_graph),
      gs_mode_(gs_mode), eta_(dgks_eta), autotune_(autotune) {
  if (P_ < 1) throw Error("ERROR: a batch needs at least one problem");
  if (values_f32 < 0 || values_f32 > 2) throw Error("ERROR: values_f32 must be 0, 1 or 2");
  for (int p = 0; p < P_; ++p)
    if (!(probs_in[p].t >= 0.0) || !std::isfinite(probs_in[p].t))
      throw Error("ERROR: batch problem " + std::to_string(p) + ": t must be finite and >= 0");
  // a failure part way releases what was already allocated (the destructor
  // does not run for an object whose constructor threw)
  try {
    MEK_HIP(hipSetDevice(device_));
    // internal order: decreasing m (stable), so Arnoldi column j concerns a
    // prefix of the problems; callers keep their own indices
    std::vector<int> order(P_);
    for (int p = 0; p < P_; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return probs_in[a].m > probs_in[c].m; });
    ext2int_.assign(P_, 0);
    for (int i = 0; i < P_; ++i) ext2int_[order[i]] = i;
    std::vector<BatchProblem> probs(P_);
    for (int i = 0; i < P_; ++i) probs[i] = probs_in[order[i]];
    // ---- host assembly: block-diagonal CSR, row blocks, streaming blocks
    row0_.assign(P_ + 1, 0);
    nrows_.assign(P_, 0);
    long long nnz_tot = 0;
    for (int p = 0; p < P_; ++p) {
      const CSRHost& A = *probs[p].A;
      if (A.rows != A.cols) throw Error("ERROR: Matrix must be square (batch problem " + std::to_string(p) + ")");
      if (probs[p].m < 1 || probs[p].m > kBatchMaxM || probs[p].m >= A.rows)
        throw Error("ERROR: batch problem " + std::to_string(p) + ": m=" + std::to_string(probs[p].m) +
                    " outside [1, min(" + std::to_string(kBatchMaxM) + ", rows-1)]");
      // an odd row count gets one empty padding row: every problem starts on
      // an even row, as the two-rows-per-lane sweeps need
      nrows_[p] = A.rows;
      row0_[p + 1] = row0_[p] + A.rows + (A.rows & 1);
      nnz_tot += A.nnz;
      M_ = std::max(M_, probs[p].m);
    }
    n_tot_ = row0_[P_];
    if (n_tot_ > 0x7fffffffLL || nnz_tot > 0x7fffffffLL) throw Error("ERROR: batch exceeds int32 rows/nnz");
    std::vector<int> row_ptr(static_cast<size_t>(n_tot_) + 1), col_idx(static_cast<size_t>(nnz_tot));
    std::vector<double> vals(static_cast<size_t>(nnz_tot));
    std::vector<int> rb_start, rb_prob, sb_row0, sb_prob, p_sb0(P_ + 1), p_rb0(P_ + 1);
    long long k = 0;
    for (int p = 0; p < P_; ++p) {
      const CSRHost& A = *probs[p].A;
      const int r0 = row0_[p];
      for (int r = 0; r < A.rows; ++r) row_ptr[r0 + r] = static_cast<int>(k + A.row_ptr[r]);
      for (long long e = 0; e < A.nnz; ++e) {
        col_idx[k + e] = A.col_idx[e] + r0;
        vals[k + e] = A.values[e];
      }
      k += A.nnz;
      if (A.rows & 1) row_ptr[r0 + A.rows] = static_cast<int>(k);  // the empty padding row
      const std::vector<int> rb = build_row_blocks(A.row_ptr.data(), A.rows, kBatchNnzCap, kSpmvRowCap);
      p_rb0[p] = static_cast<int>(rb_prob.size());
      for (size_t i = 0; i + 1 < rb.size(); ++i) {
        rb_start.push_back(r0 + rb[i]);
        rb_prob.push_back(p);
      }
      p_sb0[p] = static_cast<int>(sb_prob.size());
      for (int r = 0; r < row0_[p + 1] - r0; r += kBatchRowsPerBlock) {
        sb_row0.push_back(r0 + r);
        sb_prob.push_back(p);
      }
    }
    row_ptr[n_tot_] = static_cast<int>(k);
    rb_start.push_back(static_cast<int>(n_tot_));
    sb_row0.push_back(static_cast<int>(n_tot_));
    p_sb0[P_] = static_cast<int>(sb_prob.size());
    p_rb0[P_] = static_cast<int>(rb_prob.size());
    const int nsb = static_cast<int>(sb_prob.size());
    const int nrb = static_cast<int>(rb_prob.size());
    // per-column launch plan over the active prefix; SpMV grids are multiples
    // of 8 (one share per XCD)
    plan_.np.assign(M_, 0);
    plan_.nsb.assign(M_, 0);
    plan_.nrb.assign(M_, 0);
    plan_.spmv_grid.assign(M_, 8);
    for (int j = 0; j < M_; ++j) {
      int np = 0;
      while (np < P_ && probs[np].m > j) ++np;
      plan_.np[j] = np;
      plan_.nsb[j] = p_sb0[np];
      plan_.nrb[j] = p_rb0[np];
      const int g = std::min(plan_.nrb[j], 2048);
      plan_.spmv_grid[j] = std::max(8, (g + 7) / 8 * 8);
    }

    pint_init_.assign(static_cast<size_t>(P_) * kPInt, 0);
    norm1_.assign(P_, 0.0);
    int any_expokit = 0;
    pscal_init_.assign(static_cast<size_t>(P_) * kPScal, 0.0);
    for (int p = 0; p < P_; ++p) {
      pint_init_[p * kPInt + kBM] = probs[p].m;
      pint_init_[p * kPInt + kBMaxRestarts] = probs[p].max_restarts;
      pint_init_[p * kPInt + kBSubsteps] = probs[p].substeps;
      pint_init_[p * kPInt + kBCtrl] = probs[p].control;
      pscal_init_[p * kPScal + kBT] = probs[p].t;
      pscal_init_[p * kPScal + kBTol] = probs[p].tol;
      pscal_init_[p * kPScal + kBAnorm] = probs[p].A->norm_inf();
      norm1_[p] = probs[p].A->norm_1();  // a selector feature
      if (probs[p].control >= 1) any_expokit = 1;
    }

    // FP32 values only where that stores A exactly, unless the caller asked
    // for FP32 regardless (values_f32 == 2)
    values_f32_used_ = values_f32 == 2;
    if (values_f32 == 1) {
      values_f32_used_ = true;
      for (int p = 0; p < P_ && values_f32_used_; ++p)
        if (probs[p].A->f32_inexact(nullptr) > 0) values_f32_used_ = false;
    }

    // ---- device
    MEK_HIP(hipStreamCreate(&stream_));
    d_row_ptr_ = upload(row_ptr);
    d_col_idx_ = upload(col_idx);
    if (values_f32_used_) {
      std::vector<float> f(vals.begin(), vals.end());
      d_vals_ = upload(f);
    } else {
      d_vals_ = upload(vals);
    }
    d_rb_start_ = upload(rb_start);
    d_rb_prob_ = upload(rb_prob);
    d_sb_row0_ = upload(sb_row0);
    d_sb_prob_ = upload(sb_prob);
    d_p_sb0_ = upload(p_sb0);
    d_p_rb0_ = upload(p_rb0);
    d_np_col_ = upload(plan_.np);
    MEK_HIP(hipMalloc(&d_rb_list_, sizeof(int) * std::max(nrb, 1)));
    MEK_HIP(hipMalloc(&d_sb_list_, sizeof(int) * std::max(nsb, 1)));
    MEK_HIP(hipMalloc(&d_list_cnt_, sizeof(int) * (2 * static_cast<size_t>(M_) + 2)));
    MEK_HIP(hipMemset(d_list_cnt_, 0, sizeof(int) * (2 * static_cast<size_t>(M_) + 2)));
    d_row0_ = upload(std::vector<int>(row0_.begin(), row0_.end() - 1));
    const long long ld = (n_tot_ + 63) / 64 * 64;
    const size_t ncols = static_cast<size_t>(M_) + 2;
    MEK_HIP(hipMalloc(&d_V_, sizeof(double) * ld * ncols));
    MEK_HIP(hipMalloc(&d_w_, sizeof(double) * n_tot_));
    MEK_HIP(hipMalloc(&d_y_, sizeof(double) * n_tot_));
    MEK_HIP(hipMalloc(&d_v0_, sizeof(double) * n_tot_));
    // padding rows stay zero in every vector: the SpMV writes zero there (an
    // empty row) or leaves them alone, and the sweeps compute zero from zeros
    MEK_HIP(hipMemset(d_V_, 0, sizeof(double) * ld * ncols));
    MEK_HIP(hipMemset(d_w_, 0, sizeof(double) * n_tot_));
    MEK_HIP(hipMemset(d_y_, 0, sizeof(double) * n_tot_));
    MEK_HIP(hipMemset(d_v0_, 0, sizeof(double) * n_tot_));
    const size_t per = 3 * ncols + (ncols) * (ncols) + (M_ + 4);
    MEK_HIP(hipMalloc(&d_small_, sizeof(double) * per * P_));
    MEK_HIP(hipMemset(d_small_, 0, sizeof(double) * per * P_));
    MEK_HIP(hipMalloc(&d_pscal_, sizeof(double) * kPScal * P_));
    MEK_HIP(hipMalloc(&d_pint_, sizeof(int) * kPInt * P_));
    MEK_HIP(hipMalloc(&d_partials_, sizeof(double) * (ncols + 1) * nsb));
    const int ord = any_expokit ? M_ + 2 : M_;  // largest exponential order
    if (ord > kExpmMaxKLds) MEK_HIP(hipMalloc(&d_expm_ws_, sizeof(double) * 6 * ord * ord * P_));
    MEK_HIP(hipMalloc(&d_counters_, sizeof(int) * 4));
    MEK_HIP(hipMemset(d_counters_, 0, sizeof(int) * 4));
    MEK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_counters_), sizeof(int) * 4));

    auto* b = new BatchBufs{};
    b->P = P_;
    b->M = M_;
    b->row_ptr = d_row_ptr_;
    b->col_idx = d_col_idx_;
    b->vals = d_vals_;
    b->vals_f32 = values_f32_used_ ? 1 : 0;
    b->rb_start = d_rb_start_;
    b->rb_prob = d_rb_prob_;
    b->nrb = nrb;
    b->sb_prob = d_sb_prob_;
    b->sb_row0 = d_sb_row0_;
    b->nsb = nsb;
    b->p_sb0 = d_p_sb0_;
    b->p_rb0 = d_p_rb0_;
    b->np_col = d_np_col_;
    b->rb_list = d_rb_list_;
    b->sb_list = d_sb_list_;
    b->list_cnt = d_list_cnt_;
    b->V = d_V_;
    b->ld = ld;
    b->w = d_w_;
    b->y = d_y_;
    b->colscale = d_small_;
    b->h1 = b->colscale + ncols * P_;
    b->h2 = b->h1 + ncols * P_;
    b->H = b->h2 + ncols * P_;
    b->wH = b->H + ncols * ncols * P_;
    b->pscal = d_pscal_;
    b->pint = d_pint_;
    b->partials = d_partials_;
    b->expm_ws = d_expm_ws_;
    b->counters = d_counters_;
    b->any_expokit = any_expokit;
    bufs_ = b;
  } catch (...) {
    destroy();
    throw;
  }
}

BatchRunner::~BatchRunner() { destroy(); }

void BatchRunner::upload_v(const std::vector<const double*>& v) {
  if (static_cast<int>(v.size()) != P_) throw Error("ERROR: upload_v needs one vector per problem");
  for (int p = 0; p < P_; ++p)
    MEK_HIP(hipMemcpyAsync(d_v0_ + row_offset(p), v[p], sizeof(double) * rows(p), hipMemcpyHostToDevice,
                           stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
}

void BatchRunner::run_cycle_eager(int mcols) {
  batch_launch_cycle(*static_cast<BatchBufs*>(bufs_), plan_, mcols, gs_mode_, eta_, stream_);
  MEK_HIP_LAUNCHED();
}

hipGraphExec_t BatchRunner::cycle_graph(int mcols) {
  if (graph_exec_.empty()) {
    graph_exec_.assign(M_ + 1, nullptr);
    graph_.assign(M_ + 1, nullptr);
  }
  if (!graph_exec_[mcols]) {
    MEK_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    batch_launch_cycle(*static_cast<BatchBufs*>(bufs_), plan_, mcols, gs_mode_, eta_, stream_);
    MEK_HIP(hipStreamEndCapture(stream_, &graph_[mcols]));
    MEK_HIP(hipGraphInstantiate(&graph_exec_[mcols], graph_[mcols], nullptr, nullptr, 0));
  }
  return graph_exec_[mcols];
}

// Every problem back at its start: V[:, 0] = v, controller state from the
// problem descriptions, start norms and first steps on the device.
void BatchRunner::reset_state() {
  const BatchBufs& b = *static_cast<BatchBufs*>(bufs_);
  MEK_HIP(hipMemcpyAsync(d_V_, d_v0_, sizeof(double) * n_tot_, hipMemcpyDeviceToDevice, stream_));
  // y = v to start with: a problem with t = 0 never becomes active and keeps
  // exp(0 A) v = v; every active problem overwrites its rows in the lift
  MEK_HIP(hipMemcpyAsync(d_y_, d_v0_, sizeof(double) * n_tot_, hipMemcpyDeviceToDevice, stream_));
  MEK_HIP(hipMemcpyAsync(d_pint_, pint_init_.data(), sizeof(int) * pint_init_.size(), hipMemcpyHostToDevice, stream_));
  MEK_HIP(hipMemcpyAsync(d_pscal_, pscal_init_.data(), sizeof(double) * pscal_init_.size(),
                         hipMemcpyHostToDevice, stream_));
  batch_prepare();
  batch_launch_init(b, stream_);
  MEK_HIP_LAUNCHED();
}

double BatchRunner::time_spmv(int reps) {
  if (reps < 1) throw Error("ERROR: time_spmv needs reps >= 1");
  MEK_HIP(hipSetDevice(device_));
  reset_state();
  const BatchBufs& b = *static_cast<BatchBufs*>(bufs_);
  std::vector<double> us;
  time_candidates(
      {plan_.spmv_grid[0]}, [&](int g) { batch_launch_spmv(b, 0, plan_.nrb[0], g, stream_); }, stream_, reps,
      &us);
  return us[0];
}

void BatchRunner::solve_device() {
  MEK_HIP(hipSetDevice(device_));
  const double t0 = now_ms();
  const BatchBufs& b = *static_cast<BatchBufs*>(bufs_);
  reset_state();
  if (autotune_ && tuned_spmv_grid_ == 0) {
    // every problem is active with its start block: time the column-0 SpMV
    // over all row blocks for each grid and scale the per-column grids
    const std::vector<int> grids = {1024, 2048, 4096, 8192};
    std::vector<double> us;
    const int i = time_candidates(
        grids, [&](int g) { batch_launch_spmv(b, 0, plan_.nrb[0], g, stream_); }, stream_, 5, &us);
    tuned_spmv_grid_ = grids[i];
    for (int j = 0; j < M_; ++j)
      plan_.spmv_grid[j] = std::max(8, (std::min(plan_.nrb[j], tuned_spmv_grid_) + 7) / 8 * 8);
  }
  cycles_ = 0;
  int mcols = M_;
  const int max_cycles = 1000000;
  for (;;) {
    if (use_graph_) MEK_HIP(hipGraphLaunch(cycle_graph(mcols), stream_));
    else run_cycle_eager(mcols);
    MEK_HIP(hipMemcpyAsync(h_counters_, d_counters_, sizeof(int) * 3, hipMemcpyDeviceToHost, stream_));
    MEK_HIP(hipStreamSynchronize(stream_));
    ++cycles_;
    if (h_counters_[0] == 0) break;
    if (cycles_ >= max_cycles) throw Error("ERROR: batch did not finish within the cycle limit");
    mcols = std::min(std::max(h_counters_[2], 1), M_);
  }
  solve_ms_ = now_ms() - t0;
}

void BatchRunner::download_y(const std::vector<double*>& y) const {
  if (static_cast<int>(y.size()) != P_) throw Error("ERROR: download_y needs one vector per problem");
  for (int p = 0; p < P_; ++p)
    MEK_HIP(hipMemcpyAsync(y[p], d_y_ + row_offset(p), sizeof(double) * rows(p), hipMemcpyDeviceToHost,
                           stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
}

BatchProblemStats BatchRunner::stats(int p_ext) const {
  const int p = ext2int_[p_ext];
  int it[kPInt];
  double s[kPScal];
  MEK_HIP(hipMemcpy(it, d_pint_ + static_cast<size_t>(p) * kPInt, sizeof(it), hipMemcpyDeviceToHost));
  MEK_HIP(hipMemcpy(s, d_pscal_ + static_cast<size_t>(p) * kPScal, sizeof(s), hipMemcpyDeviceToHost));
  BatchProblemStats r{};
  r.converged = it[kBConverged];
  r.accepted = it[kBAccepted];
  r.attempts = it[kBAttempts];
  r.rejected = it[kBRejTotal];
  r.columns = it[kBColumns];
  r.aborted = it[kBAborted];
  r.reorth_columns = it[kBReorthCount];
  r.t_done = s[kBTDone];
  r.last_residual = s[kBLastRes];
  return r;
}

void BatchRunner::features(float* out) {
  MEK_HIP(hipSetDevice(device_));
  const BatchBufs& b = *static_cast<BatchBufs*>(bufs_);
  // kept for later calls (and freed by destroy()): 8 statistics per
  // streaming block, t, tol and ||A_p||_1, true row ends, and the features
  if (!d_feat_dbl_) MEK_HIP(hipMalloc(&d_feat_dbl_, sizeof(double) * (8 * static_cast<size_t>(b.nsb) + 3 * P_)));
  if (!d_feat_int_) MEK_HIP(hipMalloc(&d_feat_int_, sizeof(int) * P_));
  if (!d_feat_out_) MEK_HIP(hipMalloc(&d_feat_out_, sizeof(float) * kNumFeatures * P_));
  std::vector<double> tt(3 * static_cast<size_t>(P_));
  std::vector<int> row1(P_);
  for (int p = 0; p < P_; ++p) {
    tt[p] = pscal_init_[p * kPScal + kBT];
    tt[P_ + p] = pscal_init_[p * kPScal + kBTol];
    tt[2 * P_ + p] = norm1_[p];
    row1[p] = row0_[p] + nrows_[p];  // without the padding row
  }
  double* d_stats = d_feat_dbl_;
  double* d_tt = d_feat_dbl_ + 8 * static_cast<size_t>(b.nsb);
  MEK_HIP(hipMemcpyAsync(d_tt, tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice, stream_));
  MEK_HIP(hipMemcpyAsync(d_feat_int_, row1.data(), sizeof(int) * P_, hipMemcpyHostToDevice, stream_));
  // the streaming blocks are the feature slabs: kBatchRowsPerBlock rows from
  // each problem's first row, as a problem alone gets in features_csr
  static_assert(kBatchRowsPerBlock == kFeatSlabRows, "streaming blocks double as feature slabs");
  FeatureLayout L{b.sb_row0, b.sb_prob, b.p_sb0, d_row0_, d_feat_int_, b.nsb, P_};
  feature_launch(d_row_ptr_, d_col_idx_, d_vals_, b.vals_f32, L, d_tt + 2 * P_, d_stats, d_tt, d_tt + P_,
                 d_feat_out_, stream_);
  MEK_HIP_LAUNCHED();
  std::vector<float> f_int(static_cast<size_t>(kNumFeatures) * P_);
  MEK_HIP(hipMemcpyAsync(f_int.data(), d_feat_out_, sizeof(float) * f_int.size(), hipMemcpyDeviceToHost,
                         stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
  for (int p = 0; p < P_; ++p)
    std::memcpy(out + static_cast<size_t>(p) * kNumFeatures,
                f_int.data() + static_cast<size_t>(ext2int_[p]) * kNumFeatures, sizeof(float) * kNumFeatures);
}

// Releases whatever is allocated (also after a constructor that failed part
// way); safe to call more than once.
void BatchRunner::destroy() {
  (void)hipSetDevice(device_);
  for (hipGraphExec_t g : graph_exec_)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipGraph_t g : graph_)
    if (g) (void)hipGraphDestroy(g);
  graph_exec_.clear();
  graph_.clear();
  auto release = [](auto*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  release(d_row_ptr_);
  release(d_col_idx_);
  release(d_rb_start_);
  release(d_rb_prob_);
  release(d_sb_prob_);
  release(d_sb_row0_);
  release(d_p_sb0_);
  release(d_p_rb0_);
  release(d_np_col_);
  release(d_rb_list_);
  release(d_sb_list_);
  release(d_list_cnt_);
  release(d_pint_);
  release(d_counters_);
  release(d_vals_);
  release(d_V_);
  release(d_w_);
  release(d_y_);
  release(d_v0_);
  release(d_small_);
  release(d_partials_);
  release(d_expm_ws_);
  release(d_pscal_);
  release(d_row0_);
  release(d_feat_dbl_);
  release(d_feat_int_);
  release(d_feat_out_);
  if (h_counters_) (void)hipHostFree(h_counters_);
  h_counters_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  delete static_cast<BatchBufs*>(bufs_);
  bufs_ = nullptr;
}

}  // namespace mek
