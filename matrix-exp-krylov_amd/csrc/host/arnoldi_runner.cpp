// ArnoldiRunner: the adaptive time-stepping Arnoldi exp(t*A)*v solve on one
// MI355X.  reference: main.cu:1252-2081 (controller main.cu:1411-1571).
//
// Per step attempt the host enqueues all m columns (three sweeps and three
// finalisers each, see kernels/arnoldi.hip), the on-device Pade exp(dt*H)e1
// with the residual estimate, and the lift, then reads back one small status
// block (err, k_used, breakdown, ||y||) and runs the reference's accept /
// reject logic.  A rejected attempt keeps its Krylov basis: the basis depends
// only on the start vector, so recomputing it (as the reference does) would
// reproduce the same V and H; only exp(dt*H) and the lift are redone.
#include "mek/arnoldi_runner.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <exception>
#include <iostream>
#include <limits>
#include <thread>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../kernels/launch.hpp"
#include "hip_check.hpp"

namespace {

// roctx range with the reference's NVTX names (main.cu:38-41, :1413, :1437,
// :1575, :1770, :1955), visible in rocprofv3 --marker-trace.
struct Range {
  explicit Range(const char* name) { roctxRangePush(name); }
  ~Range() { roctxRangePop(); }
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Round up to two significant digits, as Expokit does for step sizes.
// A == A^T (structure and values), by a counting-sort transpose compared
// row by row against each row of A sorted by column.
bool csr_is_symmetric(const CSRHost& A) {
  if (A.rows != A.cols) return false;
  const int n = A.rows;
  const long long nnz = A.nnz;
  std::vector<long long> tp(static_cast<size_t>(n) + 1, 0);
  for (long long k = 0; k < nnz; ++k) tp[A.col_idx[k] + 1]++;
  for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
  std::vector<long long> pos(tp.begin(), tp.end() - 1);
  std::vector<int> tc(static_cast<size_t>(nnz));
  std::vector<double> tv(static_cast<size_t>(nnz));
  for (int i = 0; i < n; ++i)
    for (int k = A.row_ptr[i]; k < A.row_ptr[i + 1]; ++k) {
      const long long p = pos[A.col_idx[k]]++;
      tc[p] = i;
      tv[p] = A.values[k];
    }
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < n; ++i) {
    const int k0 = A.row_ptr[i], k1 = A.row_ptr[i + 1];
    if (tp[i + 1] - tp[i] != k1 - k0) return false;
    row.clear();
    for (int k = k0; k < k1; ++k) row.emplace_back(A.col_idx[k], A.values[k]);
    std::sort(row.begin(), row.end());
    for (int q = 0; q < k1 - k0; ++q)
      if (row[q].first != tc[tp[i] + q] || row[q].second != tv[tp[i] + q]) return false;
  }
  return true;
}

// s Chebyshev points of the Gershgorin interval of A (the real parts of its
// eigenvalues lie there) in Leja order: the largest in magnitude first, then
// each maximising the product of distances to those already taken, so every
// prefix is a good shift set for a shorter block.
std::vector<double> sstep_shifts(const CSRHost& A, int s) {
  double lo = HUGE_VAL, hi = -HUGE_VAL;
  for (int i = 0; i < A.rows; ++i) {
    double d = 0.0, r = 0.0;
    for (int k = A.row_ptr[i]; k < A.row_ptr[i + 1]; ++k) {
      if (A.col_idx[k] == i) d += A.values[k];
      else r += std::abs(A.values[k]);
    }
    lo = std::min(lo, d - r);
    hi = std::max(hi, d + r);
  }
  if (!(lo <= hi)) lo = hi = 0.0;
  std::vector<double> pts(s), out;
  for (int i = 0; i < s; ++i) pts[i] = 0.5 * (lo + hi) + 0.5 * (hi - lo) * std::cos(M_PI * (2 * i + 1) / (2.0 * s));
  std::vector<bool> used(s, false);
  for (int n = 0; n < s; ++n) {
    int best = -1;
    double bv = -1.0;
    for (int i = 0; i < s; ++i) {
      if (used[i]) continue;
      double v = 1.0;
      if (out.empty()) v = std::abs(pts[i]);
      else
        for (double x : out) v *= std::abs(pts[i] - x);
      if (v > bv) {
        bv = v;
        best = i;
      }
    }
    used[best] = true;
    out.push_back(pts[best]);
  }
  return out;
}

// Bytes one SpMV streams: the matrix in its layout (CSR: values, column
// indices, row pointers; sliced ELL: padded values and column indices, slice
// offsets; diagonal: n values per diagonal and the offsets; constant
// diagonals: one presence mask per row, the offsets and values), x once and
// w.  csr = true: the CSR figure whatever the layout.
double spmv_bytes(const DeviceContext& dc, bool csr = false) {
  const double vb = dc.values_f32 ? 4.0 : 8.0;
  const double n = dc.local_rows;
  if (dc.n_diags > 0 && !csr && dc.d_dia_mask)
    return (dc.n_diags <= 8 ? 1.0 : 2.0) * n + 12.0 * dc.n_diags + 16.0 * n;
  if (dc.n_diags > 0 && !csr)
    return static_cast<double>(dc.n_diags) * n * vb + 4.0 * dc.n_diags + 16.0 * n;
  if (dc.n_slices > 0 && !csr)
    return static_cast<double>(dc.sell_entries) * (vb + 4.0) + 8.0 * (dc.n_slices + 1.0) + 16.0 * n;
  return static_cast<double>(dc.local_nnz) * (vb + 4.0) + 4.0 * (n + 1.0) + 16.0 * n;
}

double round_step(double x) {
  const double s = std::pow(10.0, std::floor(std::log10(x)) - 1.0);
  return std::ceil(x / s) * s;
}

}  // namespace

ArnoldiRunner::ArnoldiRunner(int n_gpus, const ArnoldiParams& p, int total_rows, int total_cols,
                             long long total_nnz)
    : num_gpus(n_gpus), params(p), global_rows(total_rows), global_cols(total_cols),
      global_nnz(total_nnz) {
  device_contexts.reserve(num_gpus);
  for (int i = 0; i < num_gpus; ++i) device_contexts.emplace_back(i);
  // reference: main.cu:1292-1305
  ts_params.dt_max = params.t;
  ts_params.dt_min = std::max(1e-12, params.t * 1e-6);
  const double denom = (params.max_restarts > 0) ? params.max_restarts : 1;
  ts_params.dt_init = params.t / denom;
  ts_params.tol_step = params.tol;
  ts_params.max_steps = (params.max_restarts > 0) ? params.max_restarts : 1000;
  ts_params.max_rejects = 20;
  ts_params.safety = 0.8;
  ts_params.grow = 1.5;
  ts_params.shrink = 0.5;
  ts_params.adapt_m = false;
  ts_params.m_max = params.m;
  ts_params.m_step = 0;
}

void ArnoldiRunner::init_all_devices(const std::vector<CSRHost>& partitions,
                                     const std::vector<int>& device_ids,
                                     const std::vector<CSRHost::GhostMap>& ghost_maps_in) {
  ghost_maps = ghost_maps_in;
  if (static_cast<int>(device_ids.size()) != num_gpus)
    throw mek::Error("ERROR: device_ids size (" + std::to_string(device_ids.size()) +
                     ") must match num_gpus (" + std::to_string(num_gpus) + ")");
  if (num_gpus != 1)
    throw mek::Error("ERROR: the solve runs on one GPU (num_gpus=" +
                     std::to_string(num_gpus) + ")");
  init_one(partitions[0], device_ids[0]);
}

void ArnoldiRunner::init_device(const CSRHost& A, int device_id) {
  if (num_gpus != 1)
    throw mek::Error("ERROR: the solve runs on one GPU (num_gpus=" + std::to_string(num_gpus) + ")");
  if (A.rows != global_rows || A.cols != global_cols || A.nnz != global_nnz)
    throw mek::Error("ERROR: init_device: the matrix does not have the runner's rows, cols and nnz");
  ghost_maps.assign(1, CSRHost::GhostMap{});
  init_one(A, device_id);
}

void ArnoldiRunner::init_one(const CSRHost& part, int device_id) {
  if (options.lanczos && !csr_is_symmetric(part))
    throw mek::Error("ERROR: the Lanczos path needs a symmetric matrix");
  if (params.m >= global_rows)
    throw mek::Error("ERROR: Requested Arnoldi m=" + std::to_string(params.m) +
                     " must be less than global_rows=" + std::to_string(global_rows));
  if (global_rows != global_cols)
    throw mek::Error("ERROR: Matrix must be square. rows=" + std::to_string(global_rows) +
                     " cols=" + std::to_string(global_cols));
  if (params.m < 1 || params.m > mek::kMaxKrylov)
    throw mek::Error("ERROR: Arnoldi m=" + std::to_string(params.m) + " outside [1, " +
                     std::to_string(mek::kMaxKrylov) + "]");
  DeviceContext& dc = device_contexts[0];
  dc.device_id = device_id;
  dc.init_handles();
  dc.create_streams();
  dc.log = options.verbose;
  const int alloc_m = std::max(params.m, ts_params.m_max > 0 ? ts_params.m_max : params.m);
  if (options.values_f32 < 0 || options.values_f32 > 2)
    throw mek::Error("ERROR: values_f32 must be 0, 1 or 2");
  values_f32_used = options.values_f32 == 2 ||
                    (options.values_f32 == 1 && part.f32_inexact(nullptr) == 0);
  dc.alloc_from(part, global_cols, alloc_m, values_f32_used);
  static const CSRHost::GhostMap kEmpty;
  dc.create_descriptors(part, global_cols, ghost_maps.empty() ? kEmpty : ghost_maps[0], 0, alloc_m);
  dc.build_spmv_layout(part, options.spmv_format);
  if (options.autotune) tuning = mek::autotune_single(dc, part, alloc_m);
  anorm_inf = part.norm_inf();
  if (options.sstep < 0 || options.sstep == 1 || options.sstep > mek::kSsMaxS)
    throw mek::Error("ERROR: sstep must be 0 or in [2, " + std::to_string(mek::kSsMaxS) + "]");
  // the block path covers bases of up to 64 vectors; beyond that (or with
  // the Lanczos recurrence) the columns are built one by one
  ss_active_ = options.sstep >= 2 && !options.lanczos && !options.fuse_spmv_dot &&
               params.m + 1 <= mek::kSsMaxCols;
  if (ss_active_) {
    ss_shifts_ = sstep_shifts(part, options.sstep);
    dc.alloc_sstep(ss_shifts_);
    if (options.autotune) {
      // block sweep grid: the dots and the update of the largest block of a
      // cycle, timed on the basis buffer (its contents are scratch here)
      const int ss = std::min(options.sstep, params.m);
      const int k = params.m + 1 - ss;
      std::vector<int> grids;
      const int ntiles = static_cast<int>((dc.local_rows + 127) / 128);
      for (int g : {256, 512, 1024, mek::kSsMaxGrid})
        if (g <= std::max(8, ntiles)) grids.push_back(g);
      if (grids.empty()) grids.push_back(dc.ss_grid);
      mek::SsArgs a = dc.ss_args();
      std::vector<double> us;
      const int i = mek::time_candidates(
          grids,
          [&](int g) {
            a.grid = g;
            mek::launch_ss_dots(a, k, ss, dc.stream_compute);
            mek::launch_ss_update(a, 0, k + ss, k, ss, -1, dc.stream_compute);
          },
          dc.stream_compute, 5, &us);
      dc.ss_grid = grids[i];
      tuning.sstep_grid = grids[i];
      tuning.sstep_us = us[i];
      // the first block of a cycle, when its window fits two 16-column tiles:
      // its sweeps move fewer bytes per 128-row tile, and more workgroups
      // (the kernels hold fewer registers at that width) keep more tiles in
      // flight (twice the large-window grid: 2 % faster solves of the 4096^2
      // Laplacian)
      const int s0 = options.sstep_first >= 0 ? std::min(options.sstep_first, params.m) : ss;
      const int k1 = s0 + 1, ss1 = std::min(ss, params.m + 1 - k1);
      if (k1 + ss1 <= 32 && ss1 >= 2) {
        std::vector<double> us1;
        const int i1 = mek::time_candidates(
            grids,
            [&](int g) {
              a.grid = g;
              mek::launch_ss_dots(a, k1, ss1, dc.stream_compute);
              mek::launch_ss_update(a, 0, k1 + ss1, k1, ss1, -1, dc.stream_compute);
            },
            dc.stream_compute, 5, &us1);
        dc.ss_grid_small = grids[i1];
        tuning.sstep_grid_small = grids[i1];
      }
      MEK_HIP(hipMemsetAsync(dc.d_V_m, 0, sizeof(double) * static_cast<size_t>(dc.ld) * (dc.allocated_m_max + 2),
                             dc.stream_compute));
      MEK_HIP(hipStreamSynchronize(dc.stream_compute));
    }
  }
}

// ------------------------------------------------------------ timing

void ArnoldiRunner::begin_timed(int kind) {
  if (!options.timing) return;
  if (ev_used_ == ev_pool_.size()) {
    EventPair p{};
    MEK_HIP(hipEventCreate(&p.a));
    MEK_HIP(hipEventCreate(&p.b));
    ev_pool_.push_back(p);
  }
  EventPair& p = ev_pool_[ev_used_];
  p.kind = kind;
  MEK_HIP(hipEventRecord(p.a, device_contexts[0].stream_compute));
}

void ArnoldiRunner::end_timed(int kind, double bytes) {
  if (!options.timing) return;
  EventPair& p = ev_pool_[ev_used_];
  p.kind = kind;
  p.bytes = bytes;
  MEK_HIP(hipEventRecord(p.b, device_contexts[0].stream_compute));
  ++ev_used_;
}

void ArnoldiRunner::flush_timing() {
  if (!options.timing || ev_used_ == 0) return;
  MEK_HIP(hipStreamSynchronize(device_contexts[0].stream_compute));
  for (size_t i = 0; i < ev_used_; ++i) {
    float ms = 0.f;
    MEK_HIP(hipEventElapsedTime(&ms, ev_pool_[i].a, ev_pool_[i].b));
    mek::KernelTiming& kt = stats.kernels[ev_pool_[i].kind];
    kt.ms += ms;
    kt.launches += 1;
    kt.bytes += ev_pool_[i].bytes;
  }
  ev_used_ = 0;
}

// ------------------------------------------------------------ device steps

int ArnoldiRunner::read_status() {
  DeviceContext& dc = device_contexts[0];
  const size_t bytes = sizeof(double) * (20 + dc.allocated_m_max + 4);
  MEK_HIP(hipMemcpyAsync(dc.h_status, dc.d_small, bytes, hipMemcpyDeviceToHost, dc.stream_compute));
  MEK_HIP(hipStreamSynchronize(dc.stream_compute));
  flush_timing();
  return reinterpret_cast<const int*>(dc.h_status + 16)[mek::kKUsed];
}

void ArnoldiRunner::init_q1_device(const double* d_v, const double* h_v) {
  Range r("init_q1");
  DeviceContext& dc = device_contexts[0];
  const mek::ArnoldiBufs b = dc.bufs();
  hipStream_t s = dc.stream_compute;
  const size_t bytes = sizeof(double) * dc.local_rows;
  if (d_v == dc.d_V_m) {
    // v is already in V[:, 0] (the pipelined host staging puts it there)
  } else if (d_v) {
    MEK_HIP(hipMemcpyAsync(dc.d_V_m, d_v, bytes, hipMemcpyDeviceToDevice, s));
  } else {
    MEK_HIP(hipMemcpyAsync(dc.d_V_m, h_v, bytes, hipMemcpyHostToDevice, s));
  }
  mek::launch_norm2(b, dc.d_V_m, s);
  mek::launch_fin_start(b, s);
  MEK_HIP_LAUNCHED();
}

void ArnoldiRunner::run_columns(int m_cols) {
  if (options.lanczos) {
    run_lanczos_columns(m_cols);
    return;
  }
  const int s0 = options.sstep_first >= 0 ? std::min(options.sstep_first, m_cols) : std::min(options.sstep, m_cols);
  if (!ss_active_ || s0 >= m_cols) {
    run_columns_range(0, m_cols);
    return;
  }
  run_columns_range(0, s0);
  for (int k = s0 + 1; k <= m_cols;) {
    const int ss = std::min(options.sstep, m_cols + 1 - k);
    run_sstep_block(k, ss);
    k += ss;
  }
  DeviceContext& dc = device_contexts[0];
  read_status();
  if (reinterpret_cast<const int*>(dc.h_status + 16)[mek::kSStepFail]) {
    // a block could not be completed: the blocks before it stand (orthonormal
    // columns, their H columns, k_used); clear the flag and build the
    // remaining columns one by one from q_{k-1} of the failed block
    int fk = 0;
    MEK_HIP(hipMemcpy(&fk, dc.d_ss_flags + mek::kSsFailK, sizeof(int), hipMemcpyDeviceToHost));
    const int j0 = (fk >= s0 + 1 && fk <= m_cols) ? fk - 1 : s0;
    const int zero = 0;
    MEK_HIP(hipMemcpyAsync(reinterpret_cast<int*>(dc.d_small + 16) + mek::kSStepFail, &zero, sizeof(int),
                           hipMemcpyHostToDevice, dc.stream_compute));
    MEK_HIP(hipStreamSynchronize(dc.stream_compute));
    stats.spmv_count -= m_cols - j0;
    stats.sstep_fallbacks++;
    run_columns_range(j0, m_cols);
  }
}

// One s-step block: columns k .. k+ss-1 (launch sequence in launch.hpp).
void ArnoldiRunner::run_sstep_block(int k, int ss) {
  Range r("orthogonalize_sstep");
  DeviceContext& dc = device_contexts[0];
  const mek::ArnoldiBufs b = dc.bufs();
  mek::SsArgs a = dc.ss_args();
  // windows of up to 32 columns (two 16-column tiles) run their own grid
  if (k + ss <= 32 && dc.ss_grid_small > 0) a.grid = dc.ss_grid_small;
  hipStream_t s = dc.stream_compute;
  const double rows = static_cast<double>(a.rows);
  const double bspmv = spmv_bytes(dc);
  const int gs_mode = (options.gs_mode >= 0) ? options.gs_mode : (params.reorthogonalize ? 1 : 0);
  const int force2 = gs_mode == 1 ? 1 : 0;
  for (int i = 0; i < ss; ++i) {
    begin_timed(mek::kKSpmv);
    mek::launch_ss_spmv(b, a, k, i, s);
    end_timed(mek::kKSpmv, bspmv);
    mek::launch_ss_sigma(b, a, k, i, s);
  }
  begin_timed(mek::kKSsDots);
  mek::launch_ss_dots(a, k, ss, s);
  end_timed(mek::kKSsDots, 8.0 * rows * (k + ss));
  mek::launch_ss_reduce(a, k + ss, ss, -1, s);
  mek::launch_ss_plan(a, mek::kSsPhaseA, 0, k, ss, force2, -1, s);
  begin_timed(mek::kKSsUpdate);
  mek::launch_ss_update(a, 0, k + ss, k, ss, -1, s);
  end_timed(mek::kKSsUpdate, 8.0 * rows * (k + 2 * ss));
  mek::launch_ss_reduce(a, k + ss, ss, -1, s);
  mek::launch_ss_plan(a, mek::kSsPhaseB, 0, k, ss, force2, -1, s);
  // second update, only when the side dots asked for it (no bytes counted)
  begin_timed(mek::kKUpdateNorm);
  mek::launch_ss_update(a, 0, k + ss, k, ss, mek::kSsPass2, s);
  end_timed(mek::kKUpdateNorm, 0.0);
  mek::launch_ss_reduce(a, k + ss, ss, mek::kSsPass2, s);
  mek::launch_ss_plan(a, mek::kSsPhaseB, 1, k, ss, force2, mek::kSsPass2, s);
  // normalising sweep over the block's own columns, only when the fused
  // update did not already write Q_new (no bytes counted: it may return at once)
  begin_timed(mek::kKUpdateNorm);
  mek::launch_ss_update(a, k, ss, k, ss, mek::kSsNormReady, s);
  end_timed(mek::kKUpdateNorm, 0.0);
  mek::launch_ss_reduce(a, ss, ss, mek::kSsNormReady, s);
  mek::launch_ss_plan(a, mek::kSsPhaseD, 0, k, ss, force2, mek::kSsNormReady, s);
  begin_timed(mek::kKUpdateNorm);
  mek::launch_ss_update(a, k, ss, k, ss, mek::kSsNorm2, s);
  end_timed(mek::kKUpdateNorm, 0.0);
  mek::launch_ss_reduce(a, ss, ss, mek::kSsNorm2, s);
  mek::launch_ss_plan(a, mek::kSsPhaseD, 1, k, ss, force2, mek::kSsNorm2, s);
  MEK_HIP_LAUNCHED();
  stats.spmv_count += ss;
}

void ArnoldiRunner::run_columns_range(int j0, int j1) {
  Range r("orthogonalize_mgs");
  DeviceContext& dc = device_contexts[0];
  const mek::ArnoldiBufs b = dc.bufs();
  hipStream_t s = dc.stream_compute;
  const double n = dc.local_rows;
  const double bspmv = spmv_bytes(dc);
  const double bspmv_csr = spmv_bytes(dc, true);  // the fused kernel reads the CSR row blocks
  const int gs_mode = (options.gs_mode >= 0) ? options.gs_mode : (params.reorthogonalize ? 1 : 0);
  // quantity slot for the SpMV's ||w||^2 partial when the dots kernel
  // computes the norm itself (partials then share one block stride)
  const int spare_q = dc.allocated_m_max + 3;
  for (int j = j0; j < j1; ++j) {
    const int nq = j + 1;
    if (nq <= mek::kMaxFusedDots && options.fuse_spmv_dot) {
      begin_timed(mek::kKSpmv);
      mek::launch_spmv_dots(b, j, s);
      end_timed(mek::kKSpmv, bspmv_csr + 8.0 * n * nq);
      mek::launch_fin_h1(b, nq, b.spmv_grid, s);
    } else if (mek::spmv_dia_fuses_dots(b, nq)) {
      // stencils: the dots against q_0 .. q_j formed while w is in registers
      // (q_j is the SpMV's x, already counted: j more columns read)
      begin_timed(mek::kKSpmv);
      mek::launch_spmv_dia_col_dots(b, j, s);
      end_timed(mek::kKSpmv, bspmv + 8.0 * n * j);
      mek::launch_fin_h1(b, nq, b.spmv_grid, s);
    } else {
      begin_timed(mek::kKSpmv);
      mek::launch_spmv_col(b, j, spare_q, s);
      end_timed(mek::kKSpmv, bspmv);
      for (int c0 = 0; c0 < nq; c0 += mek::kMaxFusedDots) {
        const int cnt = std::min(mek::kMaxFusedDots, nq - c0);
        begin_timed(mek::kKDots);
        mek::launch_dots(b, dc.d_w, c0, cnt, (c0 + cnt == nq) ? 1 : 0, c0, s);
        end_timed(mek::kKDots, 8.0 * n * (cnt + 1));
      }
      mek::launch_fin_h1(b, nq, b.grid, s);
    }
    if (nq <= mek::kMaxFusedDots) {
      begin_timed(mek::kKUpdateDot);
      mek::launch_update_dot(b, j, s);
      end_timed(mek::kKUpdateDot, 8.0 * n * (nq + 2));
    } else {
      // generic path for j+1 > 64: update, then dots in chunks of 64 columns
      double* vn = dc.d_V_m + static_cast<long long>(j + 1) * dc.ld;
      mek::launch_update(b, dc.d_w, vn, b.h1, nq, s);
      for (int c0 = 0; c0 < nq; c0 += mek::kMaxFusedDots) {
        const int cnt = std::min(mek::kMaxFusedDots, nq - c0);
        mek::launch_dots(b, vn, c0, cnt, (c0 + cnt == nq) ? 1 : 0, c0, s);
      }
    }
    mek::launch_fin_h2(b, j, s);
    begin_timed(mek::kKUpdateNorm);
    mek::launch_update_norm(b, j, gs_mode, options.dgks_eta, s);
    // a DGKS column that needs no second pass exits at once: count no bytes
    end_timed(mek::kKUpdateNorm, gs_mode == 2 ? 0.0 : 8.0 * n * (nq + 2));
    mek::launch_fin_col(b, j, gs_mode, options.dgks_eta, s);
    MEK_HIP_LAUNCHED();
  }
  stats.spmv_count += j1 - j0;  // launched; columns after a breakdown exit at once
}

// Symmetric Lanczos: per column one SpMV (with the fused alpha partial), its
// finaliser, one three-term update sweep and its finaliser.
void ArnoldiRunner::run_lanczos_columns(int m_cols) {
  Range r("lanczos");
  DeviceContext& dc = device_contexts[0];
  const mek::ArnoldiBufs b = dc.bufs();
  hipStream_t s = dc.stream_compute;
  const double n = dc.local_rows;
  const double bspmv = spmv_bytes(dc);
  for (int j = 0; j < m_cols; ++j) {
    begin_timed(mek::kKSpmv);
    mek::launch_spmv_lanczos(b, j, s);
    end_timed(mek::kKSpmv, bspmv + 8.0 * n);  // + q_j for the fused dot
    mek::launch_fin_h1(b, 1, b.spmv_grid, s);
    begin_timed(mek::kKUpdateDot);
    mek::launch_lanczos_update(b, j, s);
    end_timed(mek::kKUpdateDot, 8.0 * n * (j > 0 ? 4 : 3));
    mek::launch_lanczos_fin(b, j, s);
    MEK_HIP_LAUNCHED();
  }
  stats.spmv_count += m_cols;
}

void ArnoldiRunner::launch_expm(double dt, int mode) {
  Range r("small_expm_and_lift");
  DeviceContext& dc = device_contexts[0];
  const mek::ArnoldiBufs b = dc.bufs();
  mek::ExpmArgs a{};
  a.H = b.H;
  a.ldh = b.ldh;
  a.iscal = b.iscal;
  a.iscal_out = b.iscal;
  a.scal = b.scal;
  a.out = b.wH;
  a.out_len = dc.allocated_m_max + 4;
  a.dt = dt;
  a.mode = mode;
  a.ws = dc.d_expm_ws;
  a.k_override = 0;
  a.breakdown_override = 0;
  a.full_out = nullptr;
  begin_timed(mek::kKExpm);
  mek::launch_expm(a, dc.allocated_m_max + 2, dc.stream_compute);
  end_timed(mek::kKExpm, 0.0);
  MEK_HIP_LAUNCHED();
}

void ArnoldiRunner::lift(int ncoef, double skip_below) {
  DeviceContext& dc = device_contexts[0];
  const mek::ArnoldiBufs b = dc.bufs();
  begin_timed(mek::kKLift);
  mek::launch_lift(b, ncoef, skip_below, dc.stream_compute);
  end_timed(mek::kKLift, 8.0 * dc.local_rows * (ncoef + 1));
  mek::launch_fin_norm(b, mek::kYNorm, b.grid, dc.stream_compute);
  MEK_HIP_LAUNCHED();
}

void ArnoldiRunner::restart_from_y() {
  DeviceContext& dc = device_contexts[0];
  mek::launch_restart(dc.bufs(), dc.stream_compute);
  MEK_HIP_LAUNCHED();
}

// ------------------------------------------------------------ controllers

// reference: compute_expmv's adaptive time stepping, main.cu:1425-1518.
void ArnoldiRunner::reference_controller() {
  DeviceContext& dc = device_contexts[0];
  const double total_t = params.t;
  double dt = std::min(std::max(ts_params.dt_init, ts_params.dt_min), ts_params.dt_max);
  double t_done = 0.0;
  int accepted = 0;
  converged = false;
  bool basis_ready = false;
  while (t_done < total_t && accepted < ts_params.max_steps) {
    dt = std::min(dt, total_t - t_done);
    int rejects = 0;
    int m_current = params.m;
    for (;;) {
      Range r_step("arnoldi_step_attempt");
      if (m_current > dc.allocated_m_max)
        throw mek::Error("ERROR: m_current exceeds allocated_m_max on GPU 0");
      if (!basis_ready) {
        run_columns(m_current);
        basis_ready = true;
      }
      current_segment_t = dt;
      launch_expm(dt, 0);
      lift(m_current + 1, 1e-15);
      const int k_used = read_status();
      const double* sc = dc.h_status;
      const int breakdown = reinterpret_cast<const int*>(sc + 16)[mek::kBreakdown];
      const double err = sc[mek::kErr];  // 0 on breakdown (set on the device)
      if (options.verbose) {
        for (int j = 0; j < k_used; ++j)
          std::cout << "    H_m column " << j << " stored. Current H_m size: " << (m_current + 1)
                    << "x" << m_current << "\n";
        std::cout << "[STEP] t=" << t_done << " dt=" << dt << " err=" << err << " m_used=" << k_used
                  << " accepts=" << accepted << " rejects=" << rejects << std::endl;
      }
      stats.steps.push_back({t_done, dt, err, k_used, accepted, rejects});
      stats.attempts++;
      if (err <= ts_params.tol_step || breakdown) {
        t_done += dt;
        accepted++;
        last_residual = err;
        if (t_done >= total_t) {
          converged = true;
          break;
        }
        restart_from_y();
        basis_ready = false;
        if (err < 0.1 * ts_params.tol_step) dt = std::min(dt * ts_params.grow * ts_params.safety, ts_params.dt_max);
        break;
      }
      rejects++;
      stats.rejected++;
      if (rejects >= ts_params.max_rejects) {
        std::cerr << "[STEP] rejected too many times, aborting" << std::endl;
        converged = false;
        last_residual = err;
        stats.aborted = true;
        t_done = total_t;
        break;
      }
      if (ts_params.adapt_m && (m_current + ts_params.m_step) <= ts_params.m_max && ts_params.m_step > 0) {
        m_current += ts_params.m_step;
        basis_ready = false;
      } else {
        dt = std::max(dt * ts_params.shrink * ts_params.safety, ts_params.dt_min);
        m_current = params.m;
      }
    }
  }
  restarts_done = accepted;
  stats.accepted = accepted;
  stats.t_done = t_done;
}

// Sidje 1998 (Expokit, ACM TOMS 24(1):130-156) step-size control: local error
// estimate from the augmented (m+2) x (m+2) exponential, err_loc <=
// delta * t_step * tol per step, step sizes rounded to two digits.  The
// Krylov basis is built once per step; a rejection only recomputes the small
// exponential with a shorter step.
void ArnoldiRunner::expokit_controller() {
  DeviceContext& dc = device_contexts[0];
  const mek::ArnoldiBufs b = dc.bufs();
  const int m = params.m;
  const double tol = params.tol;
  const double t_out = std::abs(params.t);
  const double anorm = (anorm_inf > 0.0) ? anorm_inf : 1.0;
  const int mxrej = 10;
  const double btol = 1.0e-7, gamma = 0.9, delta = 1.2;
  const double rndoff = anorm * std::numeric_limits<double>::epsilon();
  read_status();
  double beta = dc.h_status[mek::kVNorm];
  const double fact = std::pow((m + 1) / std::exp(1.0), m + 1) * std::sqrt(2.0 * M_PI * (m + 1));
  const double tol0 = (options.control == mek::StepControl::ExpokitRelative) ? tol * beta : tol;
  double t_new = (1.0 / anorm) * std::pow((fact * tol0) / (4.0 * beta * anorm), 1.0 / m);
  t_new = round_step(t_new);
  double t_now = 0.0, s_error = 0.0, err_loc = 0.0;
  int nstep = 0;
  converged = false;
  const int max_steps = (params.max_restarts > 0) ? std::max(params.max_restarts, 100000) : 100000;
  while (t_now < t_out && nstep < max_steps) {
    ++nstep;
    double t_step = std::min(t_out - t_now, t_new);
    run_columns(m);
    int k_used = read_status();
    const int breakdown = reinterpret_cast<const int*>(dc.h_status + 16)[mek::kBreakdown];
    int k1 = 2, mb = m;
    double avnorm = 0.0;
    if (breakdown) {
      k1 = 0;
      mb = k_used;
      t_step = t_out - t_now;
    } else {
      // ||A v_{m+1}||
      begin_timed(mek::kKSpmv);
      mek::launch_spmv_plain(b, dc.d_V_m + static_cast<long long>(m) * dc.ld, b.colscale + m, dc.d_w,
                             dc.stream_compute);
      end_timed(mek::kKSpmv, spmv_bytes(dc));
      mek::launch_fin_norm(b, mek::kAvNorm, b.spmv_grid, dc.stream_compute);
      MEK_HIP_LAUNCHED();
      stats.extra_spmv++;
    }
    int ireject = 0;
    double xm = 1.0 / m;
    const bool relative = options.control == mek::StepControl::ExpokitRelative;
    double tol_eff = tol;
    for (;;) {
      launch_expm(t_step, 1);
      read_status();
      const double* F = dc.h_status + 20;  // first column of exp(t_step * H_aug)
      if (relative) {
        // local error against the smaller of ||w|| at the step's start and
        // its end, ||w_end|| = beta ||F(0:mb)|| (orthonormal basis)
        double f2 = 0.0;
        for (int i = 0; i < mb + std::max(0, k1 - 1); ++i) f2 += F[i] * F[i];
        tol_eff = tol * beta * std::min(1.0, std::sqrt(f2));
      }
      if (!breakdown) avnorm = dc.h_status[mek::kAvNorm];
      stats.attempts++;
      if (k1 == 0) {
        err_loc = btol;
        break;
      }
      const double phi1 = std::abs(beta * F[m]);
      const double phi2 = std::abs(beta * F[m + 1] * avnorm);
      if (phi1 > 10.0 * phi2) {
        err_loc = phi2;
        xm = 1.0 / m;
      } else if (phi1 > phi2) {
        err_loc = (phi1 * phi2) / (phi1 - phi2);
        xm = 1.0 / m;
      } else {
        err_loc = phi1;
        xm = 1.0 / (m - 1);
      }
      if (err_loc <= delta * t_step * tol_eff) break;
      if (ireject == mxrej) throw mek::Error("ERROR: The requested tolerance is too high.");
      t_step = round_step(gamma * t_step * std::pow(t_step * tol_eff / err_loc, xm));
      ++ireject;
      stats.rejected++;
    }
    const int mx = mb + std::max(0, k1 - 1);
    lift(mx, -1.0);
    restart_from_y();
    read_status();
    beta = dc.h_status[mek::kVNorm];
    if (options.verbose) {
      std::cout << "[STEP] t=" << t_now << " dt=" << t_step << " err=" << err_loc << " m_used=" << mb
                << " accepts=" << (nstep - 1) << " rejects=" << ireject << std::endl;
    }
    stats.steps.push_back({t_now, t_step, err_loc, mb, nstep - 1, ireject});
    t_now += t_step;
    t_new = round_step(gamma * t_step * std::pow(t_step * tol_eff / err_loc, xm));
    err_loc = std::max(err_loc, rndoff);
    s_error += err_loc;
  }
  converged = (t_now >= t_out);
  stats.err_sum = s_error;
  restarts_done = nstep;
  last_residual = err_loc;
  stats.accepted = nstep;
  stats.t_done = t_now;
}

// ------------------------------------------------------------ entry points

void ArnoldiRunner::solve_on_device(const double* d_v, const double* h_v) {
  Range r("total_compute_expmv");
  DeviceContext& dc = device_contexts[0];
  MEK_HIP(hipSetDevice(dc.device_id));
  stats = mek::SolveStats();
  ev_used_ = 0;
  if (mode != ExpmvMode::AdaptiveTimeStepping) {
    std::cerr << "Expmv mode not implemented" << std::endl;
    return;
  }
  // the reference's controller never runs for t <= 0 and returns whatever
  // d_y held (main.cu:1431, :1521-1529); here t < 0 is refused and
  // exp(0 A) v = v is returned exactly
  if (!(params.t >= 0.0) || !std::isfinite(params.t))
    throw mek::Error("ERROR: t must be finite and >= 0 (t=" + std::to_string(params.t) + ")");
  if (params.t == 0.0) {
    const size_t bytes = sizeof(double) * dc.local_rows;
    if (d_v) MEK_HIP(hipMemcpyAsync(dc.d_y, d_v, bytes, hipMemcpyDeviceToDevice, dc.stream_compute));
    else MEK_HIP(hipMemcpyAsync(dc.d_y, h_v, bytes, hipMemcpyHostToDevice, dc.stream_compute));
    MEK_HIP(hipStreamSynchronize(dc.stream_compute));
    converged = true;
    restarts_done = 0;
    last_residual = 0.0;
    return;
  }
  init_q1_device(d_v, h_v);
  if (options.control != mek::StepControl::Reference) expokit_controller();
  else reference_controller();
  MEK_HIP(hipStreamSynchronize(dc.stream_compute));
  flush_timing();
  stats.reorth_columns = reinterpret_cast<const int*>(dc.h_status + 16)[mek::kReorthCount];
  v_norm = dc.h_status[mek::kVNorm];
  if (ss_active_) {
    // block counters are kept across solves: report this solve's share
    int fl[mek::kSsNumFlags];
    MEK_HIP(hipMemcpy(fl, dc.d_ss_flags, sizeof(fl), hipMemcpyDeviceToHost));
    stats.sstep_blocks = fl[mek::kSsBlocks] - ss_blocks_before_;
    stats.sstep_second = fl[mek::kSsSecond] - ss_second_before_;
    stats.sstep_fused = fl[mek::kSsFusedBlocks] - ss_fused_before_;
    ss_blocks_before_ = fl[mek::kSsBlocks];
    ss_second_before_ = fl[mek::kSsSecond];
    ss_fused_before_ = fl[mek::kSsFusedBlocks];
  }
}

// Host copies to and from the pinned staging buffer: one thread per 16 MB
// (at most 8), so the copy runs near the host's memory bandwidth instead of
// one core's.
void ArnoldiRunner::stage_copy(double* dst, const double* src, size_t n) {
  const size_t per = size_t(2) << 20;  // doubles per thread at least
  const size_t nt = std::min<size_t>(8, std::max<size_t>(1, n / per));
  if (nt == 1) {
    std::copy(src, src + n, dst);
    return;
  }
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t) {
    const size_t a = n * t / nt, b = n * (t + 1) / nt;
    th.emplace_back([=] { std::copy(src + a, src + b, dst + a); });
  }
  for (auto& x : th) x.join();
}

// Copies src to dst in nchunk chunks with up to 8 host threads, each thread
// taking its slice of every chunk.  With `after`, the calling thread waits
// until every thread has copied chunk c and then calls after(begin, end) (the
// DMA of that chunk) while the threads go on with the next chunks; with
// `ready`, each thread waits for event ready[c] before copying chunk c.
void ArnoldiRunner::stage_pipelined(double* dst, const double* src, size_t n, size_t nchunk,
                                    const std::function<void(size_t, size_t)>& after,
                                    const hipEvent_t* ready) {
  const size_t nt = std::min<size_t>(8, std::max<size_t>(1, n / nchunk / (size_t(1) << 18)));
  std::vector<std::atomic<int>> done(nchunk);
  for (auto& d : done) d.store(0, std::memory_order_relaxed);
  std::atomic<int> failed{0};
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t) {
    th.emplace_back([&, t] {
      for (size_t c = 0; c < nchunk; ++c) {
        const size_t a = n * c / nchunk, b = n * (c + 1) / nchunk;
        const size_t lo = a + (b - a) * t / nt, hi = a + (b - a) * (t + 1) / nt;
        if (ready && hipEventSynchronize(ready[c]) != hipSuccess) failed.store(1);
        if (!failed.load()) std::copy(src + lo, src + hi, dst + lo);
        done[c].fetch_add(1, std::memory_order_release);
      }
    });
  }
  std::exception_ptr err;
  if (after) {
    for (size_t c = 0; c < nchunk; ++c) {
      while (done[c].load(std::memory_order_acquire) < static_cast<int>(nt)) std::this_thread::yield();
      if (!err) {
        try {
          after(n * c / nchunk, n * (c + 1) / nchunk);
        } catch (...) {
          err = std::current_exception();
        }
      }
    }
  }
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
  if (failed.load()) throw mek::Error("ERROR: waiting for a staged copy of y failed");
}

void ArnoldiRunner::compute_expmv(const std::vector<double>& v_host, std::vector<double>& y_host) {
  DeviceContext& dc = device_contexts[0];
  if (static_cast<int>(v_host.size()) < dc.local_rows || static_cast<int>(y_host.size()) < dc.local_rows)
    throw mek::Error("ERROR: v and y must hold " + std::to_string(dc.local_rows) + " entries");
  compute_expmv_host(v_host.data(), y_host.data());
}

void ArnoldiRunner::compute_expmv_host(const double* v_host, double* y_host) {
  const double t0 = now_ms();
  DeviceContext& dc = device_contexts[0];
  const size_t n = static_cast<size_t>(dc.local_rows);
  MEK_HIP(hipSetDevice(dc.device_id));
  if (h_stage_n_ < n) {
    if (h_stage_) (void)hipHostFree(h_stage_);
    h_stage_ = nullptr;
    h_stage_n_ = 0;
    MEK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_stage_), sizeof(double) * std::max<size_t>(n, 1)));
    h_stage_n_ = n;
  }
  hipStream_t s = dc.stream_compute;
  const size_t nchunk = n >= kStageChunkMin * 2 ? std::min<size_t>(kStageChunks, n / kStageChunkMin) : 1;
  if (nchunk == 1) {
    stage_copy(h_stage_, v_host, n);
    solve_on_device(nullptr, h_stage_);
    MEK_HIP(hipMemcpyAsync(h_stage_, dc.d_y, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    MEK_HIP(hipStreamSynchronize(s));
    stage_copy(y_host, h_stage_, n);
  } else {
    // v: the host threads fill the pinned buffer chunk by chunk while the DMA
    // engine moves the chunks already filled into V[:, 0]
    stage_pipelined(h_stage_, v_host, n, nchunk, [&](size_t a, size_t b) {
      MEK_HIP(hipMemcpyAsync(dc.d_V_m + a, h_stage_ + a, sizeof(double) * (b - a), hipMemcpyHostToDevice, s));
    }, nullptr);
    solve_on_device(dc.d_V_m, nullptr);
    // y: every chunk's DMA is enqueued with an event after it; the host
    // threads empty each chunk of the pinned buffer as soon as it has landed
    if (stage_ev_.size() < nchunk) {
      const size_t have = stage_ev_.size();
      stage_ev_.resize(nchunk);
      for (size_t c = have; c < nchunk; ++c)
        MEK_HIP(hipEventCreateWithFlags(&stage_ev_[c], hipEventDisableTiming));
    }
    for (size_t c = 0; c < nchunk; ++c) {
      const size_t a = n * c / nchunk, b = n * (c + 1) / nchunk;
      MEK_HIP(hipMemcpyAsync(h_stage_ + a, dc.d_y + a, sizeof(double) * (b - a), hipMemcpyDeviceToHost, s));
      MEK_HIP(hipEventRecord(stage_ev_[c], s));
    }
    stage_pipelined(y_host, h_stage_, n, nchunk, nullptr, stage_ev_.data());
  }
  if (options.verbose) {
    // reference's analytic FLOP model (main.cu:1538-1565)
    const double N = global_rows, nnz = static_cast<double>(global_nnz), mv = params.m,
                 S = restarts_done;
    const double seg = 2.0 * nnz * mv + 4.0 * N * mv * (mv + 1.0) + 5.0 * N * mv + 2.0 * N;
    const double tot = 3.0 * N * S + S * seg;
    const double cseg = 2.0 * 1.0 * mv * mv * mv + 2.0 * mv * mv;
    std::cout << "[FLOPs] N=" << N << " nnz=" << nnz << " m=" << mv << " steps=" << S << std::endl;
    std::cout << "[FLOPs] GPU segment: " << seg << " | GPU total: " << tot << std::endl;
    std::cout << "[FLOPs] CPU segment: " << cseg << " | CPU total: " << S * cseg << " (C_exp=" << 1.0
              << ")" << std::endl;
  }
  stats.solve_ms = now_ms() - t0;
  if (options.verbose) std::cout << "METRIC total_compute_expmv_ms=" << stats.solve_ms << std::endl;
}

void ArnoldiRunner::compute_expmv_device(const double* d_v, double* d_y) {
  const double t0 = now_ms();
  DeviceContext& dc = device_contexts[0];
  solve_on_device(d_v, nullptr);
  if (d_y != dc.d_y)
    MEK_HIP(hipMemcpyAsync(d_y, dc.d_y, sizeof(double) * dc.local_rows, hipMemcpyDeviceToDevice,
                           dc.stream_compute));
  MEK_HIP(hipStreamSynchronize(dc.stream_compute));
  stats.solve_ms = now_ms() - t0;
}

int ArnoldiRunner::arnoldi_decomposition(const std::vector<double>& v_host, std::vector<double>& H_out) {
  DeviceContext& dc = device_contexts[0];
  MEK_HIP(hipSetDevice(dc.device_id));
  stats = mek::SolveStats();
  init_q1_device(nullptr, v_host.data());
  run_columns(params.m);
  const int k_used = read_status();
  const mek::ArnoldiBufs b = dc.bufs();
  std::vector<double> Hd(static_cast<size_t>(b.ldh) * b.ldh);
  MEK_HIP(hipMemcpy(Hd.data(), b.H, sizeof(double) * Hd.size(), hipMemcpyDeviceToHost));
  const int m = params.m;
  H_out.assign(static_cast<size_t>(m + 1) * m, 0.0);
  for (int c = 0; c < k_used; ++c)
    for (int r = 0; r <= std::min(c + 1, m); ++r)
      H_out[static_cast<size_t>(c) * (m + 1) + r] = Hd[static_cast<size_t>(c) * b.ldh + r];
  return k_used;
}

void ArnoldiRunner::destroy() {
  if (h_stage_) (void)hipHostFree(h_stage_);
  h_stage_ = nullptr;
  h_stage_n_ = 0;
  for (auto& e : stage_ev_) (void)hipEventDestroy(e);
  stage_ev_.clear();
  for (auto& p : ev_pool_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  ev_pool_.clear();
  for (auto& dc : device_contexts) dc.destroy();
}
