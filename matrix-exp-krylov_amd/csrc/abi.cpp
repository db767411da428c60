// extern "C" boundary of libmek.so (declared in include/mek.h).  Every entry
// point catches exceptions and returns a status; the message is kept per
// thread for mek_last_error().
#include "mek.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host/hip_check.hpp"
#include "kernels/launch.hpp"
#include "mek/arnoldi_runner.hpp"
#include "mek/batch_runner.hpp"
#include "mek/block_runner.hpp"
#include "kernels/block.hpp"
#include "mek/csr_host.hpp"
#include "mek/mlp.hpp"

struct mek_csr {
  CSRHost A;
};

struct mek_runner {
  std::unique_ptr<ArnoldiRunner> R;
};

struct mek_batch {
  std::unique_ptr<mek::BatchRunner> B;
};

struct mek_block {
  std::unique_ptr<mek::BlockRunner> B;
};

struct mek_mlp {
  std::unique_ptr<mek::Mlp> N;
};

struct mek_spmv_plan {
  mek::ArnoldiBufs b{};
  int* d_blocks = nullptr;
  double* d_partials = nullptr;
  int* d_iscal = nullptr;
  long long* d_sell_ptr = nullptr;
  int* d_sell_col = nullptr;
  void* d_sell_val = nullptr;
  void* d_dia_val = nullptr;
  int* d_dia_off = nullptr;
  void* d_dia_mask = nullptr;
  double* d_dia_cval = nullptr;
};

namespace {

thread_local std::string g_err;

// A caller's argument failed a check at the boundary (MEK_E_INVALID).
struct InvalidArgument : mek::Error {
  using mek::Error::Error;
};

int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <typename F>
int guarded(F&& f) {
  try {
    g_err.clear();
    f();
    return MEK_OK;
  } catch (const mek::LoadError& e) {
    return set_err(MEK_E_LOAD, e.what());
  } catch (const InvalidArgument& e) {
    return set_err(MEK_E_INVALID, e.what());
  } catch (const mek::Error& e) {
    const std::string m = e.what();
    return set_err(m.rfind("HIP error", 0) == 0 ? MEK_E_HIP : MEK_E_SOLVER, m);
  } catch (const std::bad_alloc&) {
    return set_err(MEK_E_INTERNAL, "out of host memory");
  } catch (const std::exception& e) {
    return set_err(MEK_E_INTERNAL, e.what());
  } catch (...) {
    return set_err(MEK_E_INTERNAL, "unknown error");
  }
}

void need(bool ok, const char* what) {
  if (!ok) throw InvalidArgument(std::string("invalid argument: ") + what);
}

}  // namespace

extern "C" {

const char* mek_last_error(void) { return g_err.c_str(); }
const char* mek_version(void) { return "mek 0.1 (gfx950)"; }

int mek_device_count(int* count) {
  return guarded([&] {
    need(count, "count");
    MEK_HIP(hipGetDeviceCount(count));
  });
}
int mek_device_synchronize(void) { return guarded([&] { MEK_HIP(hipDeviceSynchronize()); }); }
int mek_malloc(void** p, size_t bytes) {
  return guarded([&] {
    need(p, "ptr");
    MEK_HIP(hipMalloc(p, bytes ? bytes : 1));
  });
}
int mek_free(void* p) { return guarded([&] { if (p) MEK_HIP(hipFree(p)); }); }
int mek_memcpy_htod(void* d, const void* s, size_t n) {
  return guarded([&] { MEK_HIP(hipMemcpy(d, s, n, hipMemcpyHostToDevice)); });
}
int mek_memcpy_dtoh(void* d, const void* s, size_t n) {
  return guarded([&] { MEK_HIP(hipMemcpy(d, s, n, hipMemcpyDeviceToHost)); });
}
int mek_memcpy_dtod(void* d, const void* s, size_t n) {
  return guarded([&] { MEK_HIP(hipMemcpy(d, s, n, hipMemcpyDeviceToDevice)); });
}

// ------------------------------------------------------------------ CSR

int mek_csr_create(int rows, int cols, long long nnz, const int* row_ptr, const int* col_idx,
                   const double* values, mek_csr** out) {
  return guarded([&] {
    need(out && row_ptr && rows >= 0 && cols >= 0 && nnz >= 0, "csr arrays");
    need(nnz == 0 || (col_idx && values), "csr arrays");
    need(row_ptr[0] == 0 && row_ptr[rows] == nnz, "row_ptr[0] == 0 and row_ptr[rows] == nnz");
    for (int r = 0; r < rows; ++r) need(row_ptr[r] <= row_ptr[r + 1], "row_ptr non-decreasing");
    for (long long k = 0; k < nnz; ++k) need(col_idx[k] >= 0 && col_idx[k] < cols, "column index in range");
    auto h = std::make_unique<mek_csr>();
    h->A.rows = rows;
    h->A.cols = cols;
    h->A.nnz = nnz;
    h->A.row_ptr.assign(row_ptr, row_ptr + rows + 1);
    h->A.col_idx.assign(col_idx, col_idx + nnz);
    h->A.values.assign(values, values + nnz);
    *out = h.release();
  });
}

int mek_csr_load_bincsr(const char* path, mek_csr** out) {
  return guarded([&] {
    need(path && out, "path");
    auto h = std::make_unique<mek_csr>();
    h->A.load_binary_csr(path);
    *out = h.release();
  });
}

int mek_csr_load_mtx(const char* path, mek_csr** out) {
  return guarded([&] {
    need(path && out, "path");
    auto h = std::make_unique<mek_csr>();
    h->A.load_matrix_market(path);
    *out = h.release();
  });
}

int mek_csr_info(const mek_csr* A, int* rows, int* cols, long long* nnz, int* value_dtype) {
  return guarded([&] {
    need(A, "csr");
    if (rows) *rows = A->A.rows;
    if (cols) *cols = A->A.cols;
    if (nnz) *nnz = A->A.nnz;
    if (value_dtype) *value_dtype = A->A.source_value_dtype;
  });
}

int mek_csr_f32_inexact(const mek_csr* A, long long* count, double* max_rel, long long* nonfinite) {
  return guarded([&] {
    need(A && count && max_rel, "matrix, outputs");
    *count = A->A.f32_inexact(max_rel, nonfinite);
  });
}

int mek_csr_diagonals(const mek_csr* A, int* offsets, int* count) {
  return guarded([&] {
    need(A && offsets && count, "matrix, outputs");
    const std::vector<int> offs =
        mek::dia_offsets(A->A.row_ptr.data(), A->A.col_idx.data(), A->A.values.data(), A->A.rows);
    *count = static_cast<int>(offs.size());
    for (size_t i = 0; i < offs.size(); ++i) offsets[i] = offs[i];
  });
}

int mek_csr_diagonal_values(const mek_csr* A, int f32, double* values, int* count) {
  return guarded([&] {
    need(A && values && count, "matrix, outputs");
    const std::vector<int> offs =
        mek::dia_offsets(A->A.row_ptr.data(), A->A.col_idx.data(), A->A.values.data(), A->A.rows);
    const std::vector<double> cv =
        mek::dia_constants(A->A.row_ptr.data(), A->A.col_idx.data(), A->A.values.data(), A->A.rows, offs, f32 != 0);
    *count = static_cast<int>(cv.size());
    for (size_t i = 0; i < cv.size(); ++i) values[i] = cv[i];
  });
}

int mek_csr_export(const mek_csr* A, int* row_ptr, int* col_idx, double* values) {
  return guarded([&] {
    need(A, "csr");
    if (row_ptr) std::memcpy(row_ptr, A->A.row_ptr.data(), sizeof(int) * A->A.row_ptr.size());
    if (col_idx) std::memcpy(col_idx, A->A.col_idx.data(), sizeof(int) * A->A.col_idx.size());
    if (values) std::memcpy(values, A->A.values.data(), sizeof(double) * A->A.values.size());
  });
}

void mek_csr_destroy(mek_csr* A) { delete A; }

// --------------------------------------------------------------- runner

void mek_solver_options_default(mek_solver_options* o) {
  if (!o) return;
  const mek::SolverOptions d;
  o->step_control = static_cast<int>(d.control);
  o->gs_mode = d.gs_mode;
  o->dgks_eta = d.dgks_eta;
  o->values_f32 = d.values_f32;
  o->verbose = d.verbose;
  o->timing = d.timing;
  o->fuse_spmv_dot = d.fuse_spmv_dot;
  o->lanczos = d.lanczos;
  o->autotune = d.autotune;
  o->sstep = d.sstep;
  o->sstep_first = d.sstep_first;
  o->spmv_format = d.spmv_format;
}

int mek_runner_create(const mek_csr* A, const mek_arnoldi_params* p, const mek_solver_options* o,
                      int device_id, mek_runner** out) {
  return guarded([&] {
    need(A && p && out, "runner arguments");
    need(device_id >= 0, "device_id");
    need(std::isfinite(p->t) && p->t >= 0.0, "t finite and >= 0");
    ArnoldiParams params(p->m, p->t, p->tol, p->max_restarts, p->reorthogonalize != 0);
    auto h = std::make_unique<mek_runner>();
    h->R = std::make_unique<ArnoldiRunner>(1, params, A->A.rows, A->A.cols, A->A.nnz);
    if (o) {
      mek::SolverOptions& so = h->R->options;
      need(o->step_control >= 0 && o->step_control <= 2, "step_control");
      need(o->gs_mode >= -1 && o->gs_mode <= 2, "gs_mode");
      so.control = static_cast<mek::StepControl>(o->step_control);
      so.gs_mode = o->gs_mode;
      so.dgks_eta = o->dgks_eta;
      need(o->values_f32 >= 0 && o->values_f32 <= 2, "values_f32 in {0, 1, 2}");
      so.values_f32 = o->values_f32;
      so.verbose = o->verbose != 0;
      so.timing = o->timing != 0;
      so.fuse_spmv_dot = o->fuse_spmv_dot != 0;
      so.lanczos = o->lanczos != 0;
      so.autotune = o->autotune != 0;
      need(o->sstep == 0 || (o->sstep >= 2 && o->sstep <= mek::kSsMaxS), "sstep 0 or in [2, 16]");
      need(o->sstep_first >= -1, "sstep_first >= -1");
      so.sstep = o->sstep;
      so.sstep_first = o->sstep_first;
      need(o->spmv_format >= 0 && o->spmv_format <= 4, "spmv_format in {0, 1, 2, 3, 4}");
      so.spmv_format = o->spmv_format;
    }
    try {
      h->R->init_device(A->A, device_id);  // reads the matrix in place
    } catch (...) {
      h->R->destroy();
      throw;
    }
    *out = h.release();
  });
}

int mek_runner_expmv(mek_runner* R, const double* v, double* y) {
  return guarded([&] {
    need(R && v && y, "runner, v, y");
    R->R->compute_expmv_host(v, y);
  });
}

int mek_runner_expmv_device(mek_runner* R, const double* d_v, double* d_y) {
  return guarded([&] {
    need(R && d_v && d_y, "runner, d_v, d_y");
    R->R->compute_expmv_device(d_v, d_y);
  });
}

int mek_runner_set_t(mek_runner* R, double t) {
  return guarded([&] {
    need(R, "runner");
    need(std::isfinite(t) && t >= 0.0, "t finite and >= 0");
    ArnoldiRunner& r = *R->R;
    r.params.t = t;
    r.ts_params.dt_max = t;
    r.ts_params.dt_min = std::max(1e-12, t * 1e-6);
    const double denom = (r.params.max_restarts > 0) ? r.params.max_restarts : 1;
    r.ts_params.dt_init = t / denom;
  });
}

int mek_runner_stats(const mek_runner* R, mek_solve_stats* o) {
  return guarded([&] {
    need(R && o, "runner, out");
    const ArnoldiRunner& r = *R->R;
    o->converged = r.converged ? 1 : 0;
    o->restarts_done = r.restarts_done;
    o->last_residual = r.last_residual;
    o->t_done = r.stats.t_done;
    o->attempts = r.stats.attempts;
    o->accepted = r.stats.accepted;
    o->rejected = r.stats.rejected;
    o->spmv_count = r.stats.spmv_count;
    o->extra_spmv = r.stats.extra_spmv;
    o->reorth_columns = r.stats.reorth_columns;
    o->aborted = r.stats.aborted ? 1 : 0;
    o->solve_ms = r.stats.solve_ms;
    o->sstep_blocks = r.stats.sstep_blocks;
    o->sstep_second = r.stats.sstep_second;
    o->sstep_fallbacks = r.stats.sstep_fallbacks;
    o->sstep_fused = r.stats.sstep_fused;
  });
}

int mek_runner_steps(const mek_runner* R, double* tde, int* mar, int cap, int* n) {
  return guarded([&] {
    need(R && n, "runner, n");
    const auto& st = R->R->stats.steps;
    *n = static_cast<int>(st.size());
    for (int i = 0; i < cap && i < static_cast<int>(st.size()); ++i) {
      if (tde) {
        tde[3 * i] = st[i].t;
        tde[3 * i + 1] = st[i].dt;
        tde[3 * i + 2] = st[i].err;
      }
      if (mar) {
        mar[3 * i] = st[i].m_used;
        mar[3 * i + 1] = st[i].accepts;
        mar[3 * i + 2] = st[i].rejects;
      }
    }
  });
}

int mek_runner_kernel_time(const mek_runner* R, int kind, double* ms, long long* launches, double* bytes) {
  return guarded([&] {
    need(R && kind >= 0 && kind < mek::kNumKinds, "runner, kind");
    const mek::KernelTiming& k = R->R->stats.kernels[kind];
    if (ms) *ms = k.ms;
    if (launches) *launches = k.launches;
    if (bytes) *bytes = k.bytes;
  });
}

int mek_runner_arnoldi(mek_runner* R, const double* v, double* H, int* k_used) {
  return guarded([&] {
    need(R && v && H, "runner, v, H");
    const int n = R->R->global_rows;
    std::vector<double> vh(v, v + n), Hh;
    const int k = R->R->arnoldi_decomposition(vh, Hh);
    std::memcpy(H, Hh.data(), sizeof(double) * Hh.size());
    if (k_used) *k_used = k;
  });
}

int mek_runner_spmv_format(const mek_runner* R, int* format, long long* entries, int* width) {
  return guarded([&] {
    need(R && R->R && format, "runner, format");
    const DeviceContext& dc = R->R->device_contexts[0];
    if (dc.n_diags > 0) {
      *format = dc.d_dia_mask ? 4 : 3;
      if (entries) *entries = static_cast<long long>(dc.n_diags) * dc.local_rows;
      if (width) *width = dc.n_diags;
      return;
    }
    *format = dc.n_slices > 0 ? 2 : 1;
    if (entries) *entries = dc.sell_entries;
    if (width) *width = dc.sell_width;
  });
}

int mek_runner_sstep_grid(const mek_runner* R, int* grid, double* us) {
  return guarded([&] {
    need(R && grid && us, "runner, outputs");
    *grid = R->R->device_contexts[0].ss_grid;
    *us = R->R->tuning.sstep_us;
  });
}

int mek_runner_tuning(const mek_runner* R, int* nnz_cap, int* spmv_grid, int* sweep_grid, double* spmv_us,
                      double* sweep_us) {
  return guarded([&] {
    need(R && nnz_cap && spmv_grid && sweep_grid && spmv_us && sweep_us, "runner, outputs");
    const mek::TuneChoice& t = R->R->tuning;
    *nnz_cap = t.nnz_cap;
    *spmv_grid = t.spmv_grid;
    *sweep_grid = t.sweep_grid;
    *spmv_us = t.spmv_us;
    *sweep_us = t.sweep_us;
  });
}

int mek_runner_values_f32(const mek_runner* R, int* used) {
  return guarded([&] {
    need(R && used, "runner, out");
    *used = R->R->values_f32_used ? 1 : 0;
  });
}

int mek_runner_stream(const mek_runner* R, void** stream) {
  return guarded([&] {
    need(R && stream, "runner, stream");
    *stream = R->R->device_contexts[0].stream_compute;
  });
}

void mek_runner_destroy(mek_runner* R) {
  if (!R) return;
  if (R->R) R->R->destroy();
  delete R;
}

// ------------------------------------------------------------------ batch

int mek_batch_create(int nprob, const mek_csr* const* A, const double* t, const int* m,
                     const double* tol, const int* max_restarts, const int* substeps, int step_control,
                     int values_f32, int gs_mode, double dgks_eta, int use_graph, int autotune,
                     int device_id, mek_batch** out) {
  return guarded([&] {
    need(nprob >= 1 && A && t && m && tol && out, "batch arguments");
    need(gs_mode >= 0 && gs_mode <= 2, "gs_mode");
    need(step_control >= 0 && step_control <= 2, "step_control");
    std::vector<mek::BatchProblem> probs(nprob);
    need(values_f32 >= 0 && values_f32 <= 2, "values_f32 in {0, 1, 2}");
    for (int p = 0; p < nprob; ++p) {
      need(A[p] != nullptr, "matrix");
      need(std::isfinite(t[p]) && t[p] >= 0.0, "t finite and >= 0");
      probs[p] = {&A[p]->A, t[p], m[p], tol[p], max_restarts ? max_restarts[p] : 10,
                  substeps ? substeps[p] : 0, step_control};
    }
    auto h = std::make_unique<mek_batch>();
    h->B = std::make_unique<mek::BatchRunner>(device_id, probs, values_f32, gs_mode, dgks_eta,
                                             use_graph != 0, autotune != 0);
    *out = h.release();
  });
}

int mek_batch_values_f32(const mek_batch* B, int* used) {
  return guarded([&] {
    need(B && used, "batch, out");
    *used = B->B->values_f32_used() ? 1 : 0;
  });
}

int mek_batch_tuning(const mek_batch* B, int* spmv_grid) {
  return guarded([&] {
    need(B && spmv_grid, "batch, out");
    *spmv_grid = B->B->spmv_grid();
  });
}

int mek_batch_time_spmv(mek_batch* B, int reps, double* us) {
  return guarded([&] {
    need(B && us && reps >= 1, "batch, reps >= 1, out");
    *us = B->B->time_spmv(reps);
  });
}

int mek_batch_upload_v(mek_batch* B, const double* const* v) {
  return guarded([&] {
    need(B && v, "batch, v");
    std::vector<const double*> vv(v, v + B->B->num_problems());
    B->B->upload_v(vv);
  });
}

int mek_batch_solve(mek_batch* B) {
  return guarded([&] {
    need(B, "batch");
    B->B->solve_device();
  });
}

int mek_batch_download_y(mek_batch* B, double* const* y) {
  return guarded([&] {
    need(B && y, "batch, y");
    std::vector<double*> yy(y, y + B->B->num_problems());
    B->B->download_y(yy);
  });
}

int mek_batch_stats(const mek_batch* B, int p, mek_solve_stats* o) {
  return guarded([&] {
    need(B && o && p >= 0 && p < B->B->num_problems(), "batch, problem index, out");
    const mek::BatchProblemStats s = B->B->stats(p);
    std::memset(o, 0, sizeof(*o));
    o->converged = s.converged;
    o->restarts_done = s.accepted;
    o->accepted = s.accepted;
    o->attempts = s.attempts;
    o->rejected = s.rejected;
    o->spmv_count = s.columns;
    o->reorth_columns = s.reorth_columns;
    o->aborted = s.aborted;
    o->t_done = s.t_done;
    o->last_residual = s.last_residual;
    o->solve_ms = B->B->solve_ms();
  });
}

int mek_batch_cycles(const mek_batch* B, int* cycles) {
  return guarded([&] {
    need(B && cycles, "batch, cycles");
    *cycles = B->B->cycles();
  });
}

void mek_batch_destroy(mek_batch* B) { delete B; }

// ------------------------------------------------------- learned selector

int mek_features(const mek_csr* A, double t, double tol, float* out16, int device_id) {
  return guarded([&] {
    need(A && out16, "csr, out");
    mek::features_csr(A->A, t, tol, out16, device_id);
  });
}

int mek_batch_features(mek_batch* B, float* out) {
  return guarded([&] {
    need(B && out, "batch, out");
    B->B->features(out);
  });
}

int mek_mlp_param_count(void) { return mek::Mlp::param_count(); }

int mek_mlp_create(unsigned seed, int device_id, mek_mlp** out) {
  return guarded([&] {
    need(out != nullptr, "out");
    auto h = std::make_unique<mek_mlp>();
    h->N = std::make_unique<mek::Mlp>(seed, device_id);
    *out = h.release();
  });
}

int mek_mlp_set_params(mek_mlp* N, const float* p) {
  return guarded([&] {
    need(N && p, "mlp, params");
    N->N->set_params(p);
  });
}

int mek_mlp_get_params(const mek_mlp* N, float* p) {
  return guarded([&] {
    need(N && p, "mlp, params");
    N->N->get_params(p);
  });
}

int mek_mlp_forward(mek_mlp* N, const float* X, int B, float* Y) {
  return guarded([&] {
    need(N && X && Y && B >= 1, "mlp, X, Y, B");
    N->N->forward(X, B, Y);
  });
}

int mek_mlp_loss_grad(mek_mlp* N, const float* X, const float* T, int B, float* loss, float* grads) {
  return guarded([&] {
    need(N && X && T && B >= 1, "mlp, X, T, B");
    const float l = N->N->loss_grad(X, T, B, grads);
    if (loss) *loss = l;
  });
}

int mek_mlp_train(mek_mlp* N, const float* X, const float* T, int n, int batch, int epochs, float lr,
                  unsigned seed, float* loss) {
  return guarded([&] {
    need(N && X && T && n >= 1 && batch >= 1 && epochs >= 0, "mlp training arguments");
    const float l = N->N->train(X, T, n, batch, epochs, lr, seed);
    if (loss) *loss = l;
  });
}

void mek_mlp_destroy(mek_mlp* N) { delete N; }

// ----------------------------------------------------------- kernel level

int mek_spmv_plan_create(int n, long long nnz, const int* d_row_ptr, const int* d_col_idx,
                         const void* d_vals, int vals_f32, const int* h_row_ptr, mek_spmv_plan** out) {
  return guarded([&] {
    need(out && d_row_ptr && h_row_ptr && n >= 0 && nnz >= 0, "plan arguments");
    need(nnz == 0 || (d_col_idx && d_vals), "plan arrays");
    need(h_row_ptr[n] == nnz, "h_row_ptr[n] == nnz");
    auto P = std::make_unique<mek_spmv_plan>();
    const std::vector<int> blocks = mek::build_row_blocks(h_row_ptr, n, mek::kSpmvNnzCap, mek::kSpmvRowCap);
    const int grid = std::max(1, std::min(static_cast<int>(blocks.size()) - 1, mek::kMaxGrid));
    MEK_HIP(hipMalloc(&P->d_blocks, sizeof(int) * blocks.size()));
    MEK_HIP(hipMemcpy(P->d_blocks, blocks.data(), sizeof(int) * blocks.size(), hipMemcpyHostToDevice));
    MEK_HIP(hipMalloc(&P->d_partials, sizeof(double) * grid));
    MEK_HIP(hipMalloc(&P->d_iscal, sizeof(int) * mek::kNumInts));
    MEK_HIP(hipMemset(P->d_iscal, 0, sizeof(int) * mek::kNumInts));
    mek::ArnoldiBufs& b = P->b;
    b.n = n;
    b.row_ptr = d_row_ptr;
    b.col_idx = d_col_idx;
    b.vals = d_vals;
    b.vals_f32 = vals_f32 ? 1 : 0;
    b.row_blocks = P->d_blocks;
    b.n_row_blocks = static_cast<int>(blocks.size()) - 1;
    b.iscal = P->d_iscal;
    b.partials = P->d_partials;
    b.grid = grid;
    b.spmv_grid = grid;
    *out = P.release();
  });
}

int mek_spmv_execute(mek_spmv_plan* P, const double* d_x, double* d_y, void* stream) {
  return guarded([&] {
    need(P && d_x && d_y, "plan, x, y");
    mek::launch_spmv_plain(P->b, d_x, nullptr, d_y, static_cast<hipStream_t>(stream));
    MEK_HIP_LAUNCHED();
  });
}

int mek_spmv_plan_use_sell(mek_spmv_plan* P, const int* h_row_ptr, long long* entries, int* width) {
  return guarded([&] {
    need(P && h_row_ptr, "plan, h_row_ptr");
    mek::ArnoldiBufs& b = P->b;
    need(P->d_sell_ptr == nullptr, "the plan already uses the sliced-ELL copy");
    std::vector<long long> ptr;
    int wmax = 0;
    const long long ne = mek::sell_layout(h_row_ptr, b.n, ptr, &wmax);
    const size_t nes = static_cast<size_t>(std::max<long long>(ne, 1));
    const int ns = static_cast<int>(ptr.size()) - 1;
    const int grid = std::max(1, std::min(mek::kMaxGrid, (ns + 15) / 16));
    if (grid > b.spmv_grid) {  // partials for the larger of the two grids
      double* d_part = nullptr;
      MEK_HIP(hipMalloc(&d_part, sizeof(double) * grid));
      if (P->d_partials) (void)hipFree(P->d_partials);
      P->d_partials = d_part;
      b.partials = d_part;
    }
    MEK_HIP(hipMalloc(&P->d_sell_ptr, sizeof(long long) * ptr.size()));
    MEK_HIP(hipMalloc(&P->d_sell_col, sizeof(int) * nes));
    MEK_HIP(hipMalloc(&P->d_sell_val, (b.vals_f32 ? sizeof(float) : sizeof(double)) * nes));
    MEK_HIP(hipMemcpy(P->d_sell_ptr, ptr.data(), sizeof(long long) * ptr.size(), hipMemcpyHostToDevice));
    mek::launch_sell_build(b.row_ptr, b.col_idx, b.vals, b.vals_f32, b.n, P->d_sell_ptr, P->d_sell_col,
                           P->d_sell_val, nullptr);
    MEK_HIP_LAUNCHED();
    MEK_HIP(hipDeviceSynchronize());
    b.sell_ptr = P->d_sell_ptr;
    b.sell_col = P->d_sell_col;
    b.sell_val = P->d_sell_val;
    b.n_slices = ns;
    b.spmv_grid = grid;
    if (entries) *entries = ne;
    if (width) *width = wmax;
  });
}

static int plan_use_dia(mek_spmv_plan* P, const int* h_row_ptr, const int* h_col_idx, const double* h_vals,
                        int* n_diags, bool constant) {
  return guarded([&] {
    need(P && h_row_ptr && (P->b.n == 0 || (h_col_idx && h_vals)), "plan, h_row_ptr, h_col_idx, h_vals");
    mek::ArnoldiBufs& b = P->b;
    need(b.n_diags == 0, "the plan already uses a diagonal layout");
    const std::vector<int> offs = mek::dia_offsets(h_row_ptr, h_col_idx, h_vals, b.n);
    if (offs.empty() || b.n <= 0)
      throw InvalidArgument("the matrix has no diagonal form (more than 16 diagonals, columns not ascending "
                                 "within a row, or NaN values)");
    std::vector<double> cv;
    if (constant) {
      cv = mek::dia_constants(h_row_ptr, h_col_idx, h_vals, b.n, offs, b.vals_f32 != 0);
      if (cv.empty()) throw InvalidArgument("the matrix has no constant-diagonal form (a diagonal holds two values)");
    }
    const int nd = static_cast<int>(offs.size());
    const long long ldd = (static_cast<long long>(b.n) + mek::kWave - 1) / mek::kWave * mek::kWave;
    const long long slices = ldd / mek::kWave;
    const int grid = static_cast<int>(std::max<long long>(1, std::min<long long>(mek::kMaxGrid, (slices + 15) / 16)));
    if (grid > b.spmv_grid) {  // partials for the larger of the two grids
      double* d_part = nullptr;
      MEK_HIP(hipMalloc(&d_part, sizeof(double) * grid));
      if (P->d_partials) (void)hipFree(P->d_partials);
      P->d_partials = d_part;
      b.partials = d_part;
    }
    MEK_HIP(hipMalloc(&P->d_dia_off, sizeof(int) * nd));
    MEK_HIP(hipMemcpy(P->d_dia_off, offs.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
    if (constant) {
      MEK_HIP(hipMalloc(&P->d_dia_mask, (nd <= 8 ? 1 : 2) * static_cast<size_t>(ldd)));
      MEK_HIP(hipMalloc(&P->d_dia_cval, sizeof(double) * nd));
      MEK_HIP(hipMemcpy(P->d_dia_cval, cv.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
      mek::launch_dia_mask_build(b.row_ptr, b.col_idx, b.n, P->d_dia_off, nd, ldd, P->d_dia_mask, nullptr);
    } else {
      MEK_HIP(hipMalloc(&P->d_dia_val,
                        (b.vals_f32 ? sizeof(float) : sizeof(double)) * static_cast<size_t>(ldd) * nd));
      mek::launch_dia_build(b.row_ptr, b.col_idx, b.vals, b.vals_f32, b.n, P->d_dia_off, nd, ldd, P->d_dia_val,
                            nullptr);
    }
    MEK_HIP_LAUNCHED();
    MEK_HIP(hipDeviceSynchronize());
    b.dia_val = P->d_dia_val;
    b.dia_off = P->d_dia_off;
    b.dia_mask = P->d_dia_mask;
    b.dia_cval = P->d_dia_cval;
    b.n_diags = nd;
    b.dia_ld = ldd;
    b.spmv_grid = grid;
    if (n_diags) *n_diags = nd;
  });
}

int mek_spmv_plan_use_dia(mek_spmv_plan* P, const int* h_row_ptr, const int* h_col_idx, const double* h_vals,
                          int* n_diags) {
  return plan_use_dia(P, h_row_ptr, h_col_idx, h_vals, n_diags, false);
}

int mek_spmv_plan_use_dia_const(mek_spmv_plan* P, const int* h_row_ptr, const int* h_col_idx,
                                const double* h_vals, int* n_diags) {
  return plan_use_dia(P, h_row_ptr, h_col_idx, h_vals, n_diags, true);
}

void mek_spmv_plan_destroy(mek_spmv_plan* P) {
  if (!P) return;
  if (P->d_dia_val) (void)hipFree(P->d_dia_val);
  if (P->d_dia_off) (void)hipFree(P->d_dia_off);
  if (P->d_dia_mask) (void)hipFree(P->d_dia_mask);
  if (P->d_dia_cval) (void)hipFree(P->d_dia_cval);
  if (P->d_sell_ptr) (void)hipFree(P->d_sell_ptr);
  if (P->d_sell_col) (void)hipFree(P->d_sell_col);
  if (P->d_sell_val) (void)hipFree(P->d_sell_val);
  if (P->d_blocks) (void)hipFree(P->d_blocks);
  if (P->d_partials) (void)hipFree(P->d_partials);
  if (P->d_iscal) (void)hipFree(P->d_iscal);
  delete P;
}

int mek_expm_dense(int k, const double* d_A, double* d_E) {
  return guarded([&] {
    need(k >= 1 && k <= mek::kMaxKrylov + 2 && d_A && d_E, "k in [1, 502], A, E");
    double* d_scr = nullptr;
    double* d_ws = nullptr;
    MEK_HIP(hipMalloc(&d_scr, sizeof(double) * (32 + k + 4)));
    if (k > mek::kExpmMaxKLds) {
      const hipError_t e = hipMalloc(&d_ws, sizeof(double) * 6 * static_cast<size_t>(k) * k);
      if (e != hipSuccess) {
        (void)hipFree(d_scr);
        MEK_HIP(e);
      }
    }
    mek::ExpmArgs a{};
    a.H = d_A;
    a.ldh = k;
    a.iscal = reinterpret_cast<int*>(d_scr + 16);
    a.iscal_out = reinterpret_cast<int*>(d_scr + 16);
    a.scal = d_scr;
    a.out = d_scr + 24;
    a.out_len = k;
    a.dt = 1.0;
    a.mode = 2;  // plain exponential: no residual
    a.ws = d_ws;
    a.k_override = k;
    a.breakdown_override = 0;
    a.full_out = d_E;
    mek::launch_expm(a, k, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_scr);
    if (d_ws) (void)hipFree(d_ws);
    MEK_HIP(e);
  });
}

int mek_expm_dense_noalloc(int k, const double* d_A, double* d_E, void* stream) {
  return guarded([&] {
    need(d_A && d_E, "A, E");
    need(k >= 1 && k <= mek::kExpmNoAllocMaxK, "k in [1, 128]");
    mek::expm_dense_noalloc(k, d_A, d_E, static_cast<hipStream_t>(stream));
  });
}

// ------------------------------------------------------- block of RHS

int mek_block_create(const mek_csr* A, int m, double t, double tol, int substeps, int gs_mode,
                     int device_id, int timing, int autotune, mek_block** out) {
  return guarded([&] {
    need(A && out, "matrix, out");
    need(substeps >= 0, "substeps >= 0");
    need(gs_mode >= 0 && gs_mode <= 2, "gs_mode");
    auto h = std::make_unique<mek_block>();
    h->B = std::make_unique<mek::BlockRunner>(A->A, m, t, tol, substeps, device_id, timing != 0, gs_mode, 0.1,
                                             autotune != 0);
    *out = h.release();
  });
}

int mek_block_expmv(mek_block* B, const double* V, double* Y) {
  return guarded([&] {
    need(B && V && Y, "block, V, Y");
    B->B->expmv(V, Y);
  });
}

int mek_block_expmv_device(mek_block* B, const double* d_V, double* d_Y) {
  return guarded([&] {
    need(B && d_V && d_Y, "block, V, Y");
    B->B->expmv_device(d_V, d_Y);
  });
}

int mek_block_stats_get(const mek_block* B, mek_block_stats* o) {
  return guarded([&] {
    need(B && o, "block, out");
    const mek::BlockStats& s = B->B->stats();
    std::memset(o, 0, sizeof(*o));
    o->converged = s.converged;
    o->accepted = s.accepted;
    o->attempts = s.attempts;
    o->rejected = s.rejected;
    o->block_steps = s.block_steps;
    o->squarings = s.squarings;
    o->second_passes = s.second_passes;
    o->spmm_grid = B->B->spmm_grid();
    o->update_grid = B->B->update_grid();
    o->t_done = s.t_done;
    o->last_residual = s.last_residual;
    o->solve_ms = s.solve_ms;
    for (int k = 0; k < 5; ++k) {
      o->kernel_ms[k] = s.kernel_ms[k];
      o->kernel_bytes[k] = s.kernel_bytes[k];
      o->launches[k] = s.launches[k];
    }
  });
}

void mek_block_destroy(mek_block* B) { delete B; }

int mek_spmm(const mek_csr* A, const double* X, double* Y, int device_id) {
  return guarded([&] {
    need(A && X && Y, "matrix, X, Y");
    const CSRHost& M = A->A;
    need(M.rows == M.cols, "square matrix");
    MEK_HIP(hipSetDevice(device_id));
    const long long n_pad = (static_cast<long long>(M.rows) + mek::kBlkRows - 1) / mek::kBlkRows * mek::kBlkRows;
    const size_t blk = sizeof(double) * static_cast<size_t>(n_pad) * mek::kRhs;
    const size_t used = sizeof(double) * static_cast<size_t>(M.rows) * mek::kRhs;
    std::vector<void*> mem;
    auto alloc = [&](size_t bytes) {
      void* p = nullptr;
      MEK_HIP(hipMalloc(&p, std::max<size_t>(bytes, 8)));
      mem.push_back(p);
      return p;
    };
    try {
      mek::BlockBufs b{};
      b.n = M.rows;
      b.n_pad = n_pad;
      auto* rp = static_cast<int*>(alloc(sizeof(int) * (M.rows + 1)));
      auto* ci = static_cast<int*>(alloc(sizeof(int) * M.nnz));
      auto* va = static_cast<double*>(alloc(sizeof(double) * M.nnz));
      auto* dx = static_cast<double*>(alloc(blk));
      auto* dy = static_cast<double*>(alloc(blk));
      b.flags = static_cast<int*>(alloc(sizeof(int) * 4));
      MEK_HIP(hipMemset(b.flags, 0, sizeof(int) * 4));
      MEK_HIP(hipMemset(dx, 0, blk));
      MEK_HIP(hipMemcpy(rp, M.row_ptr.data(), sizeof(int) * (M.rows + 1), hipMemcpyHostToDevice));
      if (M.nnz > 0) {
        MEK_HIP(hipMemcpy(ci, M.col_idx.data(), sizeof(int) * M.nnz, hipMemcpyHostToDevice));
        MEK_HIP(hipMemcpy(va, M.values.data(), sizeof(double) * M.nnz, hipMemcpyHostToDevice));
      }
      MEK_HIP(hipMemcpy(dx, X, used, hipMemcpyHostToDevice));
      b.row_ptr = rp;
      b.col_idx = ci;
      b.vals = va;
      mek::block_launch_spmm(b, dx, dy, nullptr);
      MEK_HIP_LAUNCHED();
      MEK_HIP(hipDeviceSynchronize());
      MEK_HIP(hipMemcpy(Y, dy, used, hipMemcpyDeviceToHost));
    } catch (...) {
      for (void* p : mem) (void)hipFree(p);
      throw;
    }
    for (void* p : mem) (void)hipFree(p);
  });
}

int mek_gemm_dense(int N, const double* A, const double* B, double* C, int device_id) {
  return guarded([&] {
    need(N >= mek::kGemmTile && N % mek::kGemmTile == 0 && A && B && C, "N a positive multiple of 64, A, B, C");
    MEK_HIP(hipSetDevice(device_id));
    const size_t bytes = sizeof(double) * static_cast<size_t>(N) * N;
    double *da = nullptr, *db = nullptr, *dc = nullptr;
    MEK_HIP(hipMalloc(&da, bytes));
    hipError_t e = hipMalloc(&db, bytes);
    if (e == hipSuccess) e = hipMalloc(&dc, bytes);
    if (e == hipSuccess) e = hipMemcpy(da, A, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, B, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      mek::dense_gemm(N, da, N, db, N, 1.0, nullptr, N, 0.0, dc, N, nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(C, dc, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dc) (void)hipFree(dc);
    MEK_HIP(e);
  });
}

}  // extern "C"
