// BlockRunner: block Arnoldi exp(t A) V for 32 right-hand sides (block.hip).
#include "mek/block_runner.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "../kernels/block.hpp"
#include "hip_check.hpp"
#include "mek/autotune.hpp"

namespace mek {

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Taylor degree 16 of exp, evaluated by Paterson-Stockmeyer on X, X^2, X^3,
// X^4 after scaling so that ||X||_1 <= kTheta: the remainder is below
// kTheta^17/17! (1 + kTheta/18 + ...) < 7e-17.
constexpr double kTheta = 0.8;
constexpr int kTaylor = 16;

// Expokit's step-size constants (Sidje 1998): safety factor and acceptance slack.
constexpr double kGamma = 0.9;
constexpr double kDelta = 1.2;
constexpr int kMaxRejects = 20;

enum Kind { kSpmm = 0, kDots = 1, kUpdate = 2, kExpm = 3, kLift = 4 };

}  // namespace

BlockRunner::BlockRunner(const CSRHost& A, int m, double t, double tol, int substeps, int device_id,
                         bool timing, int gs_mode, double eta, bool autotune)
    : n_(A.rows), nnz_(A.nnz), m_(m), t_(t), tol_(tol), substeps_(substeps), device_(device_id),
      timing_(timing), gs_mode_(gs_mode), eta_(eta), autotune_(autotune) {
  if (gs_mode < 0 || gs_mode > 2) throw Error("ERROR: gs_mode must be 0, 1 or 2");
  if (A.rows != A.cols) throw Error("ERROR: Matrix must be square");
  if (m < 1 || m > kBlkMaxSteps) throw Error("ERROR: block steps m=" + std::to_string(m) + " outside [1, " +
                                             std::to_string(kBlkMaxSteps) + "]");
  if (static_cast<long long>(kRhs) * (m + 1) >= A.rows)
    throw Error("ERROR: 32 (m + 1) must be below the matrix size for a block Krylov basis");
  if (!(t > 0.0)) throw Error("ERROR: t must be positive");
  // a failure part way releases what was already allocated (the destructor
  // does not run for an object whose constructor threw)
  try {
    MEK_HIP(hipSetDevice(device_));
    MEK_HIP(hipStreamCreate(&stream_));
    anorm_ = std::max(A.norm_inf(), 1e-300);
    n_pad_ = (static_cast<long long>(n_) + kBlkRows - 1) / kBlkRows * kBlkRows;
    const int nchunks = static_cast<int>(n_pad_ / kBlkRows);
    const size_t blk = static_cast<size_t>(n_pad_) * kRhs;
    MEK_HIP(hipMalloc(&d_row_ptr_, sizeof(int) * (n_ + 1)));
    MEK_HIP(hipMalloc(&d_col_idx_, sizeof(int) * std::max<long long>(nnz_, 1)));
    MEK_HIP(hipMalloc(&d_vals_, sizeof(double) * std::max<long long>(nnz_, 1)));
    MEK_HIP(hipMemcpy(d_row_ptr_, A.row_ptr.data(), sizeof(int) * (n_ + 1), hipMemcpyHostToDevice));
    if (nnz_ > 0) {
      MEK_HIP(hipMemcpy(d_col_idx_, A.col_idx.data(), sizeof(int) * nnz_, hipMemcpyHostToDevice));
      MEK_HIP(hipMemcpy(d_vals_, A.values.data(), sizeof(double) * nnz_, hipMemcpyHostToDevice));
    }
    MEK_HIP(hipMalloc(&d_P_, sizeof(double) * blk * (m + 1)));
    MEK_HIP(hipMemset(d_P_, 0, sizeof(double) * blk * (m + 1)));
    MEK_HIP(hipMalloc(&d_W_, sizeof(double) * blk));
    MEK_HIP(hipMemset(d_W_, 0, sizeof(double) * blk));
    const int ldh = kRhs * (m + 1);
    const size_t b32 = kRhs * kRhs;
    const size_t small = 3 * b32 * (m + 1) + static_cast<size_t>(ldh) * ldh + 2 * b32 + 16;
    MEK_HIP(hipMalloc(&d_small_, sizeof(double) * small));
    MEK_HIP(hipMemset(d_small_, 0, sizeof(double) * small));
    MEK_HIP(hipMalloc(&d_partials_, sizeof(double) * b32 * nchunks * (m + 1)));
    MEK_HIP(hipMalloc(&d_flags_, sizeof(int) * 4));
    MEK_HIP(hipMemset(d_flags_, 0, sizeof(int) * 4));
    npad_max_ = (kRhs * m + kGemmTile - 1) / kGemmTile * kGemmTile;
    MEK_HIP(hipMalloc(&d_mat_, sizeof(double) * 7 * static_cast<size_t>(npad_max_) * npad_max_));
    MEK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_pinned_), sizeof(double) * 16));
    MEK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_flags_), sizeof(int) * 4));

    auto* b = new BlockBufs{};
    b->n = n_;
    b->n_pad = n_pad_;
    b->row_ptr = d_row_ptr_;
    b->col_idx = d_col_idx_;
    b->vals = d_vals_;
    b->P = d_P_;
    b->W = d_W_;
    b->Y = d_W_;
    b->S = d_small_;
    b->G = b->S + b32 * (m + 1);
    b->C = b->G + b32 * (m + 1);
    b->Hb = b->C + b32 * (m + 1);
    b->ldh = ldh;
    b->R0 = b->Hb + static_cast<size_t>(ldh) * ldh;
    b->Gm = b->R0 + b32;
    b->scal = b->Gm + b32;
    b->partials = d_partials_;
    b->nchunks = nchunks;
    b->flags = d_flags_;
    b->spmm_grid = 4096;
    b->update_grid = 4096;
    bufs_ = b;
  } catch (...) {
    destroy();
    throw;
  }
}

// Time the SpMM and update-sweep grids on this matrix (the start block is in
// P_0; W is scratch until the first SpMM) and keep the fastest.
void BlockRunner::autotune() {
  BlockBufs& b = *static_cast<BlockBufs*>(bufs_);
  const std::vector<int> grids = {1024, 2048, 4096, 8192};
  std::vector<double> us;
  int i = time_candidates(
      grids,
      [&](int g) {
        b.spmm_grid = g;
        block_launch_spmm(b, d_P_, d_W_, stream_);
      },
      stream_, 5, &us);
  b.spmm_grid = grids[i];
  tuned_spmm_grid_ = grids[i];
  i = time_candidates(
      grids,
      [&](int g) {
        b.update_grid = g;
        block_launch_update(b, 1, d_W_, d_W_, 0, stream_);
      },
      stream_, 5, &us);
  b.update_grid = grids[i];
  tuned_update_grid_ = grids[i];
  MEK_HIP(hipStreamSynchronize(stream_));
}

BlockRunner::~BlockRunner() { destroy(); }

void BlockRunner::timed_begin(int kind) {
  if (!timing_) return;
  if (ev_used_ == ev_.size()) {
    Ev e{};
    MEK_HIP(hipEventCreate(&e.a));
    MEK_HIP(hipEventCreate(&e.b));
    ev_.push_back(e);
  }
  ev_[ev_used_].kind = kind;
  MEK_HIP(hipEventRecord(ev_[ev_used_].a, stream_));
}

void BlockRunner::timed_end(int kind, double bytes) {
  if (!timing_) return;
  Ev& e = ev_[ev_used_++];
  e.kind = kind;
  e.bytes = bytes;
  MEK_HIP(hipEventRecord(e.b, stream_));
}

void BlockRunner::flush_timing() {
  if (!timing_) return;
  MEK_HIP(hipStreamSynchronize(stream_));
  for (size_t i = 0; i < ev_used_; ++i) {
    float ms = 0.f;
    MEK_HIP(hipEventElapsedTime(&ms, ev_[i].a, ev_[i].b));
    stats_.kernel_ms[ev_[i].kind] += ms;
    stats_.kernel_bytes[ev_[i].kind] += ev_[i].bytes;
    stats_.launches[ev_[i].kind] += 1;
  }
  ev_used_ = 0;
}

// P_0 = V (rows < n; padding rows stay zero), its Gram block, Cholesky:
// V = Q_0 R0.
void BlockRunner::start_block(const double* d_src, bool device_src, const double* h_src) {
  const BlockBufs& b = *static_cast<BlockBufs*>(bufs_);
  MEK_HIP(hipMemsetAsync(d_flags_, 0, sizeof(int) * 4, stream_));
  const size_t bytes = sizeof(double) * static_cast<size_t>(n_) * kRhs;
  if (device_src) MEK_HIP(hipMemcpyAsync(d_P_, d_src, bytes, hipMemcpyDeviceToDevice, stream_));
  else MEK_HIP(hipMemcpyAsync(d_P_, h_src, bytes, hipMemcpyHostToDevice, stream_));
  const double blk = 8.0 * static_cast<double>(n_pad_) * kRhs;
  timed_begin(kDots);
  block_launch_dots_partial(b, d_P_, 1, d_P_, 0, 0, stream_);
  timed_end(kDots, blk);  // X1 = X2 = P_0: one block read
  block_launch_dots_fin(b, 1, 0, 0, stream_);
  block_launch_chol(b, -1, stream_);
  MEK_HIP_LAUNCHED();
}

// m block Arnoldi steps from Q_0 (BCGS2 + Cholesky QR); returns the number
// of basis blocks k the exponential uses (m, or fewer after a breakdown).
int BlockRunner::build_basis() {
  const BlockBufs& b = *static_cast<BlockBufs*>(bufs_);
  const double blk = 8.0 * static_cast<double>(n_pad_) * kRhs;
  const double bspmm = 12.0 * nnz_ + 4.0 * (n_ + 1) + 2.0 * 8.0 * n_ * kRhs;
  const size_t bstride = static_cast<size_t>(n_pad_) * kRhs;
  for (int j = 0; j < m_; ++j) {
    const int nb = j + 1;
    double* Pj = d_P_ + bstride * j;
    double* Pn = d_P_ + bstride * (j + 1);
    timed_begin(kSpmm);
    block_launch_spmm(b, Pj, d_W_, stream_);
    timed_end(kSpmm, bspmm);
    // first pass: Q^T Z and Z^T Z in one sweep, Z' = Z - Q (Q^T Z) into P_{j+1}
    timed_begin(kDots);
    block_launch_dots_partial(b, d_P_, nb, d_W_, 1, 0, stream_);
    timed_end(kDots, blk * (nb + 1));
    block_launch_dots_fin(b, nb, 1, 0, stream_);
    block_launch_coeffs(b, j, nb, 1, stream_);
    block_launch_gram(b, j, gs_mode_, eta_, stream_);
    timed_begin(kUpdate);
    block_launch_update(b, nb, d_W_, Pn, 0, stream_);
    timed_end(kUpdate, blk * (nb + 2));
    // second pass and the new block's Gram matrix, when k_bgram asked for it
    // (these launches return at once otherwise and are not timed)
    block_launch_dots(b, d_P_, nb, Pn, 0, 1, stream_);
    block_launch_coeffs(b, j, nb, 2, stream_);
    block_launch_update(b, nb, Pn, Pn, 1, stream_);
    block_launch_dots(b, Pn, 1, Pn, 0, 1, stream_);
    block_launch_chol(b, j, stream_);
    MEK_HIP_LAUNCHED();
  }
  stats_.block_steps += m_;
  MEK_HIP(hipMemcpyAsync(h_flags_, d_flags_, sizeof(int) * 4, hipMemcpyDeviceToHost, stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
  const int built = h_flags_[1];  // blocks P_0 .. P_{built-1}
  return std::max(1, std::min(built, m_));
}

// exp(dt H_k) by Taylor degree 16 (Paterson-Stockmeyer) with scaling and
// squaring; the result is left in F_ (column-major, ld F_ld_).
double BlockRunner::exponential(double dt, int k) {
  const BlockBufs& b = *static_cast<BlockBufs*>(bufs_);
  const int kk = kRhs * k;
  const int N = (kk + kGemmTile - 1) / kGemmTile * kGemmTile;
  const size_t M2 = static_cast<size_t>(npad_max_) * npad_max_;
  double* X = d_mat_;
  double* X2 = X + M2;
  double* X3 = X2 + M2;
  double* X4 = X3 + M2;
  double* Bq = X4 + M2;
  double* T = Bq + M2;
  double* T2 = T + M2;
  timed_begin(kExpm);
  dense_load(N, kk, b.Hb, b.ldh, dt, X, stream_);
  dense_norm1(N, X, N, b.scal + 3, stream_);
  MEK_HIP_LAUNCHED();
  MEK_HIP(hipMemcpyAsync(h_pinned_ + 3, b.scal + 3, sizeof(double), hipMemcpyDeviceToHost, stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
  const double nrm = h_pinned_[3];
  if (!std::isfinite(nrm)) throw Error("ERROR: non-finite block Hessenberg matrix");
  int s = 0;
  if (nrm > kTheta) s = static_cast<int>(std::ceil(std::log2(nrm / kTheta)));
  if (s > 0) dense_combine(N, N, X, X, std::ldexp(1.0, -s), nullptr, 0, nullptr, 0, nullptr, 0, 0.0, stream_);
  double c[kTaylor + 1];
  c[0] = 1.0;
  for (int i = 1; i <= kTaylor; ++i) c[i] = c[i - 1] / i;
  dense_gemm(N, X, N, X, N, 1.0, nullptr, N, 0.0, X2, N, stream_);
  dense_gemm(N, X2, N, X, N, 1.0, nullptr, N, 0.0, X3, N, stream_);
  dense_gemm(N, X2, N, X2, N, 1.0, nullptr, N, 0.0, X4, N, stream_);
  // T = B3 + c16 X4, then T <- B_q + X4 T for q = 2, 1, 0, with
  // B_q = c_{4q} I + c_{4q+1} X + c_{4q+2} X^2 + c_{4q+3} X^3
  dense_combine(N, N, T, X, c[13], X2, c[14], X3, c[15], X4, c[16], c[12], stream_);
  double* cur = T;
  double* nxt = T2;
  for (int q = 2; q >= 0; --q) {
    dense_combine(N, N, Bq, X, c[4 * q + 1], X2, c[4 * q + 2], X3, c[4 * q + 3], nullptr, 0, c[4 * q], stream_);
    dense_gemm(N, X4, N, cur, N, 1.0, Bq, N, 1.0, nxt, N, stream_);
    std::swap(cur, nxt);
  }
  for (int i = 0; i < s; ++i) {
    dense_gemm(N, cur, N, cur, N, 1.0, nullptr, N, 0.0, nxt, N, stream_);
    std::swap(cur, nxt);
  }
  MEK_HIP_LAUNCHED();
  const double gemms = 6.0 + s;
  timed_end(kExpm, gemms * 3.0 * 8.0 * N * N);
  stats_.squarings = s;
  F_ = cur;
  F_ld_ = N;
  return nrm;
}

void BlockRunner::expmv(const double* h_V, double* h_Y) {
  MEK_HIP(hipSetDevice(device_));
  const size_t bytes = sizeof(double) * static_cast<size_t>(n_) * kRhs;
  double *dv = nullptr, *dy = nullptr;
  MEK_HIP(hipMalloc(&dv, bytes));
  hipError_t e = hipMalloc(&dy, bytes);
  if (e == hipSuccess) e = hipMemcpy(dv, h_V, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(dv);
    if (dy) (void)hipFree(dy);
    MEK_HIP(e);
  }
  try {
    expmv_device(dv, dy);
    MEK_HIP(hipMemcpy(h_Y, dy, bytes, hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(dv);
    (void)hipFree(dy);
    throw;
  }
  (void)hipFree(dv);
  (void)hipFree(dy);
}

void BlockRunner::expmv_device(const double* d_V, double* d_Y) {
  MEK_HIP(hipSetDevice(device_));
  const double t0 = now_ms();
  const BlockBufs& b = *static_cast<BlockBufs*>(bufs_);
  stats_ = BlockStats();
  ev_used_ = 0;
  start_block(d_V, true, nullptr);
  if (autotune_ && tuned_spmm_grid_ == 0) autotune();
  const double blk = 8.0 * static_cast<double>(n_pad_) * kRhs;
  const size_t full = sizeof(double) * static_cast<size_t>(n_pad_) * kRhs;
  double t_now = 0.0;
  double dt;
  if (substeps_ > 0) {
    dt = t_ / substeps_;
  } else {
    // Expokit's first step from ||A||_inf (relative form: beta cancels)
    const int m = m_;  // polynomial degree of the block Krylov space
    const double fact = std::pow((m + 1) / std::exp(1.0), m + 1) * std::sqrt(2.0 * M_PI * (m + 1));
    dt = std::min(t_, (1.0 / anorm_) * std::pow((fact * tol_) / (4.0 * anorm_), 1.0 / m));
  }
  while (t_now < t_) {
    dt = std::min(dt, t_ - t_now);
    const int k = build_basis();
    const bool brk = h_flags_[0] != 0;
    int rejects = 0;
    double err = 0.0, tol_eff = 0.0;
    for (;;) {
      stats_.attempts++;
      exponential(dt, k);
      block_launch_residual(b, k, F_, F_ld_, stream_);
      MEK_HIP_LAUNCHED();
      MEK_HIP(hipMemcpyAsync(h_pinned_, b.scal, sizeof(double) * 3, hipMemcpyDeviceToHost, stream_));
      MEK_HIP(hipStreamSynchronize(stream_));
      const double fn = h_pinned_[0], beta = h_pinned_[2];
      err = brk ? 0.0 : h_pinned_[1];
      tol_eff = kDelta * (dt / t_) * tol_ * std::min(beta, fn);
      if (!std::isfinite(err)) throw Error("ERROR: non-finite block residual estimate");
      if (err <= tol_eff) break;
      stats_.rejected++;
      if (++rejects > kMaxRejects) throw Error("ERROR: The requested tolerance is too high.");
      dt = std::min(0.5 * dt, kGamma * dt * std::pow(tol_eff / err, 1.0 / m_));
    }
    // Y = Q_{0..k-1} (E R0); a breakdown flag must not stop the lift
    if (brk) MEK_HIP(hipMemsetAsync(d_flags_, 0, sizeof(int) * 3, stream_));
    block_launch_lift_coeffs(b, k, F_, F_ld_, stream_);
    timed_begin(kLift);
    block_launch_update(b, k, nullptr, d_W_, 0, stream_);
    timed_end(kLift, blk * (k + 1));
    MEK_HIP_LAUNCHED();
    t_now += dt;
    stats_.accepted++;
    stats_.last_residual = err;
    const double grow = (err > 0.0) ? kGamma * std::pow(tol_eff / err, 1.0 / m_) : 5.0;
    dt = dt * std::min(5.0, std::max(grow, 0.2));
    if (t_now < t_) {
      MEK_HIP(hipMemcpyAsync(d_P_, d_W_, full, hipMemcpyDeviceToDevice, stream_));
      MEK_HIP(hipMemsetAsync(d_flags_, 0, sizeof(int) * 3, stream_));
      timed_begin(kDots);
      block_launch_dots_partial(b, d_P_, 1, d_P_, 0, 0, stream_);
      timed_end(kDots, blk);
      block_launch_dots_fin(b, 1, 0, 0, stream_);
      block_launch_chol(b, -1, stream_);
      MEK_HIP_LAUNCHED();
    }
  }
  MEK_HIP(hipMemcpyAsync(d_Y, d_W_, sizeof(double) * static_cast<size_t>(n_) * kRhs, hipMemcpyDeviceToDevice,
                         stream_));
  MEK_HIP(hipStreamSynchronize(stream_));
  flush_timing();
  MEK_HIP(hipMemcpy(h_flags_, d_flags_, sizeof(int) * 4, hipMemcpyDeviceToHost));
  stats_.second_passes = h_flags_[3];
  stats_.converged = (t_now >= t_) ? 1 : 0;
  stats_.t_done = t_now;
  stats_.solve_ms = now_ms() - t0;
}

// Releases whatever is allocated (also after a constructor that failed part
// way); safe to call more than once.
void BlockRunner::destroy() {
  (void)hipSetDevice(device_);
  for (Ev& e : ev_) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  ev_.clear();
  auto release = [](auto*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  release(d_row_ptr_);
  release(d_col_idx_);
  release(d_vals_);
  release(d_P_);
  release(d_W_);
  release(d_small_);
  release(d_partials_);
  release(d_flags_);
  release(d_mat_);
  if (h_pinned_) (void)hipHostFree(h_pinned_);
  if (h_flags_) (void)hipHostFree(h_flags_);
  h_pinned_ = nullptr;
  h_flags_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  delete static_cast<BlockBufs*>(bufs_);
  bufs_ = nullptr;
}

}  // namespace mek
