/*
 * C ABI of the MI355X Arnoldi exp(t*A)*v library (libmek.so).
 *
 * The reference (kol-di/matrix-exp-krylov) has no FFI: its interface is the
 * C++ class surface of main.cu (CSRHost, DeviceContext, ArnoldiRunner) plus
 * the cuSPARSE/cuBLAS calls its hot loop makes.  Each entry point below names
 * the reference interface it stands for.  Plain C types only; no exit() and no
 * C++ exception crosses this boundary.
 *
 * Conventions
 *   - Every function returns MEK_OK (0) or a negative MEK_E_* status; the
 *     message of the last failure on the calling thread is mek_last_error().
 *   - Pointers named d_* are device pointers (hipMalloc'd, or from
 *     mek_malloc); everything else is host memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Kernel-level entry points are asynchronous on that stream.
 */
#ifndef MEK_H
#define MEK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MEK_OK 0
#define MEK_E_INVALID (-1)   /* bad argument                         */
#define MEK_E_LOAD (-2)      /* matrix file failed validation / I/O  */
#define MEK_E_HIP (-3)       /* HIP runtime error                    */
#define MEK_E_SOLVER (-4)    /* solver refused the problem           */
#define MEK_E_INTERNAL (-5)

/* Message of the last failed call on this thread ("" if none). */
const char* mek_last_error(void);
/* Library version string. */
const char* mek_version(void);

/* ---------------------------------------------------------------- device */

int mek_device_count(int* count);           /* cudaGetDeviceCount, main.cu:2146 */
int mek_device_synchronize(void);
int mek_malloc(void** d_ptr, size_t bytes);
int mek_free(void* d_ptr);
int mek_memcpy_htod(void* d_dst, const void* src, size_t bytes);
int mek_memcpy_dtoh(void* dst, const void* d_src, size_t bytes);
int mek_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);

/* ---------------------------------------------------------------- CSRHost */

typedef struct mek_csr mek_csr;

/* CSRHost built from caller arrays (copied).  reference: CSRHost fields,
 * main.cu:86-95. */
int mek_csr_create(int rows, int cols, long long nnz, const int* row_ptr, const int* col_idx,
                   const double* values, mek_csr** out);
/* reference: CSRHost::from_binary_csr, main.cu:265-523 (docs/binary_csr_v1.md) */
int mek_csr_load_bincsr(const char* path, mek_csr** out);
/* reference: CSRHost::from_matrix_market, main.cu:97-245 (also reads .mtx.gz) */
int mek_csr_load_mtx(const char* path, mek_csr** out);
/* Dimensions; value_dtype is 1 (f32) or 2 (f64) as stored in the source file. */
int mek_csr_info(const mek_csr* A, int* rows, int* cols, long long* nnz, int* value_dtype);
/* Copy the arrays out (row_ptr: rows+1, col_idx/values: nnz entries). */
int mek_csr_export(const mek_csr* A, int* row_ptr, int* col_idx, double* values);
/* How many values FP32 storage (values_f32) would change, and the largest
 * relative change among the finite ones (0 when FP32 storage is lossless).
 * *nonfinite (may be NULL) counts the values that are NaN or infinite or lie
 * beyond FP32's range; they are included in *count.  No reference
 * counterpart: the check the solvers make for values_f32 = 1. */
int mek_csr_f32_inexact(const mek_csr* A, long long* count, double* max_rel, long long* nonfinite);
/* The diagonals of A for the diagonal-copy SpMV: *count (0..16) distinct
 * column - row offsets in ascending order into offsets[16], or *count = 0
 * when A has more than 16 of them, a row whose columns are not strictly
 * ascending, or a NaN value.  Host only.  No reference counterpart: the test
 * the runner makes for spmv_format 0 and 3. */
int mek_csr_diagonals(const mek_csr* A, int* offsets, int* count);
/* The value of each of those diagonals into values[16] (offset order), when
 * every entry on a diagonal has the same bits (after rounding to FP32 when
 * f32 != 0), else *count = 0; the test the runner makes for spmv_format 0 and
 * 4.  Host only.  No reference counterpart. */
int mek_csr_diagonal_values(const mek_csr* A, int f32, double* values, int* count);
void mek_csr_destroy(mek_csr* A);

/* ---------------------------------------------------------------- solver */

/* reference: ArnoldiParams(m, t, tol, max_restarts, reorthogonalize),
 * main.cu:1221-1230 */
typedef struct {
    int m;
    double t;
    double tol;
    int max_restarts;
    int reorthogonalize;
} mek_arnoldi_params;

/* Additive options (all-zero except gs_mode = -1, verbose = 1 reproduces the
 * reference; see mek_solver_options_default). */
typedef struct {
    int step_control;   /* 0 reference controller, 1 Expokit heuristic,
                           2 Expokit with the local error relative to ||w|| */
    int gs_mode;        /* -1 follow reorthogonalize, 0 never, 1 always,
                           2 conditional (DGKS, ||w'|| < dgks_eta ||w||)      */
    double dgks_eta;
    int values_f32;     /* matrix values on the device: 0 FP64; 1 FP32 when
                           every value is exact in FP32, else FP64; 2 FP32
                           always (the caller accepts the perturbation)       */
    int verbose;        /* print the reference's per-column / [STEP] / METRIC */
    int timing;         /* HIP-event timing of every kernel launch            */
    int fuse_spmv_dot;  /* first-pass dot products inside the SpMV kernel     */
    int lanczos;        /* symmetric A: Lanczos three-term recurrence        */
    int autotune;       /* time SpMV / sweep launch shapes at create, keep best */
    int sstep;          /* s-step blocks of this many basis vectors (2..16;
                           0: column by column); needs m + 1 <= 64        */
    int sstep_first;    /* columns of each cycle built one by one before the
                           blocks (-1: sstep)                               */
    int spmv_format;    /* 0 diagonal layout for entries on <= 16 diagonals
                           with little padding (constant diagonals when each
                           holds one value), else sliced ELL when its
                           padding is small, else CSR row blocks; 1 CSR row
                           blocks; 2 sliced ELL; 3 diagonal copy with
                           per-entry values; 4 constant diagonals: one value
                           per diagonal and a presence mask per row          */
} mek_solver_options;

void mek_solver_options_default(mek_solver_options* opt);

typedef struct {
    int converged;      /* ArnoldiRunner::converged      */
    int restarts_done;  /* ArnoldiRunner::restarts_done  */
    double last_residual;
    double t_done;
    int attempts;
    int accepted;
    int rejected;
    int spmv_count;
    int extra_spmv;
    int reorth_columns;
    int aborted;
    double solve_ms;
    int sstep_blocks;     /* s-step blocks accepted in the last solve        */
    int sstep_second;     /* of them, with a second update or Cholesky pass  */
    int sstep_fallbacks;  /* cycles rebuilt column by column                 */
    int sstep_fused;      /* blocks done by the fused update alone (Cholesky
                             QR from the dots, no normalising sweep)        */
} mek_solve_stats;

typedef struct mek_runner mek_runner;

/* ArnoldiRunner(1, params, rows, cols, nnz) + init_all_devices on device
 * `device_id` with the whole matrix (one GPU).  reference: main.cu:1286-1409.
 * params->t must be finite and >= 0 (t = 0 gives y = v; the reference's
 * controller never runs for t <= 0 and leaves y unset). */
int mek_runner_create(const mek_csr* A, const mek_arnoldi_params* params,
                      const mek_solver_options* options, int device_id, mek_runner** out);
/* y = exp(t A) v, host vectors of n doubles.  reference:
 * ArnoldiRunner::compute_expmv, main.cu:1411-1571 */
int mek_runner_expmv(mek_runner* R, const double* v, double* y);
/* Same with device-resident v and y (n doubles each). */
int mek_runner_expmv_device(mek_runner* R, const double* d_v, double* d_y);
/* Change t (and the derived TimeSteppingParams) between solves; t finite
 * and >= 0. */
int mek_runner_set_t(mek_runner* R, double t);
/* 1 when the runner stores the matrix values in FP32 (values_f32 = 2, or 1
 * with every value exact in FP32), else 0. */
int mek_runner_values_f32(const mek_runner* R, int* used);
int mek_runner_stats(const mek_runner* R, mek_solve_stats* out);
/* Per-attempt [STEP] records of the last solve: t, dt, err (3 doubles) and
 * m_used, accepts, rejects (3 ints) per record; returns the count in *n. */
int mek_runner_steps(const mek_runner* R, double* tde, int* mar, int cap, int* n);
/* Event-timed kernel classes of the last solve (options.timing):
 * kind 0 SpMV(+dots), 1 update+dot, 2 update+norm (and the s-step blocks'
 * conditional passes, counted with no bytes), 3 dots, 4 finalisers, 5 expm,
 * 6 lift, 7 other, 8 s-step block dots, 9 s-step block update.  Summed ms,
 * launches and algorithmic bytes. */
int mek_runner_kernel_time(const mek_runner* R, int kind, double* ms, long long* launches,
                           double* bytes);
/* One Arnoldi decomposition from v: H ((m+1) x m, column-major).  Returns the
 * number of columns built (< m after a breakdown) in *k_used. */
int mek_runner_arnoldi(mek_runner* R, const double* v, double* H, int* k_used);
/* Launch shapes chosen by the autotuner (autotune option; zeros otherwise):
 * SpMV row-block nonzero cap and grid, sweep grid, and their timings (us). */
int mek_runner_tuning(const mek_runner* R, int* nnz_cap, int* spmv_grid, int* sweep_grid, double* spmv_us,
                      double* sweep_us);
/* Workgroups of the s-step block sweeps (0 when the block path is off) and,
 * with autotune, the timed dots + update pair of the chosen grid (us). */
int mek_runner_sstep_grid(const mek_runner* R, int* grid, double* us);
/* SpMV layout in use: *format 1 CSR row blocks, 2 sliced ELL, 3 diagonal
 * copy, 4 constant diagonals; for sliced ELL the padded entry count and the
 * widest row, for the diagonal layouts rows x diagonals and the diagonal
 * count (else 0). */
int mek_runner_spmv_format(const mek_runner* R, int* format, long long* entries, int* width);
/* The runner's stream (hipStream_t). */
int mek_runner_stream(const mek_runner* R, void** stream);
void mek_runner_destroy(mek_runner* R);

/* ---------------------------------------------------------------- batch */

/* Many independent problems y_p = exp(t_p A_p) v_p of uneven size, solved in
 * lockstep on one GPU (one block-diagonal CSR, per-problem reductions, the
 * step controller on the device, one hipGraph per cycle).  No reference
 * counterpart: the reference solves one problem per process.  m[p] in
 * [1, 62].  step_control 0: the reference controller of main.cu:1425-1518
 * (substeps[p] > 0 sets the first sub-step to t/substeps with no step cap;
 * 0 keeps max_restarts).  step_control 1: Expokit step control (first step
 * from Expokit's formula, or t/substeps[p] when substeps[p] > 0, e.g. from
 * the learned selector); 2: the same with the local error measured relative
 * to the norm of each step's start vector.  t[p] finite and >= 0 (t[p] = 0
 * gives y_p = v_p).  values_f32: 0 FP64 values; 1 FP32 when every value of
 * every problem is exact in FP32, else FP64; 2 FP32 always. */
typedef struct mek_batch mek_batch;
int mek_batch_create(int nprob, const mek_csr* const* A, const double* t, const int* m,
                     const double* tol, const int* max_restarts, const int* substeps,
                     int step_control, int values_f32, int gs_mode, double dgks_eta, int use_graph,
                     int autotune, int device_id, mek_batch** out);
/* SpMV grid the autotuner chose at the first solve (0 without autotune). */
int mek_batch_tuning(const mek_batch* B, int* spmv_grid);
/* 1 when the batch stores the values in FP32 (values_f32 = 2, or 1 with
 * every value of every problem exact in FP32), else 0. */
int mek_batch_values_f32(const mek_batch* B, int* used);
/* Median microseconds of one batched SpMV over every problem's start vector
 * (all problems active, as in the first Arnoldi column of a solve), over
 * reps launches timed with HIP events on the batch's stream.  Call after
 * mek_batch_upload_v; leaves the batch at its start state.  No reference
 * counterpart (measurement of the batched form). */
int mek_batch_time_spmv(mek_batch* B, int reps, double* us);
/* v[p]: host arrays of each problem's rows, copied to the device. */
int mek_batch_upload_v(mek_batch* B, const double* const* v);
/* Solve every problem from the uploaded v; y stays in device memory. */
int mek_batch_solve(mek_batch* B);
int mek_batch_download_y(mek_batch* B, double* const* y);
/* Per-problem statistics (converged, accepted, attempts, spmv_count =
 * columns built, t_done, last_residual); solve_ms is the whole batch. */
int mek_batch_stats(const mek_batch* B, int p, mek_solve_stats* out);
int mek_batch_cycles(const mek_batch* B, int* cycles);
void mek_batch_destroy(mek_batch* B);

/* ------------------------------------------------------ learned selector */

/* No reference counterpart (the reference fixes m and the first sub-step by
 * hand, main.cu:1292-1305).  Features: 16 floats per problem computed on the
 * device (Gershgorin and 1-/inf-norm estimates, nnz-per-row statistics, t,
 * tol).  MLP: 16 -> 64 -> 64 -> 2, FP32, GELU; forward, backward and Adam on
 * the device. */
int mek_features(const mek_csr* A, double t, double tol, float* out16, int device_id);
/* Features of every problem of a batch (out: nprob x 16). */
int mek_batch_features(mek_batch* B, float* out);
typedef struct mek_mlp mek_mlp;
int mek_mlp_param_count(void);
int mek_mlp_create(unsigned seed, int device_id, mek_mlp** out);
int mek_mlp_set_params(mek_mlp* N, const float* params);
int mek_mlp_get_params(const mek_mlp* N, float* params);
/* Y (B x 2) = MLP(X (B x 16)) */
int mek_mlp_forward(mek_mlp* N, const float* X, int B, float* Y);
/* MSE loss against T (B x 2) and its gradient (param_count floats, may be NULL). */
int mek_mlp_loss_grad(mek_mlp* N, const float* X, const float* T, int B, float* loss, float* grads);
/* Mini-batch Adam on the device; *loss = MSE over all N samples afterwards. */
int mek_mlp_train(mek_mlp* N, const float* X, const float* T, int n, int batch, int epochs, float lr,
                  unsigned seed, float* loss);
void mek_mlp_destroy(mek_mlp* N);

/* ---------------------------------------------------------- kernel level */

typedef struct mek_spmv_plan mek_spmv_plan;

/* Adaptive CSR SpMV over device arrays; h_row_ptr (host copy of row_ptr) is
 * used once to build the row blocks.  Replaces cusparseCreateCsr +
 * cusparseSpMV_bufferSize (main.cu:987-1053).  vals_f32: values are float. */
int mek_spmv_plan_create(int n, long long nnz, const int* d_row_ptr, const int* d_col_idx,
                         const void* d_vals, int vals_f32, const int* h_row_ptr,
                         mek_spmv_plan** out);
/* d_y = A * d_x (beta = 0).  Replaces cusparseSpMV, main.cu:1743-1745. */
int mek_spmv_execute(mek_spmv_plan* P, const double* d_x, double* d_y, void* stream);
/* Switch the plan to a sliced-ELL copy of its matrix (64-row slices, one row
 * per lane, each slice padded to its longest row), built on the device from
 * the plan's CSR arrays; h_row_ptr as at create.  The same y bits as the CSR
 * row blocks.  *entries and *width (may be null) receive the padded entry
 * count and the widest row.  No reference counterpart (cuSPARSE chooses its
 * own algorithm). */
int mek_spmv_plan_use_sell(mek_spmv_plan* P, const int* h_row_ptr, long long* entries, int* width);
/* Switch the plan to a diagonal copy of its matrix (diagonal d holds the
 * entries (r, r + offset_d); no column indices are read), built on the device
 * from the plan's CSR arrays.  The matrix must have at most 16 distinct
 * column - row offsets, strictly ascending columns within each row and no NaN
 * value (else MEK_E_INVALID); h_row_ptr, h_col_idx, h_vals as at create.  The
 * same y bits as the CSR row blocks.  *n_diags (may be null) receives the
 * diagonal count.  No reference counterpart (cuSPARSE chooses its own
 * algorithm). */
int mek_spmv_plan_use_dia(mek_spmv_plan* P, const int* h_row_ptr, const int* h_col_idx, const double* h_vals,
                          int* n_diags);
/* The same with the constant-diagonal form (one value per diagonal, a
 * presence mask per row; no per-entry values are read): refused when a
 * diagonal holds two values.  Same y bits.  No reference counterpart. */
int mek_spmv_plan_use_dia_const(mek_spmv_plan* P, const int* h_row_ptr, const int* h_col_idx,
                                const double* h_vals, int* n_diags);
void mek_spmv_plan_destroy(mek_spmv_plan* P);

/* d_E = exp(d_A) for a dense k x k column-major matrix on the device (one
 * workgroup, FP64 Pade scaling and squaring).  Replaces the host Eigen
 * A.exp() of main.cu:33-36.  Synchronous. */
int mek_expm_dense(int k, const double* d_A, double* d_E);
/* The same exponential, allocating nothing: asynchronous on `stream`, with
 * its working memory (scalars, the first column and, for k > 62, six k x k
 * matrices) in a static device array of the library, 770 KiB per GPU, held
 * until the process ends.  k in [1, 128]; larger k is refused.  Calls that
 * overlap on one GPU share that array and race.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc). */
int mek_expm_dense_noalloc(int k, const double* d_A, double* d_E, void* stream);

/* ------------------------------------------------------- block of RHS */

/* Y = exp(t A) V for a block of 32 right-hand sides on one shared A (V, Y:
 * n x 32, row-major) by block Arnoldi: SpMM, block classical Gram-Schmidt
 * with a second pass on FP64 MFMA, Cholesky QR inside each block, and a
 * Taylor / scaling-and-squaring exponential of the block Hessenberg matrix
 * on an FP64 MFMA GEMM.  No reference counterpart: the reference solves one
 * vector per process (main.cu:1411-1571).  m: block steps (basis of 32 m
 * vectors, 32 (m + 1) < n, m <= 48); substeps > 0 fixes the first sub-step
 * to t/substeps, 0 uses Expokit's first-step formula; each sub-step passes
 * ||H(k,k-1) E_k R0||_F <= 1.2 (dt/t) tol min(||V||_F, ||Y||_F).  gs_mode:
 * 1 second block Gram-Schmidt pass always, 2 only when a column of the new
 * block keeps less than 0.1 of its norm in the first pass, 0 never. */
typedef struct mek_block mek_block;
typedef struct {
    int converged;
    int accepted;
    int attempts;
    int rejected;
    int block_steps;
    int squarings;
    int second_passes;  /* block steps that took the second Gram-Schmidt pass */
    int spmm_grid;      /* autotuned SpMM / update grids (0 without autotune) */
    int update_grid;
    double t_done;
    double last_residual;
    double solve_ms;
    double kernel_ms[5];     /* spmm, dots, update, expm, lift (timing on) */
    double kernel_bytes[5];  /* algorithmic HBM bytes of the same          */
    long long launches[5];
} mek_block_stats;
int mek_block_create(const mek_csr* A, int m, double t, double tol, int substeps, int gs_mode,
                     int device_id, int timing, int autotune, mek_block** out);
int mek_block_expmv(mek_block* B, const double* V, double* Y);
int mek_block_expmv_device(mek_block* B, const double* d_V, double* d_Y);
int mek_block_stats_get(const mek_block* B, mek_block_stats* out);
void mek_block_destroy(mek_block* B);

/* Y = A X for 32 columns (X, Y: n x 32 row-major host arrays), each column
 * summed in CSR storage order as a sequential loop does. */
int mek_spmm(const mek_csr* A, const double* X, double* Y, int device_id);
/* C = A B for N x N column-major host matrices (N a multiple of 64) on the
 * FP64 MFMA GEMM of the block exponential. */
int mek_gemm_dense(int N, const double* A, const double* B, double* C, int device_id);

#ifdef __cplusplus
}
#endif

#endif /* MEK_H */
