/*
 * CPU oracle for the Arnoldi exp(t*A)*v path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference CPU solver `matrix_exp_cpu`
 * (kol-di/matrix-exp-krylov, main_cpu.cpp:388-700) plus its own FP64 Pade
 * scaling-and-squaring matrix exponential (Higham 2005, SIAM J. Matrix Anal.
 * Appl. 26(4):1179-1193, Algorithm 2.3), which stands in for Eigen's
 * MatrixFunctions `A.exp()` (main_cpu.cpp:19-22; Eigen 3.4.0 is not
 * available here, so the reference itself cannot be built).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  The product never links it.
 *
 * Pinning (what checks each function):
 *   oracle_spmv       dense brute force, A*1 == row sums, exact integer
 *                     stencils, Laplacian sine eigenmodes (tests/test_oracle.py)
 *   oracle_expm       scipy.linalg.expm golden fixtures (tests/golden/),
 *                     closed forms: diagonal, 2x2 Jordan block, rotation,
 *                     nilpotent shift
 *   oracle_expmv      closed forms: built-in diagonal CLI matrix, eigenvector
 *                     start (happy breakdown), 2-D Laplacian via DST-I;
 *                     scipy.sparse.linalg.expm_multiply fixtures on random
 *                     Gershgorin-negative matrices
 *   controller trace  (dt sequence, accept/reject) -- closed-form traces
 *                     with m = 1 on 2 x 2 diagonal matrices (err from h11 and
 *                     h21 in closed form; the step rules restated a second
 *                     time from main_cpu.cpp:602-680) cover the grow,
 *                     shrink-and-cap and abort branches; a happy-breakdown
 *                     trace (m = 3) the pure growth branch; for m > 1 without
 *                     breakdown, a dense numpy/scipy restatement sharing no
 *                     code with this file (tests/independent.py).
 */
#ifndef MEK_ORACLE_H
#define MEK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = A*x for CSR A (n rows); main_cpu.cpp:388-397. */
void oracle_spmv(int n, const int* row_ptr, const int* col_idx,
                 const double* vals, const double* x, double* y);

/* E = exp(A) for a dense k x k column-major matrix.  Returns the number of
 * squarings used, or -1 if A holds a non-finite entry (E is then NaN). */
int oracle_expm(int k, const double* A, double* E);

/* Pade coefficients b_0..b_q of the [q/q] approximant of exp, as exact
 * integers b_j = (2q-j)! / (j! (q-j)!) converted to double. */
void oracle_pade_coeffs(int q, double* b);

typedef struct {
    int converged;      /* ArnoldiRunnerCPU::converged          */
    int restarts_done;  /* accepted time steps                   */
    double last_residual;
    double t_done;
    int attempts;       /* step attempts (accepted + rejected)   */
    int spmv_count;     /* Arnoldi columns built (SpMV calls)    */
    int aborted;        /* 1 if the 20-reject abort fired        */
} oracle_stats;

/* One record per step attempt, as printed by the reference's [STEP] line
 * (main_cpu.cpp:640-643). */
typedef struct {
    double t;
    double dt;
    double err;
    int m_used;
    int accepts;
    int rejects;
} oracle_step;

/* y = exp(t*A)*v by the reference's adaptive time-stepping Arnoldi
 * (ArnoldiRunnerCPU::compute_expmv, main_cpu.cpp:602-700) with ArnoldiParams
 * (m, t, tol, max_restarts, reorthogonalize).  `steps` (may be NULL) receives
 * up to `max_steps_log` attempt records; the total is stats->attempts.
 * Returns 0. */
int oracle_expmv(int n, const int* row_ptr, const int* col_idx,
                 const double* vals, const double* v, double t, double tol,
                 int m, int max_restarts, int reorthogonalize, double* y,
                 oracle_stats* stats, oracle_step* steps, int max_steps_log);

/* One Arnoldi decomposition from v (init_q1, then m columns of SpMV, MGS
 * (+ second pass if reorthogonalize), normalise; main_cpu.cpp:473-546).
 * H receives (m+1) x m column-major; V, when not NULL, the basis as n x (m+1)
 * column-major (columns past k_used stay zero).  Returns the columns built
 * (k_used). */
int oracle_arnoldi(int n, const int* row_ptr, const int* col_idx, const double* vals,
                   const double* v, int m, int reorthogonalize, double* H, double* V);

/* Bounded timing sample of the reference's Arnoldi column loop: init_q1 then
 * `ncols` columns of (SpMV, two MGS passes, normalise) from v.  col_ms[j]
 * receives the wall time of column j in milliseconds.  Returns the number of
 * columns built (smaller than ncols on breakdown). */
int oracle_arnoldi_columns(int n, const int* row_ptr, const int* col_idx,
                           const double* vals, const double* v, int ncols,
                           double* col_ms);

#ifdef __cplusplus
}
#endif

#endif
