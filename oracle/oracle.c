/*
 * CPU oracle for the Arnoldi exp(t*A)*v path -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for what each function restates and what pins it.
 *
 * Matrices are dense column-major: M[col * k + row].
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ SpMV */

/* main_cpu.cpp:388-397: one accumulator per row, entries in storage order. */
void oracle_spmv(int n, const int* row_ptr, const int* col_idx,
                 const double* vals, const double* x, double* y) {
    for (int r = 0; r < n; ++r) {
        double acc = 0.0;
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) acc += vals[k] * x[col_idx[k]];
        y[r] = acc;
    }
}

/* ------------------------------------------------------------------ expm */

/* Higham 2005, Table 2.3: largest ||A||_1 for which the [q/q] Pade
 * approximant meets unit roundoff 2^-53, q = 3, 5, 7, 9, 13. */
static const double kTheta[5] = {1.495585217958292e-2, 2.539398330063230e-1,
                                 9.504178996162932e-1, 2.097847961257068e0,
                                 5.371920351148152e0};
static const int kDeg[5] = {3, 5, 7, 9, 13};

void oracle_pade_coeffs(int q, double* b) {
    /* b_j = (2q-j)!/(j!(q-j)!): b_q = 1, b_{j-1} = b_j (2q-j+1) j / (q-j+1),
     * exact in 64-bit integers for q <= 13. */
    unsigned long long bj = 1ULL;
    b[q] = 1.0;
    for (int j = q; j >= 1; --j) {
        bj = bj * (unsigned long long)(2 * q - j + 1) * (unsigned long long)j /
             (unsigned long long)(q - j + 1);
        b[j - 1] = (double)bj;
    }
}

static double norm1(int k, const double* A) {
    double best = 0.0;
    for (int c = 0; c < k; ++c) {
        double s = 0.0;
        for (int r = 0; r < k; ++r) s += fabs(A[c * k + r]);
        if (s > best || isnan(s)) best = s;
    }
    return best;
}

/* C = A * B (all k x k column-major, C distinct from A and B). */
static void matmul(int k, const double* A, const double* B, double* C) {
    for (int c = 0; c < k; ++c) {
        for (int r = 0; r < k; ++r) {
            double acc = 0.0;
            for (int l = 0; l < k; ++l) acc += A[l * k + r] * B[c * k + l];
            C[c * k + r] = acc;
        }
    }
}

/* Solve Q X = P in place (P <- X) by LU with partial pivoting (row of
 * largest |entry| in the pivot column; first such row on ties). */
static void lu_solve(int k, double* Q, double* P) {
    for (int p = 0; p < k; ++p) {
        int piv = p;
        double best = fabs(Q[p * k + p]);
        for (int r = p + 1; r < k; ++r) {
            double a = fabs(Q[p * k + r]);
            if (a > best) { best = a; piv = r; }
        }
        if (piv != p) {
            for (int c = 0; c < k; ++c) {
                double t = Q[c * k + p]; Q[c * k + p] = Q[c * k + piv]; Q[c * k + piv] = t;
                t = P[c * k + p]; P[c * k + p] = P[c * k + piv]; P[c * k + piv] = t;
            }
        }
        const double d = Q[p * k + p];
        for (int r = p + 1; r < k; ++r) {
            const double l = Q[p * k + r] / d;
            Q[p * k + r] = l;
            for (int c = p + 1; c < k; ++c) Q[c * k + r] -= l * Q[c * k + p];
            for (int c = 0; c < k; ++c) P[c * k + r] -= l * P[c * k + p];
        }
    }
    for (int c = 0; c < k; ++c) {
        for (int r = k - 1; r >= 0; --r) {
            double s = P[c * k + r];
            for (int l = r + 1; l < k; ++l) s -= Q[l * k + r] * P[c * k + l];
            P[c * k + r] = s / Q[r * k + r];
        }
    }
}

int oracle_expm(int k, const double* A_in, double* E) {
    const size_t kk = (size_t)k * (size_t)k;
    for (size_t i = 0; i < kk; ++i) {
        if (!isfinite(A_in[i])) {
            for (size_t j = 0; j < kk; ++j) E[j] = NAN;
            return -1;
        }
    }
    double* A = (double*)malloc(kk * sizeof(double));
    double* A2 = (double*)malloc(kk * sizeof(double));
    double* A4 = (double*)malloc(kk * sizeof(double));
    double* A6 = (double*)malloc(kk * sizeof(double));
    double* A8 = (double*)malloc(kk * sizeof(double));
    double* U = (double*)malloc(kk * sizeof(double));
    double* V = (double*)malloc(kk * sizeof(double));
    double* T = (double*)malloc(kk * sizeof(double));
    memcpy(A, A_in, kk * sizeof(double));

    const double nrm = norm1(k, A);
    int q = 13, s = 0;
    for (int i = 0; i < 4; ++i) {
        if (nrm <= kTheta[i]) { q = kDeg[i]; break; }
    }
    if (q == 13 && nrm > kTheta[4]) {
        /* s = ceil(log2(nrm / theta_13)), computed exactly from frexp */
        int e;
        const double f = frexp(nrm / kTheta[4], &e);
        s = (f == 0.5) ? e - 1 : e;
        if (s < 0) s = 0;
        for (size_t i = 0; i < kk; ++i) A[i] = ldexp(A[i], -s);
    }
    double b[14];
    oracle_pade_coeffs(q, b);

    matmul(k, A, A, A2);
    if (q == 13) {
        matmul(k, A2, A2, A4);
        matmul(k, A4, A2, A6);
        /* U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I] */
        for (size_t i = 0; i < kk; ++i) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
        matmul(k, A6, T, V);
        for (size_t i = 0; i < kk; ++i) V[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
        for (int d = 0; d < k; ++d) V[d * k + d] += b[1];
        matmul(k, A, V, U);
        /* V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I */
        for (size_t i = 0; i < kk; ++i) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
        matmul(k, A6, T, V);
        for (size_t i = 0; i < kk; ++i) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
        for (int d = 0; d < k; ++d) V[d * k + d] += b[0];
    } else {
        /* powers A^2, A^4, A^6, A^8 as the degree needs them */
        const double* P[5] = {NULL, A2, NULL, NULL, NULL};
        if (q >= 5) { matmul(k, A2, A2, A4); P[2] = A4; }
        if (q >= 7) { matmul(k, A4, A2, A6); P[3] = A6; }
        if (q >= 9) { matmul(k, A6, A2, A8); P[4] = A8; }
        /* T = sum_{j odd} b_j A^{j-1}, U = A T;  V = sum_{j even} b_j A^j */
        for (size_t i = 0; i < kk; ++i) { T[i] = 0.0; V[i] = 0.0; }
        for (int d = 0; d < k; ++d) { T[d * k + d] = b[1]; V[d * k + d] = b[0]; }
        for (int p = 1; 2 * p <= q; ++p) {
            const double* Ap = P[p];
            for (size_t i = 0; i < kk; ++i) {
                T[i] += b[2 * p + 1] * Ap[i];
                V[i] += b[2 * p] * Ap[i];
            }
        }
        matmul(k, A, T, U);
    }
    /* (V - U) X = (V + U) */
    for (size_t i = 0; i < kk; ++i) {
        const double u = U[i], v = V[i];
        T[i] = v - u;
        E[i] = v + u;
    }
    lu_solve(k, T, E);
    for (int i = 0; i < s; ++i) {
        matmul(k, E, E, T);
        memcpy(E, T, kk * sizeof(double));
    }
    free(A); free(A2); free(A4); free(A6); free(A8); free(U); free(V); free(T);
    return s;
}

/* ------------------------------------------------------- Arnoldi (MGS2) */

typedef struct {
    int n;
    const int* row_ptr;
    const int* col_idx;
    const double* vals;
    int m;          /* allocated Krylov dimension */
    double* V;      /* n x (m+1), column-major: V[col*n + row] (main_cpu.cpp:446) */
    double* w;
    double* y;
    double* H;      /* (m+1) x m, column-major, leading dimension m+1 */
    double v_norm;
} arnoldi_state;

#define HIDX(st, i, j) ((size_t)(j) * (size_t)((st)->m + 1) + (size_t)(i))

/* main_cpu.cpp:473-484 (guarded scale, as on the CPU path). */
static void init_q1(arnoldi_state* st, const double* v) {
    double norm_sq = 0.0;
    for (int i = 0; i < st->n; ++i) norm_sq += v[i] * v[i];
    st->v_norm = sqrt(norm_sq);
    const double scale = (st->v_norm > 1e-14) ? (1.0 / st->v_norm) : 1.0;
    for (int i = 0; i < st->n; ++i) st->V[i] = v[i] * scale;
}

/* main_cpu.cpp:486-491 */
static void spmv_qj(arnoldi_state* st, int j) {
    const double* qj = st->V + (size_t)j * st->n;
    oracle_spmv(st->n, st->row_ptr, st->col_idx, st->vals, qj, st->w);
}

/* main_cpu.cpp:493-520: MGS, then (if reorthogonalize) a second MGS pass whose
 * coefficients are added to H(i,j). */
static void orthogonalize_mgs(arnoldi_state* st, int j, int reorth) {
    const int n = st->n;
    for (int i = 0; i <= j; ++i) {
        const double* qi = st->V + (size_t)i * n;
        double h = 0.0;
        for (int k = 0; k < n; ++k) h += st->w[k] * qi[k];
        st->H[HIDX(st, i, j)] = h;
        for (int k = 0; k < n; ++k) st->w[k] -= h * qi[k];
    }
    if (reorth) {
        for (int i = 0; i <= j; ++i) {
            const double* qi = st->V + (size_t)i * n;
            double c = 0.0;
            for (int k = 0; k < n; ++k) c += st->w[k] * qi[k];
            st->H[HIDX(st, i, j)] += c;
            for (int k = 0; k < n; ++k) st->w[k] -= c * qi[k];
        }
    }
}

/* main_cpu.cpp:522-546: breakdown when the norm is < 1e-14 or not finite. */
static int normalize_new_vector(arnoldi_state* st, int j) {
    const int n = st->n;
    double norm_sq = 0.0;
    for (int i = 0; i < n; ++i) norm_sq += st->w[i] * st->w[i];
    const double nrm = sqrt(norm_sq);
    if (nrm < 1e-14 || isnan(nrm) || isinf(nrm)) {
        st->H[HIDX(st, j + 1, j)] = 0.0;
        return 0;
    }
    const double scale = 1.0 / nrm;
    double* q = st->V + (size_t)(j + 1) * n;
    for (int i = 0; i < n; ++i) q[i] = st->w[i] * scale;
    st->H[HIDX(st, j + 1, j)] = nrm;
    return 1;
}

/* exp(dt * H[0:k,0:k]) into E (k x k). */
static void small_expm(const arnoldi_state* st, double dt, int k, double* E) {
    double* Hs = (double*)malloc((size_t)k * k * sizeof(double));
    for (int c = 0; c < k; ++c)
        for (int r = 0; r < k; ++r) Hs[c * k + r] = dt * st->H[HIDX(st, r, c)];
    oracle_expm(k, Hs, E);
    free(Hs);
}

/* main_cpu.cpp:548-572 (lift) and :574-587 (residual).  The reference calls
 * the expm twice with identical inputs; one call gives the same values. */
static double expm_lift_residual(arnoldi_state* st, double dt, int k) {
    if (k <= 0) return 0.0;
    const int n = st->n;
    double* E = (double*)malloc((size_t)k * k * sizeof(double));
    small_expm(st, dt, k, E);
    for (int i = 0; i < n; ++i) st->y[i] = 0.0;
    for (int i = 0; i <= k; ++i) {
        const double wh = (i < k) ? E[i] : 0.0;  /* first column, zero-padded */
        if (fabs(wh) > 1e-15) {
            const double alpha = st->v_norm * wh;
            const double* Vi = st->V + (size_t)i * n;
            for (int r = 0; r < n; ++r) st->y[r] += alpha * Vi[r];
        }
    }
    const double ek = E[k - 1];
    const double hk = st->H[HIDX(st, k, k - 1)];
    free(E);
    return fabs(hk * ek) * st->v_norm;
}

/* main_cpu.cpp:589-600 */
static void restart_from_y(arnoldi_state* st) {
    double norm_sq = 0.0;
    for (int i = 0; i < st->n; ++i) norm_sq += st->y[i] * st->y[i];
    st->v_norm = sqrt(norm_sq);
    const double scale = (st->v_norm > 1e-14) ? (1.0 / st->v_norm) : 1.0;
    for (int i = 0; i < st->n; ++i) st->V[i] = st->y[i] * scale;
}

int oracle_expmv(int n, const int* row_ptr, const int* col_idx,
                 const double* vals, const double* v, double t, double tol,
                 int m, int max_restarts, int reorthogonalize, double* y_out,
                 oracle_stats* stats, oracle_step* steps, int max_steps_log) {
    arnoldi_state st;
    st.n = n; st.row_ptr = row_ptr; st.col_idx = col_idx; st.vals = vals; st.m = m;
    st.V = (double*)calloc((size_t)n * (size_t)(m + 1), sizeof(double));
    st.w = (double*)calloc((size_t)n, sizeof(double));
    st.y = (double*)calloc((size_t)n, sizeof(double));
    st.H = (double*)calloc((size_t)(m + 1) * (size_t)m, sizeof(double));

    /* TimeSteppingParams derivation, main_cpu.cpp:454-466 */
    const double dt_max = t;
    const double dt_min = fmax(1e-12, t * 1e-6);
    const double denom = (max_restarts > 0) ? max_restarts : 1;
    const double dt_init = t / denom;
    const double tol_step = tol;
    const int max_steps = (max_restarts > 0) ? max_restarts : 1000;
    const int max_rejects = 20;
    const double safety = 0.8, grow = 1.5, shrink = 0.5;

    oracle_stats s;
    memset(&s, 0, sizeof(s));

    init_q1(&st, v);
    const double total_t = t;
    double dt = fmin(fmax(dt_init, dt_min), dt_max);
    double t_done = 0.0;
    int accepted = 0;
    int converged = 0;

    /* main_cpu.cpp:613-685 */
    while (t_done < total_t && accepted < max_steps) {
        dt = fmin(dt, total_t - t_done);
        int rejects = 0;
        for (;;) {
            memset(st.H, 0, (size_t)(m + 1) * m * sizeof(double));
            int k_used = 0;
            int breakdown = 0;
            for (int j = 0; j < m; ++j) {
                spmv_qj(&st, j);
                orthogonalize_mgs(&st, j, reorthogonalize);
                const int ok = normalize_new_vector(&st, j);
                k_used = j + 1;
                s.spmv_count++;
                if (!ok) { breakdown = 1; break; }
            }
            double err = expm_lift_residual(&st, dt, k_used);
            if (breakdown) err = 0.0;
            if (steps && s.attempts < max_steps_log) {
                oracle_step* r = &steps[s.attempts];
                r->t = t_done; r->dt = dt; r->err = err; r->m_used = k_used;
                r->accepts = accepted; r->rejects = rejects;
            }
            s.attempts++;
            if (err <= tol_step || breakdown) {
                t_done += dt;
                accepted++;
                s.last_residual = err;
                if (t_done >= total_t) { converged = 1; break; }
                restart_from_y(&st);
                if (err < 0.1 * tol_step) dt = fmin(dt * grow * safety, dt_max);
                break;
            }
            rejects++;
            if (rejects >= max_rejects) {
                converged = 0;
                s.last_residual = err;
                s.aborted = 1;
                t_done = total_t;
                break;
            }
            dt = fmax(dt * shrink * safety, dt_min);
        }
    }
    s.converged = converged;
    s.restarts_done = accepted;
    s.t_done = t_done;
    memcpy(y_out, st.y, (size_t)n * sizeof(double));
    if (stats) *stats = s;
    free(st.V); free(st.w); free(st.y); free(st.H);
    return 0;
}

int oracle_arnoldi(int n, const int* row_ptr, const int* col_idx, const double* vals,
                   const double* v, int m, int reorthogonalize, double* H, double* V) {
    arnoldi_state st;
    st.n = n; st.row_ptr = row_ptr; st.col_idx = col_idx; st.vals = vals; st.m = m;
    st.V = (double*)calloc((size_t)n * (size_t)(m + 1), sizeof(double));
    st.w = (double*)calloc((size_t)n, sizeof(double));
    st.y = NULL;
    st.H = H;
    memset(H, 0, (size_t)(m + 1) * (size_t)m * sizeof(double));
    init_q1(&st, v);
    int k_used = 0;
    for (int j = 0; j < m; ++j) {
        spmv_qj(&st, j);
        orthogonalize_mgs(&st, j, reorthogonalize);
        const int ok = normalize_new_vector(&st, j);
        k_used = j + 1;
        if (!ok) break;
    }
    if (V) memcpy(V, st.V, (size_t)n * (size_t)(m + 1) * sizeof(double));
    free(st.V); free(st.w);
    return k_used;
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

int oracle_arnoldi_columns(int n, const int* row_ptr, const int* col_idx,
                           const double* vals, const double* v, int ncols,
                           double* col_ms) {
    arnoldi_state st;
    st.n = n; st.row_ptr = row_ptr; st.col_idx = col_idx; st.vals = vals; st.m = ncols;
    st.V = (double*)calloc((size_t)n * (size_t)(ncols + 1), sizeof(double));
    st.w = (double*)calloc((size_t)n, sizeof(double));
    st.y = NULL;
    st.H = (double*)calloc((size_t)(ncols + 1) * (size_t)ncols, sizeof(double));
    init_q1(&st, v);
    int built = 0;
    for (int j = 0; j < ncols; ++j) {
        const double t0 = now_ms();
        spmv_qj(&st, j);
        orthogonalize_mgs(&st, j, 1);
        const int ok = normalize_new_vector(&st, j);
        col_ms[j] = now_ms() - t0;
        built = j + 1;
        if (!ok) break;
    }
    free(st.V); free(st.w); free(st.H);
    return built;
}
