"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or
index, transposed operand) fails it: SciPy golden fixtures, closed forms,
invariants and brute force on tiny inputs.  The reference holds no fixtures
of its own for this path.
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from closed_forms import laplacian_2d_eigvec, laplacian_2d_expmv, rel_err
from conftest import gen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------ SpMV

def test_spmv_brute_force(oracle):
    rng = np.random.default_rng(0)
    D = rng.standard_normal((13, 13)) * (rng.random((13, 13)) < 0.3)
    A = sp.csr_matrix(D)
    x = rng.standard_normal(13)
    y = oracle.spmv(A.indptr, A.indices, A.data, x)
    np.testing.assert_allclose(y, D @ x, rtol=1e-14, atol=1e-14)


def test_spmv_row_sums_and_integer_stencil(oracle):
    rp, ci, va = gen.laplacian_2d(32)
    ones = np.ones(32 * 32)
    A = sp.csr_matrix((va, ci, rp))
    np.testing.assert_array_equal(oracle.spmv(rp, ci, va, ones), A.sum(1).A1)
    xi = np.random.default_rng(1).integers(-1000, 1000, 32 * 32).astype(float)
    np.testing.assert_array_equal(oracle.spmv(rp, ci, va, xi), A @ xi)  # exact in FP64


def test_spmv_laplacian_eigenmode(oracle):
    N = 20
    rp, ci, va = gen.laplacian_2d(N)
    u, lam = laplacian_2d_eigvec(N, 2, 7)
    np.testing.assert_allclose(oracle.spmv(rp, ci, va, u), lam * u, atol=1e-13)


# ------------------------------------------------------------------ expm

def test_pade_coefficients_known_values(oracle):
    b = oracle.pade_coeffs(13)
    assert b[13] == 1 and b[12] == 182 and b[11] == 16380 and b[0] == 64764752532480000
    assert oracle.pade_coeffs(3).tolist() == [120, 60, 12, 1]


def test_expm_scipy_fixtures(oracle):
    z = np.load(os.path.join(GOLDEN, "expm_cases.npz"))
    for k, H, E in zip(z["k"], z["H"], z["E"]):
        H, E = H[:k, :k], E[:k, :k]
        got, _ = oracle.expm(H)
        assert rel_err(got, E) < 1e-12 * max(1.0, np.abs(H).sum(0).max()), k
        if k > 2:  # a transposed operand would fail the same check
            bad, _ = oracle.expm(H.T)
            assert rel_err(bad, E) > 1e-6


def test_expm_against_torch_matrix_exp(oracle):
    """A second independent library (torch.linalg.matrix_exp, FP64 on the
    CPU) on upper Hessenberg H of the sizes and norms the solver produces."""
    import torch

    rng = np.random.default_rng(17)
    for k, scale in ((3, 1e-3), (8, 0.5), (20, 5.0), (40, 40.0), (61, 300.0)):
        H = np.triu(rng.standard_normal((k, k)), -1)
        H *= scale / np.abs(H).sum(0).max()
        ref = torch.linalg.matrix_exp(torch.from_numpy(H)).numpy()
        got, _ = oracle.expm(H)
        assert rel_err(got, ref) < 1e-12 * max(1.0, scale), (k, scale)


def test_expm_closed_forms(oracle):
    a, t = -0.7, 3.0
    E, _ = oracle.expm(np.array([[a * t, t], [0.0, a * t]]))
    np.testing.assert_allclose(E, np.exp(a * t) * np.array([[1, t], [0, 1]]), rtol=1e-14)
    w = 2.5
    E, _ = oracle.expm(np.array([[0.0, -w], [w, 0.0]]))
    np.testing.assert_allclose(E, [[np.cos(w), -np.sin(w)], [np.sin(w), np.cos(w)]], atol=1e-14)
    N = np.diag(np.ones(5), 1) * 1.5
    ref = sum(np.linalg.matrix_power(N, p) / math.factorial(p) for p in range(6))
    np.testing.assert_allclose(oracle.expm(N)[0], ref, rtol=1e-14, atol=1e-14)
    d = np.linspace(-20, 3, 9)
    E, s = oracle.expm(np.diag(d))
    np.testing.assert_allclose(E, np.diag(np.exp(d)), rtol=1e-13)
    assert s == 2  # ||A||_1 = 20 > theta_13: ceil(log2(20 / 5.3719...)) = 2
    E, s = oracle.expm(np.array([[1e-4]]))
    assert s == 0 and abs(E[0, 0] - math.exp(1e-4)) <= 2.3e-16


def test_expm_scaling_count(oracle):
    theta13 = 5.371920351148152
    for nrm, s_exp in [(theta13, 0), (2 * theta13, 1), (2 * theta13 * 1.0001, 2), (1000.0, 8)]:
        _, s = oracle.expm(np.array([[-nrm]]))
        assert s == s_exp, (nrm, s)


def test_expm_nonfinite(oracle):
    E, s = oracle.expm(np.array([[1.0, np.nan], [0.0, 1.0]]))
    assert s == -1 and np.isnan(E).all()


# ------------------------------------------------------------------ solve

def test_expmv_builtin_diagonal(oracle):
    n = 1000
    d = 1.0 + np.arange(n) / n
    y, st, _ = oracle.expmv(np.arange(n + 1), np.arange(n), d, np.full(n, 0.1), 1.0, 1e-6, 30, 10)
    assert rel_err(y, 0.1 * np.exp(d * st["t_done"])) < 1e-12


def test_expmv_eigenvector_breakdown(oracle):
    N = 30
    rp, ci, va = gen.laplacian_2d(N)
    u, lam = laplacian_2d_eigvec(N, 4, 1)
    y, st, log = oracle.expmv(rp, ci, va, 3 * u, 2.0, 1e-8, 20, 10)
    assert log[0]["m_used"] == 1 and log[0]["err"] == 0.0
    assert rel_err(y, 3 * np.exp(lam * st["t_done"]) * u) < 1e-13


def test_expmv_laplacian_dst(oracle):
    N, t = 40, 5.0
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=3)
    y, st, _ = oracle.expmv(rp, ci, va, v, t, 1e-10, 30, 0)
    assert st["converged"] == 1
    assert rel_err(y, laplacian_2d_expmv(N, t, v)) < 1e-9


def test_expmv_scipy_fixtures(oracle):
    z = np.load(os.path.join(GOLDEN, "expmv_cases.npz"))
    for i in range(int(z["count"])):
        rp, ci, va = z[f"c{i}_indptr"], z[f"c{i}_indices"], z[f"c{i}_data"]
        v, t, yref = z[f"c{i}_v"], float(z[f"c{i}_t"]), z[f"c{i}_y"]
        y, st, _ = oracle.expmv(rp, ci, va, v, t, 1e-10, 30, 0)
        assert st["converged"] == 1
        assert rel_err(y, yref) < 1e-9, i


def test_controller_abort_after_20_rejects(oracle):
    # m = 1 with an unreachable tolerance: every attempt is rejected
    rp, ci, va = gen.random_negative(200, 5e-2, seed=2)
    v = gen.normal_vector(200, seed=2)
    y, st, log = oracle.expmv(rp, ci, va, v, 1.0, 1e-300, 1, 10)
    assert st["aborted"] == 1 and st["converged"] == 0 and st["attempts"] == 20
    assert [r["rejects"] for r in log] == list(range(20))
    dts = [r["dt"] for r in log]
    assert all(abs(b - a * 0.4) <= 1e-15 * a or b == max(1e-6, 1e-12) for a, b in zip(dts, dts[1:]))


def test_controller_step_cap_leaves_t_short(oracle):
    # max_restarts caps the accepted steps: after rejections shrink dt, two
    # accepted steps cannot reach t, so the run ends with converged = no and
    # y = exp(t_done A) v for t_done < t
    import scipy.sparse.linalg as spla
    rp, ci, va = gen.random_negative(300, 2e-2, seed=5)
    v = gen.normal_vector(300, seed=5)
    y, st, log = oracle.expmv(rp, ci, va, v, 1.0, 1e-10, 6, 2)
    assert st["restarts_done"] == 2 and st["converged"] == 0
    assert st["t_done"] < 1.0 and st["attempts"] > 2
    A = sp.csr_matrix((va, ci, rp), shape=(300, 300))
    assert rel_err(y, spla.expm_multiply(st["t_done"] * A, v)) < 1e-9


def test_nan_start_vector_is_accepted(oracle):
    # a NaN column norm counts as breakdown (err = 0, step accepted); the NaN
    # coefficient fails the lift's |wH_i| > 1e-15 test, so y = 0 and the run
    # reports converged (main_cpu.cpp:529, :564, :645)
    rp, ci, va = gen.laplacian_2d(10)
    v = np.ones(100)
    v[5] = np.nan
    y, st, log = oracle.expmv(rp, ci, va, v, 1.0, 1e-6, 5, 10)
    assert log[0]["err"] == 0.0 and log[0]["m_used"] == 1
    assert st["converged"] == 1 and np.all(y == 0.0)


@pytest.mark.parametrize("case", ["grow", "shrink_then_cap", "abort"])
def test_controller_trace_closed_form(oracle, case):
    """m = 1 on a 2 x 2 diagonal matrix: every attempt's t, dt, err, accept
    and reject count, the final y and the stats follow from closed forms."""
    from closed_forms import diag2_m1_trace
    lam1, lam2, a, b, t, tol, R = {
        "grow": (-1.0, -3.0, 1.0, 1.0, 4.0, 1.0, 20),            # h11 < 0: err falls, dt grows
        "shrink_then_cap": (0.5, -0.1, 1.0, 1.0, 3.0, 0.6, 10),  # h11 > 0: err rises, rejects shrink dt
        "abort": (-1.0, -3.0, 1.0, 1.0, 4.0, 0.5, 20),           # err only rises as dt shrinks
    }[case]
    rp, ci, va = np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.array([lam1, lam2])
    v = np.array([a, b])
    recs, yref, sref = diag2_m1_trace(lam1, lam2, a, b, t, tol, R)
    assert min(r["margin"] for r in recs) > 1e-9  # no decision within rounding of its threshold
    y, st, log = oracle.expmv(rp, ci, va, v, t, tol, 1, R)
    assert len(log) == len(recs)
    for got, want in zip(log, recs):
        assert got["m_used"] == 1
        assert (got["accepts"], got["rejects"]) == (want["accepts"], want["rejects"])
        for f in ("t", "dt", "err"):
            assert got[f] == pytest.approx(want[f], rel=1e-13, abs=1e-300), (f, got, want)
    for f in ("converged", "restarts_done", "attempts", "aborted"):
        assert st[f] == sref[f], f
    np.testing.assert_allclose(y, yref, rtol=1e-13)
    if case == "grow":
        dts = [r["dt"] for r in recs]
        assert any(b2 > b1 * 1.19 for b1, b2 in zip(dts, dts[1:]))  # the growth branch ran
    if case == "shrink_then_cap":
        assert sref["converged"] == 0 and any(r["rejects"] > 0 for r in recs)


def test_controller_trace_happy_breakdown(oracle):
    """v in a 2-D invariant subspace of diag(-1, -2, -3, -4) with m = 3: every
    attempt breaks down after 2 columns with err = 0, so the trace is the pure
    growth sequence and y is exact."""
    from closed_forms import breakdown_trace
    rp, ci = np.arange(5, dtype=np.int32), np.arange(4, dtype=np.int32)
    va = np.array([-1.0, -2.0, -3.0, -4.0])
    v = np.array([0.7, -1.3, 0.0, 0.0])
    y, st, log = oracle.expmv(rp, ci, va, v, 2.0, 1e-10, 3, 10)
    want = breakdown_trace(2.0, 10)
    assert len(log) == len(want) == st["restarts_done"] and st["converged"] == 1
    for got, (t0, dt) in zip(log, want):
        assert got["m_used"] == 2 and got["err"] == 0.0 and got["rejects"] == 0
        assert got["t"] == pytest.approx(t0, rel=1e-14) and got["dt"] == pytest.approx(dt, rel=1e-14)
    np.testing.assert_allclose(y, np.exp(2.0 * va) * v, rtol=1e-13, atol=1e-300)


@pytest.mark.parametrize("seed,m,tol,t,R", [(1, 4, 1e-6, 1.0, 10), (2, 6, 1e-9, 2.0, 0),
                                            (3, 3, 1e-7, 1.0, 5), (4, 8, 1e-11, 3.0, 20)])
def test_controller_trace_independent_restatement(oracle, seed, m, tol, t, R):
    """m > 1 without breakdown: the oracle's whole step trace against a dense
    numpy / scipy restatement (CGS2 basis, scipy expm); same accept / reject
    sequence, dt to rounding, err to 1e-7 relative."""
    from independent import controller_trace
    rp, ci, va = gen.random_negative(80, 8e-2, seed=seed)
    A = np.zeros((80, 80))
    for r in range(80):
        A[r, ci[rp[r]:rp[r + 1]]] += va[rp[r]:rp[r + 1]]
    v = gen.normal_vector(80, seed=seed)
    recs, yref, sref = controller_trace(A, v, t, tol, m, R)
    assert min(r["margin"] for r in recs) > 1e-5  # decisions clear of rounding
    y, st, log = oracle.expmv(rp, ci, va, v, t, tol, m, R)
    assert len(log) == len(recs)
    for got, want in zip(log, recs):
        assert (got["m_used"], got["accepts"], got["rejects"]) == \
            (want["m_used"], want["accepts"], want["rejects"])
        assert got["t"] == pytest.approx(want["t"], rel=1e-12, abs=1e-300)
        assert got["dt"] == pytest.approx(want["dt"], rel=1e-12)
        assert got["err"] == pytest.approx(want["err"], rel=1e-7, abs=1e-300)
    for f in ("converged", "restarts_done", "attempts", "aborted"):
        assert st[f] == sref[f], f
    assert rel_err(y, yref) < 1e-10


# ------------------------------------------------------------------ Arnoldi

def _dense(rp, ci, va, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] += va[rp[r]:rp[r + 1]]
    return A


@pytest.mark.parametrize("kind", ["random", "stiff_diagonal"])
def test_arnoldi_relation_and_orthonormality(oracle, kind):
    """The decomposition itself, not only the y it feeds: A V_m = V_{m+1} Hbar_m
    to rounding, V^T V = I to rounding, H upper Hessenberg with a positive
    subdiagonal.  Without the second MGS pass the basis of the same inputs
    loses orthogonality by two orders of magnitude more, so a dropped second
    pass fails the orthonormality bound."""
    n, m = 200, 40
    if kind == "random":
        rp, ci, va = gen.random_negative(n, 5e-2, seed=4)
    else:
        S = sp.diags(np.linspace(-1e4, -1.0, n)).tocsr()
        rp, ci, va = S.indptr, S.indices, S.data
    A = _dense(rp, ci, va, n)
    v = gen.normal_vector(n, seed=5)
    for reorth, orth_bound in ((True, 5e-14), (False, None)):
        H, V, k = oracle.arnoldi(rp, ci, va, v, m, reorth, basis=True)
        assert k == m
        Vk = V[:, :k + 1]
        rel = np.linalg.norm(A @ Vk[:, :k] - Vk @ H[:k + 1, :k]) / np.linalg.norm(A)
        assert rel < 1e-14
        orth = np.linalg.norm(Vk.T @ Vk - np.eye(k + 1))
        if reorth:
            assert orth < orth_bound
            orth_two = orth
        else:
            assert orth > 20 * orth_two  # the check above can tell the passes apart
        assert np.all(np.tril(H[:k + 1, :k], -2) == 0.0)
        assert np.all(np.diag(H[:k + 1, :k], -1) > 0.0)
    np.testing.assert_allclose(V[:, 0], v / np.linalg.norm(v), rtol=1e-14, atol=1e-300)


def test_arnoldi_h_against_householder_krylov_qr(oracle):
    """H of a short decomposition against an independent construction: the
    orthonormal factor Q of the Krylov matrix [v, Av, ..., A^m v] by
    Householder QR (R's diagonal made positive, as Gram-Schmidt with positive
    norms produces), then Hbar = Q^T A Q[:, :m].  A transposed operand or a
    wrong index in the MGS loop changes H far above the 1e-9 bound."""
    n, m = 40, 6
    rp, ci, va = gen.random_negative(n, 0.15, seed=8)
    A = _dense(rp, ci, va, n)
    v = gen.normal_vector(n, seed=9)
    K = np.empty((n, m + 1))
    K[:, 0] = v / np.linalg.norm(v)
    for j in range(m):
        K[:, j + 1] = A @ K[:, j]
        K[:, j + 1] /= np.linalg.norm(K[:, j + 1])  # column scaling keeps Q unchanged
    Q, R = np.linalg.qr(K)
    Q = Q * np.sign(np.diag(R))
    Href = Q.T @ A @ Q[:, :m]
    H, V, k = oracle.arnoldi(rp, ci, va, v, m, True, basis=True)
    assert k == m
    np.testing.assert_allclose(V[:, :m + 1], Q, rtol=0, atol=1e-9)
    np.testing.assert_allclose(H, Href, rtol=0, atol=1e-9 * np.abs(Href).max())
