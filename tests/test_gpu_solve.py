"""End-to-end exp(t*A)*v parity on the GPU (C ABI), at sizes the oracle finishes
in seconds.  Bar: relative 2-norm error <= 1e-8 against the FP64 oracle or a
closed form; the controller's step sequence must match the oracle's.
"""
import os
import subprocess

import numpy as np
import scipy.sparse as sp
import pytest

from closed_forms import breakdown_trace, diag2_m1_trace, kron3_expmv, laplacian_2d_eigvec, laplacian_2d_expmv, rel_err
from conftest import gen, mek

pytestmark = pytest.mark.gpu
TOL = 1e-8  # relative 2-norm error gate for the FP64 path


def _runner(rp, ci, va, **kw):
    return mek.Runner(mek.CSR.from_arrays(rp, ci, va), **kw)


def test_random_n2000_matches_oracle(gpu, oracle):
    """n=2000 random Gershgorin-negative matrix, v=0.1, t=1, m=30, tol 1e-6, R=10."""
    rp, ci, va = gen.random_negative(2000, 5e-3, seed=42)
    v = np.full(2000, 0.1)
    R = _runner(rp, ci, va, m=30, t=1.0, tol=1e-6, max_restarts=10)
    y = R.expmv(v)
    yref, st, log = oracle.expmv(rp, ci, va, v, 1.0, 1e-6, 30, 10)
    assert rel_err(y, yref) <= TOL
    s = R.stats()
    assert s["converged"] == st["converged"]
    assert s["restarts_done"] == st["restarts_done"]
    steps = R.steps()
    assert len(steps) == len(log)
    for a, b in zip(steps, log):
        assert a["m_used"] == b["m_used"] and a["accepts"] == b["accepts"] and a["rejects"] == b["rejects"]
        assert a["dt"] == pytest.approx(b["dt"], rel=1e-15)


def test_rejections_and_small_m(gpu, oracle):
    """Small m and tight tol force rejections (the basis is reused on reject)."""
    rp, ci, va = gen.random_negative(1500, 5e-3, seed=3)
    v = gen.normal_vector(1500, seed=4)
    R = _runner(rp, ci, va, m=8, t=2.0, tol=1e-10, max_restarts=0)
    y = R.expmv(v)
    yref, st, log = oracle.expmv(rp, ci, va, v, 2.0, 1e-10, 8, 0)
    assert st["attempts"] > st["restarts_done"]  # there were rejections
    steps = R.steps()
    assert [s["rejects"] for s in steps] == [s["rejects"] for s in log]
    assert R.stats()["converged"] == st["converged"]
    assert rel_err(y, yref) <= TOL


def test_builtin_diagonal_closed_form(gpu):
    """The CLI's built-in matrix a_ii = 1 + i/n: y_i = 0.1 e^{t(1+i/n)}."""
    n = 1000
    d = 1.0 + np.arange(n) / n
    rp = np.arange(n + 1, dtype=np.int32)
    R = _runner(rp, np.arange(n, dtype=np.int32), d, m=30, t=1.0, tol=1e-6, max_restarts=10)
    y = R.expmv(np.full(n, 0.1))
    assert rel_err(y, 0.1 * np.exp(d)) <= TOL


def test_eigenvector_happy_breakdown(gpu):
    N = 40
    rp, ci, va = gen.laplacian_2d(N)
    u, lam = laplacian_2d_eigvec(N, 3, 5)
    R = _runner(rp, ci, va, m=20, t=1.5, tol=1e-8, max_restarts=10)
    y = R.expmv(2.0 * u)
    assert rel_err(y, 2.0 * np.exp(1.5 * lam) * u) <= 1e-12
    assert R.steps()[0]["m_used"] == 1  # breakdown at the first column


def _dense_csr(A):
    rows, cols = np.nonzero(A)
    rp = np.zeros(A.shape[0] + 1, dtype=np.int32)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp).astype(np.int32), cols.astype(np.int32), A[rows, cols].astype(np.float64)


def _small_negative(n, seed):
    """Dense n x n matrix, off-diagonal U[-1, 1], Gershgorin-negative diagonal."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, (n, n))
    np.fill_diagonal(A, 0.0)
    np.fill_diagonal(A, -1.0 - np.abs(A).sum(axis=1))
    return A


def test_zero_start_vector(gpu, oracle):
    """v = 0: the start scale is guarded (main_cpu.cpp:480, scale 1 when
    ||v|| <= 1e-14), so y = 0 exactly, as in the oracle."""
    rp, ci, va = gen.random_negative(500, 1e-2, seed=8)
    v = np.zeros(500)
    R = _runner(rp, ci, va, m=10, t=1.0, tol=1e-8, max_restarts=10)
    y = R.expmv(v)
    yref, st, _log = oracle.expmv(rp, ci, va, v, 1.0, 1e-8, 10, 10)
    assert np.array_equal(y, np.zeros(500)) and np.array_equal(yref, np.zeros(500))
    assert R.stats()["converged"] == st["converged"]


def test_largest_m_and_rejected_m(gpu, oracle):
    """m = n - 1, the largest the reference accepts (main.cu:1322-1326): the
    m + 1 basis vectors span the whole space.  y against the oracle and a
    dense exponential; m = n is rejected with the reference's message."""
    import scipy.linalg

    n, t, tol = 6, 0.7, 1e-8
    A = _small_negative(n, seed=5)
    rp, ci, va = _dense_csr(A)
    v = np.random.default_rng(6).standard_normal(n)
    R = _runner(rp, ci, va, m=n - 1, t=t, tol=tol, max_restarts=200)
    y = R.expmv(v)
    yref, st, _log = oracle.expmv(rp, ci, va, v, t, tol, n - 1, 200)
    assert st["converged"] == 1 and R.stats()["converged"] == 1
    assert rel_err(y, yref) <= TOL
    assert rel_err(y, scipy.linalg.expm(t * A) @ v) <= TOL
    assert all(s["m_used"] == n - 1 for s in R.steps())
    with pytest.raises(mek.MekError, match="must be less than global_rows"):
        _runner(rp, ci, va, m=n, t=t, tol=tol, max_restarts=10)


def test_two_rows_m1_aborts_like_the_oracle(gpu, oracle):
    """n = 2, m = 1: the reference's residual estimate |H(1,0)| |exp(dt h00)|
    beta does not shrink with dt, so every attempt is rejected and the
    controller gives up after 20 rejections (main_cpu.cpp:602-700); the GPU
    controller takes the same 20 steps."""
    A = np.array([[-2.0, 0.5], [1.0, -3.0]])
    rp, ci, va = _dense_csr(A)
    v = np.array([0.3, -1.1])
    R = _runner(rp, ci, va, m=1, t=1.5, tol=1e-8, max_restarts=200)
    R.expmv(v)
    _y, st, log = oracle.expmv(rp, ci, va, v, 1.5, 1e-8, 1, 200)
    assert st["aborted"] == 1 and R.stats()["aborted"] == 1
    assert R.stats()["converged"] == 0
    steps = R.steps()
    assert len(steps) == len(log) == 20
    for a, b in zip(steps, log):
        assert a["rejects"] == b["rejects"]
        assert a["dt"] == pytest.approx(b["dt"], rel=1e-15)


@pytest.mark.parametrize("case", ["grow", "shrink_then_cap", "abort"])
def test_controller_trace_closed_form(gpu, case):
    """m = 1 on a 2 x 2 diagonal matrix: the GPU controller's attempts (t, dt,
    err, accepts, rejects), stats and y against closed forms (q1 keeps its
    direction, so err = h21 e^{dt h11} beta with beta growing by e^{dt h11}
    per accepted step)."""
    lam1, lam2, a, b, t, tol, R = {
        "grow": (-1.0, -3.0, 1.0, 1.0, 4.0, 1.0, 20),
        "shrink_then_cap": (0.5, -0.1, 1.0, 1.0, 3.0, 0.6, 10),
        "abort": (-1.0, -3.0, 1.0, 1.0, 4.0, 0.5, 20),
    }[case]
    rp, ci, va = np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.array([lam1, lam2])
    recs, yref, sref = diag2_m1_trace(lam1, lam2, a, b, t, tol, R)
    print(f"[trace] {case} attempts={len(recs)}", flush=True)
    run = _runner(rp, ci, va, m=1, t=t, tol=tol, max_restarts=R)
    y = run.expmv(np.array([a, b]))
    steps = run.steps()
    assert len(steps) == len(recs)
    for got, want in zip(steps, recs):
        assert got["m_used"] == 1
        assert (got["accepts"], got["rejects"]) == (want["accepts"], want["rejects"])
        for f in ("t", "dt", "err"):
            assert got[f] == pytest.approx(want[f], rel=1e-13, abs=1e-300), (f, got, want)
    st = run.stats()
    assert (st["converged"], st["restarts_done"], st["aborted"]) == \
        (sref["converged"], sref["restarts_done"], sref["aborted"])
    np.testing.assert_allclose(y, yref, rtol=1e-13)
    run.close()


def test_controller_trace_happy_breakdown(gpu):
    """v in a 2-D invariant subspace of diag(-1, -2, -3, -4) with m = 3: every
    attempt breaks down after 2 columns with err = 0, so the GPU trace is the
    pure growth sequence and y is exact."""
    rp, ci = np.arange(5, dtype=np.int32), np.arange(4, dtype=np.int32)
    va = np.array([-1.0, -2.0, -3.0, -4.0])
    v = np.array([0.7, -1.3, 0.0, 0.0])
    run = _runner(rp, ci, va, m=3, t=2.0, tol=1e-10, max_restarts=10)
    y = run.expmv(v)
    want = breakdown_trace(2.0, 10)
    steps = run.steps()
    assert len(steps) == len(want) == run.stats()["restarts_done"] and run.stats()["converged"] == 1
    for got, (t0, dt) in zip(steps, want):
        assert got["m_used"] == 2 and got["err"] == 0.0 and got["rejects"] == 0
        assert got["t"] == pytest.approx(t0, rel=1e-14) and got["dt"] == pytest.approx(dt, rel=1e-14)
    np.testing.assert_allclose(y, np.exp(2.0 * va) * v, rtol=1e-13, atol=1e-300)
    run.close()


@pytest.mark.parametrize("seed,m,tol,t,R", [(2, 6, 1e-9, 2.0, 0), (4, 8, 1e-11, 3.0, 20)])
def test_controller_trace_independent_restatement(gpu, seed, m, tol, t, R):
    """m > 1 without breakdown: the GPU's step trace against a dense numpy /
    scipy restatement of the reference controller (rejections, growth and,
    for R = 20, the step cap)."""
    from independent import controller_trace
    n = 80
    rp, ci, va = gen.random_negative(n, 8e-2, seed=seed)
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] += va[rp[r]:rp[r + 1]]
    v = gen.normal_vector(n, seed=seed)
    recs, yref, sref = controller_trace(A, v, t, tol, m, R)
    run = _runner(rp, ci, va, m=m, t=t, tol=tol, max_restarts=R)
    y = run.expmv(v)
    steps = run.steps()
    assert len(steps) == len(recs)
    for got, want in zip(steps, recs):
        assert (got["m_used"], got["accepts"], got["rejects"]) == \
            (want["m_used"], want["accepts"], want["rejects"])
        assert got["dt"] == pytest.approx(want["dt"], rel=1e-12)
        assert got["err"] == pytest.approx(want["err"], rel=1e-7, abs=1e-300)
    st = run.stats()
    assert (st["converged"], st["restarts_done"]) == (sref["converged"], sref["restarts_done"])
    assert rel_err(y, yref) < 1e-10
    run.close()


@pytest.mark.parametrize("control", ["reference", "expokit"])
def test_laplacian_dst(gpu, control):
    N, t = 128, 10.0
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=1)
    R = _runner(rp, ci, va, m=40, t=t, tol=1e-10, max_restarts=0, step_control=control)
    y = R.expmv(v)
    assert rel_err(y, laplacian_2d_expmv(N, t, v)) <= TOL
    assert R.stats()["converged"] == 1


@pytest.mark.parametrize("f32", [False, True])
def test_convdiff_kronecker(gpu, f32):
    """Nonsymmetric 3-D upwind convection-diffusion (Pe=50); FP32 storage is
    lossless for these coefficients, so both forms meet the FP64 gate."""
    N, t = 24, 0.5
    rp, ci, va = gen.convdiff_3d(N)
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    v = gen.normal_vector(N ** 3, seed=2)
    R = _runner(rp, ci, va, m=50, t=t, tol=1e-10, max_restarts=0, step_control="expokit",
                values_f32=f32)
    y = R.expmv(v)
    ref = kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                      gen.tridiag_1d(N, lx, dx, ux), t, v)
    assert rel_err(y, ref) <= TOL


def test_dgks_conditional_reorth(gpu, oracle):
    rp, ci, va = gen.random_negative(2000, 5e-3, seed=42)
    v = gen.normal_vector(2000, seed=8)
    R = _runner(rp, ci, va, m=30, t=1.0, tol=1e-10, max_restarts=0, gs_mode=2, dgks_eta=0.1)
    y = R.expmv(v)
    yref, _, _ = oracle.expmv(rp, ci, va, v, 1.0, 1e-10, 30, 0)
    assert rel_err(y, yref) <= TOL


def test_large_m_generic_path(gpu, oracle):
    """m > 64 takes the chunked dot path and the global-memory expm."""
    rp, ci, va = gen.random_rows(3000, 5, 30, seed=5)
    v = gen.normal_vector(3000, seed=6)
    R = _runner(rp, ci, va, m=70, t=0.5, tol=1e-10, max_restarts=0)
    y = R.expmv(v)
    yref, st, _ = oracle.expmv(rp, ci, va, v, 0.5, 1e-10, 70, 0)
    assert rel_err(y, yref) <= TOL


def test_device_resident_entry(gpu, oracle):
    rp, ci, va = gen.random_negative(2000, 5e-3, seed=42)
    v = gen.normal_vector(2000, seed=9)
    R = _runner(rp, ci, va, m=30, t=1.0, tol=1e-8, max_restarts=10)
    d_v = mek.DeviceBuffer.from_array(v)
    d_y = mek.DeviceBuffer(8 * 2000)
    R.expmv_device(d_v, d_y)
    yref, _, _ = oracle.expmv(rp, ci, va, v, 1.0, 1e-8, 30, 10)
    assert rel_err(d_y.to_array(), yref) <= TOL


def test_cli_builtin_matrix(gpu, tmp_path, oracle):
    out = tmp_path / "y.txt"
    env = dict(os.environ, OUTPUT_Y_FILE=str(out))
    p = subprocess.run([mek.CLI_PATH, "500", "--t", "1.0"], env=env, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Arnoldi summary: m=30, max_restarts=10" in p.stdout
    assert "METRIC total_compute_expmv_ms=" in p.stdout
    lines = out.read_text().splitlines()
    assert len(lines) == 500 and lines[0].count("e") == 1
    y = np.array([float(s) for s in lines])
    d = 1.0 + np.arange(500) / 500
    rp = np.arange(501, dtype=np.int32)
    yref, st, _ = oracle.expmv(rp, np.arange(500, dtype=np.int32), d, np.full(500, 0.1), 1.0, 1e-6, 30, 10)
    assert rel_err(y, yref) <= TOL
    assert ("converged=yes" in p.stdout) == bool(st["converged"])


@pytest.mark.parametrize("f32", [False, True])
def test_fused_spmv_dots(gpu, f32):
    """The first Gram-Schmidt pass fused into the SpMV (fuse_spmv_dot) gives
    the same solve as the separate sweep: the 2-D Laplacian against its DST-I
    closed form and against the unfused path, for basis sizes across every
    column bucket (m = 60)."""
    N, t = 96, 10.0
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=7)
    ys = {}
    for fused in (False, True):
        R = _runner(rp, ci, va, m=60, t=t, tol=1e-10, max_restarts=0, step_control="expokit",
                    gs_mode=2, values_f32=f32, fuse_spmv_dot=fused)
        ys[fused] = R.expmv(v)
        R.close()
    assert rel_err(ys[True], laplacian_2d_expmv(N, t, v)) <= TOL
    assert rel_err(ys[True], ys[False]) <= 1e-12


def test_fused_spmv_dots_random_rows(gpu, oracle):
    """Fused path on a nonsymmetric matrix with ragged rows (1 to 40
    nonzeros, one row longer than an SpMV block) against the oracle."""
    rp, ci, va = gen.random_rows(5000, 1, 40, seed=9)
    # one long row (5000 nonzeros > 4096) takes the whole-workgroup path
    n = rp.size - 1
    dense_row = np.arange(n, dtype=np.int32)
    rows = [np.asarray(ci[rp[i]:rp[i + 1]]) for i in range(n)]
    vals = [np.asarray(va[rp[i]:rp[i + 1]]) for i in range(n)]
    rows[17] = dense_row
    vals[17] = np.where(dense_row == 17, -float(n), 0.5 / n)
    rp = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    va = np.concatenate(vals)
    v = gen.normal_vector(n, seed=10)
    R = _runner(rp, ci, va, m=30, t=0.5, tol=1e-10, max_restarts=0, step_control="expokit", gs_mode=2,
                fuse_spmv_dot=True)
    y = R.expmv(v)
    R.close()
    yref, st, _log = oracle.expmv(rp, ci, va, v, 0.5, 1e-12, 30, 200)
    assert st["converged"] == 1
    assert rel_err(y, yref) <= TOL


def test_cli_bincsr_and_mtx_files(gpu, tmp_path, oracle):
    """The reference's CPU configuration as a file: n = 2000 random
    Gershgorin-negative matrix (seed 42), v = 0.1, t = 1 clamped by the CLI to
    10 / max|a_ij| (main.cu:2228-2242), m = 30, tol 1e-6, 10 restarts; the
    same matrix as .bincsr and as MatrixMarket gives the same y, which meets
    the oracle at the clamped t."""
    rp, ci, va = gen.random_negative(2000, 5e-3, seed=42)
    n = rp.size - 1
    gen.write_bincsr(tmp_path / "A.bincsr", rp, ci, va)
    rows = np.repeat(np.arange(n), np.diff(rp))
    with open(tmp_path / "A.mtx", "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% written by the test\n")
        f.write(f"{n} {n} {ci.size}\n")
        for r, c, v in zip(rows, ci, va):
            f.write(f"{r + 1} {c + 1} {float(v)!r}\n")  # shortest round-trip text of the double
    ys = {}
    for name, banner in (("A.bincsr", "Detected Binary CSR v1 input"), ("A.mtx", "Detected MatrixMarket input")):
        out = tmp_path / (name + ".y")
        env = dict(os.environ, OUTPUT_Y_FILE=str(out))
        p = subprocess.run([mek.CLI_PATH, str(tmp_path / name), "--t", "1.0"], env=env, capture_output=True,
                           text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        assert banner in p.stdout and f"Loaded matrix: {n}x{n} with {ci.size} non-zeros" in p.stdout
        assert "Clamping t to" in p.stdout
        ys[name] = np.array([float(x) for x in out.read_text().splitlines()])
    np.testing.assert_array_equal(ys["A.bincsr"], ys["A.mtx"])
    t = min(1.0, 10.0 / float(np.abs(va).max()))
    yref, st, _ = oracle.expmv(rp, ci, va, np.full(n, 0.1), t, 1e-6, 30, 10)
    assert rel_err(ys["A.bincsr"], yref) <= TOL


@pytest.mark.parametrize("control", ["expokit", "reference"])
def test_lanczos_laplacian_dst(gpu, control):
    """Symmetric Lanczos path on the 2-D Laplacian against the DST-I closed
    form, and against the Arnoldi path on the same input."""
    N = 96
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=3)
    t = 5.0
    ref = laplacian_2d_expmv(N, t, v)
    ys = {}
    for lz in (True, False):
        R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=30, t=t, tol=1e-10, max_restarts=10,
                       step_control=control, gs_mode=2, lanczos=lz)
        ys[lz] = R.expmv(v)
        R.close()
    assert rel_err(ys[True], ref) <= 1e-8, rel_err(ys[True], ref)
    assert rel_err(ys[True], ys[False]) <= 1e-8


def test_lanczos_rejects_nonsymmetric(gpu):
    rp, ci, va = gen.convdiff_3d(8)
    with pytest.raises(mek.MekError, match="symmetric"):
        mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=10, t=0.1, tol=1e-8, lanczos=True)


def test_autotune_shapes_and_result(gpu):
    """The autotuner picks shapes from its candidate sets and the tuned solve
    gives the untuned result to rounding (same SpMV sums, other reduction
    partitions) and the closed form."""
    N = 512
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=4)
    ys = {}
    for at in (True, False):
        R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=30, t=3.0, tol=1e-10, max_restarts=0,
                       step_control="expokit", gs_mode=2, autotune=at)
        ys[at] = R.expmv(v)
        tn = R.tuning()
        R.close()
        if at:
            print(f"[autotune] {tn}", flush=True)
            assert tn["nnz_cap"] in (512, 1024, 4096)
            assert tn["spmv_grid"] in (8, 1024, 2048, 4096) and tn["sweep_grid"] in (1024, 2048, 4096)
            assert tn["spmv_us"] > 0 and tn["sweep_us"] > 0
        else:
            assert tn["nnz_cap"] == 0
    assert rel_err(ys[True], ys[False]) <= 1e-12
    assert rel_err(ys[True], laplacian_2d_expmv(N, 3.0, v)) <= 1e-8


@pytest.mark.parametrize("gs_mode", [1, 2])
def test_odd_row_count(gpu, oracle, gs_mode):
    """Odd n: the sweeps read rows in pairs and the last pair includes the
    zero padding row; results must not depend on it."""
    N = 45  # n = 2025
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=6)
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=30, t=2.0, tol=1e-10, max_restarts=0,
                   step_control="expokit", gs_mode=gs_mode)
    y = R.expmv(v)
    R.close()
    assert rel_err(y, laplacian_2d_expmv(N, 2.0, v)) <= 1e-8
    rp, ci, va = gen.random_rows(3001, 4, 20, seed=8)
    v = gen.normal_vector(3001, seed=9)
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=20, t=0.5, tol=1e-10, max_restarts=10, gs_mode=gs_mode)
    y = R.expmv(v)
    R.close()
    yref, st, _log = oracle.expmv(rp, ci, va, v, 0.5, 1e-10, 20, 10)
    assert rel_err(y, yref) <= 1e-8


@pytest.mark.parametrize("sstep", [0, 12])
def test_spmv_formats_agree(gpu, sstep):
    """The automatic SpMV layout is the constant-diagonal form for a stencil
    with constant coefficients, the diagonal copy for one whose diagonals
    vary, the sliced-ELL copy for a band of more than 16 diagonals, and the
    CSR row blocks for rows of 5-30 random nonzeros (slices padded to their
    longest row would carry ~70 % padding); forcing any layout gives the same
    solve to rounding, within 1e-8 of the DST-I closed form, and the two
    diagonal forms give the same bits."""
    N, t = 80, 10.0
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=4)
    kw = dict(m=40, t=t, tol=1e-10, max_restarts=0, step_control="expokit", gs_mode=2, sstep=sstep)
    ys, infos = {}, {}
    for fmt in ("auto", "csr", "sell", "dia", "dia_const"):
        R = _runner(rp, ci, va, spmv_format=fmt, **kw)
        infos[fmt] = R.spmv_format()
        ys[fmt] = R.expmv(v)
        R.close()
    assert infos["auto"]["format"] == "dia_const" and infos["csr"]["format"] == "csr"
    assert infos["sell"]["format"] == "sell" and infos["dia"]["format"] == "dia"
    assert infos["dia_const"]["format"] == "dia_const" and infos["dia_const"]["width"] == 5
    np.testing.assert_array_equal(ys["dia_const"], ys["dia"])
    lens = np.diff(rp)
    assert infos["sell"]["width"] == 5
    assert infos["sell"]["entries"] == sum(64 * int(lens[64 * q:64 * (q + 1)].max())
                                           for q in range((N * N + 63) // 64))
    assert infos["dia"]["width"] == 5 and infos["dia"]["entries"] == 5 * N * N
    ref = laplacian_2d_expmv(N, t, v)
    for fmt, y in ys.items():
        assert rel_err(y, ref) <= TOL, fmt
    assert rel_err(ys["sell"], ys["csr"]) <= 1e-12
    assert rel_err(ys["dia"], ys["csr"]) <= 1e-12
    np.testing.assert_array_equal(ys["auto"], ys["dia"])
    # a varying coefficient on one diagonal: the automatic layout keeps the values
    va2 = va.copy()
    va2[rp[N + 5]] = 1.0 + 2.0 ** -20
    R = _runner(rp, ci, va2, m=20, t=0.5, tol=1e-10, max_restarts=0, step_control="expokit")
    assert R.spmv_format()["format"] == "dia"
    R.close()
    with pytest.raises(mek.MekError):
        _runner(rp, ci, va2, m=20, t=0.5, tol=1e-10, max_restarts=0, spmv_format="dia_const")
    rp, ci, va = gen.random_rows(4000, 5, 30, seed=3)
    R = _runner(rp, ci, va, m=20, t=0.5, tol=1e-10, max_restarts=0, step_control="expokit")
    assert R.spmv_format()["format"] == "csr"
    R.close()
    with pytest.raises(mek.MekError):
        _runner(rp, ci, va, m=20, t=0.5, tol=1e-10, max_restarts=0, spmv_format="dia")
    rp, ci, va = _band(3000, list(range(-10, 11)))
    R = _runner(rp, ci, va, m=20, t=0.5, tol=1e-10, max_restarts=0, step_control="expokit")
    assert R.spmv_format()["format"] == "sell"
    R.close()


def _band(n, offsets, seed=6):
    """A full band: row r holds the columns r + o (o in offsets, in range),
    values U[-1,1] with a Gershgorin-negative diagonal."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for o in sorted(offsets):
        r = np.arange(max(0, -o), min(n, n - o))
        rows.append(r)
        cols.append(r + o)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    va = rng.uniform(-1, 1, r.size)
    diag = r == c
    off = np.zeros(n)
    np.add.at(off, r[~diag], np.abs(va[~diag]))
    va[diag] = -1.0 - off[r[diag]]
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32), va


@pytest.mark.parametrize("N", [40, 1800])
def test_host_staging_matches_device_solve(gpu, N):
    """compute_expmv on host vectors (pinned staging; at n >= 2^21 in 8
    chunks whose host copies overlap the DMA of the chunk before, both ways)
    gives the bits of the device-resident solve; a second call with another
    v reuses the staging buffer and its events.  N = 1800: n = 3,240,000
    rows, the chunked path; N = 40: one piece."""
    rp, ci, va = gen.laplacian_2d(N)
    n = N * N
    R = _runner(rp, ci, va, m=12, t=0.05, tol=1e-10, max_restarts=0, step_control="expokit")
    for seed in (3, 4):
        v = gen.normal_vector(n, seed=seed)
        print(f"[staging] N={N} seed={seed}", flush=True)
        y_host = R.expmv(v)
        d_v = mek.DeviceBuffer.from_array(v)
        d_y = mek.DeviceBuffer(8 * n)
        R.expmv_device(d_v, d_y)
        np.testing.assert_array_equal(y_host.view(np.int64), d_y.to_array().view(np.int64))
        assert rel_err(y_host, laplacian_2d_expmv(N, 0.05, v)) <= TOL
    R.close()


@pytest.mark.parametrize("which", ["laplacian", "convdiff", "tridiag"])
def test_fused_first_column_dots_give_the_hessenberg_matrix(gpu, which):
    """On a 3-, 5- or 7-diagonal matrix the column path forms the Gram-Schmidt
    dots of columns j <= 7 inside the diagonal SpMV (8-, 4- and 2-dot
    kernels); later columns and the CSR layout use the separate dot sweep.
    The Hessenberg matrices of the two layouts agree to rounding over 12
    columns, and both satisfy the Arnoldi relation."""
    if which == "laplacian":
        rp, ci, va = gen.laplacian_2d(48)
    elif which == "convdiff":
        rp, ci, va = gen.convdiff_3d(13)
    else:  # a 1-D convection-diffusion operator, nonsymmetric
        n = 3001
        T = sp.diags([np.full(n - 1, 1.3), np.full(n, -2.0), np.full(n - 1, 0.7)], [-1, 0, 1], format="csr")
        T.sort_indices()
        rp, ci, va = T.indptr.astype(np.int32), T.indices.astype(np.int32), T.data.astype(np.float64)
    n = rp.size - 1
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    v = gen.normal_vector(n, seed=7)
    Hs = {}
    for fmt in ("dia_const", "csr"):
        R = _runner(rp, ci, va, m=12, t=1.0, gs_mode=2, spmv_format=fmt)
        assert R.spmv_format()["format"] == fmt
        H, k = R.arnoldi(v)
        R.close()
        assert k == 12
        Hs[fmt] = H[:13, :12]
    scale = np.abs(Hs["csr"]).max()
    np.testing.assert_allclose(Hs["dia_const"], Hs["csr"], rtol=0, atol=1e-12 * scale)
    # Arnoldi relation with the column path's basis rebuilt from H: A V_12 = V_13 H
    V = np.zeros((n, 13))
    V[:, 0] = v / np.linalg.norm(v)
    H = Hs["dia_const"]
    for j in range(12):
        w = A @ V[:, j] - V[:, :j + 1] @ H[:j + 1, j]
        V[:, j + 1] = w / H[j + 1, j]
    assert np.abs(V.T @ V - np.eye(13)).max() < 1e-9
