"""Batched lockstep solver on the GPU (C ABI) against the oracle, problem by
problem: every problem follows the reference controller, so its y, accepted
step count and attempt count must match the oracle's single solve."""
import numpy as np
import pytest

from closed_forms import rel_err
from conftest import gen, mek

pytestmark = pytest.mark.gpu


def _problems(P=12, seed=0):
    rng = np.random.default_rng(seed)
    mats, ts, ms, vs = [], [], [], []
    for p in range(P):
        n = int(np.exp(rng.uniform(np.log(300), np.log(6000))))
        mats.append(gen.random_rows(n, 5, 30, seed=seed * 1000 + p))
        ts.append(float(rng.uniform(0.1, 2.0)))
        ms.append(int(rng.integers(8, 41)))
        vs.append(gen.normal_vector(n, seed=seed * 1000 + p))
    return mats, ts, ms, vs


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_oracle(gpu, oracle, use_graph):
    mats, ts, ms, vs = _problems()
    tol = 1e-9
    B = mek.Batch(mats, ts, ms, tol, max_restarts=10, use_graph=use_graph)
    B.upload(vs)
    B.solve()
    ys = B.download()
    print(f"[batch] cycles={B.cycles()} graph={use_graph}", flush=True)
    for p, ((rp, ci, va), t, m, v, y) in enumerate(zip(mats, ts, ms, vs, ys)):
        yref, st, log = oracle.expmv(rp, ci, va, v, t, tol, m, 10)
        s = B.stats(p)
        assert s["accepted"] == st["restarts_done"], p
        assert s["attempts"] == st["attempts"], p
        assert s["converged"] == st["converged"], p
        assert rel_err(y, yref) <= 1e-8, (p, rel_err(y, yref))
    B.close()


def test_batch_replays_are_identical(gpu):
    """The captured graph replays must give the eager result on every solve."""
    mats, ts, ms, vs = _problems(P=6, seed=3)
    Bg = mek.Batch(mats, ts, ms, 1e-8, max_restarts=0, use_graph=True)
    Be = mek.Batch(mats, ts, ms, 1e-8, max_restarts=0, use_graph=False)
    Be.upload(vs)
    Be.solve()
    ye = Be.download()
    Bg.upload(vs)
    for rep in range(3):
        Bg.solve()
        yg = Bg.download()
        for a, b in zip(yg, ye):
            np.testing.assert_array_equal(a, b)


def test_batch_f32_stencils_and_substeps(gpu, oracle):
    """FP32-stored stencils (lossless) with a fixed first sub-step count."""
    mats = [gen.convdiff_3d(12), gen.laplacian_2d(40), gen.convdiff_3d(9)]
    ts = [0.05, 1.0, 0.02]
    ms = [30, 20, 25]
    vs = [gen.normal_vector(r[0].size - 1, seed=i) for i, r in enumerate(mats)]
    B = mek.Batch(mats, ts, ms, 1e-10, substeps=[4, 2, 1], values_f32=True)
    B.upload(vs)
    B.solve()
    for (rp, ci, va), t, m, v, y in zip(mats, ts, ms, vs, B.download()):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        A = sp.csr_matrix((va, ci, rp))
        assert rel_err(y, spla.expm_multiply(t * A, v)) <= 1e-8


@pytest.mark.parametrize("values_f32", [False, True])
def test_batch_expokit_matches_single_runner(gpu, values_f32):
    """Batched Expokit step control takes the single runner's steps (same
    first step from ||A||_inf, same accept/reject rule) and gives its y."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    mats, ts, ms, vs = _problems(P=10, seed=5)
    if values_f32:  # FP32 storage is lossless for values exactly representable in FP32
        mats = [(rp, ci, va.astype(np.float32).astype(np.float64)) for rp, ci, va in mats]
    tol = 1e-9
    B = mek.Batch(mats, ts, ms, tol, step_control="expokit", gs_mode=1, values_f32=values_f32)
    B.upload(vs)
    B.solve()
    ys = B.download()
    print(f"[batch-expokit] cycles={B.cycles()} f32={values_f32}", flush=True)
    for p, ((rp, ci, va), t, m, v, y) in enumerate(zip(mats, ts, ms, vs, ys)):
        R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=m, t=t, tol=tol, max_restarts=0,
                       step_control="expokit", gs_mode=1, values_f32=values_f32)
        y1 = R.expmv(v)
        st = R.stats()
        R.close()
        s = B.stats(p)
        assert s["converged"] == 1, p
        assert s["accepted"] == st["accepted"], (p, s, st)
        assert s["attempts"] == st["attempts"], (p, s, st)
        assert rel_err(y, y1) <= 1e-11, (p, rel_err(y, y1))
        A = sp.csr_matrix((va, ci, rp))
        assert rel_err(y, spla.expm_multiply(t * A, v)) <= 1e-8, p
    B.close()


def test_batch_autotune_same_result(gpu):
    """The autotuned SpMV grid changes only the launch shape: identical y."""
    mats, ts, ms, vs = _problems(P=8, seed=7)
    ys = {}
    for at in (False, True):
        B = mek.Batch(mats, ts, ms, 1e-9, step_control="expokit", gs_mode=2, autotune=at)
        B.upload(vs)
        B.solve()
        ys[at] = B.download()
        if at:
            assert B.spmv_grid() in (1024, 2048, 4096, 8192)
        B.close()
    for a, b in zip(ys[False], ys[True]):
        np.testing.assert_array_equal(a, b)


def _rows_edited(n, seed, dense=(), empty=()):
    """random_rows with some rows replaced by a dense row (n nonzeros, small
    values) or by an empty row."""
    rp, ci, va = gen.random_rows(n, 5, 30, seed=seed)
    rows = [list(zip(ci[rp[r]:rp[r + 1]], va[rp[r]:rp[r + 1]])) for r in range(n)]
    rng = np.random.default_rng(seed)
    for r in dense:
        rows[r] = [(c, float(x)) for c, x in zip(range(n), rng.uniform(-0.01, 0.01, n))]
    for r in empty:
        rows[r] = []
    new_rp = np.zeros(n + 1, np.int32)
    new_rp[1:] = np.cumsum([len(r) for r in rows])
    new_ci = np.array([c for r in rows for c, _ in r], np.int32)
    new_va = np.array([x for r in rows for _, x in r], np.float64)
    return new_rp, new_ci, new_va


def test_batch_odd_rows_long_and_empty_rows(gpu, oracle):
    """Odd row counts (each problem is followed by one empty padding row in
    the batch's row space), a dense last row (a long-row SpMV block that ends
    on the padding row), a dense row inside, and an empty last row."""
    mats = [_rows_edited(2501, 1, dense=[2500]),
            _rows_edited(1999, 2, empty=[1998]),
            _rows_edited(3001, 3, dense=[7], empty=[11]),
            _rows_edited(1200, 4)]
    ts = [0.2, 0.5, 0.3, 1.0]
    ms = [30, 20, 25, 12]
    vs = [gen.normal_vector(r[0].size - 1, seed=10 + i) for i, r in enumerate(mats)]
    tol = 1e-9
    B = mek.Batch(mats, ts, ms, tol, max_restarts=10)
    B.upload(vs)
    B.solve()
    ys = B.download()
    for p, ((rp, ci, va), t, m, v, y) in enumerate(zip(mats, ts, ms, vs, ys)):
        assert y.shape == v.shape, p
        yref, st, _ = oracle.expmv(rp, ci, va, v, t, tol, m, 10)
        assert B.stats(p)["converged"] == st["converged"] == 1, p
        assert rel_err(y, yref) <= 1e-8, (p, rel_err(y, yref))
    B.close()


def test_batch_time_spmv_leaves_solve_unchanged(gpu):
    """The SpMV timing entry runs between solves: it reports a positive time
    and the next solve gives the same bits as the one before it."""
    mats, ts, ms, vs = _problems(P=5, seed=11)
    B = mek.Batch(mats, ts, ms, 1e-9)
    B.upload(vs)
    B.solve()
    y0 = B.download()
    us = B.time_spmv(5)
    print(f"[batch] spmv {us:.1f} us", flush=True)
    assert us > 0.0
    B.solve()
    for a, b in zip(B.download(), y0):
        np.testing.assert_array_equal(a, b)
    B.close()
