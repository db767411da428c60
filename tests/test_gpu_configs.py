"""The batched and block workloads at their full configured sizes on one GPU.

* 256 independent problems (n log-uniform in [1e3, 2e5], random rows with
  5-30 nonzeros, values U[-1,1] in FP64 -- not rounded to FP32 --, a
  Gershgorin-negative diagonal, t ~ U[0.1, 5], m ~ U{20..60}): every
  problem's y from the batched FP64 solve against a tight-tolerance FP64
  solve of that problem alone; eight problems spanning n against the oracle
  under the reference controller (same accepted and attempted steps); the
  FP32-value request falls back to FP64 storage (these values are not exact in
  FP32) and gives the FP64 bits, while forced FP32 storage is shown to miss the
  1e-8 gate.
* The 2-D Laplacian on a 2048 x 2048 grid with a block of 32 N(0,1)
  right-hand sides (block Arnoldi, 16 block steps, t = 10, tol 1e-8): every
  column against the DST-I closed form.
* t = 0 gives y = v exactly (runner and batch).
"""
import numpy as np
import pytest

from closed_forms import laplacian_2d_expmv_block, rel_err
from conftest import gen, mek

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch256():
    print("[configs] generating 256 problems (FP64 values)", flush=True)
    probs = gen.batch_problems(P=256, seed=0, f32_exact=False)
    return probs


def _tight_refs(csrs, probs):
    """Each problem alone: m = 60, tol 1e-13 per unit time relative to ||w||,
    second Gram-Schmidt pass always."""
    refs = []
    for c, q in zip(csrs, probs):
        n = q["rp"].size - 1
        R = mek.Runner(c, m=min(60, n - 1), t=q["t"], tol=1e-13, max_restarts=0,
                       step_control="expokit_rel", gs_mode=1)
        refs.append(R.expmv(q["v"]))
        R.close()
    return refs


def test_batch256_fp64_values(gpu, batch256):
    probs = batch256
    P = len(probs)
    csrs = [mek.CSR.from_arrays(q["rp"], q["ci"], q["va"]) for q in probs]
    inexact = sum(c.f32_inexact()[0] for c in csrs)
    nnz = sum(int(q["rp"][-1]) for q in probs)
    print(f"[configs] {P} problems, {sum(q['rp'].size - 1 for q in probs)} rows, {nnz} nonzeros, "
          f"{inexact} values not exact in FP32", flush=True)
    assert inexact > 0.99 * nnz
    ts = [q["t"] for q in probs]
    ms = [q["m"] for q in probs]
    vs = [q["v"] for q in probs]
    tol = 1e-8
    print("[configs] tight FP64 references", flush=True)
    refs = _tight_refs(csrs, probs)
    ys = {}
    for f32 in (False, True, "lossy"):
        print(f"[configs] batched solve values_f32={f32}", flush=True)
        B = mek.Batch(csrs, ts, ms, [tol / t for t in ts], values_f32=f32, gs_mode=2,
                      step_control="expokit_rel")
        assert B.values_f32_used() == (f32 == "lossy")
        B.upload(vs)
        B.solve()
        ys[f32] = B.download()
        assert all(B.stats(p)["converged"] == 1 for p in range(P)), f32
        B.close()
    errs = [rel_err(y, r) for y, r in zip(ys[False], refs)]
    print(f"[configs] FP64 batch: max rel err {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-8, (int(np.argmax(errs)), max(errs))
    # the FP32 request fell back to FP64 storage: the same bits
    for a, b in zip(ys[True], ys[False]):
        np.testing.assert_array_equal(a, b)
    # FP32 storage perturbs A by up to 2^-24 relative: y moves far beyond 1e-8
    lossy = max(rel_err(y, r) for y, r in zip(ys["lossy"], refs))
    print(f"[configs] forced FP32 storage: max rel err {lossy:.2e}", flush=True)
    assert lossy > 1e-8


def test_batch256_reference_controller_vs_oracle(gpu, oracle, batch256):
    """All 256 problems in one batch under the reference controller
    (main.cu:1425-1518: first step t/10, tol 1e-9 per step, the second
    Gram-Schmidt pass always); eight problems spanning n take the oracle's
    steps and give its y."""
    probs = batch256
    tol = 1e-9
    B = mek.Batch([(q["rp"], q["ci"], q["va"]) for q in probs], [q["t"] for q in probs],
                  [q["m"] for q in probs], tol, max_restarts=10, gs_mode=1)
    B.upload([q["v"] for q in probs])
    print("[configs] reference-controller batch", flush=True)
    B.solve()
    ys = B.download()
    ns = np.array([q["rp"].size - 1 for q in probs])
    pick = np.argsort(ns)[np.linspace(0, len(probs) - 1, 8).astype(int)]
    for p in pick:
        q = probs[p]
        yref, st, _ = oracle.expmv(q["rp"], q["ci"], q["va"], q["v"], q["t"], tol, q["m"], 10)
        s = B.stats(int(p))
        e = rel_err(ys[p], yref)
        print(f"[configs] problem {p} n={ns[p]} m={q['m']} steps={s['accepted']}/{s['attempts']} "
              f"rel err vs oracle {e:.2e}", flush=True)
        assert (s["accepted"], s["attempts"], s["converged"]) == \
            (st["restarts_done"], st["attempts"], st["converged"]), p
        assert e <= 1e-8, (p, e)
    B.close()


def test_block2048_32_rhs_closed_form(gpu):
    N, t = 2048, 10.0
    print("[configs] 2048^2 Laplacian, 32 right-hand sides", flush=True)
    rp, ci, va = gen.laplacian_2d(N)
    V = np.random.default_rng(0).standard_normal((N * N, 32))
    V /= np.linalg.norm(V, axis=0)
    B = mek.Block(mek.CSR.from_arrays(rp, ci, va), m=16, t=t, tol=1e-8, gs_mode=2)
    Y = B.expmv(V)
    st = B.stats()
    print(f"[configs] block: {st['accepted']} sub-steps, {st['attempts']} attempts, "
          f"{st['second_passes']} second passes", flush=True)
    assert st["converged"] == 1
    B.close()
    ref = laplacian_2d_expmv_block(N, t, V)
    errs = [rel_err(Y[:, c], ref[:, c]) for c in range(32)]
    print(f"[configs] block: max column rel err {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-8


def test_t_zero_returns_v(gpu):
    rp, ci, va = gen.random_rows(500, 5, 30, seed=3)
    v = gen.normal_vector(500, seed=4)
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=20, t=0.0, tol=1e-8)
    np.testing.assert_array_equal(R.expmv(v), v)
    assert R.stats()["converged"] == 1
    R.set_t(0.5)
    assert rel_err(R.expmv(v), v) > 1e-3
    R.set_t(0.0)
    np.testing.assert_array_equal(R.expmv(v), v)
    R.close()
    mats = [(rp, ci, va), gen.random_rows(300, 5, 30, seed=5)]
    vs = [v, gen.normal_vector(300, seed=6)]
    B = mek.Batch(mats, [0.0, 0.3], [20, 20], 1e-9, step_control="expokit_rel")
    B.upload(vs)
    for _ in range(2):  # a second solve starts from the same state
        B.solve()
        ys = B.download()
        np.testing.assert_array_equal(ys[0], vs[0])
        assert rel_err(ys[1], vs[1]) > 1e-3
    B.close()


def test_f32_request_falls_back_for_inexact_values(gpu):
    """values_f32=True stores FP32 only when it is lossless; forced FP32
    ("lossy") stores it anyway."""
    rp, ci, va = gen.random_rows(2000, 5, 30, seed=8)
    v = gen.normal_vector(2000, seed=9)
    A = mek.CSR.from_arrays(rp, ci, va)
    out = {}
    for f32 in (False, True, "lossy"):
        R = mek.Runner(A, m=30, t=1.0, tol=1e-10, max_restarts=0, step_control="expokit_rel", gs_mode=1,
                       values_f32=f32)
        assert R.values_f32_used() == (f32 == "lossy")
        out[f32] = R.expmv(v)
        R.close()
    np.testing.assert_array_equal(out[True], out[False])
    assert rel_err(out["lossy"], out[False]) > 1e-9
    A32 = mek.CSR.from_arrays(rp, ci, va.astype(np.float32).astype(np.float64))
    R = mek.Runner(A32, m=30, t=1.0, tol=1e-10, values_f32=True)
    assert R.values_f32_used()
    R.close()
