"""Learned selector kernels on the GPU: MLP forward/backward against the FP64
numpy oracle (FP32 tolerances), training convergence, and device features
against a host restatement of their definitions."""
import numpy as np
import pytest
import scipy.sparse as sp

import mlp_ref
from conftest import gen, mek

pytestmark = pytest.mark.gpu


def test_mlp_forward_and_gradient_vs_oracle(gpu):
    net = mek.Mlp(seed=3)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((37, 16)).astype(np.float32)
    T = rng.standard_normal((37, 2)).astype(np.float32)
    Y = net.forward(X)
    Yref, _ = mlp_ref.forward(p, X.astype(np.float64))
    np.testing.assert_allclose(Y, Yref, rtol=1e-4, atol=1e-5)
    loss, g = net.loss_grad(X, T)
    lref, gref = mlp_ref.loss_grad(p, X, T)
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    assert np.linalg.norm(g - gref) <= 1e-4 * np.linalg.norm(gref)


def test_mlp_training_reduces_loss(gpu):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((2048, 16)).astype(np.float32)
    T = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), 0.3 * X[:, 2] * X[:, 3]], 1).astype(np.float32)
    net = mek.Mlp(seed=0)
    l0 = mlp_ref.loss(net.params(), X, T)
    l1 = net.train(X, T, batch=128, epochs=60, lr=2e-3, seed=1)
    assert l1 < 0.1 * l0, (l0, l1)
    # the device loss is the FP64 oracle's loss of the trained parameters
    assert abs(l1 - mlp_ref.loss(net.params(), X, T)) <= 1e-4 * max(l1, 1e-6)


def _host_features(rp, ci, va, t, tol):
    A = sp.csr_matrix((va, ci, rp))
    n = A.shape[0]
    nnz = np.diff(rp).astype(float)
    d = A.diagonal()
    off = np.abs(A).sum(1).A1 - np.abs(d)
    ninf = np.abs(A).sum(1).max()
    n1 = np.abs(A).sum(0).max()
    right = (d + off).max()
    left = -(d - off).min()
    return np.array([np.log10(n), np.log10(nnz.sum()), nnz.mean(), nnz.max() / nnz.mean(),
                     nnz.std() / nnz.mean(), np.log10(ninf), np.log10(n1), np.log10(t),
                     np.log10(t * ninf), np.log10(t * n1), t * right, np.log10(t * left),
                     np.log10(t * np.abs(d).max()), np.log10(np.mean(np.abs(d) / off)),
                     np.log10(tol), np.log10(0.5 * t * (right + left))])


def test_features_match_definition(gpu):
    for (rp, ci, va), t in [(gen.random_rows(3000, 5, 30, seed=2), 0.7), (gen.convdiff_3d(10), 0.05)]:
        f = mek.features(mek.CSR.from_arrays(rp, ci, va), t, 1e-8)
        np.testing.assert_allclose(f, _host_features(rp, ci, va, t, 1e-8), rtol=2e-5, atol=2e-5)


def test_batch_features_equal_single(gpu):
    mats = [gen.random_rows(2000, 5, 30, seed=4), gen.laplacian_2d(30)]
    B = mek.Batch(mats, [0.5, 2.0], [20, 20], 1e-8)
    fb = B.features()
    for i, (rp, ci, va) in enumerate(mats):
        fs = mek.features(mek.CSR.from_arrays(rp, ci, va), [0.5, 2.0][i], 1e-8)
        np.testing.assert_allclose(fb[i], fs, rtol=1e-5, atol=1e-6)


def test_trained_selector_gains_over_heuristic(gpu):
    """The committed selector (trained on the GPU by train_selector.py) against
    the method without it, on problems it has not seen: the same batched
    mixed-precision Expokit solve with (m = 30, Expokit's first step) and with
    the predicted (m, first step).  The learned choice must move at least 10 %
    fewer modelled HBM bytes (the quantity it was trained to minimise),
    converge everywhere and keep the relative error within 1e-8 of a tight
    FP64 solve."""
    import importlib.util
    import os
    here = os.path.join(os.path.dirname(__file__), "..", "matrix-exp-krylov_amd")
    spec = importlib.util.spec_from_file_location("mek_selector", os.path.join(here, "selector.py"))
    sel = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(sel)
    S = sel.Selector.load()
    tol = 1e-8
    probs = sel.training_set(48, seed=77, n_lo=5000, n_hi=60000)
    mats = [(rp, ci, va) for rp, ci, va, t, kind in probs]
    ts = [t for *_x, t, kind in probs]
    vs = [gen.normal_vector(rp.size - 1, seed=500 + i) for i, (rp, *_r) in enumerate(mats)]
    csrs = [mek.CSR.from_arrays(*A) for A in mats]
    F = np.array([mek.features(c, t, tol) for c, t in zip(csrs, ts)])
    learned = S.predict_features(F)
    refs = []
    for c, t, v, (rp, *_r) in zip(csrs, ts, vs, mats):
        R = mek.Runner(c, m=60, t=t, tol=1e-13, max_restarts=0, step_control="expokit_rel", gs_mode=1)
        refs.append(R.expmv(v))
        R.close()
    res = {}
    for name, choice in (("heuristic", [(30, 0)] * len(mats)), ("learned", learned)):
        B = mek.Batch(csrs, ts, [m for m, s in choice], [tol / t for t in ts],
                      substeps=[s for m, s in choice], values_f32=True, gs_mode=2,
                      step_control="expokit_rel")
        B.upload(vs)
        B.solve()
        ys = B.download()
        cols = by = 0.0
        errs = []
        for p, ((rp, ci, va), (m, s)) in enumerate(zip(mats, choice)):
            st = B.stats(p)
            assert st["converged"] == 1, (name, p)
            cols += st["spmv_count"]
            by += sel.modelled_bytes(m, 1, rp.size - 1, int(rp[-1])) * st["spmv_count"] / m
            r = refs[p]
            if np.linalg.norm(r) >= 1e-3:
                errs.append(np.linalg.norm(ys[p] - r) / np.linalg.norm(r))
        res[name] = (cols, by, max(errs))
        B.close()
    print(f"[selector] heuristic cols={res['heuristic'][0]:.0f} bytes={res['heuristic'][1]:.3e} "
          f"err={res['heuristic'][2]:.2e}; learned cols={res['learned'][0]:.0f} "
          f"bytes={res['learned'][1]:.3e} err={res['learned'][2]:.2e}", flush=True)
    assert res["learned"][1] < 0.9 * res["heuristic"][1]
    assert res["learned"][2] <= 1e-8 and res["heuristic"][2] <= 1e-8
