"""Learned (m, substeps) selector for the Arnoldi exp(t*A)*v solver.

Pipeline (every GPU step runs through libmek.so):

1. ``training_set``: synthetic matrices (2-D Laplacians, 3-D upwind
   convection-diffusion, random Gershgorin-negative sparse rows) with t chosen
   so that ||tA||_1 is log-uniform in [1, 1e4].
2. ``label``: one m=60 Arnoldi decomposition of each problem on the GPU; for
   every m on a grid the largest first sub-step dt with the reference's
   residual estimate |h_{m+1,m} [exp(dt H_m) e1]_m| <= tol (bisection), hence
   substeps s_m = ceil(t/dt).  The label is the m (and its s_m) that minimises
   the modelled HBM bytes of the whole solve, s_m * (m * B_spmv + 3 sweeps of
   Gram-Schmidt); the SpMV-minimal choice is recorded beside it.
3. ``train``: features on the GPU (``mek.features``), standardised; the MLP
   (16 -> 64 -> 64 -> 2) is trained on the GPU by the library's own kernels.
4. ``Selector.predict``: (m, substeps) per problem; the batched solver runs
   them with the reference controller, whose a-posteriori residual estimate
   judges every step (a rejected step shrinks dt), so a poor prediction costs
   time, not accuracy.

The heuristic it is compared with: m = 30 (the reference CLI's default) and
the number of sub-steps from the Expokit initial step (Sidje 1998):
dt = (1/||A||) ((fact tol) / (4 beta ||A||))^(1/m), fact = ((m+1)/e)^(m+1)
sqrt(2 pi (m+1)).
"""
from __future__ import annotations

import math
import os

import numpy as np
import scipy.linalg

M_GRID = (10, 15, 20, 25, 30, 35, 40, 45, 50, 55, 60)
M_MAX = 60
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "selector_weights.npz")


def _mods():
    import importlib.util
    import sys
    out = []
    for name, fn in (("mek", "__init__.py"), ("mek_generate", "generate.py")):
        if name not in sys.modules:
            spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, fn))
            mod = importlib.util.module_from_spec(spec)
            sys.modules[name] = mod
            spec.loader.exec_module(mod)
        out.append(sys.modules[name])
    return out


def norm1(rp, ci, va, n):
    return float(np.bincount(ci, weights=np.abs(va), minlength=n).max())


def training_set(count: int, seed: int, n_lo: int = 1000, n_hi: int = 30000):
    """List of (row_ptr, col_idx, values, t, kind) with ||tA||_1 log-uniform in [1, 1e4]."""
    _, gen = _mods()
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = rng.choice(["lap2d", "cd3d", "rows"], p=[0.3, 0.3, 0.4])
        if kind == "lap2d":
            N = int(round(math.sqrt(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))))))
            rp, ci, va = gen.laplacian_2d(max(N, 8))
        elif kind == "cd3d":
            N = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))) ** (1 / 3)))
            pe = float(math.exp(rng.uniform(math.log(0.5), math.log(100.0))))
            rp, ci, va = gen.convdiff_3d(max(N, 5), pe=pe)
        else:
            n = int(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))))
            rp, ci, va = gen.random_rows(n, 5, 30, seed=seed * 100003 + i)
        n = rp.size - 1
        tau = math.exp(rng.uniform(0.0, math.log(1e4)))
        t = tau / norm1(rp, ci, va, n)
        out.append((rp, ci, va, t, str(kind)))
    return out


def modelled_bytes(m: int, s: int, n: int, nnz: int) -> float:
    """HBM bytes of s sub-steps with an m-column basis: SpMV (FP64 values,
    i32 indices) + three Gram-Schmidt sweeps per column + lift."""
    bspmv = 12.0 * nnz + 4.0 * (n + 1) + 16.0 * n
    gs = 3 * 8.0 * n * m * (m + 1) / 2 + 5 * 8.0 * n * m
    return s * (m * bspmv + gs + 8.0 * n * (m + 2))


def max_substep(Hm: np.ndarray, h_next: float, t: float, tol: float) -> float:
    """Largest dt in (0, t] with |h_next [exp(dt H_m) e1]_{m-1}| <= tol (beta = 1)."""
    def err(dt):
        return abs(h_next * scipy.linalg.expm(dt * Hm)[-1, 0])
    if err(t) <= tol:
        return t
    lo, hi = 0.0, t
    for _ in range(40):
        mid = 0.5 * (lo + hi) if lo > 0 else hi * 1e-3
        if err(mid) <= tol:
            lo = mid
        else:
            hi = mid
        if lo > 0 and hi / lo < 1.02:
            break
    return max(lo, t * 1e-9)


def label(rp, ci, va, t, tol=1e-8, seed=0, device=0):
    """(m, substeps) minimising modelled bytes, and the SpMV-minimal (m, s)."""
    mek, gen = _mods()
    n = rp.size - 1
    nnz = int(rp[-1])
    m_top = min(M_MAX, n - 1)
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=m_top, t=t, tol=tol, device=device)
    H, k = R.arnoldi(gen.normal_vector(n, seed=seed))
    R.close()
    best_b = best_s = None
    for m in M_GRID:
        if m > k:
            break
        Hm = H[:m, :m]
        h_next = H[m, m - 1] if m < H.shape[0] else 0.0
        if m == k and k < m_top:  # happy breakdown: exact in one step
            s = 1
        else:
            s = max(1, math.ceil(t / max_substep(Hm, h_next, t, tol) - 1e-12))
        cb = modelled_bytes(m, s, n, nnz)
        if best_b is None or cb < best_b[2]:
            best_b = (m, s, cb)
        if best_s is None or m * s < best_s[0] * best_s[1]:
            best_s = (m, s)
    return best_b[0], best_b[1], best_s


def heuristic(anorm: float, t: float, tol: float = 1e-8, m: int = 30, beta: float = 1.0):
    """(m, substeps) from the Expokit initial step size."""
    fact = ((m + 1) / math.e) ** (m + 1) * math.sqrt(2 * math.pi * (m + 1))
    dt = (1.0 / anorm) * ((fact * tol) / (4.0 * beta * anorm)) ** (1.0 / m)
    return m, max(1, math.ceil(t / dt))


def encode(m, s):
    return np.array([(m - 35.0) / 25.0, math.log2(s) / 8.0], np.float32)


def decode(y):
    m = int(min(max(round(y[0] * 25.0 + 35.0), 10), 60))
    s = int(min(max(round(2.0 ** (y[1] * 8.0)), 1), 1 << 20))
    return m, s


class Selector:
    """Standardisation + the trained MLP."""

    def __init__(self, params, mean, std, device=0):
        mek, _ = _mods()
        self.net = mek.Mlp(seed=0, device=device)
        self.net.set_params(params)
        self.mean = np.asarray(mean, np.float32)
        self.std = np.asarray(std, np.float32)

    @classmethod
    def load(cls, path=WEIGHTS, device=0):
        z = np.load(path)
        return cls(z["params"], z["mean"], z["std"], device=device)

    def save(self, path=WEIGHTS):
        np.savez(path, params=self.net.params(), mean=self.mean, std=self.std)

    def predict_features(self, F):
        Y = self.net.forward((np.asarray(F, np.float32) - self.mean) / self.std)
        return [decode(y) for y in Y]


def build_dataset(problems, tol=1e-8, device=0, log=None):
    mek, _ = _mods()
    X, T, meta = [], [], []
    for i, (rp, ci, va, t, kind) in enumerate(problems):
        m, s, (ms, ss) = label(rp, ci, va, t, tol, seed=i, device=device)
        X.append(mek.features(mek.CSR.from_arrays(rp, ci, va), t, tol, device=device))
        T.append(encode(m, s))
        meta.append((kind, rp.size - 1, t, m, s, ms, ss))
        if log and (i + 1) % 100 == 0:
            log(f"labelled {i + 1}/{len(problems)}")
    return np.array(X, np.float32), np.array(T, np.float32), meta


def train(X, T, epochs=300, batch=128, lr=2e-3, seed=0, device=0):
    mek, _ = _mods()
    mean = X.mean(0)
    std = X.std(0) + 1e-6
    net = mek.Mlp(seed=seed, device=device)
    loss = net.train((X - mean) / std, T, batch=batch, epochs=epochs, lr=lr, seed=seed)
    sel = Selector(net.params(), mean, std, device=device)
    net.close()
    return sel, loss
