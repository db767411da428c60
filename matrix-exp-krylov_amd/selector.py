"""Learned (m, substeps) selector for the Arnoldi exp(t*A)*v solver.

Pipeline (every GPU step runs through libmek.so):

1. ``training_set``: synthetic matrices (2-D Laplacians, 3-D upwind
   convection-diffusion, random Gershgorin-negative sparse rows) with t chosen
   so that ||tA||_1 is log-uniform in [1, 1e4].
2. ``label``: one m=61 Arnoldi decomposition of each problem on the GPU; for
   every m on a grid the largest first sub-step dt that Expokit's step
   control accepts (local error of the augmented exponential, err_loc <=
   1.2 dt tol; bisection), hence substeps s_m = ceil(t/dt).  Every (m, s_m)
   is solved by the batched mixed-precision solver, and the label is the m
   (and its s_m) whose solve takes the least modelled time: the HBM bytes of
   its m + 1 SpMVs per sub-step at the SpMV's bandwidth (gather-bound for
   random rows, HBM-bound for stencils), of its 2 Gram-Schmidt sweeps per
   column and its lift at the sweeps' bandwidth.  The SpMV-minimal choice and
   every candidate's column count are recorded beside it.
3. ``train``: features on the GPU (``mek.features``), standardised; the MLP
   (16 -> 64 -> 64 -> 2) is trained on the GPU by the library's own kernels.
4. ``Selector.predict``: (m, substeps) per problem; the batched solver runs
   them under Expokit step control with t/substeps as the first step, so the
   local error estimate checks every step and a rejected step falls back to
   Expokit's own step formula: a poor prediction costs time, not accuracy.

The heuristic it is compared with: the problem's own m and Expokit's initial
step (Sidje 1998): dt = (1/||A||) ((fact tol) / (4 beta ||A||))^(1/m),
fact = ((m+1)/e)^(m+1) sqrt(2 pi (m+1)), then Expokit's step control.
"""
from __future__ import annotations

import json
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


def _gen():
    """The matrix generator alone (no GPU library), for CPU worker processes."""
    import importlib.util
    import sys
    if "mek_generate" not in sys.modules:
        spec = importlib.util.spec_from_file_location("mek_generate", os.path.join(HERE, "generate.py"))
        mod = importlib.util.module_from_spec(spec)
        sys.modules["mek_generate"] = mod
        spec.loader.exec_module(mod)
    return sys.modules["mek_generate"]


def norm1(rp, ci, va, n):
    return float(np.bincount(ci, weights=np.abs(va), minlength=n).max())


def problem_specs(count: int, seed: int, n_lo: int = 1000, n_hi: int = 30000):
    """Light descriptions (index, kind, size, Pe, ||tA||_1) of a training set,
    drawn in a fixed order from one generator; ``make_problem`` builds each.
    Sets of large matrices are built a part at a time from these."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = str(rng.choice(["lap2d", "cd3d", "rows"], p=[0.3, 0.3, 0.4]))
        pe = 0.0
        if kind == "lap2d":
            size = int(round(math.sqrt(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))))))
        elif kind == "cd3d":
            size = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))) ** (1 / 3)))
            pe = float(math.exp(rng.uniform(math.log(0.5), math.log(100.0))))
        else:
            size = int(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))))
        tau = math.exp(rng.uniform(0.0, math.log(1e4)))
        out.append((i, kind, size, pe, tau))
    return out


def make_problem(spec, seed: int, f32_exact: bool = True):
    """(row_ptr, col_idx, values, t, kind) of one spec; t makes ||tA||_1 = tau.
    With f32_exact the values are rounded to FP32 first (the labelled
    training sets were built so); without it they stay as generated (random
    rows: U[-1,1] in FP64)."""
    gen = _gen()
    i, kind, size, pe, tau = spec
    if kind == "lap2d":
        rp, ci, va = gen.laplacian_2d(max(size, 8))
    elif kind == "cd3d":
        rp, ci, va = gen.convdiff_3d(max(size, 5), pe=pe)
    else:
        rp, ci, va = gen.random_rows(size, 5, 30, seed=seed * 100003 + i)
    if f32_exact:
        va = va.astype(np.float32).astype(np.float64)
    n = rp.size - 1
    return (rp, ci, va, tau / norm1(rp, ci, va, n), kind)


def spec_rows(spec) -> int:
    """Rows of the matrix a spec builds."""
    _, kind, size, _, _ = spec
    if kind == "lap2d":
        return max(size, 8) ** 2
    if kind == "cd3d":
        return max(size, 5) ** 3
    return size


def training_set(count: int, seed: int, n_lo: int = 1000, n_hi: int = 30000, f32_exact: bool = True):
    """List of (row_ptr, col_idx, values, t, kind) with ||tA||_1 log-uniform in
    [1, 1e4] (2-D Laplacians, 3-D upwind convection-diffusion with Pe in
    [0.5, 100], random rows with 5-30 nonzeros)."""
    return [make_problem(sp, seed, f32_exact) for sp in problem_specs(count, seed, n_lo, n_hi)]


def modelled_bytes(m: int, s: int, n: int, nnz: int, value_bytes: int = 4, sweeps: int = 2) -> float:
    """HBM bytes of s sub-steps with an m-column basis: m + 1 SpMVs (i32
    indices, values of value_bytes), `sweeps` Gram-Schmidt sweeps per column
    (two: one dot sweep and one update sweep when the conditional second pass
    is not taken) and the lift."""
    bspmv = (value_bytes + 4.0) * nnz + 4.0 * (n + 1) + 16.0 * n
    gs = sweeps * 8.0 * n * m * (m + 1) / 2 + (2 * sweeps + 1) * 8.0 * n * m
    return s * ((m + 1) * bspmv + gs + 8.0 * n * (m + 2))


# Bandwidths of the batched mixed-precision solver's kernels on one MI355X
# (rocprofv3 kernel trace of the 256-problem batch): the SpMV of rows with
# random column indices is bound by its x gathers, a banded stencil's SpMV and
# the Gram-Schmidt sweeps by HBM.
BW_SPMV_GATHER = 2.0e12
BW_SPMV_BANDED = 4.0e12
BW_SWEEP = 4.3e12


def modelled_seconds(m: int, s: int, n: int, nnz: int, banded: bool, value_bytes: int = 4,
                     sweeps: int = 2) -> float:
    """The bytes of modelled_bytes, each kernel's at its own bandwidth: the
    SpMV's at BW_SPMV_BANDED or BW_SPMV_GATHER, the sweeps' and the lift's at
    BW_SWEEP."""
    bspmv = (value_bytes + 4.0) * nnz + 4.0 * (n + 1) + 16.0 * n
    gs = sweeps * 8.0 * n * m * (m + 1) / 2 + (2 * sweeps + 1) * 8.0 * n * m
    bw = BW_SPMV_BANDED if banded else BW_SPMV_GATHER
    return s * ((m + 1) * bspmv / bw + (gs + 8.0 * n * (m + 2)) / BW_SWEEP)


COSTS = os.path.join(HERE, "selector_costs.json")


def load_step_costs(path=COSTS):
    """Measured step costs of the batched solver (train_selector.py
    --calibrate), or None when the file is absent."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def step_seconds(costs, kind: str, m: int, n: int) -> float:
    """Measured seconds of one m-column step of a problem of `kind` with n
    rows: the per-row cost at the two calibrated sizes, linear in m between
    grid points and in log n between the sizes (clamped outside both)."""
    tab = costs[kind]
    ms = sorted(int(k) for k in tab)
    n0, n1 = costs["n"]
    w = min(max((math.log(n) - math.log(n0)) / (math.log(n1) - math.log(n0)), 0.0), 1.0)
    c0 = float(np.interp(m, ms, [tab[str(k)][0] for k in ms]))
    c1 = float(np.interp(m, ms, [tab[str(k)][1] for k in ms]))
    return ((1.0 - w) * c0 + w * c1) * n


def label_cost(m: int, cols: int, n: int, nnz: int, kind: str, costs=None) -> float:
    """Seconds of a solve that built `cols` Krylov columns with an m-column
    basis (cols / m steps): each step at its measured cost when step costs
    are given, else each kernel's bytes at its bandwidth (modelled_seconds)."""
    if costs is not None:
        return step_seconds(costs, kind, m, n) * cols / m
    return modelled_seconds(m, 1, n, nnz, kind != "rows") * cols / m


def relabel(z, costs):
    """New targets of a labelled set from its recorded candidates (m, first
    sub-steps, columns built): the least label_cost under `costs`.  Returns
    (T, m_b, s_b)."""
    P = len(z["n"])
    T = np.zeros((P, 2), np.float32)
    m_b = np.zeros(P, np.int32)
    s_b = np.zeros(P, np.int64)
    for p in range(P):
        best = None
        for m, s, c in zip(z["cand_m"][p], z["cand_s"][p], z["cand_cols"][p]):
            if m < 0:
                continue
            cost = label_cost(int(m), int(c), int(z["n"][p]), 0, str(z["kind"][p]), costs)
            if best is None or cost < best[0]:
                best = (cost, int(m), int(s))
        _, m_b[p], s_b[p] = best
        T[p] = encode(m_b[p], s_b[p])
    return T, m_b, s_b


DELTA = 1.2  # Expokit's acceptance safety factor: err_loc <= DELTA * dt * tol


def expokit_err(H: np.ndarray, m: int, dt: float, avnorm: float, beta: float = 1.0):
    """Expokit's local error estimate of one step of length dt with the first m
    Arnoldi columns of H ((k+1) x k, k > m): exponential of the augmented
    (m+2) matrix, phi1 = |beta F_m|, phi2 = |beta F_{m+1} ||A v_{m+1}|||.
    Returns (err_loc, ||F(0:m)||), the second being ||w_end|| / beta."""
    aug = np.zeros((m + 2, m + 2))
    aug[:m + 1, :m] = H[:m + 1, :m]
    aug[m + 1, m] = 1.0
    F = scipy.linalg.expm(dt * aug)[:, 0]
    phi1 = abs(beta * F[m])
    phi2 = abs(beta * F[m + 1] * avnorm)
    fn = float(np.linalg.norm(F[:m + 1]))
    if phi1 > 10.0 * phi2:
        return phi2, fn
    if phi1 > phi2:
        return phi1 * phi2 / (phi1 - phi2), fn
    return phi1, fn


def max_substep(H: np.ndarray, m: int, t: float, tol: float) -> float:
    """Largest dt in (0, t] that Expokit's step control in relative mode
    accepts as a first step with an m-column basis (geometric bisection to 2
    %): err_loc <= 1.2 dt tol min(1, ||w_end||) for ||v|| = 1."""
    avnorm = float(np.linalg.norm(H[:m + 2, m]))  # ||A v_{m+1}|| = ||H[:, m]|| (orthonormal V)

    def ok(dt):
        err, fn = expokit_err(H, m, dt, avnorm)
        return err <= DELTA * dt * tol * min(1.0, fn)
    if ok(t):
        return t
    lo, hi = t * 1e-9, t
    if not ok(lo):
        return lo
    while hi / lo > 1.02:
        mid = math.sqrt(lo * hi)
        if ok(mid):
            lo = mid
        else:
            hi = mid
    return lo


STEP_SAFETY = 0.8  # predicted first step = 0.8 x the largest accepted one


def substeps_per_m(H: np.ndarray, k: int, t: float, tol: float) -> dict:
    """{m: substeps} for every m of the grid below the Krylov dimension k
    reached by the decomposition H; s_m = ceil(t / (0.8 dt_m)) with dt_m the
    largest first step Expokit accepts at the per-unit-time tolerance tol/t.
    A happy breakdown (k below the decomposition's size) gives {k: 1}."""
    out = {}
    for m in M_GRID:
        if m >= k:
            out[k] = 1
            break
        dt = max_substep(H, m, t, tol / t)
        out[m] = max(1, math.ceil(t / (STEP_SAFETY * dt) - 1e-12))
    return out


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


def build_dataset(problems, tol=1e-8, device=0, log=None, chunk=256):
    """Features and labels of every problem.  For each m of the grid the
    first sub-step count s_m comes from one m=61 Arnoldi decomposition
    (substeps_per_m); then every (m, s_m) is actually solved, a chunk of
    problems at a time, by the batched mixed-precision solver under Expokit
    step control (tol/t), and the label is the m whose solve takes the least
    modelled time (label_cost: each kernel's bytes at its bandwidth), with
    its s_m.  The fewest-SpMV choice is recorded beside it, and so is every
    candidate (m, s_m, columns built)."""
    mek, gen = _mods()
    costs = load_step_costs()
    X, T, meta = [], [], []
    for c0 in range(0, len(problems), chunk):
        part = problems[c0:c0 + chunk]
        csrs, vs, cand = [], [], []
        for i, (rp, ci, va, t, kind) in enumerate(part):
            n = rp.size - 1
            A = mek.CSR.from_arrays(rp, ci, va)
            v = gen.normal_vector(n, seed=c0 + i)
            R = mek.Runner(A, m=min(M_MAX + 1, n - 1), t=t, tol=tol, device=device)
            H, k = R.arnoldi(v)
            R.close()
            csrs.append(A)
            vs.append(v)
            cand.append(substeps_per_m(H, k, t, tol))
        ts = [p[3] for p in part]
        X.extend(batch_features(csrs, ts, tol, device))
        cols = [dict() for _ in part]
        for m in M_GRID:
            idx = [i for i in range(len(part)) if m in cand[i]]
            if not idx:
                continue
            B = mek.Batch([csrs[i] for i in idx], [ts[i] for i in idx], [m] * len(idx),
                          [tol / ts[i] for i in idx], substeps=[cand[i][m] for i in idx],
                          values_f32=True, gs_mode=2, step_control="expokit_rel", device=device)
            B.upload([vs[i] for i in idx])
            B.solve()
            for q, i in enumerate(idx):
                cols[i][m] = B.stats(q)["spmv_count"]
            B.close()
        for i, (rp, ci, va, t, kind) in enumerate(part):
            n, nnz = rp.size - 1, int(rp[-1])
            # breakdown-sized candidates (m not on the grid) are solved in one step
            for m in cand[i]:
                cols[i].setdefault(m, m)
            cost = {m: label_cost(m, c, n, nnz, kind, costs) for m, c in cols[i].items()}
            m_b = min(cost, key=lambda mm: cost[mm])
            m_s = min(cols[i], key=lambda mm: cols[i][mm])
            T.append(encode(m_b, cand[i][m_b]))
            # every candidate (m, first sub-steps, columns built) travels with
            # the label, so other cost models can relabel without solving again
            meta.append((kind, n, t, m_b, cand[i][m_b], m_s, cand[i][m_s],
                         [(mm, cand[i][mm], cols[i][mm]) for mm in sorted(cols[i])]))
        for A in csrs:
            A.close()
        if log:
            log(f"labelled {min(c0 + chunk, len(problems))}/{len(problems)}")
    return np.array(X, np.float32), np.array(T, np.float32), meta


def batch_features(csrs, ts, tol=1e-8, device=0):
    """The 16 features of every (A_p, t_p) in one device pass (a batch of the
    already-built CSR handles; the features do not depend on m)."""
    mek, _ = _mods()
    B = mek.Batch(csrs, ts, [min(30, c.info()["rows"] - 1) for c in csrs], tol, values_f32=True,
                  device=device)
    try:
        return B.features()
    finally:
        B.close()


def train(X, T, epochs=300, batch=128, lr=2e-3, seed=0, device=0):
    mek, _ = _mods()
    mean = X.mean(0)
    std = X.std(0) + 1e-6
    net = mek.Mlp(seed=seed, device=device)
    loss = net.train((X - mean) / std, T, batch=batch, epochs=epochs, lr=lr, seed=seed)
    sel = Selector(net.params(), mean, std, device=device)
    net.close()
    return sel, loss
