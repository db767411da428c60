"""Throughput of the MI355X Arnoldi exp(t*A)*v solver on one GPU.

A step is one full solve y = exp(t*A)*v (ArnoldiRunner::compute_expmv on
device-resident v and y).  Default workload: 2-D 5-point Laplacian on a
4096 x 4096 grid (n = 16,777,216, nnz = 83,869,696, FP64 values, i32
indices), v ~ N(0,1) normalised (seed 0), t = 10, m = 40, tol 1e-8, sub-steps
by the Expokit heuristic.  FP64 Arnoldi with the basis grown by s-step
blocks of 16 vectors after the first 8 columns of each cycle (--sstep 0 gives
the reference's column-by-column Gram-Schmidt).

Prints one JSON line with
* value: solves/s of that workload, v and y resident in HBM, timed with no
  per-kernel events in the loop; beside it the same solve on the reference's
  METRIC span (host v in, host y out: main.cu:1412-1570);
* rel_err against the DST-I closed form;
* roofline: the dominant kernel's algorithmic bytes per launch over its
  HIP-event launch time on the solver's own stream (a separate pass with one
  event pair per launch), and the SpMV's;
* learned_mixed_batched: 256 independent problems (n log-uniform in
  [1e3, 2e5], random rows, U[-1,1] FP64 values) as plain FP64 solves one at a
  time, as one batched FP64 solve, with the FP32-value request (which falls
  back to FP64 storage because these values are not exact in FP32), with the
  learned (m, first sub-step) choice (selection time included), and with FP32
  storage forced (reported, not a 1e-8 form); speed-up over plain FP64;
* cpu_baseline: the oracle (one thread) on one full sub-step of the same
  solve (40 columns, exponential, lift), timed, times the sub-step count;
  and the reference's own CPU case (n = 2000 random matrix, m = 30, t = 1,
  tol 1e-6, 10 restarts), median of 3, beside the GPU on the same case.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0
METRIC = "exp(tA)v solves/s + SpMV HBM GB/s at rel err <=1e-8; learned+mixed vs plain FP64"


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


WORKLOADS = {
    "laplacian4096": dict(kind="laplacian", N=4096, t=10.0, m=40, tol=1e-8, f32=False),
    "batch256": dict(kind="batch", P=256, tol=1e-8),
    "block2048": dict(kind="block", N=2048, t=10.0, m=16, tol=1e-8),
    "convdiff256_f32": dict(kind="convdiff", N=256, t=0.5, m=50, tol=1e-8, f32=True),
    "convdiff256_f64": dict(kind="convdiff", N=256, t=0.5, m=50, tol=1e-8, f32=False),
}


def _finite(x):
    """JSON-safe float (None for inf / NaN)."""
    return float(x) if x is not None and math.isfinite(x) else None


def reference_solution(w, v, gen):
    import scipy.fft  # noqa: F401
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from closed_forms import kron3_expmv, laplacian_2d_expmv
    if w["kind"] == "laplacian":
        return laplacian_2d_expmv(w["N"], w["t"], v)
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    N = w["N"]
    return kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                       gen.tridiag_1d(N, lx, dx, ux), w["t"], v)


def measured_traffic(workload):
    """HBM bytes per launch by kernel kind from the committed counter runs
    (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated), or {} when the workload
    has none."""
    fn = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(fn):
        return {}
    with open(fn) as f:
        return json.load(f).get("bytes_per_launch", {})


def _host_desc(core):
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return f"1 thread pinned to core {core} of {os.cpu_count()} ({model or 'unknown CPU'})"


class _one_core:
    """Pin the calling thread to one core (the reference's CPU binary is
    single-threaded), restoring the affinity afterwards."""

    def __enter__(self):
        self.prev = os.sched_getaffinity(0)
        self.core = min(self.prev)
        os.sched_setaffinity(0, {self.core})
        return self.core

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.prev)


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    return pyoracle


def cpu_baseline_substep(rp, ci, va, v, m, dt, substeps):
    """The oracle's reference solver (main_cpu.cpp:602-700 restated in C: MGS
    with two passes, one thread) on one full sub-step of the GPU solve: one
    attempt of length dt from the same start vector (m columns, the
    exponential, the lift; the tolerance is set so that the attempt is
    accepted), timed whole.  The solve is `substeps` such sub-steps of the same
    cost, so its rate is 1 / (substeps x that time)."""
    pyoracle = _oracle()
    with _one_core() as core:
        t0 = time.perf_counter()
        _y, st, _log = pyoracle.expmv(rp, ci, va, v, dt, 1e300, m, 1)
        wall = time.perf_counter() - t0
    assert st["attempts"] == 1 and st["spmv_count"] == m, st
    return {
        "value": 1.0 / (substeps * wall),
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "host": _host_desc(core),
        "sample": (f"oracle solver (main_cpu.cpp:602-700 restated in C: SpMV, MGS with two passes, "
                   f"Pade expm, lift) on one full sub-step of the same solve (dt = {dt:.4g}, {m} columns "
                   f"from the same start vector), {wall:.1f} s wall; the GPU solve is {substeps} such "
                   f"sub-steps"),
    }


def cpu_baseline_batch(mats, ts, ms, vs, tol, seconds_budget):
    """The oracle's reference solver (MGS2, reference step controller with the
    first step t, tol absolute per step) on the batch's problems in turn, one
    thread, until the budget is spent: problems/s on that sample."""
    pyoracle = _oracle()
    with _one_core() as core:
        t0 = time.perf_counter()
        k = conv = 0
        for (rp, ci, va), t, m, v in zip(mats, ts, ms, vs):
            _y, st, _log = pyoracle.expmv(rp, ci, va, v, t, tol, m, 0)
            conv += int(st["converged"])
            k += 1
            if time.perf_counter() - t0 > seconds_budget:
                break
        wall = time.perf_counter() - t0
    rows = sum(int(rp.size - 1) for rp, _, _ in mats[:k])
    return {
        "value": k / wall,
        "unit": "problems/s",
        "cores": 1,
        "kind": "port",
        "host": _host_desc(core),
        "sample": (f"oracle solve (main_cpu.cpp:602-700 restated in C: MGS2, reference controller, "
                   f"first step t, tol {tol:g} per step, each problem's m) of the first {k} of the "
                   f"{len(mats)} problems ({rows} rows, {conv} converged) in turn, {wall:.1f} s wall"),
    }


def random_n2000_case(mek, gen, with_cpu):
    """The reference's own CPU-runnable case: n = 2000, int(5e-3 n^2)
    uniformly random off-diagonal draws (duplicates summed), values U[-1,1]
    from seed 42, a_ii = -1 - sum_j |a_ij|; v = 0.1 ones; t = 1.0 (exact, not
    the CLI's clamp), m = 30, tol 1e-6, 10 restarts, the reference
    controller.  GPU: median of 5 solves after one warm-up; oracle: median of
    3 runs on one core.  Relative error of the GPU y against the oracle's."""
    rp, ci, va = gen.random_negative(2000, 5e-3, seed=42)
    n = rp.size - 1
    v = np.full(n, 0.1)
    t, m, tol, R_ = 1.0, 30, 1e-6, 10
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=m, t=t, tol=tol, max_restarts=R_)
    y = R.expmv(v)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        y = R.expmv(v)
        times.append(time.perf_counter() - t0)
    st = R.stats()
    R.close()
    out = {"workload": "random n=2000, density 5e-3, seed 42; v = 0.1; t = 1, m = 30, tol 1e-6, "
                       "10 restarts, reference controller",
           "nnz": int(rp[-1]), "gpu_solves_per_s": 1.0 / statistics.median(times),
           "gpu_span": "compute_expmv, host v and y (the METRIC span)",
           "accepted_steps": st["accepted"], "attempts": st["attempts"], "converged": bool(st["converged"])}
    if with_cpu:
        pyoracle = _oracle()
        cpu = []
        with _one_core() as core:
            for _ in range(3):
                t0 = time.perf_counter()
                yo, sto, _ = pyoracle.expmv(rp, ci, va, v, t, tol, m, R_)
                cpu.append(time.perf_counter() - t0)
        out["cpu_solves_per_s"] = 1.0 / statistics.median(cpu)
        out["cpu"] = {"kind": "port", "cores": 1, "host": _host_desc(core),
                      "sample": "the whole solve, median of 3 runs"}
        out["rel_err_vs_oracle"] = float(np.linalg.norm(y - yo) / np.linalg.norm(yo))
        out["oracle_steps"] = {"accepted": sto["restarts_done"], "attempts": sto["attempts"]}
    return out


# --------------------------------------------------------------------- batch

def batch_comparison(args, mek, gen, reps, cpu_seconds):
    """256 independent problems: n_i = round(exp(U[ln 1e3, ln 2e5])), random
    rows with 5-30 nonzeros, values U[-1,1] in FP64, Gershgorin-negative
    diagonal, t_i ~ U[0.1, 5], m_i ~ U{20..60}, v_i ~ N(0,1) normalised (seed 0).

    Forms (all DGKS second pass, Expokit step control with the per-unit-time
    tolerance tol/t_i relative to each step's start norm):
    * plain_fp64: one ArnoldiRunner solve per problem in turn (m_i), v and y
      resident, column by column (the reference's Gram-Schmidt order);
    * plain_fp64_sstep: the same with s-step blocks (mek.sstep_plan(m_i));
      speedup_vs_best_plain_fp64 divides by the faster of the two;
    * batched_fp64: every problem in one lockstep solve replayed from hipGraphs;
    * batched_mixed: the same with FP32 values requested -- stored in FP32 only
      where that is exact, so here FP64 (values_f32_used says which);
    * batched_mixed_heuristic: what the selector replaces, m = 30 for every
      problem and Expokit's own first step;
    * batched_mixed_learned: (m, first sub-step) from the learned selector; its
      problems/s includes the selection (device features, MLP, decoding);
    * batched_f32_forced: FP32 storage regardless, to show what it costs in
      error (not a form that meets 1e-8).
    Errors are against a tight (1e-13, m = 60, second pass always) FP64 solve
    of each problem.  setup_s (host assembly and upload, excluded from the
    rates) is reported per form."""
    sel = _load("mek_selector", os.path.join(ROOT, "matrix-exp-krylov_amd", "selector.py"))
    tol = 1e-8
    log("batch: generating 256 problems (FP64 values)")
    probs = gen.batch_problems(P=256, seed=0, f32_exact=False)
    P = len(probs)
    mats = [(q["rp"], q["ci"], q["va"]) for q in probs]
    ts = [q["t"] for q in probs]
    ms = [min(q["m"], q["rp"].size - 2) for q in probs]
    vs = [q["v"] for q in probs]
    rows = [q["rp"].size - 1 for q in probs]
    nnz_tot = int(sum(int(q["rp"][-1]) for q in probs))
    csrs = [mek.CSR.from_arrays(*A) for A in mats]
    chk = [c.f32_inexact(nonfinite=True) for c in csrs]
    f32_check = {"inexact": int(sum(c for c, _, _ in chk)), "nonfinite": int(sum(b for _, _, b in chk)),
                 "max_rel_change": _finite(max(m for _, m, _ in chk)), "values": nnz_tot}
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    tols = [tol / t for t in ts]

    log("batch: tight FP64 reference solves (tol 1e-13)")
    refs = []
    for c, t, v, n in zip(csrs, ts, vs, rows):
        R = mek.Runner(c, m=min(60, n - 1), t=t, tol=1e-13, max_restarts=0, step_control="expokit_rel",
                       gs_mode=1)
        refs.append(R.expmv(v))
        R.close()

    def max_err(ys):
        return max(float(np.linalg.norm(y - r) / max(np.linalg.norm(r), 1e-300)) for y, r in zip(ys, refs))

    log("batch: plain FP64, one problem at a time")
    t0 = time.perf_counter()
    runners = [mek.Runner(c, m=m, t=t, tol=tl, max_restarts=0, step_control="expokit_rel", gs_mode=gs_mode)
               for c, m, t, tl in zip(csrs, ms, ts, tols)]
    dv = [mek.DeviceBuffer.from_array(v) for v in vs]
    dy = [mek.DeviceBuffer(8 * n) for n in rows]
    setup_plain = time.perf_counter() - t0
    for R, a, b in zip(runners, dv, dy):  # warm-up
        R.expmv_device(a, b)
    mek.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for R, a, b in zip(runners, dv, dy):
            R.expmv_device(a, b)
    mek.synchronize()
    plain_s = (time.perf_counter() - t0) / reps
    plain_cols = sum(R.stats()["spmv_count"] for R in runners)
    plain_err = max_err([b.to_array() for b in dy])
    for R in runners:
        R.close()

    log("batch: plain FP64 with s-step blocks, one problem at a time")
    runners = []
    for c, m, t, tl in zip(csrs, ms, ts, tols):
        ss, first = mek.sstep_plan(m)
        runners.append(mek.Runner(c, m=m, t=t, tol=tl, max_restarts=0, step_control="expokit_rel",
                                  gs_mode=gs_mode, sstep=ss, sstep_first=first))
    for R, a, b in zip(runners, dv, dy):  # warm-up
        R.expmv_device(a, b)
    mek.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for R, a, b in zip(runners, dv, dy):
            R.expmv_device(a, b)
    mek.synchronize()
    plain_ss_s = (time.perf_counter() - t0) / reps
    plain_ss_cols = sum(R.stats()["spmv_count"] for R in runners)
    plain_ss_err = max_err([b.to_array() for b in dy])
    for R in runners:
        R.close()
    del dv, dy

    def batched(name, m_list, s_list, f32, select_s=0.0, spmv_timing=False):
        log(f"batch: {name}")
        t0 = time.perf_counter()
        B = mek.Batch(csrs, ts, m_list, tols, substeps=s_list, values_f32=f32, gs_mode=gs_mode,
                      step_control="expokit_rel", autotune=args.autotune)
        B.upload(vs)
        setup = time.perf_counter() - t0
        B.solve()  # warm-up: captures the cycle graphs
        mek.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            B.solve()
        mek.synchronize()
        el = (time.perf_counter() - t0) / reps
        ys = B.download()
        st = [B.stats(p) for p in range(P)]
        out = {"problems_per_s": P / (el + select_s), "solve_s": el, "select_s": select_s, "setup_s": setup,
               "values_f32_used": B.values_f32_used(), "spmv_grid": B.spmv_grid() if args.autotune else None,
               "cycles": B.cycles(), "krylov_columns": int(sum(x["spmv_count"] for x in st)),
               "converged": int(sum(x["converged"] for x in st)), "max_rel_err": max_err(ys)}
        if spmv_timing:
            out["spmv_us"] = B.time_spmv(20)
        B.close()
        return out

    heur_s = [0] * P  # Expokit's own first step
    res = {"plain_fp64": {"problems_per_s": P / plain_s, "solve_s": plain_s, "setup_s": setup_plain,
                          "krylov_columns": plain_cols, "max_rel_err": plain_err},
           "plain_fp64_sstep": {"problems_per_s": P / plain_ss_s, "solve_s": plain_ss_s,
                                "krylov_columns": plain_ss_cols, "max_rel_err": plain_ss_err}}
    res["batched_fp64"] = batched("FP64 values, Expokit step control", ms, heur_s, False)
    res["batched_mixed"] = batched("FP32 values requested, Expokit step control", ms, heur_s, True)
    res["batched_mixed_heuristic"] = batched("FP32 values requested, m = 30, Expokit's first step",
                                             [min(30, n - 1) for n in rows], heur_s, True)
    learned = None
    if os.path.exists(sel.WEIGHTS):
        S = sel.Selector.load()
        t0 = time.perf_counter()
        FB = mek.Batch(csrs, ts, ms, tol, values_f32=True, gs_mode=gs_mode)
        fsetup = time.perf_counter() - t0
        mek.synchronize()
        t0 = time.perf_counter()
        F = FB.features()
        choice = S.predict_features(F)
        select_s = time.perf_counter() - t0
        FB.close()
        lm = [min(m, n - 1) for (m, s), n in zip(choice, rows)]
        ls = [s for m, s in choice]
        learned = batched("FP32 values requested, learned (m, sub-steps)", lm, ls, True, select_s,
                          spmv_timing=True)
        learned["setup_s"] += fsetup
        ms_, cnt = np.unique(lm, return_counts=True)
        learned["m_chosen_hist"] = {int(a): int(c) for a, c in zip(ms_, cnt)}
        learned["speedup_vs_heuristic"] = round(
            learned["problems_per_s"] / res["batched_mixed_heuristic"]["problems_per_s"], 3)
        res["batched_mixed_learned"] = learned
    forced = batched("FP32 values forced (lossy)", ms, heur_s, "lossy")
    forced["meets_1e-8"] = forced["max_rel_err"] <= 1e-8
    res["batched_f32_forced"] = forced
    head = learned if learned is not None else res["batched_mixed"]
    form = "batched_mixed_learned" if learned is not None else "batched_mixed"
    if "spmv_us" not in head:
        head["spmv_us"] = float("nan")
    # the batched SpMV: one launch over every problem's start vector, with the
    # headline form's value storage
    rows_tot = int(sum(rows))
    vb = 4 if head["values_f32_used"] else 8
    spmv_bytes = nnz_tot * (vb + 4) + 4.0 * (rows_tot + 1) + 16.0 * rows_tot
    spmv_gbs = spmv_bytes / (head["spmv_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": _finite(round(spmv_gbs, 1)), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": _finite(round(spmv_gbs / PEAK_HBM_GBS, 4)), "traffic": None, "kernel": "batch spmv",
                "bytes_per_launch": spmv_bytes, "avg_launch_us": _finite(head["spmv_us"]),
                "note": ("first-column SpMV over all problems (median of 20 HIP-event-timed launches on "
                         "the batch's stream)")}
    cpu = None
    if cpu_seconds > 0:
        log("batch: CPU baseline (oracle, one thread, bounded sample)")
        cpu = cpu_baseline_batch(mats, ts, ms, vs, tol, cpu_seconds)
    speed = head["problems_per_s"] / res["plain_fp64"]["problems_per_s"]
    speed_best = head["problems_per_s"] / max(res["plain_fp64"]["problems_per_s"],
                                              res["plain_fp64_sstep"]["problems_per_s"])
    stencil = stencil_batch_comparison(args, mek, gen, reps)
    return {
        "workload": ("batch256: 256 problems, n_i log-uniform in [1e3, 2e5], random rows with 5-30 "
                     "nonzeros, values U[-1,1] FP64, Gershgorin-negative diagonal, t_i ~ U[0.1, 5], "
                     "m_i ~ U{20..60}, seed 0"),
        "problems": P, "rows": rows_tot, "nnz": nnz_tot, "tol": tol, "gs": args.gs,
        "unit": "problems/s", "headline_form": form,
        "value": round(head["problems_per_s"], 3),
        "rel_err": head["max_rel_err"],
        "speedup_vs_plain_fp64": round(speed, 3),
        "speedup_vs_best_plain_fp64": round(speed_best, 3),
        "speedup_learned_vs_heuristic": learned["speedup_vs_heuristic"] if learned is not None else None,
        "target_speedup": 3.0, "target_met": bool(speed >= 3.0 and head["max_rel_err"] <= 1e-8),
        "values_f32_check": f32_check,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": res,
        "stencil_batch": stencil,
    }


def stencil_batch_comparison(args, mek, gen, reps):
    """Where FP32 value storage is lossless: 64 independent stencil problems
    (32 2-D Laplacians on N x N grids, N log-uniform in [40, 600], t ~ U[1, 10];
    32 3-D convection-diffusion operators on N^3 grids, N log-uniform in
    [16, 60], t ~ U[0.05, 0.5]; v ~ N(0,1) normalised; m = 30), solved as one
    batched FP64 solve and one with FP32 values (exact here: the stencil
    coefficients are FP32 numbers), each against its closed form (2-D DST-I,
    Kronecker product of 1-D exponentials) at 1e-8, with the batched SpMV
    timed over all problems in both storages."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from closed_forms import kron3_expmv, laplacian_2d_expmv
    rng = np.random.default_rng(11)
    tol = 1e-8
    mats, ts, vs, refs = [], [], [], []
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    for i in range(64):
        if i % 2 == 0:
            N = int(round(math.exp(rng.uniform(math.log(40), math.log(600)))))
            t = float(rng.uniform(1.0, 10.0))
            A = gen.laplacian_2d(N)
            v = gen.normal_vector(N * N, seed=1000 + i)
            refs.append(laplacian_2d_expmv(N, t, v))
        else:
            N = int(round(math.exp(rng.uniform(math.log(16), math.log(60)))))
            t = float(rng.uniform(0.05, 0.5))
            A = gen.convdiff_3d(N)
            v = gen.normal_vector(N ** 3, seed=1000 + i)
            refs.append(kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                                    gen.tridiag_1d(N, lx, dx, ux), t, v))
        mats.append(A)
        ts.append(t)
        vs.append(v)
    csrs = [mek.CSR.from_arrays(*A) for A in mats]
    inexact = int(sum(c.f32_inexact(nonfinite=True)[0] for c in csrs))
    rows = [A[0].size - 1 for A in mats]
    nnz = int(sum(int(A[0][-1]) for A in mats))
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    out = {"workload": ("64 stencil problems: 32 2-D Laplacians (N log-uniform in [40, 600], t ~ U[1, 10]) "
                        "and 32 3-D convection-diffusion operators (N^3, N log-uniform in [16, 60], "
                        "t ~ U[0.05, 0.5]), m = 30, tol 1e-8, seed 11"),
           "problems": len(mats), "rows": int(sum(rows)), "nnz": nnz, "values_f32_inexact": inexact}
    for name, f32 in (("batched_fp64", False), ("batched_fp32_values", True)):
        log(f"stencil batch: {name}")
        B = mek.Batch(csrs, ts, [30] * len(mats), [tol / t for t in ts], values_f32=f32, gs_mode=gs_mode,
                      step_control="expokit_rel", autotune=args.autotune)
        B.upload(vs)
        B.solve()
        mek.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            B.solve()
        mek.synchronize()
        el = (time.perf_counter() - t0) / reps
        ys = B.download()
        err = max(float(np.linalg.norm(y - r) / np.linalg.norm(r)) for y, r in zip(ys, refs))
        st = [B.stats(p) for p in range(len(mats))]
        vb = 4 if B.values_f32_used() else 8
        spmv_us = B.time_spmv(20)
        spmv_bytes = nnz * (vb + 4) + 4.0 * (sum(rows) + 1) + 16.0 * sum(rows)
        out[name] = {"problems_per_s": len(mats) / el, "solve_s": el, "values_f32_used": B.values_f32_used(),
                     "krylov_columns": int(sum(x["spmv_count"] for x in st)),
                     "converged": int(sum(x["converged"] for x in st)), "max_rel_err": err,
                     "spmv_us": spmv_us, "spmv_gbs": round(spmv_bytes / (spmv_us * 1e-6) / 1e9, 1)}
        B.close()
    for c in csrs:
        c.close()
    out["speedup_fp32_vs_fp64"] = round(out["batched_fp32_values"]["problems_per_s"] /
                                        out["batched_fp64"]["problems_per_s"], 3)
    out["spmv_speedup_fp32_vs_fp64"] = round(out["batched_fp64"]["spmv_us"] /
                                             out["batched_fp32_values"]["spmv_us"], 3)
    return out


def run_batch(args, mek, gen):
    b = batch_comparison(args, mek, gen, max(1, args.steps), args.cpu_seconds)
    head = b["forms"][b["headline_form"]]
    out = {
        "metric": METRIC,
        "value": b["value"],
        "unit": "problems/s",
        "n_gpus": 1,
        "steps": max(1, args.steps),
        "warmup": 1,
        "ms_per_step": round(1000.0 * (head["solve_s"] + head.get("select_s", 0.0)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-values/f64" if head["values_f32_used"] else "f64",
        "data": "synthetic",
        "config": {"workload": "batch256", "problems": b["problems"], "rows": b["rows"], "nnz": b["nnz"],
                   "tol": b["tol"], "gs": b["gs"], "headline_form": b["headline_form"], "seed": 0},
        "rel_err": b["rel_err"],
        "values_f32": b["values_f32_check"],
        "speedup_vs_plain_fp64": b["speedup_vs_plain_fp64"],
        "speedup_vs_best_plain_fp64": b["speedup_vs_best_plain_fp64"],
        "roofline": b["roofline"],
        "cpu_baseline": b["cpu_baseline"],
        "forms": b["forms"],
        "stencil_batch": b["stencil_batch"],
    }
    print(json.dumps(out), flush=True)


def run_block(args, mek, gen):
    """2-D Laplacian on a 2048 x 2048 grid (n = 4,194,304) with a block of 32
    N(0,1) right-hand sides (columns normalised, seed 0), t = 10, tol 1e-8:
    block Arnoldi with 16 block steps (SpMM, block Gram-Schmidt on FP64 MFMA
    with the conditional second pass, Cholesky QR, Taylor exponential of the
    512 x 512 block Hessenberg matrix on the MFMA GEMM).  Plain form: the
    single-vector solver (m = 40, DGKS, Expokit steps) on 8 of the columns in
    turn.  Errors against the DST-I closed form of every column."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from closed_forms import laplacian_2d_expmv_block
    w = WORKLOADS["block2048"]
    N, t, tol, m = w["N"], w["t"], w["tol"], args.block_m or w["m"]
    n = N * N
    log("building the 2048^2 Laplacian and 32 right-hand sides")
    rp, ci, va = gen.laplacian_2d(N)
    V = np.random.default_rng(0).standard_normal((n, 32))
    V /= np.linalg.norm(V, axis=0)
    A = mek.CSR.from_arrays(rp, ci, va)
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    B = mek.Block(A, m=m, t=t, tol=tol, gs_mode=gs_mode, timing=False, autotune=args.autotune)
    d_v = mek.DeviceBuffer.from_array(V.ravel())
    d_y = mek.DeviceBuffer(8 * V.size)
    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        B.expmv_device(d_v, d_y)
    mek.synchronize()
    log(f"timing {args.steps} block solves")
    t0 = time.perf_counter()
    for i in range(args.steps):
        B.expmv_device(d_v, d_y)
    mek.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1000.0 / args.steps
    st = B.stats()
    Y = d_y.to_array().reshape(V.shape)
    B.close()
    # per-kernel event times in a pass of their own (event pairs around every
    # launch would sit inside the timed loop otherwise)
    log("event-timed pass")
    Bt = mek.Block(A, m=m, t=t, tol=tol, gs_mode=gs_mode, timing=True, autotune=args.autotune)
    Bt.expmv_device(d_v, d_y)
    kern = {k: {"ms": 0.0, "launches": 0, "bytes": 0.0} for k in mek.BLOCK_KERNEL_KINDS}
    tsteps = 2
    for _ in range(tsteps):
        Bt.expmv_device(d_v, d_y)
        s2 = Bt.stats()
        for k in kern:
            for f in kern[k]:
                kern[k][f] += s2["kernels"][k][f]
    Bt.close()
    log("plain single-vector solves of 8 columns")
    R = mek.Runner(A, m=40, t=t, tol=tol, max_restarts=0, step_control="expokit", gs_mode=2)
    cols = [mek.DeviceBuffer.from_array(np.ascontiguousarray(V[:, c])) for c in range(8)]
    dy1 = mek.DeviceBuffer(8 * n)
    R.expmv_device(cols[0], dy1)
    mek.synchronize()
    t1 = time.perf_counter()
    for c in cols:
        R.expmv_device(c, dy1)
    mek.synchronize()
    single_ms = (time.perf_counter() - t1) * 1000.0 / len(cols)
    single_cols = R.stats()["spmv_count"]
    R.close()
    log("closed-form check of the 32 columns")
    ref = laplacian_2d_expmv_block(N, t, V)
    errs = [float(np.linalg.norm(Y[:, c] - ref[:, c]) / np.linalg.norm(ref[:, c])) for c in range(32)]

    def roof(k):
        kt = kern[k]
        if kt["launches"] == 0 or kt["ms"] <= 0:
            return None
        achieved = kt["bytes"] / (kt["ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None, "kernel": "block_" + k,
                "bytes_per_launch": kt["bytes"] / kt["launches"],
                "avg_launch_us": kt["ms"] * 1e3 / kt["launches"]}

    dominant = max(("spmm", "dots", "update", "lift"), key=lambda k: kern[k]["ms"])
    # counter traffic: SpMM and dots are one kernel each; the update sweep and
    # the lift share k_bupdate, so the profile cannot separate their bytes
    tr = measured_traffic(args.workload)
    roofs = {k: roof(k) for k in ("spmm", "dots", "update", "lift")}
    for k in ("spmm", "dots"):
        if roofs[k] is not None and ("block_" + k) in tr:
            roofs[k]["traffic"] = tr["block_" + k]
    cpu = None
    if args.cpu_seconds > 0:
        log("CPU baseline (oracle, one thread, one sub-step of one column's solve)")
        R = mek.Runner(A, m=40, t=t, tol=tol, max_restarts=0, step_control="expokit", gs_mode=2)
        R.expmv(V[:, 0].copy())
        steps = R.steps()
        R.close()
        cpu = cpu_baseline_substep(rp, ci, va, V[:, 0].copy(), 40, steps[0]["dt"], len(steps))
        cpu["unit"] = "rhs/s"  # one right-hand side per single-vector solve
        cpu["sample"] += "; one right-hand side per solve"
    rhs_per_s = 32 * 1000.0 / ms_per_step
    out = {
        "metric": METRIC,
        "value": round(rhs_per_s, 3),
        "unit": "rhs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "block2048", "n": n, "nnz": int(rp[-1]), "rhs": 32, "t": t, "block_steps": m,
                   "tol": tol, "gs": args.gs, "seed": 0},
        "rel_err": max(errs),
        "solve": {"substeps": st["accepted"], "attempts": st["attempts"], "block_steps": st["block_steps"],
                  "second_passes": st["second_passes"], "squarings": st["squarings"]},
        "launch_shapes": {"spmm_grid": st["spmm_grid"], "update_grid": st["update_grid"]},
        "plain_single_vector": {"ms_per_rhs": round(single_ms, 3), "rhs_per_s": round(1000.0 / single_ms, 3),
                                "krylov_columns": single_cols},
        "speedup_vs_single_vector": round(rhs_per_s / (1000.0 / single_ms), 3),
        "roofline": roofs[dominant],
        "spmm": roofs["spmm"],
        "kernels_ms_per_solve": {k: round(kern[k]["ms"] / tsteps, 3) for k in kern if kern[k]["launches"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def spmv_layout_comparison(mek, rp, ci, va, f32):
    """The same SpMV (w = A x, one x of N(0,1) entries) in each layout the
    matrix has, through the plan API: mean wall time per launch over 5
    batches of 20 back-to-back launches after a synchronisation (median of
    the batches), the layout's bytes, and whether every layout gave the CSR
    row blocks' bits."""
    n = rp.size - 1
    x = np.random.default_rng(5).standard_normal(n)
    d_x = mek.DeviceBuffer.from_array(x)
    d_y = mek.DeviceBuffer(8 * n)
    vb = 4 if f32 else 8
    nnz = int(rp[-1])
    out, ref = {}, None
    for name, switch in (("csr", None), ("dia", "use_dia"), ("dia_const", "use_dia_const")):
        plan = mek.SpmvPlan(rp, ci, va, values_f32=f32)
        try:
            nd = getattr(plan, switch)() if switch else 0
        except mek.MekError:
            plan.close()
            continue
        for _ in range(5):
            plan.execute(d_x, d_y)
        mek.synchronize()
        per = []
        for _ in range(5):
            t0 = time.perf_counter()
            for _ in range(20):
                plan.execute(d_x, d_y)
            mek.synchronize()
            per.append((time.perf_counter() - t0) / 20)
        y = d_y.to_array()
        plan.close()
        if name == "csr":
            nbytes = nnz * (vb + 4) + 4.0 * (n + 1) + 16.0 * n
            ref = y
        elif name == "dia":
            nbytes = nd * n * vb + 16.0 * n
        else:
            nbytes = (1 if nd <= 8 else 2) * n + 16.0 * n
        us = statistics.median(per) * 1e6
        out[name] = {"us_per_launch": round(us, 1), "layout_bytes": nbytes,
                     "layout_gbs": round(nbytes / (us * 1e-6) / 1e9, 1),
                     "same_bits_as_csr": bool(ref is not None and np.array_equal(y.view(np.int64),
                                                                              ref.view(np.int64)))}
    out["note"] = ("wall time per launch of 20 back-to-back launches (median of 5), the plan API on the "
                   "default stream, x ~ N(0,1); layout bytes as in the roofline")
    return out


def run_single(args, mek, gen):
    w = WORKLOADS[args.workload]
    log(f"building {args.workload}")
    if w["kind"] == "laplacian":
        rp, ci, va = gen.laplacian_2d(w["N"])
        n = w["N"] ** 2
    else:
        rp, ci, va = gen.convdiff_3d(w["N"])
        n = w["N"] ** 3
    nnz = int(rp[-1])
    v = gen.normal_vector(n, seed=0)
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    A = mek.CSR.from_arrays(rp, ci, va)

    sstep, sstep_first = mek.sstep_plan(w["m"], args.sstep, args.sstep_first)

    def runner(timing):
        return mek.Runner(A, m=w["m"], t=w["t"], tol=w["tol"], max_restarts=0, step_control=args.step_control,
                          gs_mode=gs_mode, values_f32=w["f32"], verbose=False, timing=timing,
                          fuse_spmv_dot=args.fuse, lanczos=args.lanczos, autotune=args.autotune,
                          sstep=sstep, sstep_first=sstep_first, spmv_format=args.spmv_format)

    R = runner(False)
    layout = R.spmv_format()
    f32_values = None
    if w["f32"]:
        c, mrel, bad = A.f32_inexact(nonfinite=True)
        f32_values = {"inexact": int(c), "nonfinite": int(bad), "max_rel_change": _finite(mrel),
                      "values_f32_used": R.values_f32_used()}
    d_v = mek.DeviceBuffer.from_array(v)
    d_y = mek.DeviceBuffer(8 * n)
    mek.synchronize()

    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        R.expmv_device(d_v, d_y)
    mek.synchronize()
    log(f"timing {args.steps} solves (no per-kernel events)")
    t0 = time.perf_counter()
    for i in range(args.steps):
        R.expmv_device(d_v, d_y)
    mek.synchronize()
    elapsed = time.perf_counter() - t0
    stats = R.stats()
    ms_per_step = elapsed * 1000.0 / args.steps
    log(f"{ms_per_step:.1f} ms per solve")
    y = d_y.to_array()

    # the reference's METRIC span: host v in, host y out
    span_steps = max(1, min(args.steps, 5))
    log(f"timing {span_steps} solves on the METRIC span (host v and y)")
    span = []
    yh = np.ones(n)  # the caller's y, allocated and touched before the span as in main.cu:2281
    for _ in range(span_steps):
        t0 = time.perf_counter()
        R.expmv(v, out=yh)
        span.append(time.perf_counter() - t0)
    assert np.array_equal(yh, y)
    tuning = R.tuning() if args.autotune else None
    steps = R.steps()  # the sub-steps of a solve, for the CPU baseline
    R.close()

    # per-kernel event timing in a pass of its own
    log("event-timed pass")
    Rt = runner(True)
    Rt.expmv_device(d_v, d_y)
    kern = {k: {"ms": 0.0, "launches": 0, "bytes": 0.0} for k in mek.KERNEL_KINDS}
    tsteps = 2
    for _ in range(tsteps):
        Rt.expmv_device(d_v, d_y)
        for k in kern:  # per-solve event totals (each solve resets them)
            kt = Rt.kernel_time(k)
            for f in kern[k]:
                kern[k][f] += kt[f]
    Rt.close()
    A.close()

    log("checking against the closed form")
    ref = reference_solution(w, v, gen)
    s = max(np.abs(ref).max(), 1e-300)
    rel = float(np.linalg.norm((y - ref) / s) / np.linalg.norm(ref / s))
    log(f"relative error {rel:.3e}")

    # The SpMV's bytes are those of the layout that runs (the runner's count:
    # the diagonal copy reads n values per diagonal and no column index, the
    # constant-diagonal form a presence mask per row and no value; the
    # sliced-ELL copy its padded entries), the least the kernel must move; the
    # CSR figure nnz (value + 4) + 4 (n + 1) + 8 n (x once) + 8 n (w) is kept
    # beside it with the bandwidth it would imply.
    f32_used = bool(w["f32"] and f32_values["values_f32_used"])
    csr_bytes = nnz * ((4 if f32_used else 8) + 4) + 4.0 * (n + 1) + 16.0 * n

    def roof(k):
        kt = kern[k]
        if kt["launches"] == 0 or kt["ms"] <= 0:
            return None
        per_launch = kt["bytes"] / kt["launches"]
        avg_us = kt["ms"] * 1e3 / kt["launches"]
        achieved = per_launch / (avg_us * 1e-6) / 1e9
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
             "frac": round(achieved / PEAK_HBM_GBS, 4),
             "traffic": None, "kernel": k,
             "bytes_per_launch": per_launch,
             "avg_launch_us": avg_us}
        if k == "spmv":
            r["layout"] = layout
            # the plain launch's layout bytes; what a launch moves beyond them
            # (the first columns' fused dots read q columns) is added to the
            # CSR figure too
            vbl = 4 if f32_used else 8
            fmt, ent, wid = layout["format"], layout["entries"], layout["width"]
            if fmt == "dia_const":
                plain = (1 if wid <= 8 else 2) * n + 16.0 * n
            elif fmt == "dia":
                plain = wid * n * vbl + 16.0 * n
            elif fmt == "sell":
                plain = ent * (vbl + 4) + 8.0 * ((n + 63) // 64 + 1) + 16.0 * n
            else:
                plain = csr_bytes
            csr_eq = csr_bytes + max(0.0, per_launch - plain)
            r["csr_bytes_per_launch"] = csr_eq
            r["csr_equivalent_gbs"] = round(csr_eq / (avg_us * 1e-6) / 1e9, 1)
        return r

    streaming = [k for k in ("spmv", "sstep_update", "sstep_dots", "update_dot", "update_norm", "dots", "lift")
                 if kern[k]["launches"]]
    dominant = max(streaming, key=lambda k: kern[k]["ms"])
    roofline = roof(dominant)
    spmv_roof = roof("spmv")
    # the committed counter runs profile the default kernels (no fused dots,
    # no Lanczos): their traffic applies only to that configuration
    tr = {} if (args.fuse or args.lanczos) else measured_traffic(args.workload)
    if roofline is not None and dominant in tr:
        roofline["traffic"] = tr[dominant]
    if spmv_roof is not None and "spmv" in tr:
        spmv_roof["traffic"] = tr["spmv"]
    solves_per_s = 1000.0 / ms_per_step
    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-values/f64" if (w["f32"] and f32_values["values_f32_used"]) else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "nnz": nnz, "t": w["t"], "m": w["m"],
                   "tol": w["tol"], "step_control": args.step_control, "gs": args.gs,
                   "fused_spmv_dot": args.fuse, "lanczos": args.lanczos, "autotune": args.autotune,
                   "sstep": sstep, "sstep_first": sstep_first, "spmv_format": layout["format"],
                   "seed": 0},
        "rel_err": rel,
        "values_f32": f32_values,
        "metric_span": {"solves_per_s": round(1.0 / statistics.median(span), 4),
                        "ms_per_solve": round(1000.0 * statistics.median(span), 3),
                        "span": "compute_expmv with host v and y (main.cu:1412-1570): the H2D of v and the "
                                "D2H of y over PCIe included; median of the solves"},
        "solve": {"substeps": stats["accepted"], "expm_attempts": stats["attempts"],
                  "krylov_columns": stats["spmv_count"], "extra_spmv": stats["extra_spmv"],
                  "converged": bool(stats["converged"]), "sstep_blocks": stats["sstep_blocks"],
                  "sstep_second": stats["sstep_second"], "sstep_fallbacks": stats["sstep_fallbacks"],
                  "sstep_fused": stats["sstep_fused"]},
        "roofline": roofline,
        "spmv": spmv_roof,
        "kernels_ms_per_solve": {k: round(kern[k]["ms"] / tsteps, 3) for k in kern if kern[k]["launches"]},
        # layout bytes over event time, per kind (a kind whose launches can return at once, such as
        # the second passes, counts no bytes)
        "kernels_gbs": {k: round(kern[k]["bytes"] / (kern[k]["ms"] * 1e-3) / 1e9, 1)
                        for k in kern if kern[k]["launches"] and kern[k]["ms"] > 0},
        # algorithmic bytes of every timed kernel of a solve / the solve's wall time
        "solve_effective_gbs": round(sum(kern[k]["bytes"] for k in kern) / tsteps / (ms_per_step * 1e-3) / 1e9, 1),
        "launch_shapes": tuning,
        "cpu_baseline": None,
    }
    if not (args.fuse or args.lanczos):
        log("SpMV layouts side by side")
        out["spmv_layouts"] = spmv_layout_comparison(mek, rp, ci, va, f32_used)
    if args.cpu_seconds > 0:
        log("CPU baseline (oracle, one thread, one full sub-step)")
        out["cpu_baseline"] = cpu_baseline_substep(rp, ci, va, v, w["m"], steps[0]["dt"], len(steps))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="laplacian4096", choices=sorted(WORKLOADS))
    ap.add_argument("--step-control", default="expokit", choices=["expokit", "reference"])
    ap.add_argument("--gs", default="dgks", choices=["always", "dgks", "never"],
                    help="second GS pass: dgks = only when ||w'|| < 0.1 ||w|| (default), always = reference MGS2 semantics")
    ap.add_argument("--fuse", action="store_true", help="first-pass dots inside the SpMV kernel")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shapes (default: time candidates when the runner is created)")
    ap.add_argument("--lanczos", action="store_true",
                    help="symmetric Lanczos recurrence instead of Gram-Schmidt (laplacian4096 only)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline budget of the batch sample (0: skip every CPU baseline)")
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="leave the batched comparison out of the default line")
    ap.add_argument("--block-m", type=int, default=0, help="block Arnoldi steps (block2048; 0: 16)")
    ap.add_argument("--sstep", type=int, default=-1,
                    help="s-step blocks of this many basis vectors (0: column by column, as the reference; "
                         "-1: equal blocks of at most 16 after the first 8 columns)")
    ap.add_argument("--spmv-format", default="auto", choices=["auto", "csr", "sell", "dia", "dia_const"],
                    help="SpMV layout: the diagonal layout for stencils (constant diagonals when each holds one "
                         "value), else sliced ELL where its padding is small (auto); CSR row blocks; sliced ELL; "
                         "diagonal copy with per-entry values; constant diagonals")
    ap.add_argument("--sstep-first", type=int, default=-1,
                    help="columns of each cycle built one by one before the blocks (-1: 8 with --sstep -1, "
                         "else --sstep)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the solver runs on one GPU; use --gpus 1 (each problem's rows would be split "
              "across GPUs with a halo exchange per SpMV and one all-reduce per Gram-Schmidt pass)",
              file=sys.stderr)
        sys.exit(2)

    mek = _load("mek", os.path.join(ROOT, "matrix-exp-krylov_amd", "__init__.py"))
    gen = _load("mek_generate", os.path.join(ROOT, "matrix-exp-krylov_amd", "generate.py"))
    if args.workload == "batch256":
        run_batch(args, mek, gen)
        return
    if args.workload == "block2048":
        run_block(args, mek, gen)
        return
    out = run_single(args, mek, gen)
    if args.workload == "laplacian4096" and not (args.fuse or args.lanczos):
        log("reference CPU case (n = 2000)")
        out["random_n2000"] = random_n2000_case(mek, gen, args.cpu_seconds > 0)
        if args.batch:
            out["learned_mixed_batched"] = batch_comparison(args, mek, gen, max(1, min(args.steps, 3)),
                                                            args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
