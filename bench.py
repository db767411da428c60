"""Throughput of the MI355X Arnoldi exp(t*A)*v solver on one GPU.

A step is one full solve y = exp(t*A)*v (ArnoldiRunner::compute_expmv on
device-resident v and y).  Default workload: 2-D 5-point Laplacian on a
4096 x 4096 grid (n = 16,777,216, nnz = 83,869,696, FP64 values, i32
indices), v ~ N(0,1) normalised (seed 0), t = 10, m = 40, tol 1e-8, sub-steps
by the Expokit heuristic.

Prints one JSON line with solves/s, the relative error against the DST-I closed
form, a roofline object for the dominant kernel (algorithmic bytes per launch /
HIP-event launch time, on the solver's own stream), the SpMV kernel's HBM
GB/s, and a CPU baseline: the oracle's single-thread Arnoldi column loop timed
on a bounded sample of the same matrix and extrapolated to the same solve.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0


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


def cpu_baseline(rp, ci, va, v, columns_per_solve, m, n, seconds_budget):
    """Time the oracle's reference Arnoldi loop (SpMV + MGS2 + normalise, one
    thread) on the first columns of the same matrix and extrapolate linearly in
    the basis size to the columns the GPU solve built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    # pick a column count that fits the time budget (first call is a probe)
    probe = pyoracle.arnoldi_columns(rp, ci, va, v, 3)
    slope = max(float(probe[2] - probe[1]), 0.0)
    ncols, total = 0, 0.0
    while ncols < m:  # largest column count whose predicted time fits the budget
        nxt = float(probe[0]) + slope * ncols
        if ncols >= 3 and total + nxt > seconds_budget * 1000.0:
            break
        total += nxt
        ncols += 1
    # one thread pinned to one core, as the reference's single-threaded CPU
    # binary runs (the calling thread's affinity; restored afterwards)
    prev = os.sched_getaffinity(0)
    core = min(prev)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.time()
        col_ms = pyoracle.arnoldi_columns(rp, ci, va, v, ncols)
        wall = time.time() - t0
    finally:
        os.sched_setaffinity(0, prev)
    j = np.arange(1, col_ms.size + 1, dtype=float)
    b, a = np.polyfit(j, col_ms, 1)  # ms per column = a + b * (basis vectors)
    per_step_ms = sum(a + b * (k + 1) for k in range(m))
    lift_ms = (m + 1) * b / 4.0  # one axpy per basis vector vs. two dots + two axpys
    solve_ms = columns_per_solve / m * (per_step_ms + lift_ms)
    return {
        "value": 1000.0 / solve_ms,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "host": _host_desc(core),
        "sample": (f"oracle Arnoldi loop (SpMV + MGS2 + normalise, main_cpu.cpp:486-546 restated in C) "
                   f"for the first {col_ms.size} columns of the same matrix and start vector "
                   f"({wall:.1f} s wall), fitted as {a:.1f} + {b:.1f}*(j+1) ms per column and "
                   f"extrapolated to the {columns_per_solve} columns (+ lifts) of one GPU solve"),
    }


def _host_desc(core):
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return f"1 thread pinned to core {core} of {os.cpu_count()} ({model or 'unknown CPU'})"


def cpu_baseline_batch(mats, ts, ms, vs, tol, seconds_budget):
    """The oracle's reference solver (MGS2, reference step controller with the
    first step t, tol absolute per step) on the batch's problems in turn, one
    thread, until the budget is spent: problems/s on that sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    prev = os.sched_getaffinity(0)
    core = min(prev)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.time()
        k = conv = 0
        for (rp, ci, va), t, m, v in zip(mats, ts, ms, vs):
            _y, st, _log = pyoracle.expmv(rp, ci, va, v, t, tol, m, 0)
            conv += int(st["converged"])
            k += 1
            if time.time() - t0 > seconds_budget:
                break
        wall = time.time() - t0
    finally:
        os.sched_setaffinity(0, prev)
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


def run_batch(args, mek, gen):
    """256 independent problems (n log-uniform in [1e3, 2e5], random rows with
    5-30 nonzeros and a Gershgorin-negative diagonal, t ~ U[0.1, 5],
    m ~ U{20..60}).  Plain form: one ArnoldiRunner solve per problem in turn
    (FP64 values, Expokit step control).  Batched forms: every problem in one
    lockstep solve replayed from a hipGraph with Expokit step control on the
    device, FP64 then FP32 values, then with (m, first sub-step) from the
    learned selector and FP32 values.  Every form uses the DGKS second pass
    and Expokit's per-unit-time tolerance tol/t, relative to the norm of each
    step's start vector.
    Errors are against a tight-tolerance (1e-13, m = 60) FP64 solve of each
    problem."""
    sel = _load("mek_selector", os.path.join(ROOT, "matrix-exp-krylov_amd", "selector.py"))
    w = WORKLOADS["batch256"]
    tol = w["tol"]
    log("generating 256 problems")
    probs = gen.batch_problems(P=w["P"], seed=0, f32_exact=True)
    P = len(probs)
    mats = [(q["rp"], q["ci"], q["va"]) for q in probs]
    ts = [q["t"] for q in probs]
    ms = [min(q["m"], q["rp"].size - 2) for q in probs]
    vs = [q["v"] for q in probs]
    rows = [q["rp"].size - 1 for q in probs]
    nnz_tot = int(sum(int(q["rp"][-1]) for q in probs))
    csrs = [mek.CSR.from_arrays(*A) for A in mats]
    # FP32 value storage is lossless only for values exact in FP32: say so
    f32_check = [c.f32_inexact() for c in csrs]
    f32_values = {"inexact": int(sum(c for c, _ in f32_check)),
                  "max_rel_change": max(m for _, m in f32_check)}
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]

    log("reference solves (tol 1e-13)")
    refs = []
    for c, t, v, n in zip(csrs, ts, vs, rows):
        R = mek.Runner(c, m=min(60, n - 1), t=t, tol=1e-13, max_restarts=0, step_control="expokit_rel",
                       gs_mode=1)
        refs.append(R.expmv(v))
        R.close()

    def max_err(ys):
        return max(float(np.linalg.norm(y - r) / max(np.linalg.norm(r), 1e-300)) for y, r in zip(ys, refs))

    # plain FP64: one runner per problem, solves in turn, v and y resident
    log("plain FP64, one problem at a time")
    # Expokit's tolerance is per unit time: tol / t bounds the summed local
    # errors of each problem by about tol, measured relative to the norm of
    # each step's start vector (relative mode)
    tols = [tol / t for t in ts]
    runners = [mek.Runner(c, m=m, t=t, tol=tl, max_restarts=0, step_control="expokit_rel", gs_mode=gs_mode)
               for c, m, t, tl in zip(csrs, ms, ts, tols)]
    dv = [mek.DeviceBuffer.from_array(v) for v in vs]
    dy = [mek.DeviceBuffer(8 * n) for n in rows]
    for R, a, b in zip(runners, dv, dy):  # warm-up
        R.expmv_device(a, b)
    mek.synchronize()
    reps = max(1, args.steps)
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
    del dv, dy

    heur_s = [0] * P  # Expokit's own first step

    def batched(name, m_list, s_list, f32, select_s=0.0, spmv_timing=False):
        log(f"batched: {name}")
        B = mek.Batch(csrs, ts, m_list, tols, substeps=s_list, values_f32=f32, gs_mode=gs_mode,
                      step_control="expokit_rel", autotune=args.autotune)
        B.upload(vs)
        B.solve()  # warm-up: captures the cycle graph
        mek.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            B.solve()
        mek.synchronize()
        el = (time.perf_counter() - t0) / reps
        ys = B.download()
        st = [B.stats(p) for p in range(P)]
        out = {"problems_per_s": P / (el + select_s), "solve_s": el, "select_s": select_s,
               "spmv_grid": B.spmv_grid() if args.autotune else None,
               "cycles": B.cycles(), "krylov_columns": int(sum(x["spmv_count"] for x in st)),
               "converged": int(sum(x["converged"] for x in st)), "max_rel_err": max_err(ys)}
        if spmv_timing:
            out["spmv_us"] = B.time_spmv(20)
        B.close()
        return out

    res = {"plain_fp64": {"problems_per_s": P / plain_s, "solve_s": plain_s, "krylov_columns": plain_cols,
                          "max_rel_err": plain_err}}
    res["batched_fp64"] = batched("FP64, Expokit step control", ms, heur_s, False)
    res["batched_mixed"] = batched("FP32 values, Expokit step control", ms, heur_s, True,
                                   spmv_timing=not os.path.exists(sel.WEIGHTS))
    learned = None
    if os.path.exists(sel.WEIGHTS):
        S = sel.Selector.load()
        FB = mek.Batch(csrs, ts, ms, tol, values_f32=True, gs_mode=gs_mode)
        t0 = time.perf_counter()
        F = FB.features()
        choice = S.predict_features(F)
        select_s = time.perf_counter() - t0
        FB.close()
        lm = [min(m, n - 1) for (m, s), n in zip(choice, rows)]
        ls = [s for m, s in choice]
        learned = batched("FP32 values, learned (m, sub-steps)", lm, ls, True, select_s, spmv_timing=True)
        res["batched_mixed_learned"] = learned
    head = learned if learned is not None else res["batched_mixed"]
    form = "batched_mixed_learned" if learned is not None else "batched_mixed"
    # the SpMV dominates the batched solve (about 60 % of its kernel time):
    # one launch over every problem's start vector, FP32 values
    rows_tot = int(sum(rows))
    spmv_bytes = nnz_tot * (4 + 4) + 4.0 * (rows_tot + 1) + 16.0 * rows_tot
    spmv_gbs = spmv_bytes / (head["spmv_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(spmv_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(spmv_gbs / PEAK_HBM_GBS, 4), "traffic": None, "kernel": "batch spmv",
                "bytes_per_launch": spmv_bytes, "avg_launch_us": head["spmv_us"],
                "note": ("first-column SpMV over all problems (median of 20 HIP-event-timed launches on "
                         "the batch's stream); bound by the rate of random-column gathers, not HBM")}
    cpu = None
    if args.cpu_seconds > 0:
        log("CPU baseline (oracle, one thread, bounded sample)")
        cpu = cpu_baseline_batch(mats, ts, ms, vs, tol, args.cpu_seconds)
    out = {
        "metric": "exp(tA)v solves/s + SpMV HBM GB/s at rel err <=1e-8; learned+mixed vs plain FP64",
        "value": round(head["problems_per_s"], 3),
        "unit": "problems/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": round(1000.0 * (head["solve_s"] + head.get("select_s", 0.0)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-values/f64",
        "data": "synthetic",
        "config": {"workload": "batch256", "problems": P, "rows": int(sum(rows)), "nnz": nnz_tot,
                   "tol": tol, "gs": args.gs, "headline_form": form, "seed": 0},
        "rel_err": head["max_rel_err"],
        "values_f32": f32_values,
        "speedup_vs_plain_fp64": round(head["problems_per_s"] / res["plain_fp64"]["problems_per_s"], 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": res,
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
    from closed_forms import laplacian_2d_expmv
    w = WORKLOADS["block2048"]
    N, t, tol, m = w["N"], w["t"], w["tol"], args.block_m or w["m"]
    n = N * N
    log("building the 2048^2 Laplacian and 32 right-hand sides")
    rp, ci, va = gen.laplacian_2d(N)
    V = np.random.default_rng(0).standard_normal((n, 32))
    V /= np.linalg.norm(V, axis=0)
    A = mek.CSR.from_arrays(rp, ci, va)
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    B = mek.Block(A, m=m, t=t, tol=tol, gs_mode=gs_mode, timing=True, autotune=args.autotune)
    d_v = mek.DeviceBuffer.from_array(V.ravel())
    d_y = mek.DeviceBuffer(8 * V.size)
    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        B.expmv_device(d_v, d_y)
    mek.synchronize()
    kern = {k: {"ms": 0.0, "launches": 0, "bytes": 0.0} for k in mek.BLOCK_KERNEL_KINDS}
    log(f"timing {args.steps} block solves")
    t0 = time.perf_counter()
    for i in range(args.steps):
        B.expmv_device(d_v, d_y)
        st = B.stats()
        for k in kern:
            for f in kern[k]:
                kern[k][f] += st["kernels"][k][f]
    mek.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1000.0 / args.steps
    Y = d_y.to_array().reshape(V.shape)
    B.close()
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
    errs = [float(np.linalg.norm(Y[:, c] - laplacian_2d_expmv(N, t, V[:, c]))
                  / np.linalg.norm(laplacian_2d_expmv(N, t, V[:, c]))) for c in range(32)]

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
        log("CPU baseline (oracle, one thread, bounded sample)")
        cpu = cpu_baseline(rp, ci, va, V[:, 0].copy(), single_cols, 40, n, args.cpu_seconds)
        cpu["unit"] = "rhs/s"  # one right-hand side per single-vector solve
        cpu["sample"] += "; one right-hand side per solve"
    rhs_per_s = 32 * 1000.0 / ms_per_step
    out = {
        "metric": "exp(tA)v solves/s + SpMV HBM GB/s at rel err <=1e-8; learned+mixed vs plain FP64",
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
        "kernels_ms_per_solve": {k: round(kern[k]["ms"] / args.steps, 3) for k in kern if kern[k]["launches"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


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
    ap.add_argument("--cpu-seconds", type=float, default=25.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--block-m", type=int, default=0, help="block Arnoldi steps (block2048; 0: 16)")
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
    R = mek.Runner(A, m=w["m"], t=w["t"], tol=w["tol"], max_restarts=0, step_control=args.step_control,
                   gs_mode=gs_mode, values_f32=w["f32"], verbose=False, timing=True,
                   fuse_spmv_dot=args.fuse, lanczos=args.lanczos, autotune=args.autotune)
    f32_values = None
    if w["f32"]:
        c, mrel = A.f32_inexact()
        f32_values = {"inexact": int(c), "max_rel_change": mrel}
    A.close()
    d_v = mek.DeviceBuffer.from_array(v)
    d_y = mek.DeviceBuffer(8 * n)
    mek.synchronize()

    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        R.expmv_device(d_v, d_y)
    mek.synchronize()
    kern = {k: {"ms": 0.0, "launches": 0, "bytes": 0.0} for k in mek.KERNEL_KINDS}
    log(f"timing {args.steps} solves")
    t0 = time.perf_counter()
    stats = None
    for i in range(args.steps):
        R.expmv_device(d_v, d_y)
        for k in kern:  # per-solve event totals (each solve resets them)
            kt = R.kernel_time(k)
            for f in kern[k]:
                kern[k][f] += kt[f]
        stats = R.stats()
    mek.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1000.0 / args.steps
    log(f"{ms_per_step:.1f} ms per solve; checking against the closed form")

    y = d_y.to_array()
    ref = reference_solution(w, v, gen)
    s = max(np.abs(ref).max(), 1e-300)
    rel = float(np.linalg.norm((y - ref) / s) / np.linalg.norm(ref / s))
    log(f"relative error {rel:.3e}")

    def roof(k):
        kt = kern[k]
        if kt["launches"] == 0 or kt["ms"] <= 0:
            return None
        achieved = kt["bytes"] / (kt["ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": None, "kernel": k,
                "bytes_per_launch": kt["bytes"] / kt["launches"],
                "avg_launch_us": kt["ms"] * 1e3 / kt["launches"]}

    streaming = [k for k in ("spmv", "update_dot", "update_norm", "dots", "lift") if kern[k]["launches"]]
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

    cpu = None
    if args.cpu_seconds > 0:
        log("CPU baseline (oracle, one thread, bounded sample)")
        cpu = cpu_baseline(rp, ci, va, v, stats["spmv_count"], w["m"], n, args.cpu_seconds)

    solves_per_s = 1000.0 / ms_per_step
    out = {
        "metric": "exp(tA)v solves/s + SpMV HBM GB/s at rel err <=1e-8; learned+mixed vs plain FP64",
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-values/f64" if w["f32"] else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "nnz": nnz, "t": w["t"], "m": w["m"],
                   "tol": w["tol"], "step_control": args.step_control, "gs": args.gs,
                   "fused_spmv_dot": args.fuse, "lanczos": args.lanczos, "autotune": args.autotune, "seed": 0},
        "rel_err": rel,
        "values_f32": f32_values,
        "solve": {"substeps": stats["accepted"], "expm_attempts": stats["attempts"],
                  "krylov_columns": stats["spmv_count"], "extra_spmv": stats["extra_spmv"],
                  "converged": bool(stats["converged"])},
        "roofline": roofline,
        "spmv": spmv_roof,
        "kernels_ms_per_solve": {k: round(kern[k]["ms"] / args.steps, 3) for k in kern if kern[k]["launches"]},
        # algorithmic bytes of every timed kernel of a solve / the solve's wall time
        "solve_effective_gbs": round(sum(kern[k]["bytes"] for k in kern) / args.steps / (ms_per_step * 1e-3) / 1e9, 1),
        "launch_shapes": R.tuning() if args.autotune else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    R.close()


if __name__ == "__main__":
    main()
