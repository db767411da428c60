"""Profiling workloads and summaries for the Arnoldi solver kernels.

Two modes:

``python matrix-exp-krylov_amd/profile.py workload [--workload W] [--gs G]``
    The program to run under ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc
    WRITE_SIZE``: first a calibration elementwise kernel (``torch.neg`` of a
    1 GiB FP64 buffer into another, 1 GiB read + 1 GiB written per launch,
    three launches), then one warm-up and one timed solve of the workload.

``python matrix-exp-krylov_amd/profile.py summarize --stats DIR --fetch DIR
--write DIR [--cache DIR] --workload W --out-md FILE --out-json FILE
[--command TEXT ...]``
    Reads the rocprofv3 CSV outputs (kernel-trace statistics, and the two
    counter runs), scales each counter by what it reads for the calibration
    kernel of known size, and writes a markdown summary plus a JSON map
    ``{kernel kind: HBM bytes per launch}`` that bench.py attaches to its
    roofline object as ``traffic``.  With ``--cache``, a third counter run
    (``--pmc TCC_REQ_sum TCC_HIT_sum TCC_MISS_sum TCP_TCC_READ_REQ_sum``) adds
    the L2 requests, hits and misses and the L1-to-L2 read requests per
    launch of each kind.
"""
from __future__ import annotations

import argparse
import csv
import glob
import importlib.util
import json
import os
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
CAL_BYTES = 1 << 30  # calibration buffer size (bytes read = bytes written)

# kernel symbol prefix -> kind reported by the runner's event timers
KINDS = [
    ("k_spmm", "block_spmm"),
    ("k_bdots_fin", "block_other"),
    ("k_bdots", "block_dots"),
    ("k_bupdate", "block_update"),
    ("k_dgemm", "block_expm_gemm"),
    ("k_b", "block_other"),
    ("k_dcombine", "block_expm_other"),
    ("k_dload", "block_expm_other"),
    ("k_dnorm1", "block_expm_other"),
    ("k_lanczos_update", "lanczos_update"),
    ("k_lanczos_fin", "finalise"),
    ("kb_spmv", "spmv"),
    ("kb_dots1", "dots"),
    ("kb_update_dot", "update_dot"),
    ("kb_update_norm", "update_norm"),
    ("kb_lift", "lift"),
    ("kb_expm", "expm"),
    ("kb_fin", "finalise"),
    ("kb_", "batch_other"),
    ("k_ss_dots", "sstep_dots"),
    ("k_ss_update", "sstep_update"),
    ("k_ss_", "finalise"),
    ("k_spmv_dot", "spmv"),
    ("k_spmv_dia", "spmv"),
    ("k_spmv_sell", "spmv"),
    ("k_update_dot", "update_dot"),
    ("k_update_norm", "update_norm"),
    ("k_dots", "dots"),
    ("k_update", "update"),
    ("k_lift", "lift"),
    ("k_expm", "expm"),
    ("k_fin_", "finalise"),
    ("k_norm2", "norm"),
    ("k_restart", "restart"),
]


def kind_of(name: str) -> str:
    base = name.split("(")[0]
    base = base.replace("void ", "").replace("mek::", "").strip()
    for prefix, kind in KINDS:
        if base.startswith(prefix):
            return kind
    return "other"


def _load(name, fn):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, fn))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


WORKLOADS = {
    "laplacian4096": dict(kind="laplacian", N=4096, t=10.0, m=40, tol=1e-8, f32=False),
    "convdiff256_f32": dict(kind="convdiff", N=256, t=0.5, m=50, tol=1e-8, f32=True),
    "convdiff256_f64": dict(kind="convdiff", N=256, t=0.5, m=50, tol=1e-8, f32=False),
}


def batch_workload(args, mek, gen):
    """256 problems of the batch benchmark (values U[-1,1] FP64, as
    configured) in one Expokit-controlled lockstep solve; the FP32-value
    request falls back to FP64 storage on these values, as in the bench (one
    warm-up solve, one measured)."""
    probs = gen.batch_problems(P=256, seed=0, f32_exact=False)
    mats = [(q["rp"], q["ci"], q["va"]) for q in probs]
    ms = [min(q["m"], q["rp"].size - 2) for q in probs]
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    B = mek.Batch(mats, [q["t"] for q in probs], ms, 1e-8, values_f32=True, gs_mode=gs_mode,
                  step_control="expokit")
    B.upload([q["v"] for q in probs])
    for _ in range(2):
        B.solve()
    mek.synchronize()
    print(f"[profile] batch256 solves done, {B.cycles()} cycles", file=sys.stderr, flush=True)
    B.close()


def block_workload(args, mek, gen):
    """2048^2 Laplacian with 32 right-hand sides, 16 block steps, t = 10 (one
    warm-up solve, one measured)."""
    import numpy as np
    N = 2048
    rp, ci, va = gen.laplacian_2d(N)
    V = np.random.default_rng(0).standard_normal((N * N, 32))
    V /= np.linalg.norm(V, axis=0)
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    B = mek.Block(mek.CSR.from_arrays(rp, ci, va), m=16, t=10.0, tol=1e-8, gs_mode=gs_mode)
    d_v = mek.DeviceBuffer.from_array(V.ravel())
    d_y = mek.DeviceBuffer(8 * V.size)
    for _ in range(2):
        B.expmv_device(d_v, d_y)
    mek.synchronize()
    print("[profile] block2048 solves done", file=sys.stderr, flush=True)
    B.close()


def workload(args):
    import torch
    mek = _load("mek", "__init__.py")
    gen = _load("mek_generate", "generate.py")
    src = torch.ones(CAL_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    del src, dst
    torch.cuda.empty_cache()
    if args.workload == "batch256":
        batch_workload(args, mek, gen)
        return
    if args.workload == "block2048":
        block_workload(args, mek, gen)
        return
    w = WORKLOADS[args.workload]
    if w["kind"] == "laplacian":
        rp, ci, va = gen.laplacian_2d(w["N"])
    else:
        rp, ci, va = gen.convdiff_3d(w["N"])
    v = gen.normal_vector(rp.size - 1, seed=0)
    gs_mode = {"always": 1, "never": 0, "dgks": 2}[args.gs]
    A = mek.CSR.from_arrays(rp, ci, va)
    # the bench's runner: autotuned launch shapes, s-step blocks as given
    sstep, sstep_first = mek.sstep_plan(w["m"], args.sstep, args.sstep_first)
    R = mek.Runner(A, m=w["m"], t=w["t"], tol=w["tol"], max_restarts=0, step_control="expokit",
                   gs_mode=gs_mode, values_f32=w["f32"], verbose=False, timing=False, autotune=True,
                   sstep=sstep, sstep_first=sstep_first)
    A.close()
    d_v = mek.DeviceBuffer.from_array(v)
    d_y = mek.DeviceBuffer(8 * v.size)
    for _ in range(2):
        R.expmv_device(d_v, d_y)
    mek.synchronize()
    print(f"[profile] {args.workload} solves done", file=sys.stderr, flush=True)
    R.close()


def _find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise SystemExit(f"no {pattern} under {d}")
    return hits


def read_stats(d):
    rows = []
    for fn in _find(d, "*kernel_stats.csv"):
        with open(fn) as f:
            rows += list(csv.DictReader(f))
    return rows


def read_counter(d, counter):
    """Per-dispatch counter values in dispatch order: list of (kernel name, value)."""
    per = defaultdict(float)
    names = {}
    for fn in _find(d, "*counter_collection.csv"):
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                key = (fn, int(r["Dispatch_Id"]))
                per[key] += float(r["Counter_Value"])
                names[key] = r["Kernel_Name"]
    return [(names[k], per[k]) for k in sorted(per)]


# A dispatch that moves less than this is a launch that returned at once (a
# conditional second pass that was not needed, work after a breakdown): it
# is left out of the per-launch averages of its kind.
WORKING_BYTES = 1 << 20
WORKING_NS = 20_000


def read_trace(d):
    """Per-dispatch durations (ns) from the kernel trace: list of (name, ns)."""
    out = []
    for fn in _find(d, "*kernel_trace.csv"):
        with open(fn) as f:
            for r in csv.DictReader(f):
                out.append((r["Kernel_Name"], int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    return out


def summarize(args):
    stats = read_stats(args.stats)
    have_counters = bool(args.fetch and args.write)
    fetch = read_counter(args.fetch, "FETCH_SIZE") if have_counters else []
    write = read_counter(args.write, "WRITE_SIZE") if have_counters else []

    def is_cal(name):
        return "neg" in name.lower()

    cal_f = [v for n, v in fetch if is_cal(n)] or [CAL_BYTES / 1024.0]
    cal_w = [v for n, v in write if is_cal(n)] or [CAL_BYTES / 1024.0]
    if have_counters and (len(cal_f) == 0 or len(cal_w) == 0):
        raise SystemExit("calibration kernel not found in the counter runs")
    # counters read KiB; scale so the calibration kernel reads CAL_BYTES
    f_scale = CAL_BYTES / (sum(cal_f) / len(cal_f) * 1024.0)
    w_scale = CAL_BYTES / (sum(cal_w) / len(cal_w) * 1024.0)

    def per_kind(vals, scale):
        out = defaultdict(list)
        for n, v in vals:
            if not is_cal(n):
                out[kind_of(n)].append(v * 1024.0 * scale)
        return out

    # the FETCH and WRITE runs execute the same launch sequence, so the i-th
    # dispatch of a kind is the same launch in both
    fk = per_kind(fetch, f_scale)
    wk = per_kind(write, w_scale)
    traffic, working, idle = {}, {}, {}
    for k in sorted(set(fk) & set(wk)):
        if len(fk[k]) != len(wk[k]):
            raise SystemExit(f"{k}: {len(fk[k])} FETCH dispatches but {len(wk[k])} WRITE dispatches")
        both = [f + w for f, w in zip(fk[k], wk[k])]
        act = [b for b in both if b >= WORKING_BYTES]
        idle[k] = len(both) - len(act)
        if act:
            traffic[k] = sum(act) / len(act)
            working[k] = len(act)
    out = {"workload": args.workload, "calibration": {
        "kernel": "torch.neg, 1 GiB FP64 read + 1 GiB written",
        "fetch_scale": f_scale, "write_scale": w_scale},
        "bytes_per_launch": traffic,
        "note": f"averaged over the launches that moved at least {WORKING_BYTES} bytes; "
                "launches that returned at once are counted under idle_launches",
        "working_launches": working, "idle_launches": idle}
    for k, v in traffic.items():
        out[k] = v
    if have_counters and args.out_json:
        with open(args.out_json, "w") as f:
            json.dump(out, f, indent=1)

    by_kind = defaultdict(lambda: [0, 0.0])
    lines = []
    for r in stats:
        name = r.get("Name") or r.get("KERNEL_NAME") or ""
        calls = int(r.get("Calls") or r.get("CALLS") or 0)
        tot = float(r.get("TotalDurationNs") or r.get("TOTAL_DURATION") or 0)
        k = kind_of(name)
        by_kind[k][0] += calls
        by_kind[k][1] += tot
        lines.append((tot, name, calls, tot / max(calls, 1)))
    lines.sort(reverse=True)
    with open(args.out_md, "w") as f:
        f.write(f"# Kernel profile: {args.workload}\n\n")
        f.write("Commands (run from the repository root on one MI355X):\n\n```\n")
        for c in args.command or []:
            f.write(c + "\n")
        f.write("```\n\n## rocprofv3 --kernel-trace --stats, by kernel kind\n\n")
        f.write("| kind | calls | total ms | avg us |\n|---|---|---|---|\n")
        for k, (c, t) in sorted(by_kind.items(), key=lambda kv: -kv[1][1]):
            f.write(f"| {k} | {c} | {t / 1e6:.3f} | {t / max(c, 1) / 1e3:.1f} |\n")
        f.write("\n## Top kernels (rocprofv3 kernel_stats.csv)\n\n")
        f.write("| kernel | calls | total ms | avg us |\n|---|---|---|---|\n")
        for tot, name, calls, avg in lines[:25]:
            short = name.split("(")[0].replace("void ", "")
            f.write(f"| `{short}` | {calls} | {tot / 1e6:.3f} | {avg / 1e3:.1f} |\n")
        trace = read_trace(args.stats)
        if trace:
            wk_t = defaultdict(list)
            for name, ns in trace:
                if ns >= WORKING_NS:
                    wk_t[kind_of(name)].append(ns)
            f.write(f"\n## Working launches (kernel trace, duration >= {WORKING_NS / 1e3:.0f} us)\n\n"
                    "Conditional launches that return at once (a second Gram-Schmidt pass that "
                    "was not needed) are left out here.\n\n")
            f.write("| kind | launches | total ms | avg us |\n|---|---|---|---|\n")
            for k, v in sorted(wk_t.items(), key=lambda kv: -sum(kv[1])):
                f.write(f"| {k} | {len(v)} | {sum(v) / 1e6:.3f} | {sum(v) / len(v) / 1e3:.1f} |\n")
        if not have_counters:
            return
        f.write("\n## HBM traffic per launch (FETCH_SIZE + WRITE_SIZE, calibrated)\n\n")
        f.write(f"Calibration: torch.neg on 1 GiB FP64 (1 GiB read, 1 GiB written); "
                f"FETCH_SIZE scale {f_scale:.4f}, WRITE_SIZE scale {w_scale:.4f} "
                f"(bytes = KiB x 1024 x scale).\n\n")
        f.write(f"Averages over the launches that moved at least {WORKING_BYTES >> 20} MiB.\n\n")
        f.write("| kind | working launches | idle launches | MB per launch (read + write) |\n"
                "|---|---|---|---|\n")
        for k, v in sorted(traffic.items(), key=lambda kv: -kv[1]):
            f.write(f"| {k} | {working[k]} | {idle[k]} | {v / 1e6:.1f} |\n")
        if args.cache:
            cache = cache_per_kind(args.cache)
            f.write("\n## L2 (TCC) and L1-to-L2 requests per launch\n\n"
                    f"Averages over the launches with at least {CACHE_MIN_REQ} L2 requests.\n\n"
                    "| kind | launches | TCC_REQ | TCC_HIT | TCC_MISS | L2 hit rate | TCP_TCC_READ_REQ |\n"
                    "|---|---|---|---|---|---|---|\n")
            for k, c in sorted(cache.items(), key=lambda kv: -kv[1]["TCC_REQ_sum"]):
                f.write(f"| {k} | {c['launches']} | {c['TCC_REQ_sum'] / 1e6:.2f}M | {c['TCC_HIT_sum'] / 1e6:.2f}M | "
                        f"{c['TCC_MISS_sum'] / 1e6:.2f}M | {c['hit_rate']:.3f} | "
                        f"{c['TCP_TCC_READ_REQ_sum'] / 1e6:.2f}M |\n")
            out["cache_per_launch"] = cache
            if args.out_json:
                with open(args.out_json, "w") as g:
                    json.dump(out, g, indent=1)
    print(json.dumps(out))


CACHE_COUNTERS = ("TCC_REQ_sum", "TCC_HIT_sum", "TCC_MISS_sum", "TCP_TCC_READ_REQ_sum")
CACHE_MIN_REQ = 8192  # L2 requests of a launch that did work (about 1 MiB)


def cache_per_kind(d):
    """Per kind: mean of each cache counter over the working launches."""
    per = {c: read_counter(d, c) for c in CACHE_COUNTERS}
    n = len(per["TCC_REQ_sum"])
    if any(len(v) != n for v in per.values()):
        raise SystemExit("cache counters: dispatch counts differ between counters")
    acc = defaultdict(lambda: defaultdict(float))
    for i in range(n):
        name, req = per["TCC_REQ_sum"][i]
        if "neg" in name.lower() or req < CACHE_MIN_REQ:
            continue
        a = acc[kind_of(name)]
        a["launches"] += 1
        for c in CACHE_COUNTERS:
            a[c] += per[c][i][1]
    out = {}
    for k, a in acc.items():
        L = a["launches"]
        out[k] = {c: a[c] / L for c in CACHE_COUNTERS}
        out[k]["launches"] = int(L)
        out[k]["hit_rate"] = a["TCC_HIT_sum"] / max(a["TCC_REQ_sum"], 1.0)
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    a = sub.add_parser("workload")
    a.add_argument("--workload", default="laplacian4096", choices=sorted(WORKLOADS) + ["batch256", "block2048"])
    a.add_argument("--gs", default="dgks", choices=["always", "dgks", "never"])
    a.add_argument("--sstep", type=int, default=-1,
                   help="s-step block size (0: column by column; -1: the bench's shape, mek.sstep_plan)")
    a.add_argument("--sstep-first", type=int, default=-1)
    s = sub.add_parser("summarize")
    s.add_argument("--stats", required=True)
    s.add_argument("--fetch", default=None)
    s.add_argument("--write", default=None)
    s.add_argument("--cache", default=None, help="directory of a TCC_REQ/HIT/MISS + TCP_TCC_READ_REQ run")
    s.add_argument("--workload", required=True)
    s.add_argument("--out-md", required=True)
    s.add_argument("--out-json", default=None)
    s.add_argument("--command", action="append")
    args = ap.parse_args()
    if args.mode == "workload":
        workload(args)
    else:
        summarize(args)


if __name__ == "__main__":
    main()
