"""Run and profile the matrix_exp command line on one MI355X.

The run step of the reference's launcher (launch.sh:271-377) for this build:
a gzip-compressed matrix is decompressed to a temporary file first
(launch.sh:271-288; a ``.mtx.gz`` is passed through, since the loader reads
gzip MatrixMarket itself), the Arnoldi flags are forwarded
(launch.sh:294-304), and the binary runs

* ``--profiler trace`` (the reference's nsys mode, launch.sh:306-357):
  under ``rocprofv3 --marker-trace --kernel-trace --stats``.  The solver opens
  roctx ranges with the reference's NVTX names (main.cu:1413, :1437, :1575,
  :1770, :1955), so the summary lists each range with its count and time,
  and the kernels by total time;
* ``--profiler counters`` (the reference's ncu mode, launch.sh:358-372, which
  collects FP64 add / mul / FMA instruction counts): under ``rocprofv3 --pmc``
  with the FP64 VALU and MFMA instruction counters, summed per kernel.
  Counters are collected in a run of their own (no marker trace);
* ``--profiler none``: directly (launch.sh:373-376).

The exit status is the program's; a profiled run that leaves no profiler
output exits with 2, as the reference's does for a missing report
(launch.sh:350-354).  ``--summary FILE`` writes the summary as JSON.

    python matrix-exp-krylov_amd/launch.py MATRIX|SIZE [--m M] [--t T]
        [--max-restarts R] [--profiler trace|counters|none] [--out DIR]
        [--summary FILE] [-- extra matrix_exp flags]
"""
from __future__ import annotations

import argparse
import csv
import glob
import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
BINARY = os.path.join(HERE, "bin", "matrix_exp")

# roctx ranges the solver opens (the reference's NVTX range names)
RANGES = ("total_compute_expmv", "arnoldi_step_attempt", "init_q1", "orthogonalize_mgs",
          "small_expm_and_lift")
# FP64 instruction counters: the counterparts of the reference's ncu metrics
# (dadd, dmul, dfma) plus the FP64 MFMA operations of the block kernels
FP64_COUNTERS = ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
                 "SQ_INSTS_VALU_MFMA_MOPS_F64")


def _rocprofv3():
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        raise SystemExit("ERROR: rocprofv3 not found")
    return exe


def _csv_rows(d, pattern):
    rows = []
    for fn in sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True)):
        with open(fn, newline="") as f:
            rows += list(csv.DictReader(f))
    return rows


def summarize_trace(d):
    """Ranges and kernels of a --marker-trace --kernel-trace run under d."""
    ranges = defaultdict(lambda: {"count": 0, "total_ms": 0.0})
    for r in _csv_rows(d, "*marker_api_trace.csv"):
        name = r.get("Function") or r.get("Message") or ""
        if not name or "Start_Timestamp" not in r:
            continue
        ns = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        ranges[name]["count"] += 1
        ranges[name]["total_ms"] += ns * 1e-6
    kernels = defaultdict(lambda: {"calls": 0, "total_ms": 0.0})
    for r in _csv_rows(d, "*kernel_trace.csv"):
        k = r["Kernel_Name"].split("(")[0]
        kernels[k]["calls"] += 1
        kernels[k]["total_ms"] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
    top = sorted(kernels.items(), key=lambda kv: -kv[1]["total_ms"])
    return {"ranges": {k: dict(v) for k, v in sorted(ranges.items())},
            "kernels": [{"kernel": k, **v} for k, v in top]}


def summarize_counters(d):
    """FP64 instruction counts per kernel (summed over dispatches) of a --pmc run."""
    per = defaultdict(lambda: defaultdict(float))
    for r in _csv_rows(d, "*counter_collection.csv"):
        per[r["Kernel_Name"].split("(")[0]][r["Counter_Name"]] += float(r["Counter_Value"])
    totals = defaultdict(float)
    for c in per.values():
        for name, v in c.items():
            totals[name] += v
    return {"counters": {k: dict(v) for k, v in sorted(per.items())}, "totals": dict(totals)}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("matrix", help="matrix file (.bincsr, .mtx, .gz) or a built-in size")
    ap.add_argument("--m", type=int)
    ap.add_argument("--t", type=float)
    ap.add_argument("--max-restarts", type=int)
    ap.add_argument("--profiler", default="trace", choices=["trace", "counters", "none"])
    ap.add_argument("--out", default="launch_out", help="rocprofv3 output directory")
    ap.add_argument("--summary", help="write the profile summary here (JSON)")
    ap.add_argument("--binary", default=BINARY)
    ap.add_argument("extra", nargs="*", help="further matrix_exp flags (after --)")
    a = ap.parse_args(argv)

    if not os.path.exists(a.binary):
        raise SystemExit(f"ERROR: {a.binary} not built (make -C {HERE})")
    tmp = None
    arg = a.matrix
    if arg.endswith(".gz") and not arg.endswith(".mtx.gz"):
        fd, tmp = tempfile.mkstemp(suffix="_" + os.path.basename(arg)[:-3])
        print(f"Decompressing matrix to {tmp} ...", flush=True)
        try:
            with gzip.open(arg, "rb") as src, os.fdopen(fd, "wb") as dst:
                shutil.copyfileobj(src, dst)
        except OSError as e:
            print(f"ERROR: Failed to decompress {arg}: {e}", flush=True)
            return 1
        arg = tmp
    cmd = [a.binary, arg]
    if a.m is not None:
        cmd += ["--m", str(a.m)]
    if a.t is not None:
        cmd += ["--t", repr(a.t)]
    if a.max_restarts is not None:
        cmd += ["--max-restarts", str(a.max_restarts)]
    cmd += a.extra

    try:
        if a.profiler == "none":
            print("Profiler disabled; running binary directly.", flush=True)
            return subprocess.call(cmd)
        os.makedirs(a.out, exist_ok=True)
        if a.profiler == "trace":
            prof = [_rocprofv3(), "--marker-trace", "--kernel-trace", "--stats", "--output-format", "csv",
                    "-d", a.out, "-o", "run", "--"]
        else:
            prof = [_rocprofv3(), "--pmc", *FP64_COUNTERS, "--output-format", "csv", "-d", a.out, "-o", "run", "--"]
        print("Profiling output will be saved to: " + a.out, flush=True)
        rc = subprocess.call(prof + cmd)
        if a.profiler == "trace":
            found = glob.glob(os.path.join(a.out, "**", "*kernel_trace.csv"), recursive=True)
        else:
            found = glob.glob(os.path.join(a.out, "**", "*counter_collection.csv"), recursive=True)
        if rc == 0 and not found:
            print(f"ERROR: rocprofv3 finished but left no output under {a.out}", flush=True)
            return 2
        if found:
            summary = summarize_trace(a.out) if a.profiler == "trace" else summarize_counters(a.out)
            summary["command"] = prof + cmd
            summary["exit_code"] = rc
            if a.profiler == "trace":
                print("roctx ranges (count, total ms):", flush=True)
                for name, r in summary["ranges"].items():
                    print(f"  {name:28s} {r['count']:6d} {r['total_ms']:12.3f}")
                print("top kernels (calls, total ms):")
                for k in summary["kernels"][:10]:
                    print(f"  {k['kernel'][:60]:60s} {k['calls']:6d} {k['total_ms']:12.3f}")
            else:
                print("FP64 instruction counts (all kernels):", flush=True)
                for name, v in sorted(summary["totals"].items()):
                    print(f"  {name:32s} {v:16.0f}")
            if a.summary:
                with open(a.summary, "w") as f:
                    json.dump(summary, f, indent=1)
        return rc
    finally:
        if tmp is not None:
            os.unlink(tmp)


if __name__ == "__main__":
    sys.exit(main())
