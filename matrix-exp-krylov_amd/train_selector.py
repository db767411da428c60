"""Train the learned (m, substeps) selector on the GPU and evaluate it against
the Expokit-style heuristic on held-out problems.

    python matrix-exp-krylov_amd/train_selector.py --train 2000 --test 500 \
        --out matrix-exp-krylov_amd/selector_weights.npz --report report.json

Large sets are labelled in parts (each part a separate run), then trained on
together:

    python matrix-exp-krylov_amd/train_selector.py --train 1200 --n-lo 1000 --n-hi 4000000 \
        --label-shard 0/4 --dataset part0.npz          # ... parts 1/4 to 3/4
    python matrix-exp-krylov_amd/train_selector.py --train-from part0.npz,part1.npz,... \
        --test 300 --n-lo 1000 --n-hi 4000000 --no-test-labels --report report.json

Evaluation runs both choices through the batched solver (Expokit step
control on the device, whose local error estimate rejects and shortens any step
that is too long) and reports total Krylov columns (SpMVs), modelled HBM
bytes, wall time (the learned form's including its selection: device
features, MLP, decoding) and the relative error against a tight-tolerance
solve of each problem.  The held-out set is built and solved a batch at a
time, so it can be large:

    python matrix-exp-krylov_amd/train_selector.py --eval-only --test 2000 \
        --test-n-lo 10000 --test-n-hi 1000000 --report heldout.json
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name, fn):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, fn))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


mek = _load("mek", "__init__.py")
gen = _load("mek_generate", "generate.py")
sel = _load("mek_selector", "selector.py")


def log(msg):
    print(f"[selector] {msg}", file=sys.stderr, flush=True)


FORMS = ("heuristic", "learned")


def evaluate_chunk(selector, problems, tol=1e-8, device=0, vseed=10_000):
    """One batch of held-out problems, solved twice by the batched solver
    (FP32 values where lossless, DGKS second pass, Expokit step control with
    the per-unit-time tolerance tol/t): with the heuristic choice (m = 30,
    Expokit's first step) and with the learned (m, first step t/substeps).

    The learned form's time includes its selection: the device features of
    the resident matrices (Batch.features), the MLP forward pass and the
    decoding.  Each form's timed solve follows an untimed one that captures
    its cycle graphs; matrix assembly and upload (setup) are timed apart for
    both forms.  Errors are against a tight FP64 solve (tol 1e-13 per unit
    time, m = 60, second pass always); the relative error is taken where
    ||y|| >= 1e-3 ||v||, the absolute error (||v|| = 1) everywhere.

    Returns per-form sums and lists, to be merged over chunks."""
    P = len(problems)
    mats = [(rp, ci, va) for rp, ci, va, t, kind in problems]
    kinds = [kind for *_x, kind in problems]
    ts = [t for *_x, t, kind in problems]
    rows = [rp.size - 1 for rp, *_r in mats]
    vs = [gen.normal_vector(n, seed=vseed + i) for i, n in enumerate(rows)]
    csrs = [mek.CSR.from_arrays(*A) for A in mats]
    B = mek.Batch(csrs, ts, [min(60, n - 2) for n in rows], [1e-13 / t for t in ts],
                  values_f32=False, gs_mode=1, step_control="expokit_rel", device=device)
    B.upload(vs)
    B.solve()
    refs = B.download()
    B.close()
    # selection on a resident copy of the matrices
    t0 = time.perf_counter()
    FB = mek.Batch(csrs, ts, [min(30, n - 2) for n in rows], tol, values_f32=True, device=device)
    feat_setup = time.perf_counter() - t0
    mek.synchronize()
    t0 = time.perf_counter()
    F = FB.features()
    learned = selector.predict_features(F)
    select_s = time.perf_counter() - t0
    FB.close()
    out = {"problems": P, "select_s": select_s, "features_setup_s": feat_setup,
           "rows": int(sum(rows)), "nnz": int(sum(int(rp[-1]) for rp, *_r in mats))}
    for name, choice in (("heuristic", [(30, 0)] * P), ("learned", learned)):
        ms = [min(m, n - 2) for (m, s), n in zip(choice, rows)]
        ss = [s for m, s in choice]
        t0 = time.perf_counter()
        B = mek.Batch(csrs, ts, ms, [tol / t for t in ts], substeps=ss, values_f32=True, gs_mode=2,
                      step_control="expokit_rel", device=device)
        B.upload(vs)
        setup = time.perf_counter() - t0
        B.solve()  # captures the cycle graphs
        mek.synchronize()
        t0 = time.perf_counter()
        B.solve()
        mek.synchronize()
        wall = time.perf_counter() - t0
        ys = B.download()
        f = {"wall_s": wall + (select_s if name == "learned" else 0.0), "solve_s": wall,
             "setup_s": setup + (feat_setup if name == "learned" else 0.0),
             "krylov_columns": 0, "modelled_bytes": 0.0, "cycles": B.cycles(), "converged": 0,
             "rejected_steps": 0, "values_f32_used": int(B.values_f32_used()),
             "rel_errs": [], "abs_errs": [], "columns_by_kind": {}, "m": ms}
        for p in range(P):
            st = B.stats(p)
            f["krylov_columns"] += st["spmv_count"]
            f["converged"] += st["converged"]
            f["rejected_steps"] += st["rejected"]
            f["columns_by_kind"][kinds[p]] = f["columns_by_kind"].get(kinds[p], 0) + st["spmv_count"]
            f["modelled_bytes"] += sel.modelled_bytes(ms[p], 1, rows[p], int(mats[p][0][-1])) * \
                st["spmv_count"] / max(ms[p], 1)
            r = refs[p]
            d = float(np.linalg.norm(ys[p] - r))
            f["abs_errs"].append(d)
            if np.linalg.norm(r) >= 1e-3:
                f["rel_errs"].append(d / float(np.linalg.norm(r)))
        B.close()
        out[name] = f
    for c in csrs:
        c.close()
    return out


def merge_eval(parts):
    """Totals over evaluated chunks, with the wall ratios that include the
    learned form's selection time."""
    out = {"chunks": len(parts), "problems": sum(q["problems"] for q in parts),
           "rows": sum(q["rows"] for q in parts), "nnz": sum(q["nnz"] for q in parts),
           "select_s": sum(q["select_s"] for q in parts),
           "features_setup_s": sum(q["features_setup_s"] for q in parts)}
    for name in FORMS:
        fs = [q[name] for q in parts]
        rel = [e for f in fs for e in f["rel_errs"]]
        ab = [e for f in fs for e in f["abs_errs"]]
        kinds = {}
        for f in fs:
            for k, c in f["columns_by_kind"].items():
                kinds[k] = kinds.get(k, 0) + c
        ms = [m for f in fs for m in f["m"]]
        out[name] = {k: sum(f[k] for f in fs) for k in ("wall_s", "solve_s", "setup_s", "krylov_columns",
                                                          "modelled_bytes", "cycles", "converged",
                                                          "rejected_steps")}
        out[name].update({"max_rel_err_y_ge_1e-3": max(rel) if rel else None,
                          "median_rel_err": float(np.median(rel)) if rel else None,
                          "max_abs_err": max(ab), "columns_by_kind": kinds,
                          "values_f32_used_chunks": sum(f["values_f32_used"] for f in fs),
                          "m_hist": {int(m): int(c) for m, c in zip(*np.unique(ms, return_counts=True))}})
    h, l_ = out["heuristic"], out["learned"]
    out["columns_ratio_heuristic_over_learned"] = h["krylov_columns"] / max(l_["krylov_columns"], 1)
    out["bytes_ratio_heuristic_over_learned"] = h["modelled_bytes"] / max(l_["modelled_bytes"], 1.0)
    out["wall_ratio_heuristic_over_learned"] = h["wall_s"] / max(l_["wall_s"], 1e-12)
    out["wall_ratio_excluding_selection"] = h["solve_s"] / max(l_["solve_s"], 1e-12)
    out["wall_includes"] = ("learned: device features of the resident matrices + MLP + decoding + the "
                            "timed solve; heuristic: the timed solve (its first step needs ||A||, "
                            "computed at setup by both)")
    return out


def evaluate(selector, specs, tol=1e-8, device=0, max_rows=40_000_000, f32_exact=False, workers=8):
    """The held-out problems of `specs`, built by CPU workers and evaluated a
    chunk of at most `max_rows` rows at a time (evaluate_chunk), merged."""
    import multiprocessing as mp

    chunks, cur, rows = [], [], 0
    for sp in specs:
        r = sel.spec_rows(sp)
        if cur and rows + r > max_rows:
            chunks.append(cur)
            cur, rows = [], 0
        cur.append(sp)
        rows += r
    if cur:
        chunks.append(cur)
    parts = []
    done = 0
    with mp.get_context("spawn").Pool(max(1, min(workers, os.cpu_count() or 1))) as pool:
        for c in chunks:
            probs = pool.starmap(_build, [(sp, 2, f32_exact) for sp in c])
            parts.append(evaluate_chunk(selector, probs, tol, device, vseed=10_000 + done))
            done += len(c)
            del probs
            q = parts[-1]
            log(f"evaluated {done}/{len(specs)}: heuristic {q['heuristic']['wall_s']:.3f} s, "
                f"learned {q['learned']['wall_s']:.3f} s (selection {q['select_s'] * 1e3:.1f} ms)")
    return merge_eval(parts)


def calibrate(tol=1e-8, device=0, log_=log):
    """Measured cost of one step of the batched solver (FP64 values, DGKS,
    Expokit control): for each problem kind, basis size m of the grid and two
    problem sizes, the wall time of a lockstep solve in which every problem
    takes exactly one accepted m-column step (t small enough that the first
    step covers it), divided by the rows solved.  Returns
    {"n": [n0, n1], kind: {m: [s/row at n0, s/row at n1]}} for
    selector.step_seconds."""
    sizes = ((25_000, 64), (250_000, 16))
    out = {"n": [n for n, _ in sizes], "values": "fp64", "gs": "dgks", "control": "expokit_rel"}
    for kind in ("lap2d", "cd3d", "rows"):
        out[kind] = {}
        for n_target, P in sizes:
            mats = []
            for i in range(P):
                if kind == "lap2d":
                    mats.append(gen.laplacian_2d(int(round(n_target ** 0.5))))
                elif kind == "cd3d":
                    mats.append(gen.convdiff_3d(int(round(n_target ** (1 / 3))), pe=10.0))
                else:
                    mats.append(gen.random_rows(n_target, 5, 30, seed=7000 + i))
            csrs = [mek.CSR.from_arrays(*A) for A in mats]
            rows = [A[0].size - 1 for A in mats]
            vs = [gen.normal_vector(n, seed=8000 + i) for i, n in enumerate(rows)]
            ts = [1e-6 / sel.norm1(*A, n) for A, n in zip(mats, rows)]
            for m in sel.M_GRID:
                B = mek.Batch(csrs, ts, [m] * P, [tol / t for t in ts], substeps=[1] * P, values_f32=False,
                              gs_mode=2, step_control="expokit_rel", device=device)
                B.upload(vs)
                B.solve()
                mek.synchronize()
                reps = 3
                t0 = time.perf_counter()
                for _ in range(reps):
                    B.solve()
                mek.synchronize()
                wall = (time.perf_counter() - t0) / reps
                cols = sum(B.stats(p)["spmv_count"] for p in range(P))
                if cols != m * P or B.cycles() != 1:
                    raise RuntimeError(f"calibration {kind} m={m}: {cols} columns in {B.cycles()} cycles")
                B.close()
                out[kind].setdefault(m, []).append(wall / sum(rows))
            for c in csrs:
                c.close()
            log_(f"calibrated {kind} at n = {rows[0]}: " +
                 ", ".join(f"m={m}: {out[kind][m][-1] * 1e12:.0f} ps/row" for m in sel.M_GRID))
    return out


def label_specs(specs, tol, chunk, seed=1):
    """Features and labels of the problems of `specs`, built and labelled
    `chunk` at a time so that only one chunk of matrices is held at once."""
    import multiprocessing as mp

    X, T, meta = [], [], []
    # matrices are built by CPU workers started fresh ("spawn": they never
    # touch the GPU) while this process labels on the GPU
    with mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        for c0 in range(0, len(specs), chunk):
            probs = pool.starmap(_build, [(sp, seed) for sp in specs[c0:c0 + chunk]])
            Xc, Tc, mc = sel.build_dataset(probs, tol, chunk=chunk)
            X.append(Xc)
            T.append(Tc)
            meta += mc
            log(f"labelled {min(c0 + chunk, len(specs))}/{len(specs)}")
    return np.concatenate(X), np.concatenate(T), meta


def _build(spec, seed, f32_exact=True):
    """make_problem in a worker process (the selector module alone)."""
    spec_mod = importlib.util.spec_from_file_location("mek_selector", os.path.join(HERE, "selector.py"))
    if "mek_selector" not in sys.modules:
        mod = importlib.util.module_from_spec(spec_mod)
        sys.modules["mek_selector"] = mod
        spec_mod.loader.exec_module(mod)
    return sys.modules["mek_selector"].make_problem(spec, seed, f32_exact)


MAX_CAND = 12  # grid m values plus one breakdown-sized candidate


def meta_arrays(meta):
    """The label metadata as plain arrays: kind, n, t, the label's m and
    first sub-steps (m_b, s_b), the fewest-SpMV m and sub-steps (m_s, s_s),
    and every candidate solved, padded with -1 (cand_m, cand_s, cand_cols:
    m, first sub-steps, Krylov columns built)."""
    cols = list(zip(*meta))
    cand = np.full((len(meta), 3, MAX_CAND), -1, np.int64)
    for p, cl in enumerate(cols[7]):
        for q, (m, s, c) in enumerate(cl[:MAX_CAND]):
            cand[p, :, q] = (m, s, c)
    return {"kind": np.array(cols[0], dtype="U8"), "n": np.array(cols[1], np.int64),
            "t": np.array(cols[2], np.float64), "m_b": np.array(cols[3], np.int32),
            "s_b": np.array(cols[4], np.int64), "m_s": np.array(cols[5], np.int32),
            "s_s": np.array(cols[6], np.int64), "cand_m": cand[:, 0], "cand_s": cand[:, 1],
            "cand_cols": cand[:, 2]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=500)
    ap.add_argument("--n-lo", type=int, default=1000)
    ap.add_argument("--n-hi", type=int, default=100000)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--out", default=sel.WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--label-shard", default=None,
                    help="i/k: label only the i-th of k contiguous parts of the training set into --dataset")
    ap.add_argument("--dataset", default=None, help="output of --label-shard (.npz)")
    ap.add_argument("--train-from", default=None, help="comma-separated --label-shard outputs to train on")
    ap.add_argument("--chunk", type=int, default=128, help="problems built and labelled at a time")
    ap.add_argument("--no-test-labels", action="store_true",
                    help="skip labelling the held-out set (its evaluation solves still run)")
    ap.add_argument("--test-n-lo", type=int, default=None, help="held-out n range (default: --n-lo)")
    ap.add_argument("--test-n-hi", type=int, default=None, help="held-out n range (default: --n-hi)")
    ap.add_argument("--eval-rows", type=int, default=40_000_000, help="rows per evaluated batch")
    ap.add_argument("--seed", type=int, default=1,
                    help="seed of the training problems (the held-out set uses 2, so it may not)")
    ap.add_argument("--calibrate", default=None,
                    help="measure the batched solver's step cost per kind and m into this JSON file and exit")
    ap.add_argument("--relabel", action="store_true",
                    help="with --train-from: recompute each label from its recorded candidates with the "
                         "measured step costs before training")
    ap.add_argument("--eval-f32-exact", action="store_true",
                    help="round the held-out values to FP32 (default: values as generated)")
    args = ap.parse_args()
    report = {"train": args.train, "test": args.test, "n_range": [args.n_lo, args.n_hi], "tol": args.tol}
    if args.seed == 2:
        raise SystemExit("--seed 2 draws the held-out problems")
    t0 = time.time()
    if args.calibrate:
        c = calibrate(args.tol)
        with open(args.calibrate, "w") as f:
            json.dump(c, f, indent=1)
        log(f"step costs written to {args.calibrate} in {time.time() - t0:.0f} s")
        return
    if args.label_shard:
        i, k = (int(x) for x in args.label_shard.split("/"))
        specs = sel.problem_specs(args.train, seed=args.seed, n_lo=args.n_lo, n_hi=args.n_hi)
        part = specs[len(specs) * i // k:len(specs) * (i + 1) // k]
        X, T, meta = label_specs(part, args.tol, args.chunk, args.seed)
        np.savez(args.dataset, X=X, T=T, **meta_arrays(meta))
        log(f"shard {i}/{k}: {len(part)} problems labelled in {time.time() - t0:.0f} s")
        return
    if args.train_from:
        parts = [np.load(f) for f in args.train_from.split(",")]
        merged = {k: np.concatenate([z[k] for z in parts]) for k in parts[0].files}
        if args.relabel:
            costs = sel.load_step_costs()
            if costs is None:
                raise SystemExit(f"--relabel needs {sel.COSTS} (train_selector.py --calibrate)")
            merged["T"], merged["m_b"], merged["s_b"] = sel.relabel(merged, costs)
            report["labels"] = "least measured step cost (selector_costs.json) over the recorded candidates"
        X, T, m_b = merged["X"], merged["T"], merged["m_b"]
        log(f"training on {len(X)} labelled problems from {len(parts)} shards")
        t1 = time.time()
        selector, loss = sel.train(X, T, epochs=args.epochs)
        report.update(train=int(len(X)), train_mse=loss, train_seconds=time.time() - t1)
        selector.save(args.out)
        report["label_m_hist"] = {int(a): int(c) for a, c in zip(*np.unique(m_b, return_counts=True))}
        # the labelled set beside the weights: features, targets and label metadata
        np.savez(os.path.join(os.path.dirname(os.path.abspath(args.out)), "selector_dataset.npz"), **merged)
    elif not args.eval_only:
        log(f"generating {args.train} training problems")
        train_p = sel.training_set(args.train, seed=args.seed, n_lo=args.n_lo, n_hi=args.n_hi)
        X, T, meta = sel.build_dataset(train_p, args.tol, log=log)
        log(f"labels done in {time.time() - t0:.0f} s; training on the GPU")
        t1 = time.time()
        selector, loss = sel.train(X, T, epochs=args.epochs)
        report["train_mse"] = loss
        report["train_seconds"] = time.time() - t1
        selector.save(args.out)
        report["label_m_hist"] = {int(a): int(c) for a, c in zip(*np.unique([mt[3] for mt in meta],
                                                                           return_counts=True))}
        np.savez(os.path.join(os.path.dirname(os.path.abspath(args.out)), "selector_dataset.npz"), X=X, T=T)
    else:
        selector = sel.Selector.load(args.out)
    n_lo = args.test_n_lo or args.n_lo
    n_hi = args.test_n_hi or args.n_hi
    log(f"held-out evaluation on {args.test} problems, n in [{n_lo}, {n_hi}]")
    test_specs = sel.problem_specs(args.test, seed=2, n_lo=n_lo, n_hi=n_hi)
    report["test_n_range"] = [n_lo, n_hi]
    report["test_values"] = "FP32-rounded" if args.eval_f32_exact else "as generated (random rows: U[-1,1] FP64)"
    if not args.no_test_labels:
        Xt, Tt, meta_t = label_specs(test_specs, args.tol, args.chunk)
        pred = selector.predict_features(Xt)
        report["heldout_label_bytes"] = float(sum(sel.modelled_bytes(mt[3], mt[4], mt[1], 0) for mt in meta_t))
        report["heldout_pred_m_abs_err_mean"] = float(np.mean([abs(p[0] - mt[3]) for p, mt in zip(pred, meta_t)]))
    report["eval"] = evaluate(selector, test_specs, args.tol, max_rows=args.eval_rows,
                              f32_exact=args.eval_f32_exact)
    report["total_seconds"] = time.time() - t0
    js = json.dumps(report, indent=1)
    print(js)
    if args.report:
        with open(args.report, "w") as f:
            f.write(js)


if __name__ == "__main__":
    main()
