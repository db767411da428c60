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
bytes, wall time and the relative error against a tight-tolerance solve of
each problem.
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


def evaluate(selector, problems, tol=1e-8, device=0):
    """Batched solves (FP32 values, DGKS second pass, Expokit step control
    with the per-unit-time tolerance tol/t) with the heuristic choice (m = 30,
    Expokit's first step) and with the learned (m, first step t/substeps).
    Errors are against a tight-tolerance FP64 solve; the relative error is
    taken over the problems whose solution keeps ||y|| >= 1e-3 ||v||, and the
    absolute error (||v|| = 1) over all."""
    P = len(problems)
    mats = [(rp, ci, va) for rp, ci, va, t, kind in problems]
    ts = [t for *_, t, kind in problems]
    vs = [gen.normal_vector(rp.size - 1, seed=10_000 + i) for i, (rp, *_rest) in enumerate(problems)]
    csrs = [mek.CSR.from_arrays(*A) for A in mats]
    # reference solutions: one FP64 lockstep solve at a tight tolerance, m = 60,
    # second Gram-Schmidt pass always
    B = mek.Batch(csrs, ts, [min(60, rp.size - 2) for rp, *_r in mats], [1e-13 / t for t in ts],
                  values_f32=False, gs_mode=1, step_control="expokit_rel", device=device)
    B.upload(vs)
    B.solve()
    refs = B.download()
    B.close()
    t0 = time.perf_counter()
    F = np.array([mek.features(c, t, tol, device=device) for c, t in zip(csrs, ts)])
    learned = selector.predict_features(F)
    select_s = time.perf_counter() - t0
    heur = [(30, 0) for _ in range(P)]
    out = {"select_s": select_s}
    for name, choice in (("heuristic", heur), ("learned", learned)):
        ms = [min(m, rp.size - 2) for (m, s), (rp, *_r) in zip(choice, mats)]
        ss = [s for m, s in choice]
        B = mek.Batch(csrs, ts, ms, [tol / t for t in ts], substeps=ss, values_f32=True, gs_mode=2,
                      step_control="expokit_rel", device=device)
        B.upload(vs)
        B.solve()  # warm-up (graph capture)
        t0 = time.perf_counter()
        B.solve()
        wall = time.perf_counter() - t0
        ys = B.download()
        cols = 0
        bytes_ = 0.0
        errs, abs_errs = [], []
        conv = rej = 0
        for p in range(P):
            st = B.stats(p)
            cols += st["spmv_count"]
            conv += st["converged"]
            rej += st["rejected"]
            n = mats[p][0].size - 1
            nnz = int(mats[p][0][-1])
            bytes_ += sel.modelled_bytes(ms[p], 1, n, nnz) * st["spmv_count"] / max(ms[p], 1)
            r = refs[p]
            d = float(np.linalg.norm(ys[p] - r))
            abs_errs.append(d)  # ||v|| = 1
            if np.linalg.norm(r) >= 1e-3:  # relative error where y has not decayed away
                errs.append(d / float(np.linalg.norm(r)))
        out[name] = {"krylov_columns": cols, "modelled_bytes": bytes_, "wall_s": wall,
                     "cycles": B.cycles(), "max_rel_err_y_ge_1e-3": max(errs),
                     "median_rel_err": float(np.median(errs)), "max_abs_err": max(abs_errs),
                     "converged": conv, "rejected_steps": rej, "problems": P}
        B.close()
    out["columns_ratio_heuristic_over_learned"] = out["heuristic"]["krylov_columns"] / max(out["learned"]["krylov_columns"], 1)
    out["bytes_ratio_heuristic_over_learned"] = out["heuristic"]["modelled_bytes"] / max(out["learned"]["modelled_bytes"], 1.0)
    out["wall_ratio_heuristic_over_learned"] = out["heuristic"]["wall_s"] / max(out["learned"]["wall_s"], 1e-12)
    return out


def label_specs(specs, tol, chunk):
    """Features and labels of the problems of `specs`, built and labelled
    `chunk` at a time so that only one chunk of matrices is held at once."""
    import multiprocessing as mp

    X, T, meta = [], [], []
    # matrices are built by CPU workers started fresh ("spawn": they never
    # touch the GPU) while this process labels on the GPU
    with mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        for c0 in range(0, len(specs), chunk):
            probs = pool.starmap(_build, [(sp, 1) for sp in specs[c0:c0 + chunk]])
            Xc, Tc, mc = sel.build_dataset(probs, tol, chunk=chunk)
            X.append(Xc)
            T.append(Tc)
            meta += mc
            log(f"labelled {min(c0 + chunk, len(specs))}/{len(specs)}")
    return np.concatenate(X), np.concatenate(T), meta


def _build(spec, seed):
    """make_problem in a worker process (the selector module alone)."""
    spec_mod = importlib.util.spec_from_file_location("mek_selector", os.path.join(HERE, "selector.py"))
    if "mek_selector" not in sys.modules:
        mod = importlib.util.module_from_spec(spec_mod)
        sys.modules["mek_selector"] = mod
        spec_mod.loader.exec_module(mod)
    return sys.modules["mek_selector"].make_problem(spec, seed)


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
    args = ap.parse_args()
    report = {"train": args.train, "test": args.test, "n_range": [args.n_lo, args.n_hi], "tol": args.tol}
    t0 = time.time()
    if args.label_shard:
        i, k = (int(x) for x in args.label_shard.split("/"))
        specs = sel.problem_specs(args.train, seed=1, n_lo=args.n_lo, n_hi=args.n_hi)
        part = specs[len(specs) * i // k:len(specs) * (i + 1) // k]
        X, T, meta = label_specs(part, args.tol, args.chunk)
        np.savez(args.dataset, X=X, T=T, **meta_arrays(meta))
        log(f"shard {i}/{k}: {len(part)} problems labelled in {time.time() - t0:.0f} s")
        return
    if args.train_from:
        parts = [np.load(f) for f in args.train_from.split(",")]
        merged = {k: np.concatenate([z[k] for z in parts]) for k in parts[0].files}
        X, T, m_b = merged["X"], merged["T"], merged["m_b"]
        log(f"training on {len(X)} labelled problems from {len(parts)} shards")
        t1 = time.time()
        selector, loss = sel.train(X, T, epochs=args.epochs)
        report.update(train=int(len(X)), train_mse=loss, train_seconds=time.time() - t1)
        selector.save(args.out)
        report["label_m_hist"] = np.bincount(m_b.astype(int), minlength=61)[10::5].tolist()
        # the labelled set beside the weights: features, targets and label metadata
        np.savez(os.path.join(os.path.dirname(os.path.abspath(args.out)), "selector_dataset.npz"), **merged)
    elif not args.eval_only:
        log(f"generating {args.train} training problems")
        train_p = sel.training_set(args.train, seed=1, n_lo=args.n_lo, n_hi=args.n_hi)
        X, T, meta = sel.build_dataset(train_p, args.tol, log=log)
        log(f"labels done in {time.time() - t0:.0f} s; training on the GPU")
        t1 = time.time()
        selector, loss = sel.train(X, T, epochs=args.epochs)
        report["train_mse"] = loss
        report["train_seconds"] = time.time() - t1
        selector.save(args.out)
        report["label_m_hist"] = np.bincount([mt[3] for mt in meta], minlength=61)[10::5].tolist()
        np.savez(os.path.join(os.path.dirname(os.path.abspath(args.out)), "selector_dataset.npz"), X=X, T=T)
    else:
        selector = sel.Selector.load(args.out)
    log(f"held-out evaluation on {args.test} problems")
    test_p = sel.training_set(args.test, seed=2, n_lo=args.n_lo, n_hi=args.n_hi)
    if not args.no_test_labels:
        Xt, Tt, meta_t = sel.build_dataset(test_p, args.tol, log=log)
        pred = selector.predict_features(Xt)
        report["heldout_label_bytes"] = float(sum(sel.modelled_bytes(mt[3], mt[4], mt[1], 0) for mt in meta_t))
        report["heldout_pred_m_abs_err_mean"] = float(np.mean([abs(p[0] - mt[3]) for p, mt in zip(pred, meta_t)]))
    report["eval"] = evaluate(selector, test_p, args.tol)
    report["total_seconds"] = time.time() - t0
    js = json.dumps(report, indent=1)
    print(js)
    if args.report:
        with open(args.report, "w") as f:
            f.write(js)


if __name__ == "__main__":
    main()
