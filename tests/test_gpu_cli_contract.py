"""The matrix_exp process contract (main.cu:2083-2324): the whole stdout,
line by line, against a golden transcript built from the oracle's step log
with iostream's default double formatting (%g, 6 significant digits), and
OUTPUT_Y_FILE byte for byte.

Golden lines, in the reference's order: the banner, the GPU count, the matrix
lines (built-in diagonal matrix, or the Binary CSR header, format and size
lines of a file), the t clamp warning, the partition line, `Arnoldi params`,
the allocation and SpMV-buffer lines, per attempt m_used H-column lines and a
[STEP] line (main.cu:1471-1474, :1951), the three [FLOPs] lines (main.cu:
1560-1565), METRIC, `Saved y to`, the ten sample values and the summary
(main.cu:2299-2313).  Integers, t and dt of each step, the FLOP model and the
sub-step counts must equal the oracle's exactly; err, y and the residual are
printed from the GPU's own values, which match the oracle's to 1e-8, and equal
those of the same solve made through the C ABI exactly.  Free fields: the GPU
count, the SpMV buffer size (this build's row-block table, not cuSPARSE's
buffer) and the METRIC time.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from closed_forms import rel_err
from conftest import gen, mek

pytestmark = pytest.mark.gpu

NUM = r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|inf|nan)"
ANY = object()


class Near:
    def __init__(self, x, rtol=1e-5):
        self.x, self.rtol = float(x), rtol


def g(x):
    """std::cout << double with the default format: %g, precision 6."""
    return "%g" % x


def _match(actual, template, values):
    parts = template.split("{}")
    pat = "^" + "".join(re.escape(p) + (f"({NUM})" if i < len(parts) - 1 else "") for i, p in enumerate(parts)) + "$"
    m = re.match(pat, actual)
    assert m, (actual, template)
    for got, want in zip(m.groups(), values):
        if want is ANY:
            continue
        if isinstance(want, Near):
            assert abs(float(got) - want.x) <= want.rtol * abs(want.x) + 1e-300, (actual, want.x)
        else:
            assert got == want, (actual, template, want)


def golden(head, n, nnz, m, t_eff, R, tol, oracle_log, oracle_stats, oracle_y, out_path):
    """[(template, values)] of the whole stdout after the matrix lines."""
    L = list(head)
    L.append(("Partitioned matrix across 1 GPUs", []))
    L.append(("Arnoldi params: m={} t={} tol={} max_restarts={}", [str(m), g(t_eff), g(tol), str(R)]))
    L.append(("GPU 0 allocating: local_rows={}, local_nnz={}", [str(n), str(nnz)]))
    L.append(("  GPU 0 SpMV buffer size: {} bytes", [ANY]))
    for s in oracle_log:
        for j in range(s["m_used"]):
            L.append((f"    H_m column {j} stored. Current H_m size: {m + 1}x{m}", []))
        L.append(("[STEP] t={} dt={} err={} m_used={} accepts={} rejects={}",
                  [g(s["t"]), g(s["dt"]), Near(s["err"]) if s["err"] > 1e-300 else ANY, str(s["m_used"]),
                   str(s["accepts"]), str(s["rejects"])]))
    S = float(oracle_stats["restarts_done"])
    N, Z, mv = float(n), float(nnz), float(m)
    seg = 2.0 * Z * mv + 4.0 * N * mv * (mv + 1.0) + 5.0 * N * mv + 2.0 * N
    tot = 3.0 * N * S + S * seg
    cseg = 2.0 * 1.0 * mv * mv * mv + 2.0 * mv * mv
    L.append(("[FLOPs] N={} nnz={} m={} steps={}", [g(N), g(Z), g(mv), g(S)]))
    L.append(("[FLOPs] GPU segment: {} | GPU total: {}", [g(seg), g(tot)]))
    L.append(("[FLOPs] CPU segment: {} | CPU total: {} (C_exp={})", [g(cseg), g(S * cseg), "1"]))
    L.append(("METRIC total_compute_expmv_ms={}", [ANY]))
    L.append((f"Saved y to {out_path}", []))
    L.append(("Computation completed. Sample results:", []))
    for i in range(min(10, n)):
        L.append((f"y[{i}] = {{}}", [Near(oracle_y[i])]))
    L.append(("Arnoldi summary: m={}, max_restarts={}, restarts_done={}, converged={}, residual={}".replace(
        "converged={}", "converged=" + ("yes" if oracle_stats["converged"] else "no")),
        [str(m), str(R), str(oracle_stats["restarts_done"]), Near(oracle_stats["last_residual"])
         if oracle_stats["last_residual"] > 1e-300 else ANY]))
    return L


def _run_cli(args, out):
    env = dict(os.environ, OUTPUT_Y_FILE=str(out))
    p = subprocess.run([mek.CLI_PATH, *args], env=env, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stderr == "", p.stderr
    return p.stdout.splitlines()


def _check(lines, gold):
    assert len(lines) == len(gold), (len(lines), len(gold), lines[-5:])
    for a, (tpl, vals) in zip(lines, gold):
        _match(a, tpl, vals)


def _api_solve(rp, ci, va, m, t, tol, R):
    """The same solve through the C ABI (reference controller, second pass
    always, FP64, default launch shapes): y and the [STEP] records."""
    Rn = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=m, t=t, tol=tol, max_restarts=R)
    y = Rn.expmv(np.full(rp.size - 1, 0.1))
    steps = Rn.steps()
    st = Rn.stats()
    Rn.close()
    return y, steps, st


def _exact_against_api(lines, y_api, steps_api, st_api, out):
    # every printed number of the GPU run equals the API solve's, formatted
    step_lines = [ln for ln in lines if ln.startswith("[STEP]")]
    assert step_lines == [f"[STEP] t={g(s['t'])} dt={g(s['dt'])} err={g(s['err'])} m_used={s['m_used']} "
                          f"accepts={s['accepts']} rejects={s['rejects']}" for s in steps_api]
    ylines = [ln for ln in lines if ln.startswith("y[")]
    assert ylines == [f"y[{i}] = {g(y_api[i])}" for i in range(len(ylines))]
    assert lines[-1].endswith(f"residual={g(st_api['last_residual'])}")
    # OUTPUT_Y_FILE: std::scientific, precision 16, one value and "\n" per row
    text = out.read_bytes()
    assert text == "".join("%.16e\n" % v for v in y_api).encode()


def test_cli_builtin_matrix_transcript(gpu, tmp_path, oracle):
    """Built-in diagonal matrix a_ii = 1 + i/n (main.cu:2198-2216), n = 200,
    --m 12, --t 8 (clamped to 10 / max|a_ii|), tol 1e-6, 10 restarts."""
    n, m, R, tol = 200, 12, 10, 1e-6
    out = tmp_path / "y.txt"
    lines = _run_cli([str(n), "--m", str(m), "--t", "8"], out)
    d = 1.0 + np.arange(n) / n
    max_abs = float(d.max())
    t_eff = 10.0 / max_abs
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    yo, st, log = oracle.expmv(rp, ci, d, np.full(n, 0.1), t_eff, tol, m, R)
    head = [("Starting matrix_exp application.", []), ("Found {} GPU(s)", [ANY]),
            (f"Matrix size (generated): {n}", []), (f"Created test matrix: {n}x{n} with {n} non-zeros", []),
            ("[WARN] Requested t={} is large for max|A|={}. Clamping t to {} to avoid overflow.",
             [g(8.0), g(max_abs), g(t_eff)])]
    _check(lines, golden(head, n, n, m, t_eff, R, tol, log, st, yo, out))
    y_api, steps_api, st_api = _api_solve(rp, ci, d, m, t_eff, tol, R)
    assert rel_err(y_api, yo) <= 1e-8
    _exact_against_api(lines, y_api, steps_api, st_api, out)


def test_cli_bincsr_transcript(gpu, tmp_path, oracle):
    """A Binary CSR v1 file (u32 indices, f64 values, sorted and duplicate-free
    flags): the reference's random Gershgorin-negative matrix at n = 300 with
    t = 0.2 (no clamp), m and max_restarts as positional arguments."""
    rp, ci, va = gen.random_negative(300, 5e-3, seed=42)
    n, nnz, m, R, tol, t = 300, int(rp[-1]), 15, 10, 1e-6, 0.2
    assert t <= 10.0 / np.abs(va).max()
    path = tmp_path / "A.bincsr"
    gen.write_bincsr(path, rp, ci, va)
    out = tmp_path / "y.txt"
    lines = _run_cli([str(path), str(m), str(R), "--t", g(t)], out)
    yo, st, log = oracle.expmv(rp, ci, va, np.full(n, 0.1), t, tol, m, R)
    head = [("Starting matrix_exp application.", []), ("Found {} GPU(s)", [ANY]),
            (f"Loading matrix from file: {path}", []),
            ("Binary CSR header: version=1 flags=0x3 index_dtype=u32 value_dtype=f64 symmetric_upper=no", []),
            ("Detected Binary CSR v1 input", []), (f"Loaded matrix: {n}x{n} with {nnz} non-zeros", [])]
    _check(lines, golden(head, n, nnz, m, t, R, tol, log, st, yo, out))
    y_api, steps_api, st_api = _api_solve(rp, ci, va, m, t, tol, R)
    assert rel_err(y_api, yo) <= 1e-8
    _exact_against_api(lines, y_api, steps_api, st_api, out)
