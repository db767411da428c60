"""Kernel-level parity on the GPU, through the C ABI of libmek.so.

SpMV: bit-exact against the oracle's sequential CSR loop for rows that fit a
short-row block (same products, same summation order); long rows (> 4096
nonzeros, reduced by a whole workgroup) within 1e-14 relative.
expm: against the oracle's Pade restatement (Higham 2005).
Gram-Schmidt: the Hessenberg matrix of one Arnoldi decomposition (CGS2 on the
device) against the oracle's MGS2.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import gen, mek

pytestmark = pytest.mark.gpu


def _long_row_matrix(n=5000, seed=3):
    """random_rows plus one dense row (row 7: n nonzeros) and two empty rows."""
    rp, ci, va = gen.random_rows(n, 5, 30, seed=seed)
    rows = [list(zip(ci[rp[r]:rp[r + 1]], va[rp[r]:rp[r + 1]])) for r in range(n)]
    rng = np.random.default_rng(seed)
    rows[7] = [(c, float(x)) for c, x in zip(range(n), rng.uniform(-1, 1, n))]
    rows[11] = []
    rows[n - 1] = []
    new_rp = np.zeros(n + 1, np.int32)
    new_rp[1:] = np.cumsum([len(r) for r in rows])
    new_ci = np.array([c for r in rows for c, _ in r], np.int32)
    new_va = np.array([x for r in rows for _, x in r], np.float64)
    return new_rp, new_ci, new_va


@pytest.mark.parametrize("which", ["laplacian", "random", "rows"])
def test_spmv_bitexact(gpu, oracle, which):
    print(f"[spmv] {which}", flush=True)
    if which == "laplacian":
        rp, ci, va = gen.laplacian_2d(64)
    elif which == "random":
        rp, ci, va = gen.random_negative(2000, 5e-3, seed=42)
    else:
        rp, ci, va = gen.random_rows(5000, 5, 30, seed=1)
    x = np.random.default_rng(7).standard_normal(rp.size - 1)
    plan = mek.SpmvPlan(rp, ci, va)
    y = plan(x)
    ref = oracle.spmv(rp, ci, va, x)
    np.testing.assert_array_equal(y, ref)


def test_spmv_long_and_empty_rows(gpu, oracle):
    rp, ci, va = _long_row_matrix()
    x = np.random.default_rng(9).standard_normal(rp.size - 1)
    y = mek.SpmvPlan(rp, ci, va)(x)
    ref = oracle.spmv(rp, ci, va, x)
    short = np.ones(rp.size - 1, bool)
    short[7] = False
    np.testing.assert_array_equal(y[short], ref[short])
    assert abs(y[7] - ref[7]) <= 1e-13 * np.abs(va[rp[7]:rp[8]] * x).sum()
    assert y[11] == 0.0 and y[-1] == 0.0


def test_spmv_f32_values_exact_stencil(gpu, oracle):
    # stencil coefficients are exact in FP32, so FP32 storage changes nothing
    rp, ci, va = gen.convdiff_3d(20)
    x = np.random.default_rng(3).standard_normal(rp.size - 1)
    y = mek.SpmvPlan(rp, ci, va, values_f32=True)(x)
    np.testing.assert_array_equal(y, oracle.spmv(rp, ci, va, x))


def _ragged_matrix(n=1000, seed=5):
    """Rows of 0..9 nonzeros (a run of empty rows included), unsorted
    columns within each row, n not a multiple of 64."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 10, n)
    lens[100:170] = 0
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([rng.choice(n, int(k), replace=False) for k in lens]).astype(np.int32)
    va = rng.uniform(-1, 1, int(rp[-1]))
    return rp, ci, va


@pytest.mark.parametrize("which", ["laplacian", "convdiff", "convdiff_f32", "ragged", "random", "tiny"])
def test_spmv_sell_bitexact(gpu, oracle, which):
    """The sliced-ELL copy gives the oracle's CSR-loop bits (and so the CSR
    row blocks' bits), including padded rows, empty rows, unsorted columns,
    FP32 values and a last slice of fewer than 64 rows."""
    print(f"[spmv-sell] {which}", flush=True)
    f32 = which == "convdiff_f32"
    if which == "laplacian":
        rp, ci, va = gen.laplacian_2d(77)
    elif which.startswith("convdiff"):
        rp, ci, va = gen.convdiff_3d(19)
    elif which == "ragged":
        rp, ci, va = _ragged_matrix()
    elif which == "random":
        rp, ci, va = gen.random_rows(3001, 5, 30, seed=2)
    else:
        rp, ci, va = np.array([0, 1, 1, 3], np.int32), np.array([2, 0, 2], np.int32), np.array([2.0, -1.0, 0.5])
    x = np.random.default_rng(8).standard_normal(rp.size - 1)
    plan = mek.SpmvPlan(rp, ci, va, values_f32=f32)
    y_csr = plan(x)
    entries, width = plan.use_sell()
    y = plan(x)
    n = rp.size - 1
    lens = np.diff(rp)
    slices = (n + 63) // 64
    assert width == lens.max()
    assert entries == sum(64 * int(lens[64 * s:64 * (s + 1)].max()) for s in range(slices))
    np.testing.assert_array_equal(y, oracle.spmv(rp, ci, va, x))
    np.testing.assert_array_equal(y, y_csr)
    plan.close()


def _banded_sparse(n, offsets, keep, seed):
    """Entries on the given diagonals, each kept with probability `keep`
    (absent entries inside the band), explicit zeros and negative zeros among
    the values, columns ascending within each row."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for o in offsets:
        r = np.arange(max(0, -o), min(n, n - o))
        r = r[rng.random(r.size) < keep]
        rows.append(r)
        cols.append(r + o)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    va = rng.uniform(-1, 1, r.size)
    va[rng.random(r.size) < 0.05] = 0.0
    va[rng.random(r.size) < 0.05] = -0.0
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32), va


@pytest.mark.parametrize("which", ["laplacian", "convdiff", "convdiff_f32", "band_holes", "nine", "sixteen",
                                   "sixteen_f32", "n65", "one_row", "zero_x", "tridiag", "tridiag_holes"])
def test_spmv_dia_bitexact(gpu, oracle, which):
    """The diagonal copy gives the oracle's CSR-loop bits: stencils with their
    absent boundary entries, FP32 values, bands with holes, explicit (and
    negative) zeros, 9 and 16 diagonals (the two kernel shapes), row counts
    that are not a multiple of 64 or 128, one row, and an x with zeros (where
    an added 0.0 product would turn a -0.0 row sum into +0.0)."""
    print(f"[spmv-dia] {which}", flush=True)
    f32 = which.endswith("_f32")
    n_diags = None
    if which == "laplacian":
        rp, ci, va = gen.laplacian_2d(77)
        n_diags = 5
    elif which.startswith("convdiff"):
        rp, ci, va = gen.convdiff_3d(19)
        n_diags = 7
    elif which == "band_holes":
        rp, ci, va = _banded_sparse(1000, [-300, -7, -3, 0, 2, 5, 11, 640], 0.6, seed=3)
    elif which == "nine":
        rp, ci, va = _banded_sparse(2000, [-40, -9, -4, -1, 0, 1, 4, 9, 40], 0.9, seed=4)
        n_diags = 9
    elif which.startswith("sixteen"):
        rp, ci, va = _banded_sparse(3001, list(range(-8, 8)), 1.0, seed=5)
        if f32:
            va = va.astype(np.float32).astype(np.float64)
        n_diags = 16
    elif which == "n65":
        rp, ci, va = _banded_sparse(65, [-1, 0, 1, 64], 0.8, seed=6)
    elif which == "one_row":
        rp, ci, va = np.array([0, 1], np.int32), np.array([0], np.int32), np.array([-2.5])
        n_diags = 1
    elif which.startswith("tridiag"):  # the exact 3-diagonal kernels
        rp, ci, va = _banded_sparse(777, [-1, 0, 1], 0.7 if which.endswith("holes") else 1.0, seed=21)
        n_diags = 3
    else:
        rp, ci, va = _banded_sparse(500, [-2, -1, 0, 1, 2], 1.0, seed=7)
        va = -np.abs(va)
    n = rp.size - 1
    x = np.random.default_rng(9).standard_normal(n)
    if which == "zero_x":
        x[::3] = 0.0
    plan = mek.SpmvPlan(rp, ci, va, values_f32=f32)
    y_csr = plan(x)
    nd = plan.use_dia()
    if n_diags is not None:
        assert nd == n_diags
    assert nd == np.unique(ci.astype(np.int64) - np.repeat(np.arange(n), np.diff(rp))).size
    y = plan(x)
    yo = oracle.spmv(rp, ci, va.astype(np.float32) if f32 else va, x)
    np.testing.assert_array_equal(y.view(np.int64), yo.view(np.int64))
    np.testing.assert_array_equal(y.view(np.int64), y_csr.view(np.int64))
    # an output that is 8- but not 16-byte aligned takes the one-row-per-lane form
    d_x = mek.DeviceBuffer.from_array(x)
    d_y = mek.DeviceBuffer(8 * (n + 1))
    mek._check(mek.lib().mek_spmv_execute(plan.h, d_x.ptr, ctypes.c_void_p(d_y.ptr.value + 8), None))
    mek.synchronize()
    y1 = d_y.to_array()[1:]
    np.testing.assert_array_equal(y1.view(np.int64), yo.view(np.int64))
    plan.close()


@pytest.mark.parametrize("which", ["unsorted", "seventeen", "nan"])
def test_spmv_dia_refused(gpu, which):
    """No diagonal copy for rows with unsorted columns (the diagonal order would
    not be the CSR order), more than 16 diagonals, or NaN values (NaN marks
    absent entries)."""
    if which == "unsorted":
        rp = np.array([0, 2, 3], np.int32)
        ci = np.array([1, 0, 1], np.int32)
        va = np.array([1.0, 2.0, 3.0])
    elif which == "seventeen":
        rp, ci, va = _banded_sparse(200, list(range(-8, 9)), 1.0, seed=8)
    else:
        rp, ci, va = gen.laplacian_2d(8)
        va = va.copy()
        va[5] = np.nan
    plan = mek.SpmvPlan(rp, ci, va)
    with pytest.raises(mek.MekError):
        plan.use_dia()
    plan.close()


def _constant_on_diagonals(rp, ci, va, seed):
    """The same pattern with one value per diagonal (drawn U[-1,1], one
    diagonal +0.0 and one -0.0 when there are at least three)."""
    n = rp.size - 1
    offs = ci.astype(np.int64) - np.repeat(np.arange(n), np.diff(rp))
    uniq = np.unique(offs)
    vals = np.random.default_rng(seed).uniform(-1, 1, uniq.size)
    if uniq.size >= 3:
        vals[0], vals[-1] = 0.0, -0.0
    return vals[np.searchsorted(uniq, offs)]


@pytest.mark.parametrize("which", ["laplacian", "convdiff", "convdiff_f32", "band_holes", "sixteen",
                                   "sixteen_f32", "n65", "one_row", "zero_x", "tridiag_holes"])
def test_spmv_dia_const_bitexact(gpu, oracle, which):
    """The constant-diagonal form (one value per diagonal, a presence mask
    per row) gives the oracle's CSR-loop bits: stencils with their absent
    boundary entries, FP32-stored values (the constants rounded as the copy
    would round them), bands with holes and with diagonals of +0.0 and -0.0,
    16 diagonals (16-bit masks), row counts that are not a multiple of 64 or
    128, one row, an x with zeros, and the one-row-per-lane form for an
    output that is not 16-byte aligned."""
    print(f"[spmv-dia-const] {which}", flush=True)
    f32 = which.endswith("_f32")
    if which == "laplacian":
        rp, ci, va = gen.laplacian_2d(77)
    elif which.startswith("convdiff"):
        rp, ci, va = gen.convdiff_3d(19)
    elif which == "band_holes":
        rp, ci, _ = _banded_sparse(1000, [-300, -7, -3, 0, 2, 5, 11, 640], 0.6, seed=3)
        va = _constant_on_diagonals(rp, ci, None, seed=13)
    elif which.startswith("sixteen"):
        rp, ci, _ = _banded_sparse(3001, list(range(-8, 8)), 0.9, seed=5)
        va = _constant_on_diagonals(rp, ci, None, seed=15)
    elif which == "n65":
        rp, ci, _ = _banded_sparse(65, [-1, 0, 1, 64], 0.8, seed=6)
        va = _constant_on_diagonals(rp, ci, None, seed=16)
    elif which == "one_row":
        rp, ci, va = np.array([0, 1], np.int32), np.array([0], np.int32), np.array([-2.5])
    elif which == "tridiag_holes":  # the exact 3-diagonal kernels
        rp, ci, _ = _banded_sparse(777, [-1, 0, 1], 0.7, seed=21)
        va = _constant_on_diagonals(rp, ci, None, seed=22)
    else:
        rp, ci, _ = _banded_sparse(500, [-2, -1, 0, 1, 2], 1.0, seed=7)
        va = -np.abs(_constant_on_diagonals(rp, ci, None, seed=17)) - 0.25
    n = rp.size - 1
    x = np.random.default_rng(9).standard_normal(n)
    if which == "zero_x":
        x[::3] = 0.0
    plan = mek.SpmvPlan(rp, ci, va, values_f32=f32)
    y_csr = plan(x)
    nd = plan.use_dia_const()
    assert nd == np.unique(ci.astype(np.int64) - np.repeat(np.arange(n), np.diff(rp))).size
    y = plan(x)
    yo = oracle.spmv(rp, ci, va.astype(np.float32) if f32 else va, x)
    np.testing.assert_array_equal(y.view(np.int64), yo.view(np.int64))
    np.testing.assert_array_equal(y.view(np.int64), y_csr.view(np.int64))
    d_x = mek.DeviceBuffer.from_array(x)
    d_y = mek.DeviceBuffer(8 * (n + 1))
    mek._check(mek.lib().mek_spmv_execute(plan.h, d_x.ptr, ctypes.c_void_p(d_y.ptr.value + 8), None))
    mek.synchronize()
    y1 = d_y.to_array()[1:]
    np.testing.assert_array_equal(y1.view(np.int64), yo.view(np.int64))
    plan.close()


@pytest.mark.parametrize("which", ["two_values", "signed_zeros", "no_diagonal_form"])
def test_spmv_dia_const_refused(gpu, which):
    """No constant-diagonal form when a diagonal holds two values (also +0.0
    beside -0.0) or the matrix has no diagonal form at all."""
    rp, ci, va = gen.laplacian_2d(8)
    va = va.copy()
    if which == "two_values":
        va[rp[20] + 1] = 0.75
    elif which == "signed_zeros":
        va = np.where(va == 1.0, 0.0, va)
        va[rp[20]] = -0.0
    else:
        rp, ci, va = gen.random_rows(400, 5, 30, seed=2)
    plan = mek.SpmvPlan(rp, ci, va)
    with pytest.raises(mek.MekError):
        plan.use_dia_const()
    plan.close()


@pytest.mark.parametrize("k", [1, 2, 5, 17, 40, 58, 62, 63, 80])
def test_expm_dense(gpu, oracle, k):
    rng = np.random.default_rng(100 + k)
    for scale in [1e-3, 0.2, 0.9, 2.0, 5.0, 30.0, 300.0]:
        print(f"[expm] k={k} scale={scale}", flush=True)
        H = np.triu(rng.standard_normal((k, k)), -1)
        H *= scale / max(np.abs(H).sum(0).max(), 1e-300)
        if scale > 5:
            H -= np.eye(k) * scale / 4  # keep exp(H) representable
        E = mek.expm_dense(H)
        R, _ = oracle.expm(H)
        err = np.linalg.norm(E - R) / np.linalg.norm(R)
        assert err < 1e-12 * max(1.0, scale), (k, scale, err)


def test_expm_closed_forms(gpu):
    # Jordan block exp(t[[a,1],[0,a]]) = e^{at}[[1,t],[0,1]]
    a, t = -0.7, 3.0
    E = mek.expm_dense(np.array([[a * t, t], [0.0, a * t]]))
    np.testing.assert_allclose(E, np.exp(a * t) * np.array([[1, t], [0, 1]]), rtol=1e-14, atol=1e-15)
    # rotation
    w = 2.5
    E = mek.expm_dense(np.array([[0.0, -w], [w, 0.0]]))
    np.testing.assert_allclose(E, [[np.cos(w), -np.sin(w)], [np.sin(w), np.cos(w)]], atol=1e-14)
    # nilpotent shift: finite series
    k = 6
    N = np.diag(np.ones(k - 1), 1) * 1.5
    ref = sum(np.linalg.matrix_power(N, p) / math.factorial(p) for p in range(k))
    np.testing.assert_allclose(mek.expm_dense(N), ref, rtol=1e-14, atol=1e-14)
    # diagonal
    d = np.linspace(-20, 3, 9)
    np.testing.assert_allclose(mek.expm_dense(np.diag(d)), np.diag(np.exp(d)), rtol=1e-13)


@pytest.mark.parametrize("which,m", [("random", 30), ("laplacian", 40), ("rows", 70)])
def test_arnoldi_hessenberg(gpu, oracle, which, m):
    print(f"[arnoldi] {which} m={m}", flush=True)
    if which == "random":
        rp, ci, va = gen.random_negative(2000, 5e-3, seed=42)
    elif which == "laplacian":
        rp, ci, va = gen.laplacian_2d(48)
    else:
        rp, ci, va = gen.random_rows(3000, 5, 30, seed=5)
    v = gen.normal_vector(rp.size - 1, seed=11, normalise=False)
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=m)
    H, k = R.arnoldi(v)
    Href, kref = oracle.arnoldi(rp, ci, va, v, m)
    assert k == kref == m
    err = np.linalg.norm(H - Href) / np.linalg.norm(Href)
    assert err < 1e-10, err
