"""The diagonal-copy eligibility test (mek_csr_diagonals, host only): the
ascending column - row offsets of a matrix whose entries lie on at most 16
diagonals, whose rows list their columns in strictly ascending order and
whose values are not NaN; nothing otherwise.  The SpMV sums a row diagonal
by diagonal, so the ascending-column rule is what makes that the CSR order."""
import numpy as np

from conftest import gen, mek


def _diags(rp, ci, va):
    A = mek.CSR.from_arrays(rp, ci, va)
    d = A.diagonals()
    A.close()
    return d


def test_stencils():
    N = 9
    assert _diags(*gen.laplacian_2d(N)) == [-N, -1, 0, 1, N]
    assert _diags(*gen.convdiff_3d(5)) == [-25, -5, -1, 0, 1, 5, 25]


def test_tridiagonal_with_absent_and_zero_entries():
    # rows: [0: cols 0,1], [1: col 2 only], [2: cols 1,2 (explicit zero)]
    rp = np.array([0, 2, 3, 5], np.int32)
    ci = np.array([0, 1, 2, 1, 2], np.int32)
    va = np.array([1.0, 2.0, 3.0, 0.0, -0.0])
    assert _diags(rp, ci, va) == [-1, 0, 1]


def test_refusals():
    rp, ci, va = gen.laplacian_2d(6)
    bad = va.copy()
    bad[7] = np.nan
    assert _diags(rp, ci, bad) == []
    # unsorted row: swap the first two entries of row 10
    ci2 = ci.copy()
    a = rp[10]
    ci2[a], ci2[a + 1] = ci2[a + 1], ci2[a]
    assert _diags(rp, ci2, va) == []
    # 17 diagonals (a full band of half-width 8)
    n = 40
    rows = [r for r in range(n) for c in range(max(0, r - 8), min(n, r + 9))]
    cols = [c for r in range(n) for c in range(max(0, r - 8), min(n, r + 9))]
    rp3 = np.searchsorted(np.array(rows), np.arange(n + 1)).astype(np.int32)
    assert _diags(rp3, np.array(cols, np.int32), np.ones(len(cols))) == []
    # 16 diagonals are taken
    rows = [r for r in range(n) for c in range(max(0, r - 8), min(n, r + 8))]
    cols = [c for r in range(n) for c in range(max(0, r - 8), min(n, r + 8))]
    rp4 = np.searchsorted(np.array(rows), np.arange(n + 1)).astype(np.int32)
    assert _diags(rp4, np.array(cols, np.int32), np.ones(len(cols))) == list(range(-8, 8))


def test_random_rows_have_no_diagonal_form():
    assert _diags(*gen.random_rows(500, 5, 30, seed=1)) == []


def _values(rp, ci, va, f32=False):
    A = mek.CSR.from_arrays(rp, ci, va)
    v = A.diagonal_values(f32)
    A.close()
    return v


def test_constant_diagonals_of_stencils():
    """The constant-diagonal form's values: one per diagonal, in offset order,
    read from the stencil's coefficients."""
    assert _values(*gen.laplacian_2d(9)) == [1.0, 1.0, -4.0, 1.0, 1.0]
    rp, ci, va = gen.convdiff_3d(5)
    coeffs = _values(rp, ci, va)
    assert len(coeffs) == 7
    offs = _diags(rp, ci, va)
    row = 2 * 25 + 2 * 5 + 2  # an interior row holds every diagonal once
    got = dict(zip((ci[rp[row]:rp[row + 1]] - row).tolist(), va[rp[row]:rp[row + 1]].tolist()))
    assert [got[o] for o in offs] == coeffs


def test_constant_diagonals_compare_bits():
    """A diagonal with two bit patterns has no constant form: a different
    value, or +0.0 beside -0.0 (the products would differ in sign); with
    f32=True values are compared after rounding to FP32."""
    rp, ci, va = gen.laplacian_2d(6)
    va2 = va.copy()
    va2[rp[10] + 1] = 1.5
    assert _values(rp, ci, va2) == []
    zeros = np.where(va == 1.0, 0.0, va)
    assert _values(rp, ci, zeros) == [0.0, 0.0, -4.0, 0.0, 0.0]
    mixed = zeros.copy()
    mixed[rp[10]] = -0.0
    assert _values(rp, ci, mixed) == []
    near = va.copy()
    near[rp[14]] = 1.0 + 2.0 ** -40  # first entry of an interior row: diagonal -N
    assert _values(rp, ci, near) == []
    assert _values(rp, ci, near, f32=True) == [1.0, 1.0, -4.0, 1.0, 1.0]


def test_no_constant_form_without_a_diagonal_form():
    assert _values(*gen.random_rows(300, 5, 30, seed=2)) == []
