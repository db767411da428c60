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
