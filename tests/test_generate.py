"""Synthetic workload generators and the Binary CSR v1 writer (CPU)."""
import numpy as np
import scipy.sparse as sp

from closed_forms import kron3_expmv, laplacian_2d_expmv, rel_err
from conftest import gen


def test_laplacian_is_kronecker_sum():
    N = 17
    rp, ci, va = gen.laplacian_2d(N)
    T = sp.diags([1.0, -2.0, 1.0], [-1, 0, 1], shape=(N, N))
    A = sp.csr_matrix((va, ci, rp), shape=(N * N, N * N))
    assert abs(A - sp.kronsum(T, T)).max() == 0
    assert A.has_sorted_indices and rp[-1] == 5 * N * N - 4 * N


def test_laplacian_4096_sizes():
    # the full-size workload: n = 16,777,216, nnz = 83,869,696
    N = 4096
    assert 5 * N * N - 4 * N == 83869696


def test_convdiff_is_kronecker_sum():
    N = 6
    rp, ci, va = gen.convdiff_3d(N)
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    assert (lx, ly, lz, ux) == (51.0, 26.0, 13.5, 1.0) and dx + dy + dz == -93.5
    Tx, Ty, Tz = (sp.csr_matrix(gen.tridiag_1d(N, *c)) for c in ((lx, dx, ux), (ly, dy, uy), (lz, dz, uz)))
    I = sp.identity(N)
    K = sp.kron(sp.kron(Tz, I), I) + sp.kron(sp.kron(I, Ty), I) + sp.kron(sp.kron(I, I), Tx)
    A = sp.csr_matrix((va, ci, rp), shape=(N ** 3, N ** 3))
    assert abs(A - K).max() == 0
    assert 7 * 256 ** 3 - 6 * 256 ** 2 == 117047296


def test_closed_forms_against_dense_expm():
    import scipy.linalg
    N = 7
    rp, ci, va = gen.laplacian_2d(N)
    A = sp.csr_matrix((va, ci, rp)).toarray()
    v = gen.normal_vector(N * N, seed=1)
    assert rel_err(laplacian_2d_expmv(N, 1.3, v), scipy.linalg.expm(1.3 * A) @ v) < 1e-13
    N = 5
    rp, ci, va = gen.convdiff_3d(N)
    A = sp.csr_matrix((va, ci, rp)).toarray()
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    v = gen.normal_vector(N ** 3, seed=2)
    ref = kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                      gen.tridiag_1d(N, lx, dx, ux), 0.05, v)
    assert rel_err(ref, scipy.linalg.expm(0.05 * A) @ v) < 1e-12


def test_random_generators():
    rp, ci, va = gen.random_negative(500, 1e-2, seed=3)
    A = sp.csr_matrix((va, ci, rp))
    d = A.diagonal()
    off = np.abs(A).sum(1).A1 - np.abs(d)
    np.testing.assert_allclose(d, -1.0 - off, rtol=1e-15)
    assert A.has_sorted_indices
    rp, ci, va = gen.random_rows(800, 5, 30, seed=4)
    A = sp.csr_matrix((va, ci, rp))
    per_row = np.diff(rp) - 1  # minus the diagonal
    assert per_row.min() >= 1 and per_row.max() <= 30
    A.sum_duplicates()
    assert A.nnz == rp[-1]  # columns distinct within rows


def test_bincsr_writer_reader_roundtrip(tmp_path):
    rp, ci, va = gen.random_rows(100, 2, 5, seed=1)
    p = tmp_path / "r.bincsr"
    gen.write_bincsr(p, rp, ci, va, index_dtype="u64")
    hdr, ip, ix, da = gen.read_bincsr(p)
    assert hdr["magic"] == gen.MAGIC and hdr["version"] == 1 and hdr["index_dtype"] == 2
    assert hdr["flags"] == 3 and hdr["nnz"] == rp[-1]
    np.testing.assert_array_equal(ip, rp)
    np.testing.assert_array_equal(ix, ci)
    np.testing.assert_array_equal(da, va)
    assert p.stat().st_size == 64 + 8 * (101 + rp[-1]) + 8 * rp[-1]
