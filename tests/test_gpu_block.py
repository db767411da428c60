"""Block of 32 right-hand sides on one shared A (C ABI): the SpMM against the
oracle's CSR loop column by column (bit for bit), the FP64 MFMA GEMM against
numpy, and the block Arnoldi solve against closed forms: the 2-D Laplacian by
its DST-I diagonalisation and the convection-diffusion operator by Kronecker
products of three dense exponentials."""
import numpy as np
import pytest

from closed_forms import kron3_expmv, laplacian_2d_expmv, rel_err
from conftest import gen, mek

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["laplacian", "random"])
def test_spmm_bitexact(gpu, oracle, kind):
    if kind == "laplacian":
        rp, ci, va = gen.laplacian_2d(70)  # 4900 rows: padding rows in the last chunk
    else:
        rp, ci, va = gen.random_rows(3000, 1, 40, seed=4)
    n = rp.size - 1
    X = np.random.default_rng(1).standard_normal((n, 32))
    Y = mek.spmm(mek.CSR.from_arrays(rp, ci, va), X)
    for c in range(32):
        np.testing.assert_array_equal(Y[:, c], oracle.spmv(rp, ci, va, X[:, c]))


@pytest.mark.parametrize("N", [64, 192])
def test_gemm_dense(gpu, N):
    rng = np.random.default_rng(N)
    A = rng.standard_normal((N, N))
    B = rng.standard_normal((N, N))
    C = mek.gemm_dense(A, B)
    ref = A @ B
    assert np.max(np.abs(C - ref)) <= 1e-13 * np.max(np.abs(ref)) * N


def _columns(f, V):
    return np.stack([f(V[:, c]) for c in range(V.shape[1])], axis=1)


@pytest.mark.parametrize("substeps,gs_mode", [(0, 2), (3, 2), (0, 1)])
def test_block_laplacian_closed_form(gpu, substeps, gs_mode):
    N = 64  # n = 4096
    rp, ci, va = gen.laplacian_2d(N)
    V = np.random.default_rng(2).standard_normal((N * N, 32))
    t = 2.0
    B = mek.Block(mek.CSR.from_arrays(rp, ci, va), m=16, t=t, tol=1e-10, substeps=substeps, gs_mode=gs_mode)
    Y = B.expmv(V)
    st = B.stats()
    print(f"[block] laplacian64 substeps={substeps} gs={gs_mode} {st['accepted']} steps, {st['attempts']} "
          f"attempts, {st['squarings']} squarings, {st['second_passes']} second passes", flush=True)
    if gs_mode == 1:
        assert st["second_passes"] == st["block_steps"]
    ref = _columns(lambda v: laplacian_2d_expmv(N, t, v), V)
    assert st["converged"] == 1
    assert rel_err(Y, ref) <= 1e-8, rel_err(Y, ref)
    for c in range(32):
        assert rel_err(Y[:, c], ref[:, c]) <= 1e-8, c
    B.close()


def test_block_convdiff_kronecker(gpu):
    N = 16  # n = 4096, nonsymmetric
    rp, ci, va = gen.convdiff_3d(N)
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    V = np.random.default_rng(3).standard_normal((N ** 3, 32))
    t = 0.05
    B = mek.Block(mek.CSR.from_arrays(rp, ci, va), m=20, t=t, tol=1e-10)
    Y = B.expmv(V)
    ref = _columns(lambda v: kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                                         gen.tridiag_1d(N, lx, dx, ux), t, v), V)
    st = B.stats()
    print(f"[block] convdiff16 {st['accepted']} steps, {st['attempts']} attempts", flush=True)
    assert st["converged"] == 1
    assert rel_err(Y, ref) <= 1e-8, rel_err(Y, ref)
    B.close()


def test_block_equals_single_vector_solves(gpu, oracle):
    """Each column of the block solve against the oracle's single-vector
    solve of the same column (tight tolerances on both sides)."""
    rp, ci, va = gen.random_rows(2500, 3, 12, seed=11)
    V = np.random.default_rng(5).standard_normal((2500, 32))
    t = 0.3
    B = mek.Block(mek.CSR.from_arrays(rp, ci, va), m=12, t=t, tol=1e-11)
    Y = B.expmv(V)
    for c in range(0, 32, 5):
        y, st, _log = oracle.expmv(rp, ci, va, V[:, c], t, 1e-12, 30, 50)
        assert rel_err(Y[:, c], y) <= 1e-8, (c, rel_err(Y[:, c], y))
    B.close()


def test_block_device_entry_and_timing(gpu):
    N = 64
    rp, ci, va = gen.laplacian_2d(N)
    V = np.random.default_rng(9).standard_normal((N * N, 32))
    B = mek.Block(mek.CSR.from_arrays(rp, ci, va), m=10, t=1.0, tol=1e-9, timing=True)
    d_v = mek.DeviceBuffer.from_array(V.ravel())
    d_y = mek.DeviceBuffer(8 * V.size)
    B.expmv_device(d_v, d_y)
    Y = d_y.to_array().reshape(V.shape)
    np.testing.assert_allclose(Y, B.expmv(V), rtol=0, atol=1e-13 * np.abs(Y).max())
    st = B.stats()
    for k in ("spmm", "dots", "update", "expm", "lift"):
        assert st["kernels"][k]["launches"] > 0 and st["kernels"][k]["ms"] > 0, k
    B.close()


def test_block_rejects_bad_sizes(gpu):
    rp, ci, va = gen.laplacian_2d(10)  # 100 rows: a 32 (m + 1) basis does not fit
    with pytest.raises(mek.MekError):
        mek.Block(mek.CSR.from_arrays(rp, ci, va), m=4, t=1.0)


def test_block_autotune_same_result(gpu):
    N = 64
    rp, ci, va = gen.laplacian_2d(N)
    V = np.random.default_rng(12).standard_normal((N * N, 32))
    A = mek.CSR.from_arrays(rp, ci, va)
    ys = {}
    for at in (False, True):
        B = mek.Block(A, m=10, t=1.0, tol=1e-9, autotune=at)
        ys[at] = B.expmv(V)
        st = B.stats()
        if at:
            assert st["spmm_grid"] in (1024, 2048, 4096, 8192) and st["update_grid"] in (1024, 2048, 4096, 8192)
        B.close()
    np.testing.assert_allclose(ys[True], ys[False], rtol=0, atol=1e-12 * np.abs(ys[False]).max())
