"""s-step (block) Arnoldi on the GPU (sstep.hip): the basis grows several
vectors at a time from a shifted (Newton) recurrence, is orthogonalised by
block Gram-Schmidt with Cholesky QR, and the Hessenberg columns come from the
change of basis.  Checked against closed forms (2-D Laplacian by DST-I,
convection-diffusion by Kronecker products), against the oracle's
column-by-column MGS solve, against the column-by-column GPU path (y and the
Hessenberg matrix itself), and on the fall-back and breakdown paths."""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from closed_forms import kron3_expmv, laplacian_2d_eigvec, laplacian_2d_expmv, rel_err
from conftest import gen, mek

pytestmark = pytest.mark.gpu


def _solve(rp, ci, va, v, **kw):
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), **kw)
    y = R.expmv(v)
    st = R.stats()
    R.close()
    return y, st


@pytest.mark.parametrize("s", [8, 12, 16])
def test_sstep_laplacian_closed_form(gpu, s):
    N, t = 96, 10.0
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=1)
    kw = dict(m=40, t=t, tol=1e-10, max_restarts=0, step_control="expokit", gs_mode=2)
    y, st = _solve(rp, ci, va, v, sstep=s, **kw)
    y0, _ = _solve(rp, ci, va, v, **kw)
    e = rel_err(y, laplacian_2d_expmv(N, t, v))
    print(f"[sstep] laplacian s={s} err={e:.2e} blocks={st['sstep_blocks']} second={st['sstep_second']} "
          f"fallbacks={st['sstep_fallbacks']}", flush=True)
    assert e <= 1e-8
    assert rel_err(y, y0) <= 1e-10
    assert st["sstep_blocks"] > 0 and st["sstep_fallbacks"] == 0


@pytest.mark.parametrize("s,f32", [(8, False), (10, False), (10, True)])
def test_sstep_convdiff_closed_form(gpu, s, f32):
    N, t = 24, 0.5
    rp, ci, va = gen.convdiff_3d(N)
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    v = gen.normal_vector(N ** 3, seed=2)
    y, st = _solve(rp, ci, va, v, m=50, t=t, tol=1e-10, max_restarts=0, step_control="expokit", gs_mode=2,
                   sstep=s, values_f32=f32)
    ref = kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                      gen.tridiag_1d(N, lx, dx, ux), t, v)
    e = rel_err(y, ref)
    print(f"[sstep] convdiff s={s} f32={f32} err={e:.2e} blocks={st['sstep_blocks']} "
          f"second={st['sstep_second']} fallbacks={st['sstep_fallbacks']}", flush=True)
    assert e <= 1e-8 and st["sstep_blocks"] > 0


def test_sstep_random_rows_vs_oracle(gpu, oracle):
    """A nonsymmetric random matrix under the reference controller: the
    oracle's accepted and attempted steps and its y."""
    rp, ci, va = gen.random_rows(3000, 5, 30, seed=4)
    v = gen.normal_vector(3000, seed=5)
    t, tol, m = 1.5, 1e-9, 40
    y, st = _solve(rp, ci, va, v, m=m, t=t, tol=tol, max_restarts=10, sstep=8)
    yo, so, _ = oracle.expmv(rp, ci, va, v, t, tol, m, 10)
    assert (st["accepted"], st["attempts"], st["converged"]) == (so["restarts_done"], so["attempts"], so["converged"])
    assert rel_err(y, yo) <= 1e-8
    assert st["sstep_blocks"] > 0


@pytest.mark.parametrize("kind", ["laplacian", "random", "convdiff"])
def test_sstep_hessenberg_matches_column_path(gpu, kind):
    """One Arnoldi decomposition (m = 40) from the same v: the block path's H
    equals the column path's with the second Gram-Schmidt pass on every
    column (the Arnoldi process fixes H for a given v; only rounding
    differs): to 1e-12 of ||H|| on the stencils, to 1e-8 on the random rows,
    whose blocks are the worst conditioned (every one takes the second update
    and the second Cholesky QR, accepted at orthogonality 1e-10).  The
    conditional second pass is no reference here: on the random rows it skips
    columns whose orthogonality then drifts, and its H leaves the others by up
    to 1e-6 in the last columns."""
    if kind == "laplacian":
        rp, ci, va = gen.laplacian_2d(80)
    elif kind == "random":
        rp, ci, va = gen.random_rows(5000, 5, 30, seed=6)
    else:
        rp, ci, va = gen.convdiff_3d(20)
    n = rp.size - 1
    v = gen.normal_vector(n, seed=7)
    Hs = {}
    for s in (0, 8):
        R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=40, t=1.0, gs_mode=1, sstep=s)
        H, k = R.arnoldi(v)
        R.close()
        assert k == 40
        Hs[s] = H
    d = np.linalg.norm(Hs[8] - Hs[0]) / np.linalg.norm(Hs[0])
    print(f"[sstep] H {kind}: rel diff {d:.2e}", flush=True)
    assert d <= (1e-8 if kind == "random" else 1e-12)
    assert np.all(np.tril(Hs[8], -2) == 0.0)


@pytest.mark.parametrize("m,s,first", [(37, 8, 8), (37, 8, 0), (37, 8, 3), (63, 16, 16), (20, 2, 4)])
def test_sstep_block_shapes(gpu, m, s, first):
    """Uneven last blocks, blocks from the first column, tiny blocks and the
    largest basis (m + 1 = 64) give the column path's y."""
    rp, ci, va = gen.random_rows(4000, 5, 30, seed=8)
    v = gen.normal_vector(4000, seed=9)
    kw = dict(m=m, t=2.0, tol=1e-10, max_restarts=0, step_control="expokit_rel", gs_mode=2)
    y, st = _solve(rp, ci, va, v, sstep=s, sstep_first=first, **kw)
    y0, _ = _solve(rp, ci, va, v, **kw)
    print(f"[sstep] m={m} s={s} first={first} blocks={st['sstep_blocks']} second={st['sstep_second']} "
          f"fallbacks={st['sstep_fallbacks']} diff={rel_err(y, y0):.2e}", flush=True)
    assert rel_err(y, y0) <= 1e-10


def test_sstep_falls_back_on_a_rank_deficient_block(gpu):
    """A diagonal matrix with three distinct eigenvalues spans a Krylov space
    of dimension 3: a block from the first column has no new directions past
    the third, its Cholesky QR fails, and the cycle is rebuilt column by
    column, ending in the reference's happy breakdown with the exact y."""
    n = 600
    d = -np.array([1.0, 2.0, 3.0])[np.arange(n) % 3]
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = gen.normal_vector(n, seed=10)
    y, st = _solve(rp, ci, d, v, m=20, t=1.3, tol=1e-10, max_restarts=0, step_control="expokit", gs_mode=2,
                   sstep=8, sstep_first=0)
    print(f"[sstep] diagonal: fallbacks={st['sstep_fallbacks']} blocks={st['sstep_blocks']}", flush=True)
    assert st["sstep_fallbacks"] >= 1
    assert rel_err(y, np.exp(1.3 * d) * v) <= 1e-12


def test_sstep_breakdown_in_the_first_columns(gpu):
    """An eigenvector start breaks down at the first column; the blocks that
    follow return at once and y = e^{t lambda} v."""
    N, t = 40, 3.0
    rp, ci, va = gen.laplacian_2d(N)
    v, lam = laplacian_2d_eigvec(N, 2, 5)
    y, st = _solve(rp, ci, va, v, m=30, t=t, tol=1e-10, max_restarts=0, step_control="expokit", gs_mode=2,
                   sstep=8)
    assert rel_err(y, np.exp(t * lam) * v) <= 1e-12
    assert st["sstep_blocks"] == 0 and st["sstep_fallbacks"] == 0


@pytest.mark.parametrize("which", ["laplacian", "convdiff", "random"])
def test_fused_cholesky_qr_matches_unfused(gpu, oracle, which):
    """The fused Cholesky QR (R from the dots by Pythagoras, Q_new written by
    the update sweep, no normalising sweep) against the unfused blocks
    (gs_mode 1: second update and normalising sweep every time) and the
    column path: the same y to 1e-10 (1e-8 on random rows) and the reference
    to 1e-8, with the fused path taken in some blocks on the stencils (DGKS
    mode) and in none with gs_mode 1."""
    if which == "laplacian":
        N, t, m = 96, 10.0, 40
        rp, ci, va = gen.laplacian_2d(N)
        v = gen.normal_vector(N * N, seed=3)
        ref = laplacian_2d_expmv(N, t, v)
    elif which == "convdiff":
        N, t, m = 24, 0.5, 50
        rp, ci, va = gen.convdiff_3d(N)
        (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
        v = gen.normal_vector(N ** 3, seed=4)
        ref = kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                          gen.tridiag_1d(N, lx, dx, ux), t, v)
    else:
        t, m = 0.5, 40
        rp, ci, va = gen.random_rows(6000, 5, 30, seed=5)
        v = gen.normal_vector(6000, seed=5)
        ref, _st, _log = oracle.expmv(rp, ci, va, v, t, 1e-12, 40, 0)
    # the bench's shape on the stencils; blocks of 8 on random rows, whose
    # Newton blocks of 16 are too ill-conditioned to pass the Cholesky test
    s, first = mek.sstep_plan(m) if which != "random" else (8, 8)
    kw = dict(m=m, t=t, tol=1e-10, max_restarts=0, step_control="expokit", sstep=s, sstep_first=first)
    yf, stf = _solve(rp, ci, va, v, gs_mode=2, **kw)
    yu, stu = _solve(rp, ci, va, v, gs_mode=1, **kw)
    yc, _ = _solve(rp, ci, va, v, m=m, t=t, tol=1e-10, max_restarts=0, step_control="expokit", gs_mode=2)
    print(f"[sstep] {which} fused: blocks={stf['sstep_blocks']} fused={stf['sstep_fused']} "
          f"second={stf['sstep_second']}; unfused: blocks={stu['sstep_blocks']} fused={stu['sstep_fused']}",
          flush=True)
    if which != "random":
        assert stf["sstep_fallbacks"] == 0 and stu["sstep_fallbacks"] == 0
    assert stu["sstep_fused"] == 0 and stu["sstep_blocks"] > 0 and stf["sstep_blocks"] > 0
    if which != "random":
        assert stf["sstep_fused"] > 0
    assert stf["sstep_fused"] <= stf["sstep_blocks"]
    # random rows give the worst-conditioned blocks: agreement with the column
    # path to 1e-8 there (as for H in test_sstep_hessenberg_matches_column_path)
    agree = 1e-8 if which == "random" else 1e-10
    for y in (yf, yu):
        assert rel_err(y, ref) <= 1e-8
        assert rel_err(y, yc) <= agree
