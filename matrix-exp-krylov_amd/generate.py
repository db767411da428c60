"""Synthetic test matrices and Binary CSR v1 files.

Generators for the workloads the solver is measured on, written directly in
CSR (no COO sort) so the 16.8M-row grids build in seconds:

* :func:`laplacian_2d`     2-D 5-point Laplacian, Dirichlet, diag -4, +1 neighbours
* :func:`convdiff_3d`      3-D 7-point upwind convection-diffusion (nonsymmetric)
* :func:`random_negative`  uniform random off-diagonals, a_ii = -1 - sum_j |a_ij|
* :func:`random_rows`      per-row 5..30 distinct random columns, Gershgorin-negative

Row i of a grid is ``i = y*N + x`` (2-D) or ``i = (z*N + y)*N + x`` (3-D);
columns are sorted within every row.

:func:`write_bincsr` / :func:`read_bincsr` follow the reference's
docs/binary_csr_v1.md (64-byte little-endian header ``<8sIIQQQIIQQ``, then
indptr, indices, data).  The reference's own writer
(tools/generate_matrices.py:100-167) cannot emit these configurations: it only
applies a positive diagonal shift and has no stencils.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"CSR\x00\x00\x00\x00\x01"
HEADER = struct.Struct("<8sIIQQQIIQQ")
FLAG_SORTED, FLAG_NO_DUP, FLAG_SYM_UPPER = 1, 2, 4


def _stencil_csr(n, offsets, coeffs, valid):
    """CSR from a stencil: offsets sorted ascending, valid[k] a boolean row mask."""
    counts = np.zeros(n, dtype=np.int64)
    for v in valid:
        counts += v
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    pos = row_ptr[:-1].copy()
    rows = np.arange(n, dtype=np.int64)
    for off, c, v in zip(offsets, coeffs, valid):
        r = rows[v]
        p = pos[v]
        col[p] = (r + off).astype(np.int32)
        val[p] = c
        pos[v] += 1
    return row_ptr.astype(np.int32), col, val


def laplacian_2d(N: int):
    """5-point Laplacian on an N x N grid (n = N^2), Dirichlet boundary."""
    n = N * N
    i = np.arange(n, dtype=np.int64)
    x = i % N
    y = i // N
    valid = [y > 0, x > 0, np.ones(n, bool), x < N - 1, y < N - 1]
    return _stencil_csr(n, [-N, -1, 0, 1, N], [1.0, 1.0, -4.0, 1.0, 1.0], valid)


def convdiff_3d_coeffs(pe: float = 50.0, h: float = 1.0, diff: float = 1.0):
    """Upwind convection-diffusion coefficients for velocity pe*(1, 0.5, 0.25).

    Returns (1-D operators per axis as (lower, diag, upper)) for axes x, y, z.
    """
    b = pe * np.array([1.0, 0.5, 0.25])
    out = []
    for bk in b:
        lower = diff / h ** 2 + bk / h
        upper = diff / h ** 2
        diag = -2.0 * diff / h ** 2 - bk / h
        out.append((lower, diag, upper))
    return out


def convdiff_3d(N: int, pe: float = 50.0, h: float = 1.0, diff: float = 1.0):
    """7-point upwind convection-diffusion on N^3 (nonsymmetric)."""
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = convdiff_3d_coeffs(pe, h, diff)
    n = N ** 3
    i = np.arange(n, dtype=np.int64)
    x = i % N
    y = (i // N) % N
    z = i // (N * N)
    valid = [z > 0, y > 0, x > 0, np.ones(n, bool), x < N - 1, y < N - 1, z < N - 1]
    offs = [-N * N, -N, -1, 0, 1, N, N * N]
    coeffs = [lz, ly, lx, dx + dy + dz, ux, uy, uz]
    return _stencil_csr(n, offs, coeffs, valid)


def tridiag_1d(N: int, lower: float, diag: float, upper: float) -> np.ndarray:
    """Dense 1-D operator with the given bands (for Kronecker closed forms)."""
    T = np.diag(np.full(N, diag))
    T += np.diag(np.full(N - 1, lower), -1)
    T += np.diag(np.full(N - 1, upper), 1)
    return T


def _finish_random(n, r, c, v, f32_exact=False):
    """Sum duplicates, add the Gershgorin-negative diagonal, build sorted CSR.
    With f32_exact every value (diagonal included) is rounded to the nearest
    FP32 number, so FP32 storage of the matrix is lossless."""
    key = r.astype(np.int64) * n + c
    order = np.argsort(key, kind="stable")
    key = key[order]
    v = v[order]
    uk, start = np.unique(key, return_index=True)
    vs = np.add.reduceat(v, start) if v.size else v
    rr = (uk // n).astype(np.int64)
    cc = (uk % n).astype(np.int64)
    if f32_exact:
        vs = vs.astype(np.float32).astype(np.float64)
    rowabs = np.bincount(rr, weights=np.abs(vs), minlength=n)
    diag = -1.0 - rowabs
    if f32_exact:  # rounding moves a_ii by < 1e-6 |a_ii|: still Gershgorin-negative
        diag = diag.astype(np.float32).astype(np.float64)
    # merge the diagonal into each row at its sorted position
    rr = np.concatenate([rr, np.arange(n)])
    cc = np.concatenate([cc, np.arange(n)])
    vv = np.concatenate([vs, diag])
    order = np.lexsort((cc, rr))
    rr, cc, vv = rr[order], cc[order], vv[order]
    row_ptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rr, minlength=n), out=row_ptr[1:])
    return row_ptr.astype(np.int32), cc.astype(np.int32), vv


def random_negative(n: int, density: float = 5e-3, seed: int = 42):
    """int(density*n^2) uniform off-diagonal draws (duplicates summed), values
    U[-1,1], a_ii = -1 - sum_{j != i} |a_ij|."""
    rng = np.random.default_rng(seed)
    cnt = int(density * n * n)
    r = rng.integers(0, n, cnt)
    c = rng.integers(0, n, cnt)
    v = rng.uniform(-1.0, 1.0, cnt)
    keep = r != c
    return _finish_random(n, r[keep], c[keep], v[keep])


def random_rows(n: int, nnz_lo: int = 5, nnz_hi: int = 30, seed: int = 0, f32_exact: bool = False):
    """Each row has U{nnz_lo..nnz_hi} nonzeros in distinct random off-diagonal
    columns (values U[-1,1]) plus a_ii = -1 - sum_{j != i} |a_ij|."""
    rng = np.random.default_rng(seed)
    k = rng.integers(nnz_lo, nnz_hi + 1, n)
    k = np.minimum(k, max(n - 1, 0))
    r = np.repeat(np.arange(n, dtype=np.int64), k)
    c = rng.integers(0, max(n - 1, 1), r.size)
    c = c + (c >= r)  # skip the diagonal
    v = rng.uniform(-1.0, 1.0, r.size)
    # one sort by (row, col); repeated columns of a row keep their first draw
    key = r * n + c
    order = np.argsort(key, kind="stable")
    key, v = key[order], v[order]
    keep = np.ones(key.size, bool)
    keep[1:] = key[1:] != key[:-1]
    key, v = key[keep], v[keep]
    rows, cols = key // n, key % n
    if f32_exact:
        v = v.astype(np.float32).astype(np.float64)
    cnt = np.bincount(rows, minlength=n)
    diag = -1.0 - np.bincount(rows, weights=np.abs(v), minlength=n)
    if f32_exact:  # rounding moves a_ii by < 1e-6 |a_ii|: still Gershgorin-negative
        diag = diag.astype(np.float32).astype(np.float64)
    # merged rows hold one extra (diagonal) entry, placed after the columns < i
    row_ptr = np.zeros(n + 1, np.int64)
    np.cumsum(cnt + 1, out=row_ptr[1:])
    off_start = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=off_start[1:])
    idx_in_row = np.arange(key.size, dtype=np.int64) - off_start[rows]
    pos = row_ptr[rows] + idx_in_row + (cols > rows)
    n_lt = np.bincount(rows[cols < rows], minlength=n)
    dpos = row_ptr[:-1] + n_lt
    nnz = int(row_ptr[-1])
    col_idx = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    col_idx[pos] = cols
    vals[pos] = v
    col_idx[dpos] = np.arange(n)
    vals[dpos] = diag
    return row_ptr.astype(np.int32), col_idx, vals


def batch_problems(P: int = 256, n_lo: float = 1e3, n_hi: float = 2e5, seed: int = 0,
                   f32_exact: bool = True):
    """P independent problems: n log-uniform in [n_lo, n_hi], random rows with
    5-30 nonzeros, Gershgorin-negative diagonal, t ~ U[0.1, 5], m ~ U{20..60},
    v ~ N(0,1) normalised.  Returns a list of dicts."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(P):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        rp, ci, va = random_rows(n, 5, 30, seed=seed * 7919 + p, f32_exact=f32_exact)
        out.append(dict(rp=rp, ci=ci, va=va, t=float(rng.uniform(0.1, 5.0)),
                        m=int(rng.integers(20, 61)), v=normal_vector(n, seed=seed * 7919 + p)))
    return out


def normal_vector(n: int, seed: int = 0, normalise: bool = True) -> np.ndarray:
    v = np.random.default_rng(seed).standard_normal(n)
    if normalise:
        v /= np.linalg.norm(v)
    return v


# ------------------------------------------------------------- Binary CSR v1

def write_bincsr(path, row_ptr, col_idx, values, ncols=None, index_dtype="u32",
                 value_dtype="f64", flags=FLAG_SORTED | FLAG_NO_DUP, version=1,
                 reserved=(0, 0), magic=MAGIC):
    """Write Binary CSR v1 (docs/binary_csr_v1.md).  The keyword arguments
    after ``flags`` exist to craft invalid files in tests."""
    row_ptr = np.asarray(row_ptr)
    nrows = row_ptr.size - 1
    ncols = nrows if ncols is None else ncols
    nnz = int(np.asarray(col_idx).size)
    it = {"u32": (1, "<u4"), "u64": (2, "<u8")}[index_dtype]
    vt = {"f32": (1, "<f4"), "f64": (2, "<f8")}[value_dtype]
    with open(path, "wb") as f:
        f.write(HEADER.pack(magic, version, flags, nrows, ncols, nnz, it[0], vt[0],
                            reserved[0], reserved[1]))
        f.write(np.asarray(row_ptr, dtype=it[1]).tobytes())
        f.write(np.asarray(col_idx, dtype=it[1]).tobytes())
        f.write(np.asarray(values, dtype=vt[1]).tobytes())


def read_bincsr(path):
    """Read a Binary CSR v1 file (no symmetric expansion); returns
    (header dict, indptr, indices, data)."""
    with open(path, "rb") as f:
        raw = f.read(HEADER.size)
        magic, ver, flags, nr, nc, nnz, idt, vdt, r0, r1 = HEADER.unpack(raw)
        it = "<u4" if idt == 1 else "<u8"
        vt = "<f4" if vdt == 1 else "<f8"
        ip = np.frombuffer(f.read((nr + 1) * np.dtype(it).itemsize), dtype=it)
        ix = np.frombuffer(f.read(nnz * np.dtype(it).itemsize), dtype=it)
        da = np.frombuffer(f.read(nnz * np.dtype(vt).itemsize), dtype=vt)
    hdr = dict(magic=magic, version=ver, flags=flags, nrows=nr, ncols=nc, nnz=nnz,
               index_dtype=idt, value_dtype=vdt, reserved0=r0, reserved1=r1)
    return hdr, ip, ix, da
