"""CSRHost loaders through the C ABI (host code only, no GPU).

Binary CSR v1 per the reference's docs/binary_csr_v1.md and its strict loader
(main.cu:265-523): one valid/invalid file per validation branch, with the
reference's error text.  MatrixMarket per main.cu:97-245: general, pattern,
symmetric mirroring, duplicate summing, comments, bounds.
"""
import gzip

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import gen, mek


def _small():
    D = np.array([[4.0, 0, 1.5, 0], [0, -2.0, 0, 0], [0.5, 0, 0, 3.0], [0, 7.0, 0, -1.0]])
    A = sp.csr_matrix(D)
    return A.indptr, A.indices, A.data, D


def _load_err(path, loader="bincsr"):
    with pytest.raises(mek.MekError) as e:
        (mek.CSR.load_bincsr if loader == "bincsr" else mek.CSR.load_mtx)(path)
    return str(e.value)


@pytest.mark.parametrize("idt,vdt", [("u32", "f64"), ("u64", "f64"), ("u32", "f32"), ("u64", "f32")])
def test_bincsr_roundtrip(tmp_path, idt, vdt):
    rp, ci, va, D = _small()
    p = tmp_path / "a.bincsr"
    gen.write_bincsr(p, rp, ci, va, index_dtype=idt, value_dtype=vdt)
    A = mek.CSR.load_bincsr(p)
    info = A.info()
    assert info == {"rows": 4, "cols": 4, "nnz": 7, "value_dtype": 1 if vdt == "f32" else 2}
    r2, c2, v2 = A.arrays()
    np.testing.assert_array_equal(r2, rp)
    np.testing.assert_array_equal(c2, ci)
    np.testing.assert_array_equal(v2, va.astype(np.float32 if vdt == "f32" else np.float64))


def test_bincsr_symmetric_upper_expansion(tmp_path):
    # upper triangle incl. diagonal, plus a duplicate (1,1) to be summed
    U = sp.csr_matrix(np.array([[2.0, 1.0, 0, -3.0], [0, 5.0, 4.0, 0], [0, 0, 6.0, 0], [0, 0, 0, 1.0]]))
    p = tmp_path / "s.bincsr"
    gen.write_bincsr(p, U.indptr, U.indices, U.data, flags=gen.FLAG_SORTED | gen.FLAG_SYM_UPPER)
    full = U.toarray() + np.triu(U.toarray(), 1).T  # hand expansion
    A = mek.CSR.load_bincsr(p)
    rp, ci, va = A.arrays()
    np.testing.assert_array_equal(sp.csr_matrix((va, ci, rp), shape=(4, 4)).toarray(), full)
    assert A.info()["nnz"] == np.count_nonzero(full)


def test_bincsr_duplicates_summed_in_symmetric_mode(tmp_path):
    rp = np.array([0, 2, 3])
    ci = np.array([1, 1, 1])
    va = np.array([1.0, 2.5, 4.0])
    p = tmp_path / "d.bincsr"
    gen.write_bincsr(p, rp, ci, va, flags=gen.FLAG_SYM_UPPER)
    r2, c2, v2 = mek.CSR.load_bincsr(p).arrays()
    # row 0: (0,1)=1+2.5; row 1: (1,0)=3.5 mirrored, (1,1)=4
    np.testing.assert_array_equal(r2, [0, 1, 3])
    np.testing.assert_array_equal(c2, [1, 0, 1])
    np.testing.assert_array_equal(v2, [3.5, 3.5, 4.0])


INVALID = [
    ("magic", dict(magic=b"CSR\x00\x00\x00\x00\x02"), "Invalid Binary CSR magic in"),
    ("version", dict(version=2), "Unsupported Binary CSR version 2 in"),
    ("reserved", dict(reserved=(1, 0)), "Reserved header fields must be zero in"),
]


@pytest.mark.parametrize("name,kw,msg", INVALID)
def test_bincsr_header_errors(tmp_path, name, kw, msg):
    rp, ci, va, _ = _small()
    p = tmp_path / f"{name}.bincsr"
    gen.write_bincsr(p, rp, ci, va, **kw)
    assert msg in _load_err(p)


def _patch(path, offset, fmt_bytes):
    b = bytearray(path.read_bytes())
    b[offset:offset + len(fmt_bytes)] = fmt_bytes
    path.write_bytes(bytes(b))


def test_bincsr_dtype_codes(tmp_path):
    rp, ci, va, _ = _small()
    p = tmp_path / "i.bincsr"
    gen.write_bincsr(p, rp, ci, va)
    _patch(p, 40, (3).to_bytes(4, "little"))
    assert "Unsupported index_dtype=3 in" in _load_err(p)
    gen.write_bincsr(p, rp, ci, va)
    _patch(p, 44, (7).to_bytes(4, "little"))
    assert "Unsupported value_dtype=7 in" in _load_err(p)


def test_bincsr_int32_limit_and_size(tmp_path):
    rp, ci, va, _ = _small()
    p = tmp_path / "big.bincsr"
    gen.write_bincsr(p, rp, ci, va)
    _patch(p, 16, (2 ** 31).to_bytes(8, "little"))
    assert "exceed int32 limits" in _load_err(p)
    gen.write_bincsr(p, rp, ci, va)
    p.write_bytes(p.read_bytes() + b"\x00")
    assert "Binary CSR size mismatch in" in _load_err(p) and "expected 168 bytes, got 169" in _load_err(p)
    p.write_bytes(b"CSR")
    assert "too small" in _load_err(p)
    assert "Could not open binary CSR file" in _load_err(tmp_path / "missing.bincsr")


@pytest.mark.parametrize("case,msg", [
    ("indptr0", "indptr[0] must be 0 in"),
    ("decreasing", "indptr is not non-decreasing at row 1 in"),
    ("last", "indptr[nrows] != nnz in"),
    ("col", "Column index out of range at k=2 in"),
    ("unsorted", "Row 0 is not sorted but sorted flag is set in"),
    ("dup", "Row 0 has duplicates but no_duplicates flag is set in"),
])
def test_bincsr_invariants(tmp_path, case, msg):
    rp = np.array([0, 2, 3, 4, 6])
    ci = np.array([0, 2, 1, 0, 1, 3])
    va = np.arange(6, dtype=float)
    if case == "indptr0":
        rp = np.array([1, 2, 3, 4, 6])
    elif case == "decreasing":
        rp = np.array([0, 3, 2, 4, 6])
    elif case == "last":
        rp = np.array([0, 2, 3, 4, 5])
    elif case == "col":
        ci = np.array([0, 2, 9, 0, 1, 3])
    elif case == "unsorted":
        ci = np.array([2, 0, 1, 0, 1, 3])
    elif case == "dup":
        ci = np.array([2, 2, 1, 0, 1, 3])
    p = tmp_path / f"{case}.bincsr"
    gen.write_bincsr(p, rp, ci, va)
    assert msg in _load_err(p)


def test_bincsr_symmetric_requires_square(tmp_path):
    p = tmp_path / "ns.bincsr"
    gen.write_bincsr(p, np.array([0, 1, 2]), np.array([0, 2]), np.array([1.0, 2.0]), ncols=3,
                     flags=gen.FLAG_SYM_UPPER)
    assert "symmetric_upper requires square matrix in" in _load_err(p)


def test_bincsr_unsorted_without_flag_is_accepted(tmp_path):
    p = tmp_path / "u.bincsr"
    gen.write_bincsr(p, np.array([0, 2, 2]), np.array([1, 0]), np.array([1.0, 2.0]), flags=0)
    r, c, v = mek.CSR.load_bincsr(p).arrays()
    np.testing.assert_array_equal(c, [1, 0])  # kept as stored


# ----------------------------------------------------------- MatrixMarket

def _mtx(tmp_path, text, name="a.mtx", gz=False):
    p = tmp_path / (name + (".gz" if gz else ""))
    if gz:
        with gzip.open(p, "wt") as f:
            f.write(text)
    else:
        p.write_text(text)
    return p


def _dense(A):
    rp, ci, va = A.arrays()
    i = A.info()
    return sp.csr_matrix((va, ci, rp), shape=(i["rows"], i["cols"])).toarray()


def test_mtx_general_with_comments_and_duplicates(tmp_path):
    text = ("% leading comment\n%%MatrixMarket matrix coordinate real general\n% c\n\n3 3 5\n"
            "1 1 2.0\n3 2 -1.5\n% inside\n1 1 0.5\n2 3 4e-1\n3 3 1\n")
    D = _dense(mek.CSR.load_mtx(_mtx(tmp_path, text)))
    np.testing.assert_array_equal(D, [[2.5, 0, 0], [0, 0, 0.4], [0, -1.5, 1.0]])


def test_mtx_symmetric_pattern_and_gzip(tmp_path):
    text = "%%MatrixMarket MATRIX Coordinate Pattern Symmetric\n3 3 3\n1 1\n3 1\n2 2\n"
    for gz in (False, True):
        D = _dense(mek.CSR.load_mtx(_mtx(tmp_path, text, gz=gz)))
        np.testing.assert_array_equal(D, [[1, 0, 1], [0, 1, 0], [1, 0, 0]])


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "Only 'coordinate' format is supported"),
    ("%%MatrixMarket vector coordinate real general\n", "MatrixMarket object must be 'matrix'"),
    ("%%MatrixMarket matrix coordinate complex general\n", "Only 'real' or 'pattern' fields are supported"),
    ("%%MatrixMarket matrix coordinate real hermitian\n", "Only 'general' or 'symmetric' symmetry types"),
    ("% no header\n2 2 1\n1 1 1\n", "Failed to parse MatrixMarket header"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "MatrixMarket index out of bounds r=3 c=1"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n", "Expected 3 entries but read 1"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n", "Failed to read value for entry 0"),
    ("%%MatrixMarket matrix coordinate real general\n% only comments\n", "Matrix dimensions line not found"),
    ("%%MatrixMarket matrix coordinate real general\n-1 2 0\n", "Invalid matrix dimensions -1 2 0"),
    ("%%MatrixMarket matrix coordinate real general\n-5 -5 1\n1 1 1.0\n", "Invalid matrix dimensions -5 -5 1"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 -1\n", "Invalid matrix dimensions 2 2 -1"),
])
def test_mtx_errors(tmp_path, text, msg):
    assert msg in _load_err(_mtx(tmp_path, text), "mtx")


def test_generated_matrix_roundtrip_through_bincsr(tmp_path):
    rp, ci, va = gen.convdiff_3d(6)
    p = tmp_path / "cd.bincsr"
    gen.write_bincsr(p, rp, ci, va, value_dtype="f32")
    r2, c2, v2 = mek.CSR.load_bincsr(p).arrays()
    np.testing.assert_array_equal(r2, rp)
    np.testing.assert_array_equal(v2, va)  # stencil coefficients are exact in f32


def test_f32_inexact_count():
    """FP32 storage check: stencil coefficients are exact, 0.1 is not (its
    FP32 rounding changes it by 1.49e-9 relative)."""
    A = mek.CSR.from_arrays(np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                            np.array([-4.0, 1.0, 13.5, -93.5]))
    assert A.f32_inexact() == (0, 0.0)
    B = mek.CSR.from_arrays(np.array([0, 2, 3], np.int32), np.array([0, 1, 1], np.int32),
                            np.array([0.1, 0.5, 1e-40]))
    n, worst = B.f32_inexact()
    assert n == 2  # 0.1, and 1e-40, which is subnormal in FP32
    r = abs(float(np.float32(0.1)) - 0.1) / 0.1
    assert worst >= r and worst <= 1.0


def test_f32_inexact_nonfinite_apart():
    """Values beyond FP32's range and non-finite values are counted apart;
    the largest relative change stays finite (it is over the finite ones)."""
    B = mek.CSR.from_arrays(np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                            np.array([1e300, 0.1, np.inf, 2.0]))
    n, worst, bad = B.f32_inexact(nonfinite=True)
    assert (n, bad) == (3, 2)
    assert np.isfinite(worst) and worst == abs(float(np.float32(0.1)) - 0.1) / 0.1
