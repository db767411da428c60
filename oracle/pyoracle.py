"""ctypes wrapper of the CPU oracle (oracle/oracle.c) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` as the checker.  The product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")


class Stats(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int),
        ("restarts_done", ctypes.c_int),
        ("last_residual", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("attempts", ctypes.c_int),
        ("spmv_count", ctypes.c_int),
        ("aborted", ctypes.c_int),
    ]


class Step(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("err", ctypes.c_double),
        ("m_used", ctypes.c_int),
        ("accepts", ctypes.c_int),
        ("rejects", ctypes.c_int),
    ]


_lib = None


def build() -> None:
    global _lib
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_spmv.argtypes = [ctypes.c_int, ip, ip, dp, dp, dp]
        L.oracle_spmv.restype = None
        L.oracle_expm.argtypes = [ctypes.c_int, dp, dp]
        L.oracle_expm.restype = ctypes.c_int
        L.oracle_pade_coeffs.argtypes = [ctypes.c_int, dp]
        L.oracle_pade_coeffs.restype = None
        L.oracle_expmv.argtypes = [ctypes.c_int, ip, ip, dp, dp, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   dp, ctypes.POINTER(Stats), ctypes.POINTER(Step), ctypes.c_int]
        L.oracle_expmv.restype = ctypes.c_int
        L.oracle_arnoldi.argtypes = [ctypes.c_int, ip, ip, dp, dp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_arnoldi.restype = ctypes.c_int
        L.oracle_arnoldi_columns.argtypes = [ctypes.c_int, ip, ip, dp, dp, ctypes.c_int, dp]
        L.oracle_arnoldi_columns.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _csr(row_ptr, col_idx, vals):
    return (np.ascontiguousarray(row_ptr, dtype=np.int32),
            np.ascontiguousarray(col_idx, dtype=np.int32),
            np.ascontiguousarray(vals, dtype=np.float64))


def spmv(row_ptr, col_idx, vals, x):
    rp, ci, va = _csr(row_ptr, col_idx, vals)
    n = rp.size - 1
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(n, dtype=np.float64)
    lib().oracle_spmv(n, _ip(rp), _ip(ci), _dp(va), _dp(x), _dp(y))
    return y


def expm(A):
    A = np.asarray(A, dtype=np.float64)
    k = A.shape[0]
    a = np.asfortranarray(A)
    e = np.empty((k, k), dtype=np.float64, order="F")
    s = lib().oracle_expm(k, _dp(a), _dp(e))
    return np.array(e), s


def pade_coeffs(q):
    b = np.zeros(q + 1, dtype=np.float64)
    lib().oracle_pade_coeffs(q, _dp(b))
    return b


def expmv(row_ptr, col_idx, vals, v, t, tol=1e-6, m=30, max_restarts=10,
          reorthogonalize=True, max_steps_log=4096):
    """Reference-semantics adaptive Arnoldi; returns (y, stats dict, steps list)."""
    rp, ci, va = _csr(row_ptr, col_idx, vals)
    n = rp.size - 1
    v = np.ascontiguousarray(v, dtype=np.float64)
    y = np.zeros(n, dtype=np.float64)
    st = Stats()
    steps = (Step * max_steps_log)()
    lib().oracle_expmv(n, _ip(rp), _ip(ci), _dp(va), _dp(v), float(t), float(tol), int(m),
                       int(max_restarts), 1 if reorthogonalize else 0, _dp(y),
                       ctypes.byref(st), steps, max_steps_log)
    stats = {f: getattr(st, f) for f, _ in Stats._fields_}
    log = [{f: getattr(steps[i], f) for f, _ in Step._fields_}
           for i in range(min(st.attempts, max_steps_log))]
    return y, stats, log


def arnoldi(row_ptr, col_idx, vals, v, m, reorthogonalize=True, basis=False):
    """(H as an (m+1) x m array, k_used) of one MGS(2) Arnoldi decomposition;
    with basis=True, (H, V, k_used) with V the n x (m+1) basis."""
    rp, ci, va = _csr(row_ptr, col_idx, vals)
    n = rp.size - 1
    v = np.ascontiguousarray(v, dtype=np.float64)
    H = np.zeros((m + 1) * m, dtype=np.float64)
    V = np.zeros((m + 1) * n, dtype=np.float64) if basis else None
    k = lib().oracle_arnoldi(n, _ip(rp), _ip(ci), _dp(va), _dp(v), int(m),
                             1 if reorthogonalize else 0, _dp(H),
                             _dp(V) if basis else None)
    H = H.reshape(m, m + 1).T.copy()
    if basis:
        return H, V.reshape(m + 1, n).T.copy(), k
    return H, k


def arnoldi_columns(row_ptr, col_idx, vals, v, ncols):
    """Per-column wall times (ms) of the reference's Arnoldi loop (bounded sample)."""
    rp, ci, va = _csr(row_ptr, col_idx, vals)
    n = rp.size - 1
    v = np.ascontiguousarray(v, dtype=np.float64)
    col_ms = np.zeros(ncols, dtype=np.float64)
    built = lib().oracle_arnoldi_columns(n, _ip(rp), _ip(ci), _dp(va), _dp(v), int(ncols), _dp(col_ms))
    return col_ms[:built]
