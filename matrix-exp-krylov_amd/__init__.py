"""Python access to libmek.so (the MI355X Arnoldi exp(t*A)*v library) over its C ABI.

The product is the C++/HIP library and the ``matrix_exp`` command line; this
module is a thin ctypes layer used by the tests, ``bench.py`` and
``__graft_entry__``.  It never falls back to a CPU path: if ``lib/libmek.so``
is missing, :func:`lib` raises.

The directory name is not a Python identifier, so load it with :func:`load`
from ``__graft_entry__`` (or ``importlib`` by path).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmek.so")
CLI_PATH = os.path.join(HERE, "bin", "matrix_exp")

# kernel classes of mek_runner_kernel_time
KERNEL_KINDS = {"spmv": 0, "update_dot": 1, "update_norm": 2, "dots": 3, "final": 4,
                "expm": 5, "lift": 6, "other": 7, "sstep_dots": 8, "sstep_update": 9}


STEP_CONTROL = {"reference": 0, "expokit": 1, "expokit_rel": 2}
SPMV_FORMAT = {"auto": 0, "csr": 1, "sell": 2, "dia": 3, "dia_const": 4}
SPMV_FORMAT_NAME = {1: "csr", 2: "sell", 3: "dia", 4: "dia_const"}


class MekError(RuntimeError):
    pass


def build(jobs: int = 8) -> None:
    """Compile libmek.so and matrix_exp for gfx950 in-tree (hipcc)."""
    subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}"], check=True)


class SolverOptions(ctypes.Structure):
    _fields_ = [("step_control", ctypes.c_int), ("gs_mode", ctypes.c_int),
                ("dgks_eta", ctypes.c_double), ("values_f32", ctypes.c_int),
                ("verbose", ctypes.c_int), ("timing", ctypes.c_int),
                ("fuse_spmv_dot", ctypes.c_int), ("lanczos", ctypes.c_int), ("autotune", ctypes.c_int),
                ("sstep", ctypes.c_int), ("sstep_first", ctypes.c_int), ("spmv_format", ctypes.c_int)]


class ArnoldiParams(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("t", ctypes.c_double), ("tol", ctypes.c_double),
                ("max_restarts", ctypes.c_int), ("reorthogonalize", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("restarts_done", ctypes.c_int),
                ("last_residual", ctypes.c_double), ("t_done", ctypes.c_double),
                ("attempts", ctypes.c_int), ("accepted", ctypes.c_int),
                ("rejected", ctypes.c_int), ("spmv_count", ctypes.c_int),
                ("extra_spmv", ctypes.c_int), ("reorth_columns", ctypes.c_int),
                ("aborted", ctypes.c_int), ("solve_ms", ctypes.c_double),
                ("sstep_blocks", ctypes.c_int), ("sstep_second", ctypes.c_int),
                ("sstep_fallbacks", ctypes.c_int), ("sstep_fused", ctypes.c_int)]


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# exported entry points and their (restype, argtypes); the header include/mek.h
# declares exactly these
SIGNATURES = {
    "mek_last_error": (ctypes.c_char_p, []),
    "mek_version": (ctypes.c_char_p, []),
    "mek_device_count": (ctypes.c_int, [_ip]),
    "mek_device_synchronize": (ctypes.c_int, []),
    "mek_malloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "mek_free": (ctypes.c_int, [_vp]),
    "mek_memcpy_htod": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "mek_memcpy_dtoh": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "mek_memcpy_dtod": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "mek_csr_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_longlong, _ip, _ip, _dp,
                                      ctypes.POINTER(_vp)]),
    "mek_csr_load_bincsr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "mek_csr_load_mtx": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "mek_csr_info": (ctypes.c_int, [_vp, _ip, _ip, ctypes.POINTER(ctypes.c_longlong), _ip]),
    "mek_csr_export": (ctypes.c_int, [_vp, _ip, _ip, _dp]),
    "mek_csr_diagonals": (ctypes.c_int, [_vp, _ip, _ip]),
    "mek_csr_diagonal_values": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _ip]),
    "mek_csr_f32_inexact": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong), _dp,
                                           ctypes.POINTER(ctypes.c_longlong)]),
    "mek_csr_destroy": (None, [_vp]),
    "mek_solver_options_default": (None, [ctypes.POINTER(SolverOptions)]),
    "mek_runner_create": (ctypes.c_int, [_vp, ctypes.POINTER(ArnoldiParams),
                                         ctypes.POINTER(SolverOptions), ctypes.c_int,
                                         ctypes.POINTER(_vp)]),
    "mek_runner_expmv": (ctypes.c_int, [_vp, _dp, _dp]),
    "mek_runner_expmv_device": (ctypes.c_int, [_vp, _vp, _vp]),
    "mek_runner_set_t": (ctypes.c_int, [_vp, ctypes.c_double]),
    "mek_runner_values_f32": (ctypes.c_int, [_vp, _ip]),
    "mek_runner_stats": (ctypes.c_int, [_vp, ctypes.POINTER(SolveStats)]),
    "mek_runner_steps": (ctypes.c_int, [_vp, _dp, _ip, ctypes.c_int, _ip]),
    "mek_runner_kernel_time": (ctypes.c_int, [_vp, ctypes.c_int, _dp,
                                              ctypes.POINTER(ctypes.c_longlong), _dp]),
    "mek_runner_arnoldi": (ctypes.c_int, [_vp, _dp, _dp, _ip]),
    "mek_runner_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "mek_runner_destroy": (None, [_vp]),
    "mek_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), _dp, _ip, _dp, _ip, _ip,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "mek_batch_tuning": (ctypes.c_int, [_vp, _ip]),
    "mek_batch_values_f32": (ctypes.c_int, [_vp, _ip]),
    "mek_batch_time_spmv": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "mek_batch_upload_v": (ctypes.c_int, [_vp, ctypes.POINTER(_dp)]),
    "mek_batch_solve": (ctypes.c_int, [_vp]),
    "mek_batch_download_y": (ctypes.c_int, [_vp, ctypes.POINTER(_dp)]),
    "mek_batch_stats": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(SolveStats)]),
    "mek_batch_cycles": (ctypes.c_int, [_vp, _ip]),
    "mek_batch_destroy": (None, [_vp]),
    "mek_features": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
    "mek_batch_features": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "mek_mlp_param_count": (ctypes.c_int, []),
    "mek_mlp_create": (ctypes.c_int, [ctypes.c_uint, ctypes.c_int, ctypes.POINTER(_vp)]),
    "mek_mlp_set_params": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "mek_mlp_get_params": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "mek_mlp_forward": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_float)]),
    "mek_mlp_loss_grad": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "mek_mlp_train": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_uint, ctypes.POINTER(ctypes.c_float)]),
    "mek_mlp_destroy": (None, [_vp]),
    "mek_spmv_plan_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, _vp, _vp, _vp,
                                            ctypes.c_int, _ip, ctypes.POINTER(_vp)]),
    "mek_spmv_execute": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "mek_spmv_plan_destroy": (None, [_vp]),
    "mek_spmv_plan_use_sell": (ctypes.c_int, [_vp, _ip, ctypes.POINTER(ctypes.c_longlong), _ip]),
    "mek_spmv_plan_use_dia": (ctypes.c_int, [_vp, _ip, _ip, _dp, _ip]),
    "mek_spmv_plan_use_dia_const": (ctypes.c_int, [_vp, _ip, _ip, _dp, _ip]),
    "mek_runner_spmv_format": (ctypes.c_int, [_vp, _ip, ctypes.POINTER(ctypes.c_longlong), _ip]),
    "mek_expm_dense": (ctypes.c_int, [ctypes.c_int, _vp, _vp]),
    "mek_expm_dense_noalloc": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp]),
    "mek_runner_tuning": (ctypes.c_int, [_vp, _ip, _ip, _ip, _dp, _dp]),
    "mek_runner_sstep_grid": (ctypes.c_int, [_vp, _ip, _dp]),
    "mek_block_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(_vp)]),
    "mek_block_expmv": (ctypes.c_int, [_vp, _dp, _dp]),
    "mek_block_expmv_device": (ctypes.c_int, [_vp, _vp, _vp]),
    "mek_block_stats_get": (ctypes.c_int, [_vp, ctypes.c_void_p]),
    "mek_block_destroy": (None, [_vp]),
    "mek_spmm": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int]),
    "mek_gemm_dense": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
}


def lib():
    """The loaded libmek.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MekError(f"{LIB_PATH} is missing: build it first (make -C matrix-exp-krylov_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        raise MekError(f"mek status {status}: {lib().mek_last_error().decode()}")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _value_storage(values_f32):
    """values_f32 argument -> the ABI's storage mode: False 0 (FP64), True 1
    (FP32 only when every value is exact in FP32, else FP64), "lossy" 2 (FP32
    always, the caller accepts the perturbation)."""
    if values_f32 == "lossy":
        return 2
    if values_f32 in (False, True, 0, 1):
        return 1 if values_f32 else 0
    raise ValueError(f"values_f32 must be False, True or 'lossy', not {values_f32!r}")


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


# ------------------------------------------------------------------ device

class DeviceBuffer:
    """Raw device allocation (hipMalloc) owned by Python."""

    def __init__(self, nbytes: int):
        p = _vp()
        _check(lib().mek_malloc(ctypes.byref(p), int(nbytes)))
        self.ptr = p
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib().mek_memcpy_htod(b.ptr, a.ctypes.data_as(_vp), a.nbytes))
        return b

    def to_array(self, dtype=np.float64, count=None) -> np.ndarray:
        count = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        _check(lib().mek_memcpy_dtoh(out.ctypes.data_as(_vp), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().mek_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sstep_plan(m: int, sstep: int = -1, first: int = -1):
    """(sstep, sstep_first) for a solve with Krylov dimension m: as given, or
    with sstep = -1 the first 8 columns of each cycle one by one (first, if
    given) and the remaining ones in the fewest equal s-step blocks of at most
    16 vectors (m = 40: 8 + 16 + 16; m = 50: 8 + 3 x 14), the best shapes
    measured on the 4096^2 Laplacian and the 256^3 convection-diffusion
    matrix.  (0, -1), the column path, when fewer than 2 columns are left."""
    if sstep != -1:
        return sstep, first
    first = min(8, m) if first < 0 else first
    rest = m - first
    if rest < 2:
        return 0, -1
    nb = -(-rest // 16)
    return -(-rest // nb), first


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().mek_device_count(ctypes.byref(n)))
    return n.value


def synchronize() -> None:
    _check(lib().mek_device_synchronize())


# ------------------------------------------------------------------ CSR

class CSR:
    """CSRHost handle (host memory)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_arrays(cls, row_ptr, col_idx, values, rows=None, cols=None):
        rp, ci, va = _i32(row_ptr), _i32(col_idx), _f64(values)
        rows = rp.size - 1 if rows is None else rows
        cols = rows if cols is None else cols
        h = _vp()
        _check(lib().mek_csr_create(rows, cols, ci.size, _pi(rp), _pi(ci), _pd(va), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load_bincsr(cls, path):
        h = _vp()
        _check(lib().mek_csr_load_bincsr(os.fsencode(path), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load_mtx(cls, path):
        h = _vp()
        _check(lib().mek_csr_load_mtx(os.fsencode(path), ctypes.byref(h)))
        return cls(h)

    def info(self):
        r, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        z = ctypes.c_longlong()
        _check(lib().mek_csr_info(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(z), ctypes.byref(d)))
        return {"rows": r.value, "cols": c.value, "nnz": z.value, "value_dtype": d.value}

    def f32_inexact(self, nonfinite=False):
        """(count, max relative change) of the values FP32 storage would change;
        with nonfinite=True also the count of values that are not finite or
        lie beyond FP32's range (included in count, not in the max)."""
        c, m, b = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_longlong()
        _check(lib().mek_csr_f32_inexact(self.h, ctypes.byref(c), ctypes.byref(m), ctypes.byref(b)))
        return (c.value, m.value, b.value) if nonfinite else (c.value, m.value)

    def diagonals(self):
        """Ascending column - row offsets of the diagonal-copy SpMV, or []
        when the matrix has no diagonal form (> 16 diagonals, columns not
        ascending within a row, or a NaN value)."""
        offs = np.zeros(16, np.int32)
        n = ctypes.c_int()
        _check(lib().mek_csr_diagonals(self.h, _pi(offs), ctypes.byref(n)))
        return [int(o) for o in offs[:n.value]]

    def diagonal_values(self, f32=False):
        """The one value of each diagonal (in diagonals() order) of the
        constant-diagonal SpMV, or [] when some diagonal holds two values (bit
        patterns, after rounding to FP32 with f32=True) or the matrix has no
        diagonal form."""
        vals = np.zeros(16, np.float64)
        n = ctypes.c_int()
        _check(lib().mek_csr_diagonal_values(self.h, 1 if f32 else 0, _pd(vals), ctypes.byref(n)))
        return [float(v) for v in vals[:n.value]]

    def arrays(self):
        i = self.info()
        rp = np.empty(i["rows"] + 1, np.int32)
        ci = np.empty(i["nnz"], np.int32)
        va = np.empty(i["nnz"], np.float64)
        _check(lib().mek_csr_export(self.h, _pi(rp), _pi(ci), _pd(va)))
        return rp, ci, va

    def close(self):
        if self.h:
            lib().mek_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ runner

class Runner:
    """ArnoldiRunner on one GPU: y = exp(t*A)*v."""

    def __init__(self, A: CSR, m=30, t=1.0, tol=1e-8, max_restarts=10, reorthogonalize=True,
                 step_control="reference", gs_mode=-1, dgks_eta=0.1, values_f32=False,
                 verbose=False, timing=False, fuse_spmv_dot=False, lanczos=False, autotune=False, device=0,
                 sstep=0, sstep_first=-1, spmv_format="auto"):
        p = ArnoldiParams(int(m), float(t), float(tol), int(max_restarts), 1 if reorthogonalize else 0)
        o = SolverOptions()
        lib().mek_solver_options_default(ctypes.byref(o))
        o.step_control = STEP_CONTROL[step_control]
        o.gs_mode = int(gs_mode)
        o.dgks_eta = float(dgks_eta)
        o.values_f32 = _value_storage(values_f32)
        o.verbose = 1 if verbose else 0
        o.timing = 1 if timing else 0
        o.fuse_spmv_dot = 1 if fuse_spmv_dot else 0
        o.lanczos = 1 if lanczos else 0
        o.autotune = 1 if autotune else 0
        o.sstep = int(sstep)
        o.sstep_first = int(sstep_first)
        o.spmv_format = SPMV_FORMAT[spmv_format]
        h = _vp()
        _check(lib().mek_runner_create(A.h, ctypes.byref(p), ctypes.byref(o), int(device), ctypes.byref(h)))
        self.h = h
        self.n = A.info()["rows"]
        self.m = int(m)

    def expmv(self, v, out=None):
        """y = exp(t A) v for a host vector v.  `out` (a C-contiguous float64
        array of n entries) receives y in place, as the reference's
        compute_expmv writes into the caller's y vector."""
        v = _f64(v)
        if out is None:
            y = np.zeros(self.n, np.float64)
        else:
            y = out
            if not (isinstance(y, np.ndarray) and y.dtype == np.float64 and y.flags.c_contiguous
                    and y.flags.writeable and y.size == self.n):
                raise ValueError(f"out must be a writeable C-contiguous float64 array of {self.n} entries")
        _check(lib().mek_runner_expmv(self.h, _pd(v), _pd(y)))
        return y

    def expmv_device(self, d_v: DeviceBuffer, d_y: DeviceBuffer):
        _check(lib().mek_runner_expmv_device(self.h, d_v.ptr, d_y.ptr))

    def set_t(self, t):
        _check(lib().mek_runner_set_t(self.h, float(t)))

    def values_f32_used(self):
        """True when the matrix values are stored in FP32 on the device."""
        u = ctypes.c_int()
        _check(lib().mek_runner_values_f32(self.h, ctypes.byref(u)))
        return bool(u.value)

    def stats(self):
        s = SolveStats()
        _check(lib().mek_runner_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in SolveStats._fields_}

    def steps(self, cap=100000):
        n = ctypes.c_int()
        _check(lib().mek_runner_steps(self.h, None, None, 0, ctypes.byref(n)))
        k = min(n.value, cap)
        tde = np.zeros(3 * max(k, 1), np.float64)
        mar = np.zeros(3 * max(k, 1), np.int32)
        _check(lib().mek_runner_steps(self.h, _pd(tde), _pi(mar), k, ctypes.byref(n)))
        return [{"t": tde[3 * i], "dt": tde[3 * i + 1], "err": tde[3 * i + 2], "m_used": int(mar[3 * i]),
                 "accepts": int(mar[3 * i + 1]), "rejects": int(mar[3 * i + 2])} for i in range(k)]

    def kernel_time(self, kind):
        ms, nb = ctypes.c_double(), ctypes.c_double()
        nl = ctypes.c_longlong()
        _check(lib().mek_runner_kernel_time(self.h, KERNEL_KINDS[kind], ctypes.byref(ms), ctypes.byref(nl),
                                            ctypes.byref(nb)))
        return {"ms": ms.value, "launches": nl.value, "bytes": nb.value}

    def tuning(self):
        """Launch shapes the autotuner chose (all zero without autotune)."""
        c, g, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().mek_runner_tuning(self.h, ctypes.byref(c), ctypes.byref(g), ctypes.byref(w),
                                       ctypes.byref(a), ctypes.byref(b)))
        sg, su = ctypes.c_int(), ctypes.c_double()
        _check(lib().mek_runner_sstep_grid(self.h, ctypes.byref(sg), ctypes.byref(su)))
        return {"nnz_cap": c.value, "spmv_grid": g.value, "sweep_grid": w.value,
                "spmv_us": a.value, "sweep_us": b.value, "sstep_grid": sg.value, "sstep_us": su.value}

    def spmv_format(self):
        """The SpMV layout in use: {"format": "csr" | "sell" | "dia", "entries":
        padded entries (sliced ELL) or rows x diagonals (diagonal copy),
        "width": widest row (sliced ELL) or diagonal count}."""
        f, w = ctypes.c_int(), ctypes.c_int()
        e = ctypes.c_longlong()
        _check(lib().mek_runner_spmv_format(self.h, ctypes.byref(f), ctypes.byref(e), ctypes.byref(w)))
        return {"format": SPMV_FORMAT_NAME[f.value], "entries": e.value, "width": w.value}

    def arnoldi(self, v):
        v = _f64(v)
        H = np.zeros((self.m + 1) * self.m, np.float64)
        k = ctypes.c_int()
        _check(lib().mek_runner_arnoldi(self.h, _pd(v), _pd(H), ctypes.byref(k)))
        return H.reshape(self.m, self.m + 1).T.copy(), k.value

    def close(self):
        if self.h:
            lib().mek_runner_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ batch

class Batch:
    """Many independent exp(t_p A_p) v_p solves in lockstep on one GPU."""

    def __init__(self, mats, t, m, tol, max_restarts=None, substeps=None, values_f32=False,
                 gs_mode=1, dgks_eta=0.1, use_graph=True, device=0, step_control="reference", autotune=False):
        P = len(mats)
        self.csrs = [A if isinstance(A, CSR) else CSR.from_arrays(*A) for A in mats]
        self.rows = [c.info()["rows"] for c in self.csrs]
        handles = (_vp * P)(*[c.h for c in self.csrs])
        t = _f64(np.broadcast_to(t, P))
        m = _i32(np.broadcast_to(m, P))
        tol = _f64(np.broadcast_to(tol, P))
        mr = _i32(np.broadcast_to(10 if max_restarts is None else max_restarts, P))
        ss = _i32(np.broadcast_to(0 if substeps is None else substeps, P))
        h = _vp()
        ctrl = STEP_CONTROL[step_control]
        _check(lib().mek_batch_create(P, handles, _pd(t), _pi(m), _pd(tol), _pi(mr), _pi(ss), ctrl,
                                      _value_storage(values_f32), int(gs_mode), float(dgks_eta),
                                      1 if use_graph else 0, 1 if autotune else 0, int(device), ctypes.byref(h)))
        self.h = h
        self.P = P

    def upload(self, vs):
        vs = [_f64(v) for v in vs]
        arr = (_dp * self.P)(*[_pd(v) for v in vs])
        _check(lib().mek_batch_upload_v(self.h, arr))

    def solve(self):
        _check(lib().mek_batch_solve(self.h))

    def download(self):
        ys = [np.zeros(n, np.float64) for n in self.rows]
        arr = (_dp * self.P)(*[_pd(y) for y in ys])
        _check(lib().mek_batch_download_y(self.h, arr))
        return ys

    def stats(self, p):
        s = SolveStats()
        _check(lib().mek_batch_stats(self.h, int(p), ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in SolveStats._fields_}

    def features(self):
        out = np.zeros((self.P, 16), np.float32)
        _check(lib().mek_batch_features(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def cycles(self):
        c = ctypes.c_int()
        _check(lib().mek_batch_cycles(self.h, ctypes.byref(c)))
        return c.value

    def values_f32_used(self):
        """True when the batch stores its values in FP32 on the device."""
        u = ctypes.c_int()
        _check(lib().mek_batch_values_f32(self.h, ctypes.byref(u)))
        return bool(u.value)

    def spmv_grid(self):
        """SpMV grid chosen by the autotuner (0 without autotune)."""
        g = ctypes.c_int()
        _check(lib().mek_batch_tuning(self.h, ctypes.byref(g)))
        return g.value

    def time_spmv(self, reps=20):
        """Median us of one SpMV over every problem (after upload)."""
        us = ctypes.c_double()
        _check(lib().mek_batch_time_spmv(self.h, int(reps), ctypes.byref(us)))
        return us.value

    def close(self):
        if getattr(self, "h", None):
            lib().mek_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- learned selector

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def features(A: CSR, t: float, tol: float = 1e-8, device: int = 0) -> np.ndarray:
    """The 16 selector features of (A, t, tol), computed on the GPU."""
    out = np.zeros(16, np.float32)
    _check(lib().mek_features(A.h, float(t), float(tol), out.ctypes.data_as(_fp), int(device)))
    return out


class Mlp:
    """16 -> 64 -> 64 -> 2 MLP trained and evaluated by the library's HIP kernels."""

    def __init__(self, seed: int = 0, device: int = 0):
        h = _vp()
        _check(lib().mek_mlp_create(int(seed), int(device), ctypes.byref(h)))
        self.h = h
        self.nparams = lib().mek_mlp_param_count()

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        _check(lib().mek_mlp_get_params(self.h, p.ctypes.data_as(_fp)))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparams
        _check(lib().mek_mlp_set_params(self.h, p.ctypes.data_as(_fp)))

    def forward(self, X) -> np.ndarray:
        X = _f32(X)
        Y = np.zeros((X.shape[0], 2), np.float32)
        _check(lib().mek_mlp_forward(self.h, X.ctypes.data_as(_fp), X.shape[0], Y.ctypes.data_as(_fp)))
        return Y

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_float()
        _check(lib().mek_mlp_loss_grad(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp), X.shape[0],
                                       ctypes.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def train(self, X, T, batch=256, epochs=100, lr=1e-3, seed=0) -> float:
        X, T = _f32(X), _f32(T)
        loss = ctypes.c_float()
        _check(lib().mek_mlp_train(self.h, X.ctypes.data_as(_fp), T.ctypes.data_as(_fp), X.shape[0],
                                   int(batch), int(epochs), float(lr), int(seed), ctypes.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            lib().mek_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ kernel level

class SpmvPlan:
    """Device CSR + adaptive row blocks; execute() computes y = A x."""

    def __init__(self, row_ptr, col_idx, values, values_f32=False):
        rp = _i32(row_ptr)
        self.n = rp.size - 1
        self.nnz = int(rp[-1])
        vals = np.ascontiguousarray(values, dtype=np.float32 if values_f32 else np.float64)
        self.d_rp = DeviceBuffer.from_array(rp)
        self.d_ci = DeviceBuffer.from_array(_i32(col_idx))
        self.d_va = DeviceBuffer.from_array(vals)
        h = _vp()
        _check(lib().mek_spmv_plan_create(self.n, self.nnz, self.d_rp.ptr, self.d_ci.ptr, self.d_va.ptr,
                                          1 if values_f32 else 0, _pi(rp), ctypes.byref(h)))
        self.h = h
        self._rp = rp
        self._ci = _i32(col_idx)
        self._va = _f64(values)

    def use_dia(self):
        """Switch to the diagonal copy of the matrix; returns the diagonal
        count (MekError when the matrix has no diagonal form)."""
        nd = ctypes.c_int()
        _check(lib().mek_spmv_plan_use_dia(self.h, _pi(self._rp), _pi(self._ci), _pd(self._va), ctypes.byref(nd)))
        return nd.value

    def use_dia_const(self):
        """Switch to the constant-diagonal form (one value per diagonal, a
        presence mask per row); returns the diagonal count (MekError when a
        diagonal holds two values or the matrix has no diagonal form)."""
        nd = ctypes.c_int()
        _check(lib().mek_spmv_plan_use_dia_const(self.h, _pi(self._rp), _pi(self._ci), _pd(self._va),
                                                 ctypes.byref(nd)))
        return nd.value

    def use_sell(self):
        """Switch to the sliced-ELL copy of the matrix; returns (padded
        entries, widest row)."""
        e, w = ctypes.c_longlong(), ctypes.c_int()
        _check(lib().mek_spmv_plan_use_sell(self.h, _pi(self._rp), ctypes.byref(e), ctypes.byref(w)))
        return e.value, w.value

    def execute(self, d_x: DeviceBuffer, d_y: DeviceBuffer, stream=None):
        _check(lib().mek_spmv_execute(self.h, d_x.ptr, d_y.ptr, stream))

    def __call__(self, x):
        d_x = DeviceBuffer.from_array(_f64(x))
        d_y = DeviceBuffer(8 * self.n)
        self.execute(d_x, d_y)
        synchronize()
        return d_y.to_array()

    def close(self):
        if getattr(self, "h", None):
            lib().mek_spmv_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BlockStats(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("accepted", ctypes.c_int), ("attempts", ctypes.c_int),
                ("rejected", ctypes.c_int), ("block_steps", ctypes.c_int), ("squarings", ctypes.c_int),
                ("second_passes", ctypes.c_int), ("spmm_grid", ctypes.c_int), ("update_grid", ctypes.c_int),
                ("t_done", ctypes.c_double), ("last_residual", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double * 5), ("kernel_bytes", ctypes.c_double * 5),
                ("launches", ctypes.c_longlong * 5)]


BLOCK_KERNEL_KINDS = ("spmm", "dots", "update", "expm", "lift")
RHS = 32


class Block:
    """Y = exp(t A) V for a block of 32 right-hand sides (block Arnoldi)."""

    def __init__(self, A: CSR, m: int, t: float, tol: float = 1e-8, substeps: int = 0, gs_mode: int = 2,
                 device: int = 0, timing: bool = False, autotune: bool = False):
        h = _vp()
        _check(lib().mek_block_create(A.h, int(m), float(t), float(tol), int(substeps), int(gs_mode),
                                      int(device), 1 if timing else 0, 1 if autotune else 0, ctypes.byref(h)))
        self.h = h
        self.n = A.info()["rows"]

    def expmv(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64)
        if V.shape != (self.n, RHS):
            raise ValueError(f"V must be {self.n} x {RHS}")
        Y = np.zeros_like(V)
        _check(lib().mek_block_expmv(self.h, _pd(V.ravel()), _pd(Y.ravel())))
        return Y

    def expmv_device(self, d_V: DeviceBuffer, d_Y: DeviceBuffer):
        _check(lib().mek_block_expmv_device(self.h, d_V.ptr, d_Y.ptr))

    def stats(self):
        s = BlockStats()
        _check(lib().mek_block_stats_get(self.h, ctypes.byref(s)))
        out = {f: getattr(s, f) for f, _ in BlockStats._fields_ if f not in ("kernel_ms", "kernel_bytes", "launches")}
        out["kernels"] = {k: {"ms": s.kernel_ms[i], "bytes": s.kernel_bytes[i], "launches": s.launches[i]}
                          for i, k in enumerate(BLOCK_KERNEL_KINDS)}
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().mek_block_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spmm(A: CSR, X, device: int = 0):
    """A X for a block of 32 columns (n x 32 row-major) on the device."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.zeros_like(X)
    _check(lib().mek_spmm(A.h, _pd(X.ravel()), _pd(Y.ravel()), int(device)))
    return Y


def gemm_dense(A, B, device: int = 0):
    """A B (N x N, N a multiple of 64) on the device FP64 MFMA GEMM."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    N = A.shape[0]
    C = np.zeros((N, N), np.float64, order="F")
    _check(lib().mek_gemm_dense(N, _pd(A.ravel(order="F")), _pd(B.ravel(order="F")),
                                C.ctypes.data_as(_dp), int(device)))
    return C


def expm_dense(A):
    """exp(A) of a small dense matrix on the device (one workgroup)."""
    A = np.asarray(A, dtype=np.float64)
    k = A.shape[0]
    d_a = DeviceBuffer.from_array(np.asfortranarray(A).ravel(order="F"))
    d_e = DeviceBuffer(8 * k * k)
    _check(lib().mek_expm_dense(k, d_a.ptr, d_e.ptr))
    return d_e.to_array().reshape(k, k, order="F")
