"""The C ABI library loads without a GPU and exports exactly what include/mek.h
declares; error reporting works through the boundary (host-only calls)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT, mek

HEADER = os.path.join(ROOT, "include", "mek.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mek_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_table():
    assert _declared() == sorted(mek.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = mek.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", mek.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (mek_\w+)", out))
    assert set(_declared()) <= exported


def test_errors_cross_the_boundary_as_status_codes(tmp_path):
    with pytest.raises(mek.MekError) as e:
        mek.CSR.load_bincsr(tmp_path / "none.bincsr")
    assert "status -2" in str(e.value) and "Could not open binary CSR file" in str(e.value)
    with pytest.raises(mek.MekError) as e:
        mek.CSR.from_arrays([0, 1, 3], [0, 5, 1], [1.0, 2.0, 3.0])  # column 5 out of range
    assert "status -1" in str(e.value) and "column index in range" in str(e.value)
    assert mek.lib().mek_version().decode().startswith("mek")


C_CALLER = r"""
#include <stdio.h>
#include <string.h>
#include "mek.h"

int main(int argc, char** argv) {
  mek_csr* A = NULL;
  mek_solver_options o;
  if (argc < 2) return 2;
  printf("%s\n", mek_version());
  if (mek_csr_load_bincsr(argv[1], &A) == MEK_OK) return 3;        /* the file does not exist */
  printf("error: %s\n", mek_last_error());
  if (mek_csr_load_bincsr(argv[2], &A) != MEK_OK) return 4;        /* a valid file */
  mek_solver_options_default(&o);
  printf("loaded; step_control=%d\n", o.step_control);
  mek_csr_destroy(A);
  return 0;
}
"""


def test_header_is_plain_c_and_links(tmp_path):
    """A C11 program includes include/mek.h, links libmek.so and calls the
    host-only entry points (no GPU needed), as a harness in another language
    would through its FFI."""
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(mek.LIB_PATH)
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", f"-I{os.path.dirname(HEADER)}",
                    str(src), "-o", str(exe), f"-L{libdir}", "-lmek", f"-Wl,-rpath,{libdir}"], check=True)
    from conftest import gen
    rp, ci, va = gen.laplacian_2d(5)
    gen.write_bincsr(tmp_path / "A.bincsr", rp, ci, va)
    p = subprocess.run([str(exe), str(tmp_path / "missing.bincsr"), str(tmp_path / "A.bincsr")],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, (p.returncode, p.stdout, p.stderr)
    lines = p.stdout.splitlines()
    assert lines[0].startswith("mek")
    assert "Could not open binary CSR file" in lines[1]
    # the loader prints the reference's header line, then the program its own
    assert lines[2].startswith("Binary CSR header: version=1")
    assert lines[-1].startswith("loaded; step_control=")


def test_cli_binary_is_built():
    assert os.access(mek.CLI_PATH, os.X_OK)


def _tiny():
    from conftest import gen
    return mek.CSR.from_arrays(*gen.laplacian_2d(3))


@pytest.mark.parametrize("t", [-1.0, float("nan"), float("inf")])
def test_negative_or_nonfinite_t_refused_before_any_device_call(t):
    """t < 0 or not finite is refused by the ABI with a status code (the
    reference's controller would never run and leave y unset); these checks
    come before any HIP call, so they run without a GPU."""
    with pytest.raises(mek.MekError) as e:
        mek.Runner(_tiny(), m=3, t=t)
    assert "status -1" in str(e.value) and "t finite and >= 0" in str(e.value)
    with pytest.raises(mek.MekError) as e:
        mek.Batch([_tiny()], [t], [3], [1e-8])
    assert "t finite and >= 0" in str(e.value)


def test_value_storage_mode_checked():
    with pytest.raises(ValueError):
        mek.Runner(_tiny(), m=3, t=1.0, values_f32="sometimes")
    o = mek.SolverOptions()
    mek.lib().mek_solver_options_default(o)
    o.values_f32 = 3
    p = mek.ArnoldiParams(3, 1.0, 1e-8, 10, 1)
    h = mek._vp()
    A = _tiny()
    import ctypes
    st = mek.lib().mek_runner_create(A.h, ctypes.byref(p), ctypes.byref(o), 0, ctypes.byref(h))
    assert st == -1 and b"values_f32" in mek.lib().mek_last_error()
