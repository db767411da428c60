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
    assert "status -4" in str(e.value) and "column index in range" in str(e.value)
    assert mek.lib().mek_version().decode().startswith("mek")


def test_cli_binary_is_built():
    assert os.access(mek.CLI_PATH, os.X_OK)
