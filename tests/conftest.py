"""Shared fixtures: the package (loaded by path), the CPU oracle and markers."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "matrix-exp-krylov_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


mek = _load("mek", os.path.join(PKG_DIR, "__init__.py"))
gen = _load("mek_generate", os.path.join(PKG_DIR, "generate.py"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libmek.so)")
    config.addinivalue_line("markers", "slow: large CPU work")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """libmek.so with a visible GPU; GPU tests fail loudly if the library is missing."""
    lib = mek.lib()
    n = mek.device_count()
    if n < 1:
        pytest.fail("no GPU visible to libmek.so")
    return mek
