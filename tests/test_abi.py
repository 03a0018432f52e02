"""CPU checks of the C-ABI library: it builds, loads, exports every declared symbol, and
rejects bad parameters with a status and a message (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    from sdc_amd import build, rbc
    if not os.path.exists(rbc.LIB_PATH):
        build.build()
    return rbc.load_library()


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            src = open(os.path.join(inc, fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(rbc_[A-Za-z_0-9]+)\s*\(", src))
    return names


def test_exports_every_declared_symbol(lib):
    decl = declared_symbols()
    assert len(decl) >= 15
    for name in sorted(decl):
        assert hasattr(lib, name), name


def test_python_binding_covers_exports(lib):
    from sdc_amd import rbc
    assert declared_symbols() <= set(rbc.EXPORTS) | {"rbc_create"}


def test_create_rejects_bad_params(lib):
    from sdc_amd import rbc
    bad = [rbc.rbc_params(30, 32, 8, 4.0, 4.0, 1e3, 1.0, 1e-3, 0, 0, 0),   # M not a power of two
           rbc.rbc_params(32, 32, 0, 4.0, 4.0, 1e3, 1.0, 1e-3, 0, 0, 0),   # L = 0
           rbc.rbc_params(32, 32, 8, -1.0, 4.0, 1e3, 1.0, 1e-3, 0, 0, 0),  # Lx < 0
           rbc.rbc_params(32, 32, 8, 4.0, 4.0, 1e3, 1.0, 1e-3, 7, 0, 0)]   # precision
    for p in bad:
        h = ctypes.c_void_p()
        rc = lib.rbc_create(ctypes.byref(p), ctypes.byref(h))
        assert rc == -1
        assert not h.value
        assert len(lib.rbc_last_error(None)) > 0


def test_null_context_calls_fail_cleanly(lib):
    assert lib.rbc_step(None, 1) == -1
    assert lib.rbc_synchronize(None) == -1
    assert lib.rbc_reset_history(None) == -1


def test_stage_names(lib):
    from sdc_amd import rbc
    names = [lib.rbc_stage_name(i).decode() for i in range(rbc.NSTAGES)]
    assert names[0] == "eval" and "rows_products" in names and lib.rbc_stage_name(99) == b""
