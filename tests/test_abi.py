"""CPU checks of the C ABI: the built library loads and exports every symbol
declared in include/fvx.h (no GPU compute)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "finitevolx_amd", "libfvx_hip.so")


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"\b(fvx_[a-z0-9_]+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "fvx_advection2d_f64" in names and "fvx_qg_step" in names
    assert len(names) > 40


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_names_exist():
    from finitevolx_amd import _lib
    import finitevolx_amd.qg  # noqa: F401  (registers signatures)
    import finitevolx_amd.solvers  # noqa: F401
    lib = ctypes.CDLL(LIB)
    missing = [n for n in _lib.exported_symbols() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_error_path_without_gpu_compute():
    """Argument validation runs on the host and reports through fvx_last_error."""
    lib = ctypes.CDLL(LIB)
    lib.fvx_last_error.restype = ctypes.c_char_p
    rc = lib.fvx_stencil2d_f64(99, None, None, None, 10, 10, ctypes.c_double(1.0),
                               ctypes.c_double(1.0), None, None)
    assert rc == -1
    assert b"unknown stencil op" in lib.fvx_last_error()
