"""The C-ABI library loads and exports every entry point include/sfx.h declares."""
import ctypes
import os
import re

from sfincs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "sfx.h")).read()
    return sorted(set(re.findall(r"\b(sfx_[a-z0-9_]+)\s*\(", txt)))


def test_all_declared_symbols_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_are_reported_without_gpu():
    L = _lib.lib()
    out = ctypes.c_void_p()
    p = _lib.default_params()
    rc = L.sfx_model_create_regular(ctypes.byref(p), 0, 4, 0.0, 0.0, 10.0, 10.0, 0.0, None, ctypes.byref(out))
    assert rc == -1
    assert "invalid" in _lib.last_error()
