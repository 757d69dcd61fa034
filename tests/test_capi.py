"""The C-ABI library loads and exports every symbol declared in include/*.h (CPU, no compute)."""

import ctypes
import os
import re

import pytest

from thick_ptycho_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|size_t|const char\*)\s+(tp_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libtp_hip.so is not built; run python -m thick_ptycho_amd.build")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_functions()
    assert declared, "no declarations parsed"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    # the Python binding covers every declared entry point
    assert declared <= set(_lib.SIGNATURES), sorted(declared - set(_lib.SIGNATURES))


def test_abi_version_and_error_path():
    lib = _lib.load()
    assert lib.tp_abi_version() == 1
    # an invalid call fails on the host before touching the GPU
    rc = lib.tp_tri_factor(None, 10, 5, None, -1.0, 1.0, 0, 0, 0, 1, 5, 3, 0, None, None, None)
    assert rc == _lib.TP_ERR_INVALID
    assert b"chunk" in lib.tp_last_error()
    rc = lib.tp_farfield(None, 100, 100, 1, None, None)
    assert rc == _lib.TP_ERR_UNSUPPORTED
    with pytest.raises(ValueError):
        _lib.check(rc, "tp_farfield")


def test_coef_bytes():
    lib = _lib.load()
    assert lib.tp_tri_coef_bytes(1000, 7, 16) == 7 * 6 * 1024 * 8
