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
        for m in re.finditer(r"^\s*(?:int|int64_t|size_t|const char\*)\s+(tp_\w+)\s*\(", text, flags=re.M):
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
    header = open(os.path.join(ROOT, "include", "thick_ptycho_amd.h")).read()
    declared = int(re.search(r"#define TP_ABI_VERSION (\d+)", header).group(1))
    assert lib.tp_abi_version() == declared == _lib.ABI_VERSION
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


def test_new_entry_points_reject_bad_arguments():
    """Argument checks run before any device work (no GPU needed)."""
    lib = _lib.load()
    # far field: widths must be multiples of 64 up to 2048
    assert lib.tp_farfield(1, 100, 100, 1, 1, None) == _lib.TP_ERR_UNSUPPORTED
    assert b"multiples of 64" in lib.tp_last_error()
    assert lib.tp_amplitude_seed(1, 1, 4096, 1, 1.0, 1, 1, None) == _lib.TP_ERR_UNSUPPORTED
    # 3D factor: band half-width and sizes
    assert lib.tp_band_factor(1, 64, 64, 8, 1, 1, -1.0, 1.0, 0, 0, 0, 1, 8, 5, 0, 1, None, 1, 1, None) \
        == _lib.TP_ERR_UNSUPPORTED
    assert lib.tp_band_factor(1, 1, 64, 8, 1, 1, -1.0, 1.0, 0, 0, 0, 1, 8, 6, 0, 1, None, 1, 1, None) \
        == _lib.TP_ERR_INVALID
    assert lib.tp_band_sweep(None, 0, 1, None) == _lib.TP_ERR_INVALID
    assert lib.tp_band_conj_transpose(None, 8, 8, 2, 4, 1, None, None) == _lib.TP_ERR_INVALID
    assert lib.tp_band_sweep_ex(None, 1, 1, 3, None) == _lib.TP_ERR_INVALID
    assert lib.tp_plane_conj_dot(None, 0, 0, 1, None, 0, 0, 1, 1, 1, 8, None, None) == _lib.TP_ERR_INVALID
    assert lib.tp_band_sigma(None, 8, 8, 2, -1.0, 1.0, 0, 0, 0, 1, 2, None, None) == _lib.TP_ERR_INVALID
    # parallel in time: null operands, x aliasing y
    assert lib.tp_pint_matvec(None, 1, 0, 1, None, 2, 8, 4, 1, 0, None) == _lib.TP_ERR_INVALID
    assert lib.tp_pint_matvec(1, 1, 0, 2, None, 2, 8, 4, 1, 0, None) == _lib.TP_ERR_INVALID
    assert b"alias" in lib.tp_last_error()
    assert lib.tp_pint_matvec(1, 1, 31, 2, None, 3, 8, 4, 1, 0, None) == _lib.TP_ERR_INVALID  # c_pstride < L * N
    assert lib.tp_pint_dots(1, 0, 0, 33, 1, 0, 10, 1, 1, 1, None) == _lib.TP_ERR_INVALID
    assert lib.tp_pie_update(1, 4, 8, 8, 1, 1, 1.0, 1.0, 1.0, 0.1, 1e-12, None) == _lib.TP_ERR_INVALID  # ld < nz
    assert lib.tp_pie_adjoint_transmission(None, 8, 8, 8, 1.0, 1.0, 1.0, 1, None) == _lib.TP_ERR_INVALID
    assert lib.tp_window_grad_accumulate(1, 1, None, 2, 4, 8, 16, 1, None) == _lib.TP_ERR_INVALID
    assert lib.tp_window_grad_accumulate(1, 1, 1, 2, 4, 32, 16, 1, None) == _lib.TP_ERR_INVALID  # W > nx
    assert lib.tp_pint_combine(None, 0, 0, 2, None, 0, None, 0, 1, 0, 10, 1, None) == _lib.TP_ERR_INVALID
    # Krylov / CNN
    assert lib.tp_bicgstab_solve(1, 1, 1, 8, 8, 2, 1, 1, 3, None, 1e-6, 10, 0, 1, None, None) == _lib.TP_ERR_INVALID
    assert lib.tp_cnn_apply(1, 1, 1, 0, 8, 8, 1, None) == _lib.TP_ERR_INVALID
    assert lib.tp_cnn_param_count() == 9 * (16 + 256 + 256 + 16)
    assert lib.tp_band_coef_bytes(128, 128, 255, 6) == 255 * 128 * 14 * 128 * 8


def test_sweep_rejects_unsupported_chunk_mix():
    """A scan may take half the launch's chunk (16 or 32) only; anything else is
    refused on the host before any device work."""
    lib = _lib.load()
    scans = (_lib.TriScan * 2)()
    for s, (n, ch) in zip(scans, ((200, 4), (1000, 0))):
        s.n, s.steps, s.probes, s.chunk = n, 10, 1, ch
    rc = lib.tp_tri_sweep(scans, 2, 16, 4, None)
    assert rc == _lib.TP_ERR_INVALID
    assert b"half the launch's chunk" in lib.tp_last_error()
    scans[0].chunk = 8
    rc = lib.tp_tri_step(scans, 2, 16, 1, None)  # the plain form takes one chunk per launch
    assert rc == _lib.TP_ERR_INVALID
