"""ctypes binding of ``libtp_hip.so`` (C ABI in ``include/thick_ptycho_amd.h``).

The library is built in-tree by ``python -m thick_ptycho_amd.build``.  There is
no CPU fallback: if the library or a GPU is missing every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtp_hip.so")

TP_OK = 0
TP_ERR_INVALID = -1
TP_ERR_HIP = -2
TP_ERR_SINGULAR = -3
TP_ERR_UNSUPPORTED = -4
TP_MAX_SCANS = 8

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t


class TriFactorJob(ctypes.Structure):
    _fields_ = [
        ("C", c_void_p),
        ("base", c_void_p),
        ("coef", c_void_p),
        ("status", c_void_p),
        ("lhs_sign", c_double),
        ("rhs_sign", c_double),
        ("n", c_int),
        ("ncol", c_int),
        ("steps", c_int),
        ("conj_c", c_int),
        ("lhs_col0", c_int),
        ("rhs_col0", c_int),
        ("col_step", c_int),
        ("with_sigma", c_int),
    ]


class TriScan(ctypes.Structure):
    _fields_ = [
        ("coef", c_void_p),
        ("u0", c_void_p),
        ("rhs", c_void_p),
        ("out", c_void_p),
        ("ckpt", c_void_p),
        ("exitp", c_void_p),
        ("u0_pstride", c_int64),
        ("rhs_pstride", c_int64),
        ("rhs_sstride", c_int64),
        ("out_pstride", c_int64),
        ("out_zstride", c_int64),
        ("ckpt_pstride", c_int64),
        ("ckpt_zstride", c_int64),
        ("exit_pstride", c_int64),
        ("n", c_int),
        ("steps", c_int),
        ("probes", c_int),
        ("out_z0", c_int),
        ("out_zdir", c_int),
        ("ckpt_every", c_int),
        ("sigma", c_int),
        ("rhs_steps", c_int),
        ("gfwd", c_void_p),
        ("grad", c_void_p),
        ("gfwd_pstride", c_int64),
        ("gfwd_zstride", c_int64),
        ("grad_zstride", c_int64),
        ("gfwd_z0", c_int),
        ("gfwd_zdir", c_int),
        ("grad_z0", c_int),
        ("grad_zdir", c_int),
        ("grad_skip_row0", c_int),
        ("coef_pstride", c_int64),
        ("chunk", c_int),
        ("reach", c_void_p),
        ("offd", ctypes.c_float * 12),
        ("has_offd", c_int),
    ]


class MsScan(ctypes.Structure):
    _fields_ = [
        ("T", c_void_p),
        ("H", c_void_p),
        ("u0", c_void_p),
        ("out", c_void_p),
        ("exitp", c_void_p),
        ("u0_pstride", c_int64),
        ("out_pstride", c_int64),
        ("out_zstride", c_int64),
        ("exit_pstride", c_int64),
        ("n", c_int),
        ("steps", c_int),
        ("probes", c_int),
        ("conj_h", c_int),
    ]


class BandScan(ctypes.Structure):
    _fields_ = [
        ("W", c_void_p),
        ("cpl", c_void_p),
        ("sigma", c_void_p),
        ("u0", c_void_p),
        ("out", c_void_p),
        ("exitp", c_void_p),
        ("ckpt", c_void_p),
        ("rhs", c_void_p),
        ("work", c_void_p),
        ("u0_pstride", c_int64),
        ("out_pstride", c_int64),
        ("out_zstride", c_int64),
        ("exit_pstride", c_int64),
        ("ckpt_pstride", c_int64),
        ("ckpt_zstride", c_int64),
        ("rhs_pstride", c_int64),
        ("rhs_sstride", c_int64),
        ("work_pstride", c_int64),
        ("nx", c_int),
        ("ny", c_int),
        ("w", c_int),
        ("steps", c_int),
        ("probes", c_int),
        ("out_z0", c_int),
        ("out_zdir", c_int),
        ("ckpt_every", c_int),
        ("rhs_steps", c_int),
        ("reserved", c_int),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "tp_abi_version": (c_int, []),
    "tp_last_error": (ctypes.c_char_p, []),
    "tp_tri_coef_bytes": (c_size_t, [c_int, c_int, c_int]),
    "tp_tri_factor": (
        c_int,
        [c_void_p, c_int, c_int, c_void_p, c_double, c_double, c_int, c_int, c_int, c_int, c_int, c_int,
         c_int, c_void_p, c_void_p, c_void_p],
    ),
    "tp_tri_sweep": (c_int, [ctypes.POINTER(TriScan), c_int, c_int, c_int, c_void_p]),
    "tp_tri_step": (c_int, [ctypes.POINTER(TriScan), c_int, c_int, c_int, c_void_p]),
    "tp_tri_coef_bytes_wide": (c_size_t, [c_int, c_int, c_int, c_int]),
    "tp_tri_factor_wide": (
        c_int,
        [c_void_p, c_int, c_int, c_void_p, c_double, c_double, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
         c_int, c_void_p, c_void_p, c_void_p],
    ),
    "tp_tri_sweep_wide": (c_int, [ctypes.POINTER(TriScan), c_int, c_int, c_int, c_void_p]),
    "tp_ms_transmission": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_double, c_void_p, c_void_p]),
    "tp_ms_kernel": (c_int, [c_int, c_double, c_double, c_double, c_void_p, c_void_p]),
    "tp_ms_sweep": (c_int, [ctypes.POINTER(MsScan), c_int, c_void_p]),
    "tp_ms_step": (c_int, [ctypes.POINTER(MsScan), c_int, c_int, c_void_p]),
    "tp_amplitude_seed": (c_int, [c_void_p, c_void_p, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p]),
    "tp_farfield": (c_int, [c_void_p, c_int64, c_int, c_int, c_void_p, c_void_p]),
    "tp_spectral_len": (c_int, [c_int]),
    "tp_spectral_work_bytes": (c_size_t, [c_int, c_int]),
    "tp_spectral_setup_bytes": (c_size_t, [c_int]),
    "tp_ms_kernel_any": (c_int, [c_int, c_double, c_double, c_double, c_void_p, c_void_p, c_void_p]),
    "tp_chirp_tables": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tp_ms_sweep_any": (c_int, [ctypes.POINTER(MsScan), c_int, c_void_p, c_void_p]),
    "tp_farfield_any": (c_int, [c_void_p, c_int64, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tp_amplitude_seed_any": (c_int, [c_void_p, c_void_p, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p]),
    "tp_dft_any": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_int64, c_int64, c_int, c_int, c_int, c_int,
                           c_void_p, c_void_p, c_void_p, c_void_p]),
    "tp_twiddle_mul": (c_int, [c_void_p, c_int, c_int, c_int, c_int64, c_int, c_void_p]),
    "tp_object_polar": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "tp_tri_residual": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int, c_int, c_int, c_void_p,
                                c_int, c_void_p, c_double, c_int, c_int, c_int, c_int, c_void_p, c_int64, c_int64,
                                c_void_p]),
    "tp_amplitude_seed_refined": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int, c_int,
                                          c_double, c_void_p, c_void_p, c_void_p]),
    "tp_band_residual": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int, c_int, c_int, c_int,
                                 c_void_p, c_int64, c_int64, c_int, c_void_p, c_void_p, c_double, c_int, c_int, c_int,
                                 c_int, c_void_p, c_int64, c_int64, c_void_p]),
    "tp_amplitude_seed2d_refined_work_bytes": (c_size_t, [c_int, c_int, c_int]),
    "tp_amplitude_seed2d_refined": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int, c_int,
                                            c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tp_object_join": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "tp_tv_denoise": (c_int, [c_void_p, c_int, c_int, c_double, c_int, c_double, c_void_p, c_void_p]),
    "tp_tv_denoise_noalloc": (c_int, [c_void_p, c_int, c_int, c_double, c_int, c_double, c_void_p]),
    "tp_lowpass_weight": (c_int, [c_void_p, c_int, c_int, c_double, c_void_p]),
    "tp_band_coef_bytes": (c_int64, [c_int, c_int, c_int, c_int]),
    "tp_band_factor": (
        c_int,
        [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_double, c_double, c_int, c_int, c_int, c_int, c_int,
         c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tp_tri_factor_batch": (c_int, [ctypes.POINTER(TriFactorJob), c_int, c_int, c_void_p]),
    "tp_band_sweep": (c_int, [ctypes.POINTER(BandScan), c_int, c_int, c_void_p]),
    "tp_band_sweep_ex": (c_int, [ctypes.POINTER(BandScan), c_int, c_int, c_int, c_void_p]),
    "tp_plane_conj_dot": (c_int, [c_void_p, ctypes.c_int64, c_int, c_int, c_void_p, ctypes.c_int64, c_int, c_int,
                                  c_int, c_int, ctypes.c_int64, c_void_p, c_void_p]),
    "tp_band_conj_transpose": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "tp_band_sigma": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, c_int, c_int, c_int, c_int, c_int,
                              c_void_p, c_void_p]),
    "tp_cnn_param_count": (c_int, []),
    "tp_cnn_weight_bytes": (c_int64, []),
    "tp_cnn_workspace_bytes": (c_int64, [c_int, c_int, c_int, c_int]),
    "tp_cnn_expand": (c_int, [c_void_p, c_void_p, c_void_p]),
    "tp_cnn_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "tp_cnn_train_step": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
         c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tp_krylov_diag": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_double, c_int, c_int, c_void_p, c_void_p]),
    "tp_bicgstab_workspace_bytes": (c_int64, [c_int, c_int, c_int, c_int, c_int]),
    "tp_bicgstab_solve": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_double, c_int,
         c_int, c_void_p, c_void_p, c_void_p],
    ),
    "tp_cn_finish": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "tp_ms2d_kernel": (c_int, [c_int, c_int, c_double, c_double, c_double, c_double, c_void_p, c_void_p]),
    "tp_ms2d_sweep": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p],
    ),
    "tp_farfield2d": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "tp_amplitude_seed2d": (
        c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tp_grad3d_accumulate": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int64, c_void_p, c_void_p]),
    "tp_pint_matvec": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                               c_void_p]),
    "tp_window_grad_accumulate": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                          c_void_p]),
    "tp_pie_adjoint_transmission": (c_int, [c_void_p, c_int64, c_int, c_int, c_double, c_double, c_double, c_void_p,
                                            c_void_p]),
    "tp_pie_update": (c_int, [c_void_p, c_int64, c_int, c_int, c_void_p, c_void_p, c_double, c_double, c_double,
                              c_double, c_double, c_void_p]),
    "tp_pint_dots": (c_int, [c_void_p, c_int64, c_int64, c_int, c_void_p, c_int64, c_int64, c_int, c_int, c_void_p,
                             c_void_p]),
    "tp_pint_combine": (c_int, [c_void_p, c_int64, c_int64, c_int, c_void_p, c_int, c_void_p, c_int64, c_void_p,
                                c_int64, c_int64, c_int, c_void_p]),
}

_lib: Optional[ctypes.CDLL] = None


class LibraryMissing(RuntimeError):
    pass


ABI_VERSION = 4  # TP_ABI_VERSION of include/thick_ptycho_amd.h


def load() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} not found: build it with `python -m thick_ptycho_amd.build` (no CPU fallback exists)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.tp_abi_version() != ABI_VERSION:
        raise LibraryMissing(
            f"{LIB_PATH} has ABI version {lib.tp_abi_version()}, these bindings expect {ABI_VERSION}: rebuild it"
        )
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    """Map a status code to the reference's exception types."""
    if rc == TP_OK:
        return
    msg = f"{what}: {load().tp_last_error().decode(errors='replace')}"
    if rc in (TP_ERR_INVALID, TP_ERR_UNSUPPORTED):
        raise ValueError(msg)
    raise RuntimeError(msg)


def ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def stream_handle(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def require_gpu() -> None:
    import torch

    load()
    if not torch.cuda.is_available():
        raise RuntimeError("thick_ptycho_amd needs a ROCm GPU (torch.cuda.is_available() is False); no CPU fallback")
