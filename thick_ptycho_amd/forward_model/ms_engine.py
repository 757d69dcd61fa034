"""Device-side plumbing for the multislice and far-field kernels."""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from .. import _lib


def transmission(n: torch.Tensor, k: float, dz: float, n_medium: float, stream=None) -> torch.Tensor:
    """T[z] = exp(i k (n[:, z] - n_medium) dz), z < nz - 1; n complex128 (nx, nz) on the GPU."""
    assert n.is_cuda and n.dtype == torch.complex128 and n.dim() == 2
    n = n.contiguous()
    nx, nz = n.shape
    T = torch.empty((nz - 1, nx), dtype=torch.complex64, device=n.device)
    _lib.check(
        _lib.load().tp_ms_transmission(_lib.ptr(n), nx, nz, float(k), float(dz), float(n_medium), _lib.ptr(T),
                                       _lib.stream_handle(stream)),
        "tp_ms_transmission",
    )
    return T


def propagator(nx: int, dx: float, k: float, dz: float, device, stream=None) -> torch.Tensor:
    """H(kx) / nx in numpy.fft.fftfreq order (global phase removed, evanescent band clipped)."""
    H = torch.empty(nx, dtype=torch.complex64, device=device)
    _lib.check(_lib.load().tp_ms_kernel(nx, float(dx), float(k), float(dz), _lib.ptr(H), _lib.stream_handle(stream)),
               "tp_ms_kernel")
    return H


@dataclass
class MsScanSpec:
    T: torch.Tensor
    H: torch.Tensor
    u0: torch.Tensor
    out: Optional[torch.Tensor] = None
    exit: Optional[torch.Tensor] = None
    conj_h: bool = False

    def struct(self) -> _lib.MsScan:
        P, n = self.u0.shape
        L = self.T.shape[0]
        s = _lib.MsScan()
        assert self.T.dtype == torch.complex64 and self.T.shape[1] == n and self.T.is_contiguous()
        assert self.H.dtype == torch.complex64 and self.H.shape == (n,)
        assert self.u0.dtype == torch.complex64 and self.u0.stride(-1) == 1
        s.T = _lib.ptr(self.T)
        s.H = _lib.ptr(self.H)
        s.u0 = _lib.ptr(self.u0)
        s.u0_pstride = self.u0.stride(0) if P > 1 else n
        if self.out is not None:
            assert self.out.shape == (P, L + 1, n) and self.out.stride(-1) == 1
            s.out = _lib.ptr(self.out)
            s.out_pstride = self.out.stride(0)
            s.out_zstride = self.out.stride(1)
        if self.exit is not None:
            assert self.exit.shape == (P, n) and self.exit.stride(-1) == 1
            s.exitp = _lib.ptr(self.exit)
            s.exit_pstride = self.exit.stride(0)
        s.n = n
        s.steps = L
        s.probes = P
        s.conj_h = int(self.conj_h)
        return s


def _structs(specs: Sequence[MsScanSpec]):
    if not 1 <= len(specs) <= _lib.TP_MAX_SCANS:
        raise ValueError(f"between 1 and {_lib.TP_MAX_SCANS} scans per launch")
    return (_lib.MsScan * len(specs))(*[s.struct() for s in specs])


def sweep(specs: Sequence[MsScanSpec], stream=None) -> None:
    arr = _structs(specs)
    _lib.check(_lib.load().tp_ms_sweep(arr, len(specs), _lib.stream_handle(stream)), "tp_ms_sweep")


def sweep_plain(specs: Sequence[MsScanSpec], stream=None) -> None:
    arr = _structs(specs)
    lib = _lib.load()
    sh = _lib.stream_handle(stream)
    L = specs[0].T.shape[0]
    for z in range(L):
        _lib.check(lib.tp_ms_step(arr, len(specs), z, sh), "tp_ms_step")


def farfield(exit_w: torch.Tensor, stream=None) -> torch.Tensor:
    """|fft(exit)|^2 per row, float32 (P, n)."""
    assert exit_w.is_cuda and exit_w.dtype == torch.complex64 and exit_w.dim() == 2 and exit_w.stride(-1) == 1
    P, n = exit_w.shape
    out = torch.empty((P, n), dtype=torch.float32, device=exit_w.device)
    _lib.check(
        _lib.load().tp_farfield(_lib.ptr(exit_w), exit_w.stride(0) if P > 1 else n, n, P, _lib.ptr(out),
                                _lib.stream_handle(stream)),
        "tp_farfield",
    )
    return out


# ---------------------------------------------------------------- two transverse axes
def propagator_2d(nx: int, ny: int, dx: float, dy: float, k: float, dz: float, device, stream=None) -> torch.Tensor:
    """H(kx, ky) / (nx ny) in numpy.fft.fftfreq order along both axes, stored transposed:
    complex64 (ny, nx), element [iy, ix] = H(kx[ix], ky[iy]) (the sweep's read order)."""
    H = torch.empty((ny, nx), dtype=torch.complex64, device=device)
    _lib.check(_lib.load().tp_ms2d_kernel(nx, ny, float(dx), float(dy), float(k), float(dz), _lib.ptr(H),
                                          _lib.stream_handle(stream)), "tp_ms2d_kernel")
    return H


def multislice_2d(n: torch.Tensor, probes: torch.Tensor, dx: float, dy: float, k: float, dz: float,
                  n_medium: float = 1.0, keep: str = "exit", conj_h: bool = False, H: Optional[torch.Tensor] = None,
                  stream=None) -> torch.Tensor:
    """Multislice with a 2D FFT per slice (the reference's rule with both transverse
    frequencies; it has no two-axis form).  n: complex128 (nx, ny, nz) on the GPU;
    probes: complex64 (P, nx, ny) or (P, nx*ny).  Returns (P, nx*ny) exit planes or
    (P, nz, nx*ny) for keep='all'."""
    assert n.is_cuda and n.dtype == torch.complex128 and n.dim() == 3
    nx, ny, nz = n.shape
    N = nx * ny
    T = transmission(n.reshape(N, nz), k, dz, n_medium, stream)
    if H is None:
        H = propagator_2d(nx, ny, dx, dy, k, dz, n.device, stream)
    u0 = probes.reshape(-1, N).to(torch.complex64).contiguous()
    P = u0.shape[0]
    work = torch.empty((3, P, N), dtype=torch.complex64, device=n.device)
    out = torch.empty((P, nz, N), dtype=torch.complex64, device=n.device) if keep == "all" else None
    ex = torch.empty((P, N), dtype=torch.complex64, device=n.device) if keep != "all" else None
    _lib.check(
        _lib.load().tp_ms2d_sweep(_lib.ptr(T), _lib.ptr(H), _lib.ptr(u0), _lib.ptr(out), _lib.ptr(ex), nx, ny,
                                  nz - 1, P, int(conj_h), _lib.ptr(work), _lib.stream_handle(stream)),
        "tp_ms2d_sweep",
    )
    return out if keep == "all" else ex


def farfield_2d(exit_w: torch.Tensor, nx: int, ny: int, stream=None) -> torch.Tensor:
    """|fft2(exit)|^2 per probe, float32 (P, nx, ny); exit complex64 (P, nx*ny) or (P, nx, ny)."""
    ex = exit_w.reshape(-1, nx * ny).to(torch.complex64).contiguous()
    P = ex.shape[0]
    out = torch.empty((P, nx, ny), dtype=torch.float32, device=ex.device)
    work = torch.empty((P, nx * ny), dtype=torch.complex64, device=ex.device)
    _lib.check(_lib.load().tp_farfield2d(_lib.ptr(ex), nx, ny, P, _lib.ptr(out), _lib.ptr(work),
                                         _lib.stream_handle(stream)), "tp_farfield2d")
    return out
