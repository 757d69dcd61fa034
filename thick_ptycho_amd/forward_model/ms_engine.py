"""Device-side plumbing for the multislice and far-field kernels.

Widths n = 64 r (r <= 32) run the one-wavefront kernels (tp_ms_sweep,
tp_farfield, tp_amplitude_seed); every other width up to 16384 runs the
any-length forms (tp_ms_sweep_any, tp_farfield_any, tp_amplitude_seed_any), a
circular convolution on a padded power-of-two grid: there the propagator is
the convolution spectrum S of tp_ms_kernel_any in place of H / n, and the
DFTs take Bluestein chirp tables, built once per width and device.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from .. import _lib


def native(n: int) -> bool:
    """True where the one-wavefront spectral kernels apply (n = 64 r, r <= 32)."""
    return 64 <= n <= 2048 and n % 64 == 0


def spectral_len(n: int) -> int:
    """Padded convolution grid of the any-length forms (a power of two >= 2n - 1)."""
    F = _lib.load().tp_spectral_len(int(n))
    if F < 0:
        raise ValueError(f"spectral length {n} outside [1, 16384]")
    return F


def propagator_len(n: int) -> int:
    """Length of the propagator buffer the sweep takes at width n."""
    return n if native(n) else spectral_len(n)


def _setup_scratch(n: int, device) -> torch.Tensor:
    nbytes = int(_lib.load().tp_spectral_setup_bytes(int(n)))
    return torch.empty(max(nbytes, 16), dtype=torch.uint8, device=device)


def _work(n: int, probes: int, device) -> Optional[torch.Tensor]:
    nbytes = int(_lib.load().tp_spectral_work_bytes(int(n), int(probes)))
    return torch.empty(nbytes, dtype=torch.uint8, device=device) if nbytes else None


_CHIRP_CACHE: dict = {}


def chirp_tables(n: int, device, stream=None):
    """(chirp complex64 (n,), spectrum complex64 (F,)) for length-n DFTs, cached
    per (n, device) (at most 16 entries)."""
    key = (int(n), str(device))
    hit = _CHIRP_CACHE.get(key)
    if hit is None:
        F = spectral_len(n)
        chirp = torch.empty(n, dtype=torch.complex64, device=device)
        S = torch.empty(F, dtype=torch.complex64, device=device)
        setup = _setup_scratch(n, device)
        _lib.check(_lib.load().tp_chirp_tables(int(n), _lib.ptr(chirp), _lib.ptr(S), _lib.ptr(setup),
                                               _lib.stream_handle(stream)), "tp_chirp_tables")
        if len(_CHIRP_CACHE) >= 16:
            _CHIRP_CACHE.pop(next(iter(_CHIRP_CACHE)))
        hit = _CHIRP_CACHE[key] = (chirp, S, setup)
    return hit[0], hit[1]


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
    """H(kx) / nx in numpy.fft.fftfreq order (global phase removed, evanescent band
    clipped); at widths the one-wavefront kernels do not take, the convolution
    spectrum of the same propagator on spectral_len(nx) points."""
    if not native(nx):
        S = torch.empty(spectral_len(nx), dtype=torch.complex64, device=device)
        setup = _setup_scratch(nx, device)
        _lib.check(_lib.load().tp_ms_kernel_any(nx, float(dx), float(k), float(dz), _lib.ptr(S), _lib.ptr(setup),
                                                _lib.stream_handle(stream)), "tp_ms_kernel_any")
        S._tp_keep = setup  # the set-up scratch lives until the kernels that read it have run
        return S
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
        assert self.H.dtype == torch.complex64 and self.H.shape == (propagator_len(n),)
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
    n = specs[0].u0.shape[-1]
    if native(n):
        _lib.check(_lib.load().tp_ms_sweep(arr, len(specs), _lib.stream_handle(stream)), "tp_ms_sweep")
        return
    work = _work(n, sum(int(s.u0.shape[0]) for s in specs), specs[0].u0.device)
    _lib.check(_lib.load().tp_ms_sweep_any(arr, len(specs), _lib.ptr(work), _lib.stream_handle(stream)),
               "tp_ms_sweep_any")


def sweep_plain(specs: Sequence[MsScanSpec], stream=None) -> None:
    if not native(specs[0].u0.shape[-1]):
        raise ValueError("the plain one-launch-per-slice multislice form is built for n = 64 r, r <= 32; other "
                         "widths step in the persistent sweep")
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
    ld = exit_w.stride(0) if P > 1 else n
    if native(n):
        _lib.check(_lib.load().tp_farfield(_lib.ptr(exit_w), ld, n, P, _lib.ptr(out), _lib.stream_handle(stream)),
                   "tp_farfield")
        return out
    chirp, S = chirp_tables(n, exit_w.device, stream)
    work = _work(n, P, exit_w.device)
    _lib.check(_lib.load().tp_farfield_any(_lib.ptr(exit_w), ld, n, P, _lib.ptr(chirp), _lib.ptr(S), _lib.ptr(out),
                                           _lib.ptr(work), _lib.stream_handle(stream)), "tp_farfield_any")
    return out


def amplitude_seed(exit_w: torch.Tensor, sqrt_i: torch.Tensor, scale: float = 1.0, stream=None):
    """(seed, loss per probe): seed = scale * dJ/d conj(exit) of J = sum (|fft(exit)| - sqrt_i)^2
    (tp_amplitude_seed; tp_amplitude_seed_any at other widths)."""
    assert exit_w.is_cuda and exit_w.dtype == torch.complex64 and exit_w.dim() == 2 and exit_w.is_contiguous()
    assert sqrt_i.dtype == torch.float32 and sqrt_i.is_contiguous() and sqrt_i.shape == exit_w.shape
    P, n = exit_w.shape
    seed = torch.empty_like(exit_w)
    loss = torch.empty(P, dtype=torch.float32, device=exit_w.device)
    sh = _lib.stream_handle(stream)
    if native(n):
        _lib.check(_lib.load().tp_amplitude_seed(_lib.ptr(exit_w), _lib.ptr(sqrt_i), n, P, float(scale),
                                                 _lib.ptr(seed), _lib.ptr(loss), sh), "tp_amplitude_seed")
        return seed, loss
    chirp, S = chirp_tables(n, exit_w.device, stream)
    work = _work(n, P, exit_w.device)
    _lib.check(_lib.load().tp_amplitude_seed_any(_lib.ptr(exit_w), _lib.ptr(sqrt_i), n, P, float(scale),
                                                 _lib.ptr(chirp), _lib.ptr(S), _lib.ptr(seed), _lib.ptr(loss),
                                                 _lib.ptr(work), sh), "tp_amplitude_seed_any")
    return seed, loss


# ---------------------------------------------------------------- DFTs of any length
def _dft(inp, in_stride, in_dist, out, out_stride, out_dist, n, count, inverse=False, power=False, stream=None):
    """One tp_dft_any launch (n <= 16384)."""
    chirp, S = chirp_tables(n, inp.device, stream)
    work = _work(n, count, inp.device)
    _lib.check(_lib.load().tp_dft_any(_lib.ptr(inp), int(in_stride), int(in_dist), _lib.ptr(out), int(out_stride),
                                      int(out_dist), int(n), int(count), int(inverse), int(power), _lib.ptr(chirp),
                                      _lib.ptr(S), _lib.ptr(work), _lib.stream_handle(stream)), "tp_dft_any")


def _split(n: int):
    """n = A * B with A, B <= 16384 and A closest to sqrt(n) (four-step DFT)."""
    best = None
    for a in range(1, int(n ** 0.5) + 1):
        if n % a == 0 and n // a <= 16384:
            best = a
    if best is None:
        raise ValueError(f"DFT length {n} above 16384 needs a factorisation into two lengths <= 16384")
    return best, n // best


def dft_rows(x: torch.Tensor, inverse: bool = False, power: bool = False, stream=None) -> torch.Tensor:
    """DFT of every row of a contiguous complex64 (R, n) tensor on the GPU
    (numpy.fft.fft convention; inverse unnormalised).  power=True returns |X|^2
    as float32.  Rows above 16384 points take a four-step split n = A B."""
    assert x.is_cuda and x.dtype == torch.complex64 and x.dim() == 2 and x.is_contiguous()
    R, n = x.shape
    out = torch.empty((R, n), dtype=torch.float32 if power else torch.complex64, device=x.device)
    if n <= 16384:
        _dft(x, 1, n, out, 1, n, n, R, inverse, power, stream)
        return out
    A, B = _split(n)
    tmp = torch.empty((A, B), dtype=torch.complex64, device=x.device)
    sign = 1 if inverse else -1
    for r in range(R):
        _dft(x[r], B, 1, tmp, B, 1, A, B, inverse, False, stream)  # over j1 of x[j2 + B j1]
        _lib.check(_lib.load().tp_twiddle_mul(_lib.ptr(tmp), A, B, 1, A * B, sign, _lib.stream_handle(stream)),
                   "tp_twiddle_mul")
        _dft(tmp, 1, B, out[r], A, 1, B, A, inverse, power, stream)  # X[k1 + A k2]
    return out


def farfield_scans_fft2(exit_w: torch.Tensor, stream=None) -> torch.Tensor:
    """|fft2|^2 over both axes of the (scans, N) exit-wave array, float32: the
    reference's far field for two transverse axes, which applies numpy's fft2 to
    the flattened (total_scans, block_size) array, i.e. across the scan axis
    (base/base_solver.py:237-242)."""
    assert exit_w.dim() == 2
    ex = exit_w.to(torch.complex64).contiguous()
    S, N = ex.shape
    rows = dft_rows(ex, stream=stream)
    out = torch.empty((S, N), dtype=torch.float32, device=ex.device)
    _dft(rows, N, 1, out, N, 1, S, N, False, True, stream)  # along the scan axis, |X|^2
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
    """|fft2(exit)|^2 per probe, float32 (P, nx, ny); exit complex64 (P, nx*ny) or (P, nx, ny).

    Planes of powers of two from 64 to 512 take tp_farfield2d, others the
    any-length DFT along y then x."""
    ex = exit_w.reshape(-1, nx * ny).to(torch.complex64).contiguous()
    P = ex.shape[0]
    out = torch.empty((P, nx, ny), dtype=torch.float32, device=ex.device)
    m2 = (64, 128, 256, 512)
    if nx not in m2 or ny not in m2:
        rows = dft_rows(ex.reshape(P * nx, ny), stream=stream).reshape(P, nx * ny)
        for p in range(P):  # along x: stride ny, one sequence per ky
            _dft(rows[p], ny, 1, out[p], ny, 1, nx, ny, False, True, stream)
        return out
    work = torch.empty((P, nx * ny), dtype=torch.complex64, device=ex.device)
    _lib.check(_lib.load().tp_farfield2d(_lib.ptr(ex), nx, ny, P, _lib.ptr(out), _lib.ptr(work),
                                         _lib.stream_handle(stream)), "tp_farfield2d")
    return out
