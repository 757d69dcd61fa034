"""FP64 restatement of the reference's multislice propagator and far field.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``thick_ptycho/forward_model/multislice/ms_solver.py:49-159``
* ``thick_ptycho/forward_model/base/base_solver.py:194-249``
"""

from __future__ import annotations

from typing import Optional

import numpy as np


def propagation_kernel(nx: int, dx: float, k: float, dz: float, remove_global_phase=True) -> np.ndarray:
    """``_get_propagation_kernels`` (ms_solver.py:49-72).

    The reference passes ``psi.size`` as ``remove_global_phase`` (always truthy,
    quirk Q6), so the global phase exp(-i k dz) is always removed and the
    evanescent band is clipped to kz = 0."""
    fx = np.fft.fftfreq(nx, d=dx)
    kx = 2 * np.pi * fx
    inside = k**2 - kx**2
    kz = np.sqrt(np.clip(inside, 0.0, None))
    H = np.exp(1j * kz * dz)
    if remove_global_phase:
        H *= np.exp(-1j * k * dz)
    return H.astype(np.complex128)


def transmission(n_slice: np.ndarray, k: float, dz: float, n_medium: float) -> np.ndarray:
    """``_object_transmission_function`` (ms_solver.py:86-88)."""
    return np.exp(1j * k * (n_slice - n_medium) * dz)


def solve_single_probe(n: np.ndarray, probe: np.ndarray, dx: float, k: float, dz: float, n_medium: float = 1.0) -> np.ndarray:
    """``MSForwardModelSolver._solve_single_probe`` (ms_solver.py:90-159): (nx, nz)."""
    nx, nz = n.shape
    H = propagation_kernel(nx, dx, k, dz, True)
    psi = np.asarray(probe, dtype=complex).copy()
    u = np.empty((nx, nz), dtype=complex)
    u[:, 0] = psi
    for z in range(nz - 1):
        psi = psi * transmission(n[:, z], k, dz, n_medium)
        psi = np.fft.ifft(np.fft.fft(psi) * H)
        u[:, z + 1] = psi
    return u


def solve_batched(n: np.ndarray, probes: np.ndarray, dx: float, k: float, dz: float, n_medium: float = 1.0, keep: str = "all") -> np.ndarray:
    """The loop above with the probe index vectorised: (P, nx, nz) or exit (P, nx)."""
    nx, nz = n.shape
    H = propagation_kernel(nx, dx, k, dz, True)
    psi = np.asarray(probes, dtype=complex).copy()
    P = psi.shape[0]
    out = None
    if keep == "all":
        out = np.empty((P, nx, nz), dtype=complex)
        out[:, :, 0] = psi
    for z in range(nz - 1):
        psi = psi * transmission(n[:, z], k, dz, n_medium)[None, :]
        psi = np.fft.ifft(np.fft.fft(psi, axis=-1) * H[None, :], axis=-1)
        if keep == "all":
            out[:, :, z + 1] = psi
    return psi if keep == "exit" else out


def backpropagate(psi: np.ndarray, dx: float, k: float, dz: float) -> np.ndarray:
    """``_backpropagate`` (ms_solver.py:80-84): ifft(fft(psi) * conj(H))."""
    H = propagation_kernel(psi.shape[-1], dx, k, dz, True)
    return np.fft.ifft(np.fft.fft(psi, axis=-1) * np.conj(H), axis=-1)


# ----------------------------------------------------------------------------
# Exit waves and far field  (base/base_solver.py:194-249)
# ----------------------------------------------------------------------------


def exit_waves(u: np.ndarray, total_scans: int, block_size: int) -> np.ndarray:
    """``get_exit_waves``: u[..., -1] reshaped to (total_scans, block_size)."""
    return u[..., -1].reshape((total_scans, block_size))


def farfield_intensities(exit_w: np.ndarray, dimension: int, total_scans: int, block_size: int,
                         rng: Optional[np.random.Generator] = None, poisson_noise: bool = False) -> np.ndarray:
    """``get_farfield_intensities``: |fft(exit)|^2, unnormalised, no shift.

    For ``dimension == 3`` the reference applies ``fft2`` over the last two
    axes of the flattened (scans, N) array, i.e. across the scan axis (quirk
    Q9); this restates that literally.  ``farfield_per_probe_2d`` is the
    per-probe variant the HIP library computes for two transverse axes."""
    if dimension == 2:
        f = np.fft.fft(exit_w)
    else:
        f = np.fft.fft2(exit_w, axes=(-2, -1))
    inten = np.abs(f) ** 2
    if poisson_noise:
        rng = rng or np.random.default_rng()
        inten = rng.poisson(inten)
    return inten.reshape((total_scans, block_size))


def farfield_per_probe_2d(exit_w: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """|fft2|^2 of each probe's (nx, ny) exit wave, flattened back to (P, nx*ny)."""
    P = exit_w.shape[0]
    f = np.fft.fft2(exit_w.reshape(P, nx, ny), axes=(-2, -1))
    return (np.abs(f) ** 2).reshape(P, nx * ny)


def propagation_kernel_2d(nx: int, ny: int, dx: float, dy: float, k: float, dz: float) -> np.ndarray:
    """Two-axis form of ``propagation_kernel`` (the reference has no two-axis
    multislice, ms_solver.py:28-31): exp(i (kz - k) dz), kz = sqrt(max(k^2 - kx^2 - ky^2, 0)),
    shape (nx, ny) in numpy.fft.fftfreq order along both axes."""
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)[:, None]
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy)[None, :]
    kz = np.sqrt(np.clip(k**2 - kx**2 - ky**2, 0.0, None))
    return np.exp(1j * (kz - k) * dz)


def solve_batched_2d(n: np.ndarray, probes: np.ndarray, dx: float, dy: float, k: float, dz: float,
                     n_medium: float = 1.0, keep: str = "all", conj_h: bool = False) -> np.ndarray:
    """psi <- ifft2(fft2(psi * T_z) * H) over z < nz-1 for probes (P, nx, ny); n (nx, ny, nz).
    Returns (P, nx, ny, nz) or the exit planes (P, nx, ny)."""
    nx, ny, nz = n.shape
    H = propagation_kernel_2d(nx, ny, dx, dy, k, dz)
    if conj_h:
        H = np.conj(H)
    psi = np.asarray(probes, dtype=complex).reshape(-1, nx, ny).copy()
    out = None
    if keep == "all":
        out = np.empty(psi.shape + (nz,), dtype=complex)
        out[..., 0] = psi
    for z in range(nz - 1):
        psi = psi * transmission(n[:, :, z], k, dz, n_medium)[None]
        psi = np.fft.ifft2(np.fft.fft2(psi, axes=(-2, -1)) * H[None], axes=(-2, -1))
        if out is not None:
            out[..., z + 1] = psi
    return out if out is not None else psi


# ----------------------------------------------------------------------------
# Exact gradient of the far-field amplitude misfit through the multislice
# (the differentiable form of ms_solver.py:143-158; the reference's MS
# reconstruction, ms_reconstructor.py:109-190, is a 3PIE update instead).
# psi_{z+1} = P (T_z psi_z) with P = F^-1 H F, so with lambda_L = dJ/d conj(psi_L):
#   mu_z = P^H lambda_{z+1} = F^-1 conj(H) F lambda_{z+1},  lambda_z = conj(T_z) mu_z,
#   dJ/dRe n_z + i dJ/dIm n_z = 2 conj(dT_z/dn_z) sum_p conj(psi_z) mu_z
#                             = -2 i k dz sum_p conj(T_z psi_z) mu_z,
# and the last column of n (never used by the loop) has zero gradient.
# ----------------------------------------------------------------------------


def _ms_gradient(n, psi0, k, dz, n_medium, H, fwd, inv, loss_fn, seed_fn, seed):
    nz = n.shape[-1]
    T = transmission(n, k, dz, n_medium)
    psi = [np.asarray(psi0, dtype=complex)]
    for z in range(nz - 1):
        psi.append(inv(fwd(psi[-1] * T[..., z]) * H))
    J = loss_fn(psi[-1])
    lam = seed_fn(psi[-1]) if seed is None else np.asarray(seed, dtype=complex).reshape(psi[-1].shape)
    grad = np.zeros(n.shape, dtype=complex)
    for z in range(nz - 2, -1, -1):
        mu = inv(fwd(lam) * np.conj(H))
        grad[..., z] = -2j * k * dz * np.sum(np.conj(T[..., z] * psi[z]) * mu, axis=0)
        lam = np.conj(T[..., z]) * mu
    return grad, J


def ms_gradient(n: np.ndarray, probes: np.ndarray, dx: float, k: float, dz: float, I_meas: np.ndarray,
                n_medium: float = 1.0, seed: Optional[np.ndarray] = None):
    """(dJ/dRe n + i dJ/dIm n of shape (nx, nz), J) for J = sum (|fft(psi_L)| - sqrt(I))^2,
    one transverse axis; ``seed`` (P, nx) replaces dJ/d conj(psi_L) when given."""
    from .gradient import amplitude_loss, amplitude_loss_seed

    nx = n.shape[0]
    H = propagation_kernel(nx, dx, k, dz, True)
    return _ms_gradient(n, np.asarray(probes).reshape(-1, nx), k, dz, n_medium, H,
                        lambda a: np.fft.fft(a, axis=-1), lambda a: np.fft.ifft(a, axis=-1),
                        lambda e: amplitude_loss(e, I_meas), lambda e: amplitude_loss_seed(e, I_meas), seed)


def ms_gradient_2d(n: np.ndarray, probes: np.ndarray, dx: float, dy: float, k: float, dz: float,
                   I_meas: np.ndarray, n_medium: float = 1.0, seed: Optional[np.ndarray] = None):
    """Two-axis form: n (nx, ny, nz), probes (P, nx, ny), the per-probe fft2 far field."""
    from .gradient import amplitude_loss_2d, amplitude_loss_seed_2d

    nx, ny, _ = n.shape
    P = probes.shape[0]
    H = propagation_kernel_2d(nx, ny, dx, dy, k, dz)
    return _ms_gradient(n, np.asarray(probes).reshape(P, nx, ny), k, dz, n_medium, H,
                        lambda a: np.fft.fft2(a, axes=(-2, -1)), lambda a: np.fft.ifft2(a, axes=(-2, -1)),
                        lambda e: amplitude_loss_2d(e, I_meas, nx, ny),
                        lambda e: amplitude_loss_seed_2d(e, I_meas, nx, ny).reshape(P, nx, ny), seed)
