"""FP64 restatement of the reference's multislice 3PIE reconstruction.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows thick_ptycho/reconstruction/ms_reconstructor.py:40-190 with the
multislice operators of ``oracle.multislice`` (ms_solver.py:49-159):

* per (projection, angle, scan) in order, the forward multislice of the current
  object on the scan's window keeps every plane psi_z (z = 0..nz-1);
* the modulus constraint replaces |fft(psi_L)| by sqrt(I) keeping the phase
  (``_apply_modulus_constraint``, :164-170);
* the corrected exit wave is back-propagated for z = nz-1 .. 0 as
  x <- conj(T_z) P^H x (P^H = ifft(conj(H) fft(.))), and slice z of n is
  updated from dpsi_z = x - psi_z with step alpha / (iteration + 1)
  (``_apply_backprop_updates``, :109-159; ``_update_object_n``, :172-190):
      n_z += step conj(psi_z) dpsi_z / (max |psi_z|^2 + 1e-12) / (i k dz);
  the backward loop runs nz steps although the forward one applies nz-1
  transmissions, so it starts from the exit plane compared with psi_{nz-1};
* the epoch loss is sqrt(mean |psi_corr - psi_L|^2) over all scans and rows.

Pinned in tests/test_oracle_pie.py: the modulus projection's invariants, the
back-propagation as the conjugate transpose of the dense propagation matrix,
and the loss falling over epochs from free space on a seeded object.
"""

from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np

from . import multislice as oms


def modulus_constraint(psi: np.ndarray, data: np.ndarray, phase_retrieval: bool = True) -> np.ndarray:
    """``_apply_modulus_constraint`` (ms_reconstructor.py:164-170)."""
    if not phase_retrieval:
        return data
    F = np.fft.fft(psi, axis=-1)
    return np.fft.ifft(np.sqrt(data) * (F / (np.abs(F) + 1e-12)))


def update_object_n(n_slice, psi_i, dpsi, alpha, k, dz, eps=1e-12):
    """``_update_object_n`` (ms_reconstructor.py:172-190)."""
    grad = np.conj(psi_i) * dpsi
    scale = np.max(np.abs(psi_i) ** 2) + eps
    return n_slice + (alpha * grad / scale) / (1j * k * dz)


def backprop_updates(psi_slices, psi_corr, n_est, alpha, it, x_min, x_max, dx, k, dz, n_medium) -> None:
    """``_apply_backprop_updates`` (ms_reconstructor.py:109-159), in place on n_est."""
    nz = n_est.shape[1]
    psi_back = psi_corr
    for z in reversed(range(nz)):
        psi_back = oms.backpropagate(psi_back, dx, k, dz)
        psi_back = psi_back * np.conj(oms.transmission(n_est[x_min:x_max, z], k, dz, n_medium))
        dpsi = psi_back - psi_slices[:, z]
        n_est[x_min:x_max, z] = update_object_n(n_est[x_min:x_max, z], psi_slices[:, z], dpsi, alpha / (it + 1), k, dz)


def reconstruct_3pie(probes: np.ndarray, data: np.ndarray, nx: int, nz: int, dx: float, k: float, dz: float,
                     n_medium: float = 1.0, max_iters: int = 10, alpha_obj: float = 1.0,
                     n0: Optional[np.ndarray] = None, windows: Optional[Sequence[Tuple[int, int]]] = None,
                     phase_retrieval: bool = True):
    """(n_est (nx, nz), loss history) of ``ReconstructorMS.reconstruct`` for scans
    ``probes`` (S, W) with measurements ``data`` (S, W) (one projection, one angle)."""
    S = probes.shape[0]
    n_est = n_medium * np.ones((nx, nz), dtype=complex) if n0 is None else np.array(n0, dtype=complex)
    hist = []
    for it in range(max_iters):
        err = []
        for s in range(S):
            a, b = windows[s] if windows is not None else (0, nx)
            psi = oms.solve_single_probe(n_est[a:b], probes[s], dx, k, dz, n_medium)
            corr = modulus_constraint(psi[:, -1], data[s], phase_retrieval)
            err.append(corr - psi[:, -1])
            backprop_updates(psi, corr, n_est, alpha_obj, it, a, b, dx, k, dz, n_medium)
        hist.append(float(np.sqrt(np.mean(np.abs(np.array(err)) ** 2))))
    return n_est, hist
