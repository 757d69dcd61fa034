"""FP64 gradients of the amplitude misfit through the PWE slice stepping.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Two gradients are restated:

* ``reference_gradient`` follows ``ReconstructorPWE.compute_grad_least_squares``
  (thick_ptycho/reconstruction/pwe_reconstructor.py:100-164) with the error of
  ``ReconstructorBase.get_error_in_exit_wave`` / ``_apply_phase_retrieval_constraint``
  (thick_ptycho/reconstruction/base_reconstructor.py:87-207) and the operator of
  ``setup_inhomogeneous_forward_model(grad=True)`` (forward_model.py:128-153).
  It keeps the reference's adjoint B index (quirk Q3), constant g0, missing
  u_0 term and 1/N FFT scale (quirk Q17).
* ``exact_gradient`` is the exact discrete adjoint of the forward map for
  J(n) = sum_p sum_q (|F u_L|_q - sqrt(I_q))^2, returned in the
  ``dJ/dRe(n) + i dJ/dIm(n)`` convention.  It is pinned by central finite
  differences (tests/test_oracle_gradient.py).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import pwe


def phase_retrieval_model(exit_w: np.ndarray, I_meas: np.ndarray) -> np.ndarray:
    """``_apply_phase_retrieval_constraint`` (base_reconstructor.py:171-207)."""
    fmodel = np.fft.fft(exit_w, axis=-1)
    magnitude = np.abs(fmodel)
    magnitude[magnitude == 0] = 1.0
    phase = fmodel / magnitude
    return np.fft.ifft(np.sqrt(I_meas) * phase, axis=-1)


def gradient_operator(g: pwe.Grid, n: np.ndarray) -> sp.spmatrix:
    """``setup_inhomogeneous_forward_model(grad=True)`` (forward_model.py:128-153)."""
    eo = pwe.object_contribution(g, n, grad=True)
    N = g.block_size
    eo = eo.reshape(N, -1)
    return sp.diags(eo.T.flatten()) + sp.diags([eo[..., 1:].T.flatten()], offsets=[-N])


def to_vector_form(u: np.ndarray) -> np.ndarray:
    """``convert_to_vector_form`` for one projection/angle: (P, N, nz) -> (P, L*N)."""
    P = u.shape[0]
    return u[:, :, 1:].transpose(0, 2, 1).reshape(P, -1)


def reference_gradient_from_states(g: pwe.Grid, adj: pwe.LUCache, uk: np.ndarray, err_last: np.ndarray, grad_A):
    """compute_grad_least_squares (pwe_reconstructor.py:100-142) for one projection
    and angle: the reference adjoint of each probe's exit-wave error (last column
    of rhs_block), g = grad_A @ u, grad_re -= Re(conj(g) back), grad_im -=
    Re(conj(i g) back).  uk: (P, L*N) vector form; err_last: (P, N)."""
    N = g.block_size
    L = g.nz - 1
    grad_re = np.zeros(N * L)
    grad_im = np.zeros(N * L)
    for p in range(uk.shape[0]):
        rhs = np.zeros((N, L), dtype=complex)
        rhs[:, -1] = err_last[p]
        back = pwe.solve_single_probe(g, adj, None, "adjoint", rhs_block=rhs)
        back = back[:, :-1].T.ravel()
        gb = grad_A @ uk[p]
        grad_re -= (gb.conj() * back).real
        grad_im -= ((1j * gb).conj() * back).real
    return grad_re, grad_im


def reference_gradient(g: pwe.Grid, bc_type, n: np.ndarray, probes: np.ndarray, I_meas: np.ndarray):
    """Return (grad_real, grad_imag), each (L*N,) step-major, as the reference's
    reconstruction loop computes them: compute_grad_least_squares is called with
    grad_Ak = -get_gradient(n) (pwe_reconstructor.py:257, 281)."""
    N = g.block_size
    P = probes.shape[0]
    fwd = pwe.build_lu_cache(g, bc_type, n, "forward")
    adj = pwe.build_lu_cache(g, bc_type, n, "adjoint")
    u = pwe.solve_batched(g, fwd, probes.reshape(P, N), "forward")
    uk = to_vector_form(u)
    exit_w = u[:, :, -1]
    err_last = exit_w - phase_retrieval_model(exit_w, I_meas)
    return reference_gradient_from_states(g, adj, uk, err_last, -gradient_operator(g, n))


def amplitude_loss(exit_w: np.ndarray, I_meas: np.ndarray) -> float:
    """J = sum (|fft(u_L)| - sqrt(I))^2 over probes and detector pixels."""
    return float(np.sum((np.abs(np.fft.fft(exit_w, axis=-1)) - np.sqrt(I_meas)) ** 2))


def amplitude_loss_seed(exit_w: np.ndarray, I_meas: np.ndarray) -> np.ndarray:
    """dJ/d conj(u_L) = N * ifft((|U| - a) U / |U|), with U = fft(u_L)."""
    U = np.fft.fft(exit_w, axis=-1)
    mag = np.abs(U)
    safe = np.where(mag == 0, 1.0, mag)
    w = np.where(mag == 0, 0.0, (mag - np.sqrt(I_meas)) / safe)
    return exit_w.shape[-1] * np.fft.ifft(w * U, axis=-1)


def exact_adjoint_states(g: pwe.Grid, bc_type, n: np.ndarray, seed: np.ndarray) -> np.ndarray:
    """lambda_s, s = 1..L, of A_s^H lambda_s = B_{s+1}^H lambda_{s+1} (+ seed at s = L).

    Returns (P, N, L+1) with column s holding lambda_s (column 0 unused)."""
    A0, B0 = pwe.step_operators(g, bc_type)
    C = pwe.object_contribution(g, n).reshape(g.block_size, -1)
    L = g.nz - 1
    P = seed.shape[0]
    lam = np.zeros((P, g.block_size, L + 1), dtype=complex)
    v = seed.T.copy()
    for s in range(L, 0, -1):
        As = (A0 - sp.diags(C[:, s - 1])).conj().T.tocsc()
        if s < L:
            Bn = (B0 + sp.diags(C[:, s])).conj().T.tocsr()
            v = Bn @ v
        v = sp.linalg.splu(As).solve(v)
        lam[:, :, s] = v.T
    return lam


def exact_gradient(g: pwe.Grid, bc_type, n: np.ndarray, probes: np.ndarray, I_meas: np.ndarray,
                   exit_data: np.ndarray = None):
    """Exact dJ/dRe(n) + i dJ/dIm(n), shape (N, nz), and J, for the amplitude loss
    or, with ``exit_data`` (P, N), for J = sum |u_L - exit_data|^2 (seed u_L - d)."""
    N = g.block_size
    L = g.nz - 1
    P = probes.shape[0]
    fwd = pwe.build_lu_cache(g, bc_type, n, "forward")
    u = pwe.solve_batched(g, fwd, probes.reshape(P, N), "forward")
    if exit_data is not None:
        seed = u[:, :, -1] - exit_data
        lam = exact_adjoint_states(g, bc_type, n, seed)
        J = float(np.sum(np.abs(seed) ** 2))
        return _assemble_exact(g, n, u, lam), J
    seed = amplitude_loss_seed(u[:, :, -1], I_meas)
    lam = exact_adjoint_states(g, bc_type, n, seed)
    J = float(np.sum((np.abs(np.fft.fft(u[:, :, -1], axis=-1)) - np.sqrt(I_meas)) ** 2))
    return _assemble_exact(g, n, u, lam), J


def _assemble_exact(g: pwe.Grid, n: np.ndarray, u: np.ndarray, lam: np.ndarray) -> np.ndarray:
    """2 conj(Gn) from the forward planes u and adjoint states lam (P, N, L+1)."""
    N, L = g.block_size, g.nz - 1
    GC = np.zeros((N, L), dtype=complex)
    for s in range(1, L + 1):
        GC[:, s - 1] = np.sum(np.conj(lam[:, :, s]) * (u[:, :, s] + u[:, :, s - 1]), axis=0)
    n = np.asarray(n).reshape(N, g.nz)
    dcoef = (g.k / 1j) * (n if g.exact_ref_coeff else np.ones_like(n))
    Gn = np.zeros((N, g.nz), dtype=complex)
    Gn[:, :-1] += GC
    Gn[:, 1:] += GC
    Gn *= (g.dz / 4) * dcoef
    return 2 * np.conj(Gn)


# ----------------------------------------------------------------------------
# Two transverse axes: amplitude misfit of the per-probe 2D far field and the
# steepest-descent reconstruction step built on it (the three-dimensional
# counterpart of compute_grad_least_squares with the exact adjoint).
# ----------------------------------------------------------------------------


def amplitude_loss_2d(exit_w: np.ndarray, I_meas: np.ndarray, nx: int, ny: int) -> float:
    """J = sum (|fft2(u_L)| - sqrt(I))^2 over probes and detector pixels, planes (nx, ny)."""
    U = np.fft.fft2(np.asarray(exit_w).reshape(-1, nx, ny), axes=(-2, -1))
    return float(np.sum((np.abs(U) - np.sqrt(np.asarray(I_meas).reshape(U.shape))) ** 2))


def amplitude_loss_seed_2d(exit_w: np.ndarray, I_meas: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """dJ/d conj(u_L) = nx ny ifft2((|U| - a) U / |U|), flattened (P, nx*ny)."""
    P = exit_w.shape[0]
    U = np.fft.fft2(np.asarray(exit_w).reshape(P, nx, ny), axes=(-2, -1))
    mag = np.abs(U)
    safe = np.where(mag == 0, 1.0, mag)
    w = np.where(mag == 0, 0.0, (mag - np.sqrt(np.asarray(I_meas).reshape(U.shape))) / safe)
    return (nx * ny * np.fft.ifft2(w * U, axes=(-2, -1))).reshape(P, nx * ny)


def exact_gradient_3d(g: pwe.Grid, bc_type, n: np.ndarray, probes: np.ndarray, I_meas: np.ndarray, seed=None):
    """Exact dJ/dRe(n) + i dJ/dIm(n), shape (nx, ny, nz), and J, for the 2D far-field
    amplitude misfit.  ``seed`` replaces the loss seed when given."""
    N = g.block_size
    L = g.nz - 1
    P = probes.shape[0]
    fwd = pwe.build_lu_cache(g, bc_type, n, "forward")
    u = pwe.solve_batched(g, fwd, probes.reshape(P, N), "forward")
    if seed is None:
        seed = amplitude_loss_seed_2d(u[:, :, -1], I_meas, g.nx, g.ny)
    lam = exact_adjoint_states(g, bc_type, n, seed)
    GC = np.zeros((N, L), dtype=complex)
    for s in range(1, L + 1):
        GC[:, s - 1] = np.sum(np.conj(lam[:, :, s]) * (u[:, :, s] + u[:, :, s - 1]), axis=0)
    dcoef = g.k / 1j
    Gn = np.zeros((N, g.nz), dtype=complex)
    Gn[:, :-1] += GC
    Gn[:, 1:] += GC
    Gn *= (g.dz / 4) * dcoef
    J = amplitude_loss_2d(u[:, :, -1], I_meas, g.nx, g.ny)
    return (2 * np.conj(Gn)).reshape(g.nx, g.ny, g.nz), J


def reconstruct_3d(g: pwe.Grid, bc_type, n0: np.ndarray, probes: np.ndarray, I_meas: np.ndarray, iters: int,
                   step: float):
    """Steepest descent n <- n - step * grad J(n) for ``iters`` iterations; returns
    (n, [J before each iteration])."""
    n = np.array(n0, dtype=complex)
    hist = []
    for _ in range(iters):
        grad, J = exact_gradient_3d(g, bc_type, n, probes, I_meas)
        hist.append(J)
        n = n - step * grad
    return n, hist
