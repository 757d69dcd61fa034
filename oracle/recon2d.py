"""FP64 restatement of the reference's one-axis NLCG reconstruction step and
object filters.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``reference_alphak``        <- ReconstructorPWE.compute_alphak
                                 (thick_ptycho/reconstruction/pwe_reconstructor.py:166-219)
* ``betak``                   <- compute_betak (:221-228)
* ``tv_denoise_amp``          <- tv_denoise_amp (:370-386)
* ``gaussian_lowpass_phase``  <- gaussian_lowpass_phase (:388-400)
* ``global_object_update``    <- global_object_update (:402-419)
* ``reference_reconstruct``   <- ReconstructorPWE.reconstruct (:230-368) with
                                 ReconstructorBase.get_error_in_exit_wave
                                 (base_reconstructor.py:87-121), one projection
                                 and one angle, full domain.
"""

from __future__ import annotations

import numpy as np

from . import pwe
from .gradient import gradient_operator, phase_retrieval_model, reference_gradient_from_states, to_vector_form


def reference_alphak(g: pwe.Grid, fwd: pwe.LUCache, probes: np.ndarray, uk: np.ndarray, grad_A, grad_E: np.ndarray,
                     d: np.ndarray) -> float:
    """compute_alphak: per probe, delta_u = forward solve of the probe with
    rhs_block = (grad_A @ u_p) as columns (step j takes column j-1); the scalar
    delta_p = -(last plane of delta_u) . (last N entries of d), no conjugate;
    alpha = -<d, grad_E> / sum |delta_p|^2 (0 when the sum vanishes)."""
    N, L = g.block_size, g.nz - 1
    den = 0.0
    for p in range(probes.shape[0]):
        perturb = grad_A @ uk[p]
        du = pwe.solve_single_probe(g, fwd, probes[p], "forward", rhs_block=perturb.reshape(L, N).T)
        du = du[:, 1:].T.ravel()
        dp = -du[-N:] @ d[-N:]
        den += np.abs(dp) ** 2
    num = np.vdot(d, grad_E)
    return float(np.real(-(num / den))) if den > 0 else 0.0


def betak(g_new: np.ndarray, g_old: np.ndarray) -> float:
    diff = g_new - g_old
    nrm = np.vdot(g_old, g_old)
    pr = np.vdot(g_new, diff) / nrm
    fr = np.vdot(g_new, g_new) / nrm
    return float(np.real(min(fr, max(pr, 0))))


def tv_denoise_amp(u: np.ndarray, lam: float = 0.01, n_iter: int = 30, eps: float = 1e-8) -> np.ndarray:
    """Chambolle-type TV iteration with periodic differences (np.roll), tau = 1/4."""
    px = np.zeros_like(u)
    py = np.zeros_like(u)
    tau = 0.25
    for _ in range(n_iter):
        u_xy = u - lam * (np.roll(px, 1, 1) - px + np.roll(py, 1, 0) - py)
        dx = np.roll(u_xy, -1, axis=1) - u_xy
        dy = np.roll(u_xy, -1, axis=0) - u_xy
        denom = 1.0 + np.sqrt(dx * dx + dy * dy + eps) / lam
        px = (px + tau * dx) / denom
        py = (py + tau * dy) / denom
    return u - lam * (np.roll(px, 1, 1) - px + np.roll(py, 1, 0) - py)


def gaussian_lowpass_phase(phi: np.ndarray, sigma: float = 0.08) -> np.ndarray:
    """real(ifft2(fft2(phi) * fftshift(W))), W = exp(-(qx^2 + qy^2) / (2 sigma^2)) on
    linspace(-1/2, 1/2, n, endpoint=False) along each axis (axis 1 is x)."""
    ny, nx = phi.shape
    qx = np.linspace(-0.5, 0.5, nx, endpoint=False)
    qy = np.linspace(-0.5, 0.5, ny, endpoint=False)
    Qx, Qy = np.meshgrid(qx, qy)
    W = np.exp(-(Qx**2 + Qy**2) / (2 * sigma**2))
    return np.real(np.fft.ifft2(np.fft.fft2(phi) * np.fft.fftshift(W)))


def global_object_update(obj: np.ndarray, lam_amp: float, sigma_phase: float) -> np.ndarray:
    A = np.abs(obj)
    Phi = np.angle(obj)
    if lam_amp > 0.0:
        A = tv_denoise_amp(A, lam=lam_amp)
    if sigma_phase > 0.0:
        Phi = gaussian_lowpass_phase(Phi, sigma=sigma_phase)
    return A * np.exp(1j * Phi)


def reference_reconstruct(g: pwe.Grid, bc_type, n0: np.ndarray, probes: np.ndarray, I_meas: np.ndarray,
                          max_iters: int, tol: float = 1e-8, fixed_step_size=None, tv_lambda_amp: float = 0.0,
                          low_pass_sigma_phase: float = 0.0, log=None):
    """(n_est, residual history, step history [(alpha_re, alpha_im)]) of the reference loop."""
    N, L, nz = g.block_size, g.nz - 1, g.nz
    P = probes.shape[0]
    nk = np.array(n0, dtype=complex)
    residual, steps = [], []
    grad_Ak = -gradient_operator(g, nk)
    update = np.zeros((N, nz), dtype=complex)
    dk_re = dk_im = g_re_old = g_im_old = None
    for i in range(max_iters):
        fwd = pwe.build_lu_cache(g, bc_type, nk, "forward")
        adj = pwe.build_lu_cache(g, bc_type, nk, "adjoint")
        u = pwe.solve_batched(g, fwd, probes.reshape(P, N), "forward")
        uk = to_vector_form(u)
        if g.exact_ref_coeff:
            grad_Ak = -gradient_operator(g, nk)
        exit_w = u[:, :, -1]
        err_last = exit_w - phase_retrieval_model(exit_w, I_meas)
        residual.append(float(np.sqrt(np.mean(np.abs(err_last) ** 2))))
        if residual[i] < tol:
            break
        g_re, g_im = reference_gradient_from_states(g, adj, uk, err_last, grad_Ak)
        if i == 0:
            dk_re, dk_im = -g_re, -g_im
            g_re_old, g_im_old = g_re, g_im
        if fixed_step_size is not None:
            a_re = a_im = -fixed_step_size / (i + 1)
        else:
            a_re = reference_alphak(g, fwd, probes, uk, grad_Ak, g_re, dk_re)
            a_im = reference_alphak(g, fwd, probes, uk, 1j * grad_Ak, g_im, dk_im)
        steps.append((a_re, a_im))
        update[:, 1:] = (a_re * dk_re).reshape(L, N).T + 1j * (a_im * dk_im).reshape(L, N).T
        nk = nk + update
        b_re, b_im = betak(g_re, g_re_old), betak(g_im, g_im_old)
        dk_re = -g_re + b_re * dk_re
        g_re_old = g_re
        dk_im = -g_im + b_im * dk_im
        g_im_old = g_im
        if tv_lambda_amp > 0.0 or low_pass_sigma_phase > 0.0:
            nk = global_object_update(nk, tv_lambda_amp, low_pass_sigma_phase)
        if log:
            log(f"iteration {i + 1}: rmse {residual[i]:.6e} alpha ({a_re:.4e}, {a_im:.4e})")
    return nk, residual, steps
