"""One-axis reconstruction drivers (NLCG / steepest descent on the CN gradient,
multislice 3PIE) and the reference's data files.

Data files (numpy ``.npz``, the keys the reference writes):

* simulated data, ``np.savez_compressed(..., intensity=..., exit_waves=...)``
  (generate_data.py:193-197);
* reconstructions, ``np.savez_compressed(..., n_est=..., wave_field=...,
  residual_history=...)`` (reconstruct_sample.py:105-108).

The driver minimises J(n) = sum_p sum (|fft(u_L^p)| - sqrt(I^p))^2 with the
exact gradient of ``gradient.amplitude_gradient`` (checkpointed forward and
exact-adjoint sweeps on the GPU).  ``method="nlcg"`` is the nonlinear conjugate
gradient iteration of the reference's ReconstructorPWE.reconstruct
(pwe_reconstructor.py:230-368): separate directions for Re n and Im n, each with
beta = min(beta_FR, max(beta_PR, 0)) (compute_betak, pwe_reconstructor.py:221-228).
The step differs: 0.5 / L with L a secant estimate of the gradient's Lipschitz
constant instead of the reference's linearised line search (compute_alphak,
which needs a forward solve per probe and direction), halved and restarted
along -grad whenever an iterate raises J.  ``method="sd"`` is steepest descent
with the same step.
"""

from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib
from . import gradient as tg


def save_simulated_data(path: str, intensity, exit_waves) -> None:
    np.savez_compressed(path, intensity=np.asarray(intensity), exit_waves=np.asarray(exit_waves))


def load_simulated_data(path: str) -> Dict[str, np.ndarray]:
    with np.load(path, allow_pickle=False) as f:
        return {"intensity": f["intensity"], "exit_waves": f["exit_waves"]}


def save_reconstruction(path: str, n_est, wave_field, residual_history) -> None:
    np.savez_compressed(path, n_est=np.asarray(n_est), wave_field=np.asarray(wave_field),
                        residual_history=np.asarray(residual_history))


def load_reconstruction(path: str) -> Dict[str, np.ndarray]:
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in ("n_est", "wave_field", "residual_history")}


def beta_rule(g_new: torch.Tensor, g_old: torch.Tensor) -> float:
    """min(FR, max(PR, 0)) for one real gradient: FR = |g_new|^2 / |g_old|^2,
    PR = g_new . (g_new - g_old) / |g_old|^2 (pwe_reconstructor.py:221-228)."""
    den = float(torch.dot(g_old.flatten(), g_old.flatten()))
    if den <= 0:
        return 0.0
    fr = float(torch.dot(g_new.flatten(), g_new.flatten())) / den
    pr = float(torch.dot(g_new.flatten(), (g_new - g_old).flatten())) / den
    return min(fr, max(pr, 0.0))


class Reconstruction2D:
    """Reconstruct n (nx, nz) from far-field intensities I (P, nx) of probes (P, nx),
    or, with ``phase_retrieval=False``, from complex exit waves (P, nx) with the
    loss sum |u_L - data|^2 (the reference's two data kinds, pwe_reconstructor.py).
    With ``windows`` [(a_s, b_s)] (reduced domain) probe s and its data are the
    window width and see rows a_s..b_s-1 of n only."""

    def __init__(self, intensity, probes, k: float, dx: float, dz: float, bc: str = "dirichlet",
                 checkpoint: int = 32, device=None, phase_retrieval: bool = True, windows=None):
        self.windows = None if windows is None else [(int(a), int(b)) for a, b in windows]
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.probes = torch.as_tensor(np.asarray(probes)).to(self.dev, torch.complex64)
        self.phase_retrieval = phase_retrieval
        if phase_retrieval:
            self.sqrt_i = torch.as_tensor(np.sqrt(np.asarray(intensity, dtype=float))).to(self.dev, torch.float32)
            self.exit_data = None
        else:
            self.exit_data = torch.as_tensor(np.asarray(intensity)).to(self.dev, torch.complex64)
            self.sqrt_i = torch.zeros(self.exit_data.shape, dtype=torch.float32, device=self.dev)
        self.k, self.dx, self.dz, self.bc, self.checkpoint = k, dx, dz, bc, checkpoint

    def gradient(self, n: torch.Tensor):
        if self.windows is not None:
            return tg.amplitude_gradient_reduced(n, self.probes, self.sqrt_i, self.windows, self.k, self.dx, self.dz,
                                                 self.bc, device=self.dev, exit_data=self.exit_data)
        return tg.amplitude_gradient(n, self.probes, self.sqrt_i, self.k, self.dx, self.dz, self.bc,
                                     self.checkpoint, device=self.dev, exit_data=self.exit_data)

    def wave_field(self, n) -> np.ndarray:
        """Forward field of every probe at slices 1..nz-1, complex64 (1, P, nx, nz-1): the
        reference's ``convert_to_tensor_form(uk)`` layout (base_reconstructor.py:123-140);
        with windows each probe's field on its window, (1, P, W, nz-1)."""
        n_t = tg._dev_tensor(n, torch.complex128, self.dev)
        if self.windows is not None:
            return np.stack([self._window_field(n_t[a:b], p) for p, (a, b) in enumerate(self.windows)])[None]
        N, nz = n_t.shape
        A, B = tg._operators(N, self.k, self.dx, self.dz, self.bc)
        sw = tg._Sweeper(tg.object_contribution(n_t, self.k, self.dz), A, B, "exact_adjoint", 1)
        ck, _ = sw.forward(self.probes.reshape(-1, N))
        return ck[:, 1:].transpose(1, 2).unsqueeze(0).cpu().numpy()

    def _window_field(self, n_w: torch.Tensor, p: int) -> np.ndarray:
        N, nz = n_w.shape
        A, B = tg._operators(N, self.k, self.dx, self.dz, self.bc)
        sw = tg._Sweeper(tg.object_contribution(n_w.contiguous(), self.k, self.dz), A, B, "exact_adjoint", 1)
        ck, _ = sw.forward(self.probes[p:p + 1].reshape(1, N).contiguous())
        return ck[0, 1:].transpose(0, 1).cpu().numpy()

    def lipschitz(self, n: torch.Tensor, h: float = 1e-5) -> float:
        _, g0 = self.gradient(n)
        d = (h / g0.abs().max()) * g0
        _, g1 = self.gradient(n - d)
        return float(torch.linalg.vector_norm(g1 - g0) / torch.linalg.vector_norm(d))

    def rmse(self, J: float) -> float:
        """sqrt(mean |e|^2) of the reference's exit-wave error (pwe_reconstructor.py:298-301):
        e = u - ifft(sqrt(I) U/|U|), ||e_p||^2 = J_p / N by Parseval, or e = u - data."""
        P, N = self.sqrt_i.shape
        return float(np.sqrt(max(J, 0.0) / (P * N * (N if self.phase_retrieval else 1))))

    def run(self, n0, iters: int, method: str = "nlcg", alpha: Optional[float] = None,
            wave_field: bool = True, tol: Optional[float] = None) -> dict:
        """Returns {'n_est': complex128 (nx, nz), 'wave_field': ``wave_field(n_est)``,
        'residual_history': [J before the first iteration and after each accepted
        one], 'rmse_history': the same as exit-wave RMSE, 'alpha': the last step};
        the three arrays are what ``save_reconstruction`` writes.  With ``tol`` the
        loop stops once the RMSE is below it (the reference's test)."""
        if method not in ("nlcg", "sd"):
            raise ValueError("method must be 'nlcg' or 'sd'")
        n = torch.as_tensor(np.asarray(n0)).to(self.dev, torch.complex128).clone()
        if alpha is None:
            alpha = 0.5 / self.lipschitz(n)
        J, g = self.gradient(n)
        d = -g
        hist: List[float] = [J]
        for _ in range(iters):
            if tol is not None and self.rmse(J) < tol:
                break
            n_new = n + alpha * d
            J_new, g_new = self.gradient(n_new)
            if J_new > J and method == "nlcg":  # safeguard: smaller step along -grad
                alpha *= 0.5
                d = -g
                continue
            if method == "nlcg":
                b_re = beta_rule(g_new.real, g.real)
                b_im = beta_rule(g_new.imag, g.imag)
                d = -g_new + torch.complex(b_re * d.real, b_im * d.imag)
            else:
                d = -g_new
            n, J, g = n_new, J_new, g_new
            hist.append(J)
        n_est = n.cpu().numpy()
        return {"n_est": n_est, "wave_field": self.wave_field(n) if wave_field else None,
                "residual_history": hist, "rmse_history": [self.rmse(v) for v in hist], "alpha": alpha}


class ReconstructorPWE:
    """The reference's entry point for one transverse axis (ReconstructorPWE,
    pwe_reconstructor.py:34-368) on ``Reconstruction2D``: ``reconstruct`` returns
    (n_est, wave_field, residual) with the reference's residual (exit-wave RMSE)
    and stops once it is below ``tol``.  ``probes`` are given explicitly (the
    reference builds them from its simulation config); ``data`` are intensities,
    or exit waves with ``phase_retrieval=False``; the step is 0.5 / Lipschitz, or
    ``fixed_step_size`` as a constant step; the TV / low-pass object filters are
    not built."""

    def __init__(self, simulation_space, data, probes, phase_retrieval: bool = True, solver_type: str = "iterative",
                 bc_type="impedance", verbose: Optional[bool] = None, device=None):
        if solver_type not in {"full", "iterative"}:
            raise ValueError(f"Invalid solver_type: {solver_type!r}")
        ss = simulation_space
        if ss.dimension != 2:
            raise NotImplementedError("two transverse axes: thick_ptycho_amd.recon3d.Reconstruction3D")
        self.ss = ss
        windows = list(getattr(ss, "_scan_windows", None) or []) if ss.solve_reduced_domain else None
        W = ss.effective_nx if windows else ss.nx
        self.rc = Reconstruction2D(np.asarray(data).reshape(-1, W), np.asarray(probes).reshape(-1, W), ss.k,
                                   ss.dx, ss.dz, bc_type, device=device, phase_retrieval=phase_retrieval,
                                   windows=windows)

    def reconstruct(self, max_iters: int = 10, tol: float = 1e-8, n_initial=None, fixed_step_size=None,
                    tv_lambda_amp: float = 0.0, low_pass_sigma_phase: float = 0.0):
        if tv_lambda_amp or low_pass_sigma_phase:
            raise NotImplementedError("TV / low-pass object filters are not built")
        n0 = n_initial if n_initial is not None else np.ones(self.ss.shape, dtype=complex) * self.ss.n_medium
        out = self.rc.run(n0, max_iters, method="nlcg", alpha=fixed_step_size, tol=tol)
        return out["n_est"], out["wave_field"], out["rmse_history"]


class ReconstructorMS:
    """Multislice 3PIE reconstruction on the GPU (ms_reconstructor.py:7-190).

    Scan positions are visited in the reference's (projection, angle, scan)
    order and each updates the object before the next is propagated.  Per scan:
    ``tp_ms_transmission`` and a ``tp_ms_sweep`` keeping every plane (forward),
    ``tp_amplitude_seed`` with scale 1/N for the modulus constraint
    (psi_corr = psi_L - seed), ``tp_pie_adjoint_transmission`` and a conj(H)
    ``tp_ms_sweep`` for the back-propagation, ``tp_pie_update`` for every
    slice.  ``probes`` (A, S, W) or (S, W) are given explicitly (the reference
    builds them from its simulation config); ``data`` (A*S, W) are intensities
    or, with ``phase_retrieval=False``, exit waves.  Reduced-domain spaces update
    each scan's window of the object."""

    def __init__(self, simulation_space, data, probes, phase_retrieval: bool = True, verbose: bool = False,
                 device=None):
        from .forward_model import ms_engine

        self._ms = ms_engine
        ss = simulation_space
        if ss.dimension != 2:
            raise NotImplementedError("3PIE runs on one transverse axis (the reference's multislice)")
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.ss, self.phase_retrieval, self.verbose = ss, phase_retrieval, verbose
        self.k, self.nz, self.dz, self.dx = ss.k, ss.nz, ss.dz, ss.dx
        n_med = complex(getattr(ss, "n_medium", 1.0))
        if n_med.imag != 0:
            raise ValueError("n_medium must be real for the multislice kernels")
        self.n_medium = n_med.real
        self.nx = ss.nx
        self.W = ss.effective_nx if ss.solve_reduced_domain else ss.nx
        pr = np.asarray(probes).reshape(-1, self.W)
        self.num_scans = ss.num_probes
        reps = ss.num_projections  # the reference visits every projection with the same object and probes
        self.probes = torch.from_numpy(np.tile(pr, (reps, 1))).to(self.dev, torch.complex64)
        d = np.asarray(data).reshape(self.probes.shape[0], self.W)
        if phase_retrieval:
            self.target = torch.from_numpy(np.sqrt(d.real.astype(float))).to(self.dev, torch.float32)
        else:
            self.target = torch.from_numpy(d.astype(complex)).to(self.dev, torch.complex64)
        self.windows = None
        if ss.solve_reduced_domain:
            self.windows = list(getattr(ss, "_scan_windows"))

    def _window(self, idx: int):
        if self.windows is None:
            return 0, self.nx
        return self.windows[idx % self.num_scans]

    def reconstruct(self, max_iters: int = 10, alpha_obj: float = 1.0, refractive_index=None):
        """(n_est complex128 (nx, nz), loss history: sqrt(mean |psi_corr - psi_L|^2) per epoch)."""
        ms = self._ms
        nz, W = self.nz, self.W
        if refractive_index is None:
            n = torch.full((self.nx, nz), self.n_medium, dtype=torch.complex128, device=self.dev)
        else:
            n = torch.as_tensor(np.asarray(refractive_index)).to(self.dev, torch.complex128).clone()
        H = ms.propagator(W, self.dx, self.k, self.dz, self.dev)
        psi = torch.empty((1, nz, W), dtype=torch.complex64, device=self.dev)
        back = torch.empty((1, nz + 1, W), dtype=torch.complex64, device=self.dev)
        Tadj = torch.empty((nz, W), dtype=torch.complex64, device=self.dev)
        lib = _lib.load()
        hist = []
        for it in range(max_iters):
            err2 = torch.zeros((), dtype=torch.float64, device=self.dev)
            for idx in range(self.probes.shape[0]):
                a, b = self._window(idx)
                nw = n[a:b]
                T = ms.transmission(nw, self.k, self.dz, self.n_medium)
                ms.sweep([ms.MsScanSpec(T, H, self.probes[idx:idx + 1], out=psi)])
                ex = psi[:, nz - 1].contiguous()
                if self.phase_retrieval:
                    seed, _ = tg.amplitude_seed(ex, self.target[idx:idx + 1], 1.0 / W)
                    corr = ex - seed
                else:
                    corr = self.target[idx:idx + 1].contiguous()
                err2 += ((corr - ex).abs() ** 2).sum().double()
                sh = _lib.stream_handle()
                _lib.check(lib.tp_pie_adjoint_transmission(_lib.ptr(nw), nz, W, nz, float(self.k), float(self.dz),
                                                           self.n_medium, _lib.ptr(Tadj), sh),
                           "tp_pie_adjoint_transmission")
                ms.sweep([ms.MsScanSpec(Tadj, H, corr, out=back, conj_h=True)])
                _lib.check(lib.tp_pie_update(_lib.ptr(nw), nz, W, nz, _lib.ptr(psi), _lib.ptr(back), float(self.k),
                                             float(self.dz), self.n_medium, alpha_obj / (it + 1), 1e-12, sh),
                           "tp_pie_update")
            hist.append(float(torch.sqrt(err2 / (self.probes.shape[0] * W))))
        return n.cpu().numpy(), hist
