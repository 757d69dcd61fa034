"""Drop-in ``MSForwardModelSolver`` running on the MI355X kernels.

Mirrors thick_ptycho/forward_model/multislice/ms_solver.py:8-159.  Each slice
is psi <- ifft(fft(psi * exp(i k (n - n_medium) dz)) * H) with
H = exp(i (sqrt(max(k^2 - kx^2, 0)) - k) dz) (global phase always removed,
quirk Q6).  Only one transverse axis, like the reference (3D raises).
"""

from __future__ import annotations

import time

import numpy as np
import torch

from . import ms_engine
from .base_solver import BaseForwardModelSolver, host_complex128


class MSForwardModelSolver(BaseForwardModelSolver):
    def __init__(self, simulation_space, ptycho_probes, results_dir="", use_logging=False, verbose=False, log=None,
                 device=None):
        super().__init__(simulation_space, ptycho_probes, results_dir=results_dir, use_logging=use_logging,
                         verbose=verbose, log=log, device=device)
        if self.simulation_space.dimension == 3:
            raise NotImplementedError("3D simulation space is not yet implemented for MSForwardModelSolver.")
        self.k = self.simulation_space.k
        self.nx_eff = self.simulation_space.effective_nx
        self.dx = self.simulation_space.dx
        self.nx = self.simulation_space.nx
        self.nz = self.simulation_space.nz
        self.dz = self.simulation_space.dz
        self.n_medium = self.simulation_space.n_medium
        if complex(self.n_medium).imag != 0:
            raise ValueError("n_medium must be real for the multislice kernels")
        self._kernel_cache = {}

    def reset_cache(self):
        self._kernel_cache = {}

    def _get_propagation_kernels(self, dz: float, remove_global_phase=True) -> torch.Tensor:
        """Device H / nx, cached by (dz, nx)."""
        dz = self.dz if dz is None else dz
        key = (dz, self.nx_eff)
        if key not in self._kernel_cache:
            self._kernel_cache[key] = ms_engine.propagator(self.nx_eff, self.dx, self.k, dz, self.device)
        return self._kernel_cache[key]

    def _n_device(self, n) -> torch.Tensor:
        if n is None:
            n = self.simulation_space.refractive_index_empty
        if isinstance(n, torch.Tensor):
            return n.to(self.device, torch.complex128).contiguous()
        return torch.from_numpy(np.ascontiguousarray(n, dtype=np.complex128)).to(self.device)

    def solve_device(self, n=None, mode: str = "forward", probes=None, keep: str = "all",
                     plain: bool = False) -> torch.Tensor:
        """All probes in one launch: complex64 (Proj, A, S, nz, nx) or exit planes (Proj, A, S, nx)."""
        assert mode in {"forward", "backward"}, f"Invalid mode: {mode}"
        if probes is None:
            probes = self.probes
        pr = probes if isinstance(probes, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(probes))
        pr = pr.to(self.device, torch.complex64).reshape(self.num_angles, self.num_probes, self.nx_eff).contiguous()
        T = ms_engine.transmission(self._n_device(n), self.k, self.dz, float(complex(self.n_medium).real))
        H = self._get_propagation_kernels(self.dz)
        shape = (self.num_projections, self.num_angles, self.num_probes, self.nz if keep == "all" else 1, self.nx_eff)
        out = torch.empty(shape, dtype=torch.complex64, device=self.device)
        specs = []
        for p in range(self.num_projections):
            for a in range(self.num_angles):
                if keep == "all":
                    specs.append(ms_engine.MsScanSpec(T, H, pr[a], out=out[p, a]))
                else:
                    specs.append(ms_engine.MsScanSpec(T, H, pr[a], exit=out[p, a, :, 0]))
        t0 = time.time()
        for i in range(0, len(specs), 8):
            (ms_engine.sweep_plain if plain else ms_engine.sweep)(specs[i:i + 8])
        if self.verbose:
            torch.cuda.synchronize()
            self._log(f"[{self.solver_type}] {len(specs) * self.num_probes} probes in {time.time() - t0:.3f}s")
        return out if keep == "all" else out[..., 0, :]

    def solve(self, n=None, mode: str = "forward", rhs_block=None, probes=None, **kwargs) -> np.ndarray:
        """Reference-compatible solve: complex128 (Proj, A, S, nx, nz) on the host.

        As in the reference, only mode='forward' gets past both mode checks
        (base/base_solver.py:113 and ms_solver.py:122-125)."""
        assert mode in {"forward", "adjoint", "reverse"}, f"Invalid mode: {mode!r}"
        u = self.solve_device(n=n, mode=mode, probes=probes, keep="all")
        return host_complex128(u.transpose(-1, -2))

    def _solve_single_probe(self, scan_idx: int = 0, proj_idx: int = 0, n=None, mode="forward", probe=None,
                            **kwargs) -> np.ndarray:
        assert mode in {"forward", "backward"}, f"Invalid mode: {mode}"
        if probe is None:
            probe = np.zeros((self.nx_eff,), dtype=complex)
        pr = torch.from_numpy(np.asarray(probe, dtype=np.complex128).reshape(1, self.nx_eff)).to(self.device,
                                                                                               torch.complex64)
        T = ms_engine.transmission(self._n_device(n), self.k, self.dz, float(complex(self.n_medium).real))
        H = self._get_propagation_kernels(self.dz)
        out = torch.empty((1, self.nz, self.nx_eff), dtype=torch.complex64, device=self.device)
        ms_engine.sweep([ms_engine.MsScanSpec(T, H, pr.contiguous(), out=out)])
        return out[0].transpose(0, 1).cpu().numpy().astype(np.complex128)
