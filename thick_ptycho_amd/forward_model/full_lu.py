"""Drop-in ``PWEFullLUSolver``: the whole space-z system solved exactly.

Mirrors thick_ptycho/forward_model/pwe/solvers/full_lu.py:40-172.  The system
matrix M (forward_model.py:108-209) is block lower bidiagonal -- diagonal
blocks A0 - C_j, sub-diagonal blocks -(B0 + C_j) -- with right-hand side
b = [B0 u_probe, 0, ...] or ``rhs_block``.  Its LU solve is therefore block
forward substitution, which is the CN slice stepping with the first step's
object term C_1 u_probe moved to the right-hand side (the whole-system form
drops it); ``mode='adjoint'`` (the reference's ``lu.solve(b, trans='H')``) is
block backward substitution with the conjugate-transposed blocks, which is the
exact-adjoint stepping with b as its per-step right-hand side.  Both run on the
stepping kernels (tp_tri_sweep / tp_band_sweep) with the per-slice factors of
``PWEIterativeLUSolver``; no whole-system factorisation is formed.
"""

from __future__ import annotations

import numpy as np
import torch

from . import band_engine, tri_engine
from .base_solver import host_complex128
from .iterative_lu import PWEIterativeLUSolver
from .operators import BoundaryType


class PWEFullLUSolver(PWEIterativeLUSolver):
    """Whole-system PWE solve (full_lu.py:40-172) on the slice-stepping kernels."""

    def __init__(self, simulation_space, ptycho_probes, bc_type=BoundaryType.IMPEDANCE, results_dir="",
                 use_logging=False, verbose=False, log=None, test_bcs=None, device=None):
        if test_bcs is not None:
            raise NotImplementedError("test_bcs right-hand sides are not built for the whole-system solver")
        super().__init__(simulation_space, ptycho_probes, bc_type=bc_type, results_dir=results_dir,
                         use_logging=use_logging, verbose=verbose, log=log, device=device)
        if self.solve_reduced_domain:
            raise NotImplementedError("the whole-system solver does not take reduced-domain windows")

    def _b0(self, probes: torch.Tensor) -> torch.Tensor:
        """B0 u_probe for every probe, complex64 (P, N) (forward_model.py:155-193)."""
        B = self.pwe_finite_differences.B
        pr = probes.to(torch.complex128)
        if self.dimension == 3:
            Bs = self.pwe_finite_differences._to_sparse(B)
            out = (Bs @ pr.cpu().numpy().T).T
            return torch.from_numpy(np.ascontiguousarray(out)).to(self.device, torch.complex64)
        d = torch.from_numpy(B.diag).to(self.device)
        lo = torch.from_numpy(B.sub).to(self.device)
        up = torch.from_numpy(B.sup).to(self.device)
        out = d * pr
        out[:, 1:] += lo[1:] * pr[:, :-1]
        out[:, :-1] += up[:-1] * pr[:, 1:]
        return out.to(torch.complex64)

    def _rhs_rows(self, rhs_block) -> torch.Tensor:
        """(L, N) complex64 rows b_j from the reference's flat block-major vector
        (or an (N, L) array)."""
        N, L = self.block_size, self.nz - 1
        r = np.asarray(rhs_block, dtype=complex)
        r = r.reshape(L, N) if r.ndim == 1 else r.reshape(N, L).T
        return torch.from_numpy(np.ascontiguousarray(r)).to(self.device, torch.complex64)

    def _construct_solve_cache(self, n=None, mode: str = "forward", scan_idx: int = 0, proj_idx: int = 0) -> None:
        """Per-slice factors of M's diagonal blocks: the forward ones for 'forward'
        (and 'reverse', which the reference solves with the same matrix), the
        exact-adjoint ones for 'adjoint' (full_lu.py:66-98 forms and factors M)."""
        fm = self.pwe_finite_differences
        C = self.object_contribution_device(n, scan_idx)
        fmode = "exact_adjoint" if mode == "adjoint" else "forward"
        if self.dimension == 3:
            fac = band_engine.factor(C, fm.A, fm.B, fmode)
        else:
            fac = self._tri_factor(C, fmode)
        cache = self.projection_cache[proj_idx].modes[mode]
        cache.factors = fac
        cache.b = np.zeros(self.block_size * (self.nz - 1), dtype=complex)
        cache.cached_n = n

    def solve_device(self, n=None, mode: str = "forward", rhs_block=None, probes=None, **kwargs) -> torch.Tensor:
        """complex64 (Proj, A, S, nz, N) on the GPU: plane 0 the probe, planes 1..L the
        solution blocks (full_lu.py:168-172)."""
        assert mode in {"forward", "adjoint", "reverse"}, f"Invalid mode: {mode!r}"
        if n is None:
            n = self.simulation_space.refractive_index_empty
        pr = self._probe_tensor(self.probes if probes is None else probes)
        N, L = self.block_size, self.nz - 1
        A, S = self.num_angles, self.num_probes
        pr = pr.reshape(A, S, N)
        rows = self._rhs_rows(rhs_block)[None] if rhs_block is not None else None
        out = torch.empty((self.num_projections, A, S, self.nz, N), dtype=torch.complex64, device=self.device)
        for p in range(self.num_projections):
            fac = self._get_or_construct_cache(n=n, mode=mode, proj_idx=p).factors
            n_p = self.get_projected_obj(n=n, proj_idx=p)
            C1 = None
            if mode != "adjoint" and rows is None:
                C1 = self.object_contribution_device(n_p)[:, 0]
            for a in range(A):
                u0 = torch.zeros((S, N), dtype=torch.complex64, device=self.device)
                if mode == "adjoint":  # M^H x = b: exact-adjoint stepping, b_j at step L + 1 - j
                    if rows is not None:
                        rhs = torch.flip(rows, dims=[1]).contiguous()
                    else:
                        rhs = torch.zeros((S, L, N), dtype=torch.complex64, device=self.device)
                        rhs[:, L - 1] = self._b0(pr[a])
                    u = torch.empty((S, self.nz, N), dtype=torch.complex64, device=self.device)
                    self._launch([self._spec(fac, u0, out=u, rhs=rhs)], probes_per_block=self.probes_per_block)
                    out[p, a, :, 1:] = torch.flip(u[:, 1:], dims=[1])
                else:  # M x = b: forward stepping; the first step lacks C_1 u_probe
                    if rows is not None:
                        spec = self._spec(fac, u0, out=out[p, a], rhs=rows)
                    else:
                        rhs = (-(C1[None, :] * pr[a].to(torch.complex128))).to(torch.complex64)[:, None]
                        spec = self._spec(fac, pr[a].contiguous(), out=out[p, a], rhs=rhs.contiguous(), rhs_steps=1)
                    self._launch([spec], probes_per_block=self.probes_per_block)
                out[p, a, :, 0] = pr[a]
        return out

    def solve(self, n=None, mode: str = "forward", rhs_block=None, probes=None, **kwargs) -> np.ndarray:
        """Reference-compatible solve: complex128 (Proj, A, S, *effective_shape) on the host."""
        u = self.solve_device(n=n, mode=mode, rhs_block=rhs_block, probes=probes)
        u = host_complex128(u.transpose(-1, -2))
        return u.reshape(self.num_projections, self.num_angles, self.num_probes, *self.effective_shape)

    def _solve_single_probe_impl(self, scan_idx: int = 0, proj_idx: int = 0, probe=None, mode: str = "forward",
                                 rhs_block=None) -> np.ndarray:
        N, L = self.block_size, self.nz - 1
        pr = np.zeros(N, dtype=complex) if probe is None else np.asarray(probe, dtype=complex).reshape(N)
        cache = self.projection_cache[proj_idx].modes[mode]
        fac = cache.factors
        prt = self._probe_tensor(pr[None])
        u0 = torch.zeros_like(prt)
        u = torch.empty((1, self.nz, N), dtype=torch.complex64, device=self.device)
        rows = self._rhs_rows(rhs_block)[None] if rhs_block is not None else None
        if mode == "adjoint":
            if rows is not None:
                rhs = torch.flip(rows, dims=[1]).contiguous()
            else:
                rhs = torch.zeros((1, L, N), dtype=torch.complex64, device=self.device)
                rhs[:, L - 1] = self._b0(prt)
            self._launch([self._spec(fac, u0, out=u, rhs=rhs)], probes_per_block=1)
            u[:, 1:] = torch.flip(u[:, 1:].clone(), dims=[1])
        elif rows is not None:
            self._launch([self._spec(fac, u0, out=u, rhs=rows)], probes_per_block=1)
        else:
            n = cache.cached_n
            C1 = self.object_contribution_device(n, scan_idx)[:, 0]
            rhs = (-(C1[None, :] * prt.to(torch.complex128))).to(torch.complex64)[:, None].contiguous()
            self._launch([self._spec(fac, prt, out=u, rhs=rhs, rhs_steps=1)], probes_per_block=1)
        u[:, 0] = prt
        return u[0].transpose(0, 1).cpu().numpy().astype(np.complex128)
