"""Drop-in ``PWEIterativeLUSolver`` running on the MI355X kernels.

Mirrors thick_ptycho/forward_model/pwe/solvers/iterative_lu.py:15-197:
  _construct_solve_cache  -> tp_tri_factor (one transverse axis) or tp_band_factor
                             (two axes); all slices, FP64 factorisation, one launch
  _solve_single_probe_impl -> tp_tri_sweep / tp_band_sweep with one probe
  solve                    -> one persistent sweep over every probe/angle/projection
"""

from __future__ import annotations

import time
from dataclasses import dataclass, fields
from typing import Optional

import numpy as np
import scipy.sparse as sp
import torch

from . import band_engine, tri_engine
from .base_solver import BasePWESolver, host_complex128
from .operators import BoundaryType, layout_for, step_operators_1d


class PWEForwardModel:
    """The parts of the reference's ``PWEForwardModel`` (forward_model.py:9-209)
    that callers use: cached step operators and the gradient operator."""

    def __init__(self, simulation_space, bc_type="impedance"):
        self.simulation_space = simulation_space
        self.bc_type = bc_type
        self.block_size = simulation_space.block_size
        ss = simulation_space
        if ss.dimension == 2:
            self.A, self.B = step_operators_1d(ss.effective_shape[0], ss.k, ss.dx, ss.dz, bc_type)
        elif ss.dimension == 3:
            nx, ny = ss.effective_shape[0], ss.effective_shape[1]
            self.A, self.B = band_engine.step_operators_3d(nx, ny, ss.k, ss.dx, ss.dy, ss.dz, bc_type)
        else:
            raise ValueError("Unsupported dimension")
        self._cache = None

    @staticmethod
    def _to_sparse(t) -> sp.csr_matrix:
        if isinstance(t, band_engine.Block3D):
            return t.to_sparse()
        n = t.diag.shape[0]
        return sp.diags([t.sub[1:], t.diag, t.sup[:-1]], offsets=[-1, 0, 1], shape=(n, n), format="csr")

    def _get_or_generate_step_matrices(self, probe=None):
        if self._cache is None:
            self._cache = (self._to_sparse(self.A), self._to_sparse(self.B), np.zeros(self.block_size, dtype=complex))
        return self._cache

    def reset_cache(self):
        self._cache = None

    def setup_inhomogeneous_forward_model(self, n=None, grad: bool = False, scan_index: int = 0):
        eo = self.simulation_space.create_object_contribution(n=n, grad=grad, scan_index=scan_index)
        return sp.diags(eo.T.flatten()) + sp.diags([eo[..., 1:].T.flatten()], offsets=[-self.block_size])


@dataclass
class PWEIterativeLUSolverCache:
    """Per (projection, mode) cache: the device factors of every slice."""

    cached_n: Optional[np.ndarray] = None
    factors: Optional[object] = None  # tri_engine.TriFactors or band_engine.BandFactors
    b: Optional[np.ndarray] = None

    def reset(self, n=None):
        for f in fields(self):
            setattr(self, f.name, None)
        self.cached_n = n


class PWEIterativeLUSolver(BasePWESolver):
    """Slice-by-slice CN propagation with per-slice LU factors on the GPU."""

    solver_cache_class = PWEIterativeLUSolverCache

    def __init__(self, simulation_space, ptycho_probes, bc_type=BoundaryType.IMPEDANCE, results_dir="",
                 use_logging=False, verbose=False, log=None, test_bcs=None, device=None, probes_per_block=None):
        super().__init__(simulation_space, ptycho_probes, bc_type=bc_type, results_dir=results_dir,
                         use_logging=use_logging, verbose=verbose, log=log, test_bcs=test_bcs, device=device)
        self.pwe_finite_differences = PWEForwardModel(simulation_space, bc_type=bc_type)
        self.probes_per_block = probes_per_block
        self.dimension = simulation_space.dimension
        # 2D stepping layout: rows per lane and wavefronts per probe (> 1 above 2048 rows)
        self.chunk, self.waves = layout_for(self.block_size) if self.dimension == 2 else (None, None)

    # ------------------------------------------------------------------ factors
    def object_contribution_device(self, n, scan_idx: int = 0) -> torch.Tensor:
        C = self.simulation_space.create_object_contribution(n=n, scan_index=scan_idx).reshape(-1, self.nz - 1)
        return torch.from_numpy(np.ascontiguousarray(C, dtype=np.complex128)).to(self.device)

    def _construct_solve_cache(self, n=None, mode: str = "forward", scan_idx: int = 0, proj_idx: int = 0) -> None:
        t0 = time.time()
        fm = self.pwe_finite_differences
        C = self.object_contribution_device(n, scan_idx)
        if self.dimension == 3:
            fac = band_engine.factor(C, fm.A, fm.B, mode)
        else:
            fac = self._tri_factor(C, mode)
        cache = self.projection_cache[proj_idx].modes[mode]
        cache.factors = fac
        cache.b = np.zeros(self.block_size, dtype=complex) * fac.b_sign
        cache.cached_n = n
        self._timed(f"factored {self.nz - 1} slices ({mode})", t0)

    def _tri_factor(self, C, mode: str, out=None):
        fm = self.pwe_finite_differences
        return tri_engine.factor(C, fm.A, fm.B, mode, chunk=self.chunk, waves=self.waves if self.waves > 1 else None,
                                 out=out)

    # ------------------------------------------------------------------ rhs handling
    def _rhs_steps(self, mode: str, rhs_block, per_probe: bool) -> Optional[torch.Tensor]:
        """Per-step extra RHS in stepping order, complex64 (P or 1, L, N).

        The reference uses rhs_block[:, j-1] at forward step j and
        rhs_block[:, -j] in adjoint mode (iterative_lu.py:186-187); test_bcs
        replaces b at every step (:189-190)."""
        L = self.nz - 1
        N = self.block_size
        if self.test_bcs is not None:
            b = np.stack([np.asarray(self.test_bcs.test_exact_impedance_rhs_step(j)).reshape(N) for j in range(1, L + 1)])
            return torch.from_numpy(b.astype(np.complex64)[None]).to(self.device)
        if rhs_block is None:
            return None
        if isinstance(rhs_block, torch.Tensor):
            r = rhs_block.to(self.device, torch.complex64)
        else:
            r = torch.from_numpy(np.ascontiguousarray(rhs_block)).to(self.device, torch.complex64)
        if not per_probe:
            r = r.reshape(1, N, L)
        r = r.transpose(-1, -2)  # (P, L, N), column j-1 -> step j
        if mode in ("adjoint", "exact_adjoint"):
            r = torch.flip(r, dims=[1])
        return r.contiguous()

    def _probe_tensor(self, probes) -> torch.Tensor:
        if isinstance(probes, torch.Tensor):
            t = probes.to(self.device, torch.complex64)
        else:
            t = torch.from_numpy(np.ascontiguousarray(probes)).to(self.device, torch.complex64)
        return t.reshape(-1, self.block_size).contiguous()

    # ------------------------------------------------------------------ single probe
    def _solve_single_probe_impl(self, scan_idx: int = 0, proj_idx: int = 0, probe=None, mode: str = "forward",
                                 rhs_block=None) -> np.ndarray:
        fac = self.projection_cache[proj_idx].modes[mode].factors
        N = self.block_size
        if probe is None:
            probe = np.zeros((N,), dtype=complex)
        u0 = self._probe_tensor(np.asarray(probe).reshape(1, N))
        out = torch.empty((1, self.nz, N), dtype=torch.complex64, device=self.device)
        spec = self._spec(fac, u0, out=out, rhs=self._rhs_steps(mode, rhs_block, False), flip=(mode == "adjoint"))
        self._launch([spec], probes_per_block=1)
        return out[0].transpose(0, 1).cpu().numpy().astype(np.complex128)

    # ------------------------------------------------------------------ engine dispatch
    def _spec(self, fac, u0, **kw):
        if self.dimension == 3:
            return band_engine.BandScanSpec(fac, u0, **kw)
        return tri_engine.TriScanSpec(fac, u0, **kw)

    def _launch(self, specs, probes_per_block=None, plain: bool = False) -> None:
        if self.dimension == 3:
            # two probes per workgroup share each band-coefficient load once the
            # launch has enough probes to keep most CUs busy with pairs
            total = sum(int(sp.u0.shape[0]) for sp in specs)
            band_engine.sweep(specs, probes_per_block=2 if total >= 384 else 1)
        elif plain:
            tri_engine.sweep_plain(specs)
        else:
            if probes_per_block == "auto":
                from .. import autotune

                probes_per_block = autotune.tune_sweep(specs)["ppb"]
            tri_engine.sweep(specs, probes_per_block=probes_per_block)

    # ------------------------------------------------------------------ batched
    def solve_device(self, n=None, mode: str = "forward", rhs_block=None, probes=None, keep: str = "all",
                     per_probe_rhs: bool = False, plain: bool = False) -> torch.Tensor:
        """Every probe of every angle and projection in one launch.

        Returns complex64 on the GPU: (Proj, A, S, nz, N) for keep='all' (plane
        index first; ``.transpose(-1, -2)`` gives the reference's (..., N, nz)
        view) or (Proj, A, S, N) for keep='exit'.  ``plain=True`` uses one
        launch per slice instead of the persistent kernel."""
        assert mode in {"forward", "adjoint", "reverse", "exact_adjoint"}, f"Invalid mode: {mode!r}"
        if n is None:
            n = self.simulation_space.refractive_index_empty
        if probes is None:
            probes = self.probes
        N = self.block_size
        A = self.num_angles
        S = self.num_probes
        pr = self._probe_tensor(probes).reshape(A, S, N)
        rhs = self._rhs_steps(mode, rhs_block, per_probe_rhs)
        shape = (self.num_projections, A, S, self.nz if keep == "all" else 1, N)
        out = torch.empty(shape, dtype=torch.complex64, device=self.device)
        if self.solve_reduced_domain:
            self._solve_reduced(n, mode, pr, rhs, per_probe_rhs, out, keep, plain)
            return out if keep == "all" else out[..., 0, :]
        specs = []
        self._last_factors = []
        for p in range(self.num_projections):
            fac = self._get_or_construct_cache(n=n, mode=mode, proj_idx=p).factors
            self._last_factors.append(fac)
            for a in range(A):
                u0 = pr[a]
                if keep == "all":
                    specs.append(self._spec(fac, u0, out=out[p, a], rhs=rhs, flip=(mode == "adjoint")))
                else:
                    specs.append(self._spec(fac, u0, exit=out[p, a, :, 0], rhs=rhs))
        t0 = time.time()
        for i in range(0, len(specs), 8):
            self._launch(specs[i:i + 8], probes_per_block=self.probes_per_block, plain=plain)
        if self.verbose:
            torch.cuda.synchronize()
            self._timed(f"stepped {len(specs) * S} probes x {self.nz - 1} slices ({mode})", t0)
        return out if keep == "all" else out[..., 0, :]

    def _solve_reduced(self, n, mode, pr, rhs, per_probe_rhs, out, keep, plain) -> None:
        """Reduced domain: scan s steps through its own window of the object, so
        every (projection, scan) has its own factors, rebuilt on every call as in
        the reference (_get_or_construct_cache, pwe/solvers/base_solver.py:114).
        The factorisations run eight per launch (tp_tri_factor_batch) into one
        buffer of per-scan factor sets, and one persistent sweep per (projection,
        angle) steps every scan through its own set (tp_tri_scan.coef_pstride)."""
        fm = self.pwe_finite_differences
        S, L = self.num_probes, self.nz - 1
        per = tri_engine.coef_elems(L, self.chunk, self.waves)
        wide = self.waves > 1
        t0 = time.time()
        windows = self._reduced_windows()
        self._last_factors = []
        ss = self.simulation_space
        device_windows = (windows is not None and not getattr(ss, "exact_ref_coeff", False)
                          and getattr(ss, "empty_space_px", 0) == 0)
        for p in range(self.num_projections):
            n_p = self.get_projected_obj(n=n, proj_idx=p)
            if device_windows:  # C of the whole object once; a window's C is a block of its rows
                nt = torch.as_tensor(np.ascontiguousarray(n_p)).to(self.device, torch.complex128)
                coef_n = (ss.k / 1j) * (nt - 1)
                C_full = ((ss.dz / 2) * (coef_n[:, :-1] + coef_n[:, 1:]) / 2).contiguous()

                def contribution(s):
                    a, b = windows[s]
                    return C_full[a:b]
            else:
                def contribution(s):
                    return self.object_contribution_device(n_p, s)
            coef = torch.empty((S, per), dtype=torch.complex64, device=self.device)
            status = tri_engine.new_status(self.device)  # shared: the sweep steps all windows' sets together
            fac = None
            for i in range(0, S, 8):
                jobs = [(contribution(s), fm.A, fm.B, mode) for s in range(i, min(i + 8, S))]
                facs = tri_engine.factor_many(jobs, self.chunk, out=[coef[s] for s in range(i, min(i + 8, S))],
                                              waves=self.waves if wide else None, status=status)
                fac = fac or facs[0]
                self._last_factors += facs
            whole = tri_engine.TriFactors(coef.reshape(-1), fac.n, L, self.chunk, mode, fac.b_sign, fac.status,
                                          fac.sigma, keep=[coef], waves=self.waves, wide=wide, plan=fac.plan)
            if plain:  # the plain form shares factors within a scan: one scan per window
                views = [tri_engine.TriFactors(coef[s], fac.n, L, self.chunk, mode, fac.b_sign, fac.status, fac.sigma,
                                               waves=self.waves, wide=wide)
                         for s in range(S)]
            for a in range(self.num_angles):
                kw = {}
                if keep == "all":
                    kw["out"], kw["flip"] = out[p, a], mode == "adjoint"
                else:
                    kw["exit"] = out[p, a, :, 0]
                if plain:
                    specs = []
                    for s in range(S):
                        r = rhs[s:s + 1] if (rhs is not None and per_probe_rhs) else rhs
                        k2 = {key: (v[s:s + 1] if isinstance(v, torch.Tensor) else v) for key, v in kw.items()}
                        specs.append(self._spec(views[s], pr[a, s:s + 1], rhs=r, **k2))
                    for i in range(0, S, 8):
                        self._launch(specs[i:i + 8], plain=True)
                else:
                    spec = self._spec(whole, pr[a], rhs=rhs, coef_pstride=per, **kw)
                    self._launch([spec], probes_per_block=1)
        if self.verbose:
            torch.cuda.synchronize()
            self._timed(f"stepped {self.num_projections * self.num_angles * S} reduced-domain probes x {L} "
                        f"slices ({mode})", t0)

    def _reduced_windows(self):
        """[(xmin, xmax)] per scan: this package's SimulationSpace keeps them in
        _scan_windows, the reference's in _scan_frame_info[s].reduced_limits_discrete."""
        ss = self.simulation_space
        w = getattr(ss, "_scan_windows", None)
        if w:
            return list(w)
        frames = getattr(ss, "_scan_frame_info", None)
        if frames:
            return [tuple(int(v) for v in f.reduced_limits_discrete.x) for f in frames]
        return None

    def solve(self, n=None, mode: str = "forward", rhs_block=None, probes=None, **kwargs) -> np.ndarray:
        """Reference-compatible solve (base/base_solver.py:84-154): complex128
        (Proj, A, S, *effective_shape) on the host."""
        assert mode in {"forward", "adjoint", "reverse"}, f"Invalid mode: {mode!r}"
        u = self.solve_device(n=n, mode=mode, rhs_block=rhs_block, probes=probes, keep="all")
        tri_engine.check_pivots(getattr(self, "_last_factors", []))  # splu raises on a singular slice
        u = host_complex128(u.transpose(-1, -2))
        return u.reshape(self.num_projections, self.num_angles, self.num_probes, *self.effective_shape)
