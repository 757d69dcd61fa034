"""A duck-typed simulation space carrying the attributes the solvers read.

The reference solvers read a fixed set of attributes and one method from
their ``simulation_space`` (BaseForwardModelSolver.__init__
base/base_solver.py:24-60, BoundaryConditions.__init__ boundary_conditions.py:9-36,
MSForwardModelSolver.__init__ ms_solver.py:11-44, and
BaseSimulationSpace.create_object_contribution base_simulation_space.py:315-383).
``SimulationSpace`` provides exactly that contract for arbitrary grid sizes
(the reference's own SimulationSpace2D forces an odd nx for even scan steps),
and the drop-in solvers also accept the reference's own simulation-space
objects.
"""

from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np


class SimulationSpace:
    def __init__(
        self,
        nx: int,
        nz: int,
        dx: float,
        dz: float,
        wavelength: float = 1.0,
        ny: Optional[int] = None,
        dy: Optional[float] = None,
        num_probes: int = 1,
        probe_angles: Sequence[float] = (0.0,),
        n_medium: complex = 1.0,
        exact_ref_coeff: bool = False,
        empty_space_px: int = 0,
        probe_type: str = "gaussian",
        results_dir: Optional[str] = None,
        reduced_windows: Optional[Sequence[Tuple[int, int]]] = None,
    ):
        self.dimension = 2 if ny is None else 3
        self.nx = nx
        self.ny = ny if ny is not None else 1
        self.nz = nz
        self.dx = float(dx)
        self.dy = float(dy if dy is not None else dx)
        self.dz = float(dz)
        self.wave_length = wavelength
        self.k = 2 * np.pi / wavelength
        self.n_medium = complex(n_medium)
        self.exact_ref_coeff = exact_ref_coeff
        self.empty_space_px = empty_space_px
        self.probe_type = probe_type
        self.results_dir = results_dir
        # reduced domain: scan s sees the object rows n[xmin_s:xmax_s] (the
        # reference's _get_reduced_sample, simulation_space_2d.py:132-135)
        self.solve_reduced_domain = reduced_windows is not None
        self._scan_windows: List[Tuple[int, int]] = [tuple(map(int, w)) for w in (reduced_windows or [])]
        if self.solve_reduced_domain:
            if ny is not None:
                raise NotImplementedError("reduced-domain solves are built for one transverse axis")
            if len(self._scan_windows) != num_probes:
                raise ValueError("reduced_windows needs one (xmin, xmax) window per scan position")
            widths = {b - a for a, b in self._scan_windows}
            if len(widths) != 1 or min(a for a, _ in self._scan_windows) < 0 or \
                    max(b for _, b in self._scan_windows) > nx:
                raise ValueError("reduced windows must share one width and lie inside [0, nx]")
        self.num_projections = 1
        self.probe_angles = tuple(probe_angles)
        self.num_angles = len(self.probe_angles)
        self.num_probes = num_probes
        self.scan_points = num_probes
        self.total_scans = self.num_angles * self.num_probes * self.num_projections
        self.effective_nx = (self._scan_windows[0][1] - self._scan_windows[0][0] if self.solve_reduced_domain
                             else nx) + 2 * empty_space_px
        self.effective_ny = self.ny
        self.z = np.arange(nz) * self.dz
        self.x = np.arange(nx) * self.dx
        if self.dimension == 2:
            self.shape = (nx, nz)
            self.effective_shape = (self.effective_nx, nz)
            self.block_size = self.effective_nx
        else:
            self.y = np.arange(self.ny) * self.dy
            self.shape = (nx, self.ny, nz)
            self.effective_shape = (nx, self.ny, nz)
            self.block_size = nx * self.ny
        self.refractive_index_empty = np.ones(self.shape, dtype=complex) * self.n_medium

    def _get_reduced_sample(self, n, scan_index):
        if not self.solve_reduced_domain:
            return n
        x_min, x_max = self._scan_windows[scan_index]
        return n[x_min:x_max, :]

    @staticmethod
    def reference_scan_windows(nx: int, scan_points: int, step_size: int, pad_discrete: int) -> List[Tuple[int, int]]:
        """The reduced windows of SimulationSpace2D._generate_scan_frames
        (simulation_space_2d.py:80-130) for scan_points > 1: window width
        W = step + pad, centres round(mid + (i - (S-1)/2) step), xmin = max(c - W//2, 0),
        xmax = min(xmin + W - 1, nx - 1); the sample is n[xmin:xmax] (W - 1 rows)."""
        W = step_size + pad_discrete
        half, mid = W // 2, (nx - 1) / 2.0
        out = []
        for i in range(scan_points):
            cx = int(round(mid + (i - (scan_points - 1) / 2) * step_size))
            xmin = max(cx - half, 0)
            out.append((xmin, min(xmin + W - 1, nx - 1)))
        return out

    def create_object_contribution(self, n=None, grad=False, scan_index=0) -> np.ndarray:
        """Half-step object coefficients C[..., j-1] = (dz/2)(coef[..., j-1] + coef[..., j])/2.

        coef = (k/i)(n - 1), or (k/2i)(n^2 - 1) with exact_ref_coeff; with
        grad=True coef = k/i (or (k/i) n).  Same semantics as
        base_simulation_space.py:315-383."""
        if n is None:
            n = self.refractive_index_empty
        n = np.asarray(n)
        if self.solve_reduced_domain:
            n = self._get_reduced_sample(n, scan_index)
        if grad:
            coef = (self.k / 1j) * np.ones_like(n) if not self.exact_ref_coeff else (self.k / 1j) * n
        else:
            coef = (self.k / 1j) * (n - 1) if not self.exact_ref_coeff else (self.k / 2j) * (n**2 - 1)
        if self.empty_space_px > 0:
            coef = np.pad(coef, ((self.empty_space_px, self.empty_space_px), (0, 0)), mode="constant",
                          constant_values=0.0)
        return (self.dz / 2) * (coef[..., :-1] + coef[..., 1:]) / 2


def space_for_scan(scan, num_probes: Optional[int] = None) -> SimulationSpace:
    """SimulationSpace matching a ``workloads.Scan2D`` / ``Scan3D``."""
    if hasattr(scan, "ny"):
        return SimulationSpace(scan.nx, scan.nz, scan.dx, scan.dz, wavelength=2 * np.pi / scan.k, ny=scan.ny,
                               dy=scan.dy, num_probes=num_probes or scan.probes.shape[0])
    return SimulationSpace(scan.nx, scan.nz, scan.dx, scan.dz, wavelength=2 * np.pi / scan.k,
                           num_probes=num_probes or scan.probes.shape[0], n_medium=getattr(scan, "n_medium", 1.0))
