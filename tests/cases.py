"""Shared test cases: the reference's analytic convergence problems.

Follows tests/test_forward_model/pwe/test_boundary_conditions_2d.py:24-106 and
test_boundary_conditions_3d.py:22-140 of the reference: k = 100, x in (0, 1)
(y in (0, 1)), z in (0, 2), one probe equal to the analytic solution at z = 0,
grids nx = step + 1 (the reference's parity rule turns step 16 into 17
points), nz = step.
"""

from __future__ import annotations

import numpy as np

from oracle import pwe

K_TEST = 100.0


def _a():
    return 1j / (2 * K_TEST)


def exact_2d(bc: str, X, Z):
    a = _a()
    if bc in ("neumann", "impedance"):
        modes = [(1, 1.0), (2, 0.5), (5, 0.2)]
        f = np.cos
    else:
        modes = [(1, 1.0), (5, 0.5), (9, 0.2)]
        f = np.sin
    return sum(w * np.exp(-a * n**2 * np.pi**2 * Z) * f(n * np.pi * X) for n, w in modes)


def exact_3d(bc: str, X, Y, Z):
    a = _a()
    if bc in ("neumann", "impedance"):
        modes = [(1, 1.0), (2, 0.5), (5, 0.2)]
        f = np.cos
    else:
        modes = [(1, 1.0), (5, 0.5), (9, 0.2)]
        f = np.sin
    u = sum(w * np.exp(-a * 2 * n**2 * np.pi**2 * Z) * f(n * np.pi * X) * f(n * np.pi * Y) for n, w in modes)
    if bc == "dirichlet":
        u[:, 0, :] = 0
        u[:, -1, :] = 0
        u[0, :, :] = 0
        u[-1, :, :] = 0
    return u


def convergence_case_2d(step: int, bc: str):
    nx = step + 1
    nz = step
    x = np.linspace(0, 1, nx)
    z = np.linspace(0, 2, nz)
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=x[1] - x[0], dz=z[1] - z[0], k=K_TEST)
    X, Z = np.meshgrid(x, z, indexing="ij")
    exact = exact_2d(bc, X, Z)
    probe = exact[:, 0].copy()
    test_bcs = pwe.exact_impedance_rhs(g, x, z) if bc == "impedance" else None
    n = np.ones((nx, nz), dtype=complex)
    return g, n, probe, exact, test_bcs, x, z


def convergence_case_3d(step: int, bc: str):
    nx = step + 1
    nz = step
    x = np.linspace(0, 1, nx)
    z = np.linspace(0, 2, nz)
    g = pwe.Grid(dimension=3, nx=nx, ny=nx, nz=nz, dx=x[1] - x[0], dy=x[1] - x[0], dz=z[1] - z[0], k=K_TEST)
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
    exact = exact_3d(bc, X, Y, Z)
    probe = exact[:, :, 0].copy()
    test_bcs = pwe.exact_impedance_rhs(g, x, z, y=x) if bc == "impedance" else None
    n = np.ones((nx, nx, nz), dtype=complex)
    return g, n, probe, exact, test_bcs


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def random_object(nx: int, nz: int, rng, amp=1e-3) -> np.ndarray:
    """Smooth-ish absorbing object n = 1 - (delta + i beta)."""
    d = amp * rng.random((nx, nz))
    b = 1e-2 * amp * rng.random((nx, nz))
    return 1.0 - (d + 1j * b)
