"""Host-side assembly of the CN step operators as diagonals.

The per-slice systems are A_j = A0 - D_j and B_j = B0 + D_j with
A0 = I - K, B0 = I + K.  K is the scaled second difference
tridiag(mu_x, -2(mu_x + mu_y), mu_x) with the boundary rules of
thick_ptycho/forward_model/pwe/operators/finite_differences/boundary_conditions.py:57-94
applied through operator_matrices.py:59-72:

  dirichlet                 K[0,0] = K[0,1] = K[N-1,N-1] = K[N-1,N-2] = 0
  neumann (and any value    K[0,1] *= 2, K[N-1,N-2] *= 2
    that is not one of the
    strings, quirk Q1)
  impedance                 neumann, then K[0,0] += beta_x, K[N-1,N-1] += beta_x
  impedance2                neumann, then K[0,0] += beta_x + gamma/dz^2, K[0,1] -= gamma/dz^2
                            (mirrored at the last row)

with mu_x = (i / 2k) dz / (2 dx^2), beta_x = 2 i k dx mu_x, gamma = -1 / (2 i k)
(boundary_conditions.py:9-55).  Diagonals are stored as three length-N arrays
(sub[i] couples row i to column i-1, sup[i] couples row i to i+1).
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum

import numpy as np


class BoundaryType(Enum):
    """Boundary condition names (thick_ptycho/forward_model/pwe/operators/utils.py:4-18).

    A plain Enum, like the reference: passing a member instead of its string
    value selects the Neumann rule on one axis and raises on two (quirk Q1)."""

    DIRICHLET = "dirichlet"
    NEUMANN = "neumann"
    IMPEDANCE = "impedance"
    IMPEDANCE2 = "impedance2"

    @classmethod
    def list(cls):
        return [bc.value for bc in cls]


@dataclass
class Tridiag:
    sub: np.ndarray
    diag: np.ndarray
    sup: np.ndarray

    def conj_transpose(self) -> "Tridiag":
        n = self.diag.shape[0]
        sub = np.zeros(n, dtype=complex)
        sup = np.zeros(n, dtype=complex)
        sub[1:] = np.conj(self.sup[:-1])
        sup[:-1] = np.conj(self.sub[1:])
        return Tridiag(sub, np.conj(self.diag), sup)


def bc_scalars(k: float, dx: float, dz: float, bc_type, dy: float | None = None) -> dict:
    a = 1j / (2 * k)
    mu_x = a * (0.5 * dz / dx**2)
    mu_y = a * (0.5 * dz / dy**2) if dy is not None else 0.0
    out = {"mu_x": mu_x, "mu_y": mu_y}
    if bc_type == "impedance" or bc_type == "impedance2":
        out["beta_x"] = mu_x * 2 * 1j * k * dx
        out["beta_y"] = mu_y * 2 * 1j * k * dy if dy is not None else 0.0
    if bc_type == "impedance2":
        out["gamma"] = -1 / (2j * k)
    return out


def laplacian_1d(n: int, k: float, dx: float, dz: float, bc_type, mu_y=0.0) -> Tridiag:
    """K with the one-axis boundary rules (see module docstring)."""
    c = bc_scalars(k, dx, dz, bc_type)
    mu_x = c["mu_x"]
    sub = np.full(n, mu_x, dtype=complex)
    sup = np.full(n, mu_x, dtype=complex)
    diag = np.full(n, -2 * (mu_x + mu_y), dtype=complex)
    sub[0] = 0
    sup[-1] = 0
    if bc_type == "dirichlet":
        sup[0] = 0
        sub[-1] = 0
        diag[0] = 0
        diag[-1] = 0
        return Tridiag(sub, diag, sup)
    sup[0] *= 2
    sub[-1] *= 2
    if bc_type == "impedance":
        diag[0] += c["beta_x"]
        diag[-1] += c["beta_x"]
    elif bc_type == "impedance2":
        g = c["gamma"] / dz**2
        diag[0] += c["beta_x"] + g
        sup[0] += -g
        diag[-1] += c["beta_x"] + g
        sub[-1] += -g
    return Tridiag(sub, diag, sup)


def step_operators_1d(n: int, k: float, dx: float, dz: float, bc_type):
    """(A0, B0) = (I - K, I + K) as diagonals."""
    K = laplacian_1d(n, k, dx, dz, bc_type)
    A = Tridiag(-K.sub, 1 - K.diag, -K.sup)
    B = Tridiag(K.sub.copy(), 1 + K.diag, K.sup.copy())
    return A, B


@dataclass
class ModePlan:
    """What the factor kernel needs for one propagation mode."""

    lhs: Tridiag
    rhs: Tridiag
    lhs_sign: float
    rhs_sign: float
    conj_c: int
    lhs_col0: int
    rhs_col0: int
    col_step: int
    b_sign: float = 1.0
    sigma: bool = False  # rhs diagonal + lhs diagonal differs from 2 (different slices)

    def check_complementary(self) -> None:
        """The kernels use rhs = diag(sigma) - lhs: the off-diagonals must cancel."""
        if not (np.allclose(self.rhs.sub[1:], -self.lhs.sub[1:], rtol=0, atol=1e-14 * (1 + np.abs(self.lhs.sub).max()))
                and np.allclose(self.rhs.sup[:-1], -self.lhs.sup[:-1], rtol=0,
                                atol=1e-14 * (1 + np.abs(self.lhs.sup).max()))):
            raise ValueError("rhs off-diagonals are not the negated lhs off-diagonals")
        if not self.sigma and not np.allclose(self.lhs.diag + self.rhs.diag, 2.0, rtol=0, atol=1e-12):
            raise ValueError("lhs + rhs diagonals differ from 2; factor with sigma")

    def base_array(self) -> np.ndarray:
        """complex128 [6][N]: lhs sub/diag/sup, rhs sub/diag/sup."""
        return np.stack([self.lhs.sub, self.lhs.diag, self.lhs.sup, self.rhs.sub, self.rhs.diag, self.rhs.sup])


MODES = ("forward", "adjoint", "reverse", "exact_adjoint")


def mode_plan(A: Tridiag, B: Tridiag, mode: str, L: int) -> ModePlan:
    """Per-mode operators following iterative_lu.py:90-128.

    forward:        lhs = A0 - D_s, rhs = B0 + D_s, D_s = diag(C[:, s-1])
    adjoint:        lhs = (A0 - D)^H, rhs = (B0 + D)^H with D = diag(C[:, L-s])
                    (the reference form, quirk Q3)
    reverse:        lhs = B0 + C[:, L-s], rhs = A0 - C[:, L-s], b -> -b
    exact_adjoint:  lhs = (A0 - D_{L-s+1})^H, rhs = (B0 + D_{L-s+2})^H; this is
                    the exact transpose of the forward map (used for gradients).
    """
    if mode == "forward":
        return ModePlan(A, B, -1.0, 1.0, 0, 0, 0, 1)
    if mode == "adjoint":
        return ModePlan(A.conj_transpose(), B.conj_transpose(), -1.0, 1.0, 1, L - 1, L - 1, -1)
    if mode == "reverse":
        return ModePlan(B, A, 1.0, -1.0, 0, L - 1, L - 1, -1, b_sign=-1.0)
    if mode == "exact_adjoint":
        return ModePlan(A.conj_transpose(), B.conj_transpose(), -1.0, 1.0, 1, L - 1, L, -1, sigma=True)
    raise AssertionError(f"Invalid mode: {mode!r}")


def chunk_for(n: int) -> int:
    """Rows per GPU lane of the one-wavefront kernels: the smallest power of two
    with 64 * chunk >= n (n <= 2048)."""
    for c in (1, 2, 4, 8, 16, 32):
        if 64 * c >= n:
            return c
    raise ValueError(
        f"transverse size {n} exceeds the 2048 rows one wavefront holds; use layout_for(n) (several wavefronts "
        "per probe)"
    )


WIDE_MAX_ROWS = 64 * 16 * 16  # 16 wavefronts of 64 lanes x 16 rows


def layout_for(n: int, wide_chunk: int = 16) -> tuple:
    """(chunk, waves) for the 2D stepping kernels: waves = 1 is one wavefront per
    probe (n <= 2048, chunk from chunk_for); above that each probe's rows are
    split over `waves` wavefronts of one workgroup with `wide_chunk` (8 or 16)
    rows per lane, up to 16 wavefronts (16384 rows at chunk 16)."""
    if n <= 64 * 32:
        return chunk_for(n), 1
    if wide_chunk not in (8, 16):
        raise ValueError(f"wide_chunk must be 8 or 16 (got {wide_chunk})")
    waves = -(-n // (64 * wide_chunk))
    if waves > 16:
        if wide_chunk == 8:
            return layout_for(n, 16)
        raise ValueError(
            f"transverse size {n} exceeds the {WIDE_MAX_ROWS} rows of the widest 2D stepping layout "
            "(16 wavefronts x 64 lanes x 16 rows)"
        )
    return wide_chunk, waves
