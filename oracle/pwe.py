"""FP64 restatement of the reference's PWE (paraxial wave equation) path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Never imported by the
product package.

Everything here is complex128 numpy/scipy and follows the reference files
cited per function (paths relative to the reference repository root):

* operator assembly: ``thick_ptycho/forward_model/pwe/operators/finite_differences/
  boundary_conditions.py`` and ``operator_matrices.py``
* object term: ``thick_ptycho/simulation/simulation_space/base_simulation_space.py``
* slice stepping: ``thick_ptycho/forward_model/pwe/solvers/iterative_lu.py``
* whole-system solve: ``.../pwe/solvers/full_lu.py`` + ``.../finite_differences/forward_model.py``
* space-z mat-vec: ``.../pwe/solvers/utils/_pint_utils.py``
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

MODES = ("forward", "adjoint", "reverse")


@dataclass
class Grid:
    """Minimal geometry used by the oracle (duck-typed simulation space).

    ``dimension`` follows the reference convention: 2 = one transverse axis,
    3 = two transverse axes.  ``nx``/``ny`` are the transverse sizes of the
    solved (effective) domain.
    """

    dimension: int
    nx: int
    nz: int
    dx: float
    dz: float
    k: float
    ny: int = 1
    dy: float = 1.0
    exact_ref_coeff: bool = False
    empty_space_px: int = 0
    n_medium: float = 1.0

    @property
    def block_size(self) -> int:
        return self.nx * (self.ny if self.dimension == 3 else 1)


# ----------------------------------------------------------------------------
# Boundary-condition coefficients  (boundary_conditions.py:9-55)
# ----------------------------------------------------------------------------


def bc_coefficients(g: Grid, bc_type) -> dict:
    """mu_x, mu_y, beta_x, beta_y, gamma exactly as ``BoundaryConditions.__init__``.

    ``bc_type`` is compared with ``==`` against strings, so a non-string value
    (the reference's plain ``Enum``) matches no branch (quirk Q1)."""
    a = 1j / (2 * g.k)
    mu_x = a * (0.5 * g.dz / g.dx**2)
    mu_y = a * (0.5 * g.dz / g.dy**2) if g.dimension == 3 else 0
    out = {"a": a, "mu_x": mu_x, "mu_y": mu_y}
    if bc_type == "impedance" or bc_type == "impedance2":
        out["beta_x"] = mu_x * 2 * 1j * g.k * g.dx
        out["beta_y"] = mu_y * 2 * 1j * g.k * g.dy if g.dimension == 3 else 0
    if bc_type == "impedance2":
        out["gamma"] = -1 / (2j * g.k)
    return out


# ----------------------------------------------------------------------------
# Operators  (operator_matrices.py:36-126, boundary_conditions.py:57-130)
# ----------------------------------------------------------------------------


def _laplacian_1d(nx: int, mu_x, mu_y) -> sp.lil_matrix:
    # operator_matrices.py:36-45
    e = np.ones(nx, dtype=complex)
    return sp.diags(
        [mu_x * e, -2 * (mu_x + mu_y) * e, mu_x * e],
        offsets=[-1, 0, 1],
        shape=(nx, nx),
        dtype=complex,
    ).tolil()


def _apply_bc_1d(K: sp.lil_matrix, bc_type, c: dict, dz: float) -> sp.lil_matrix:
    # operator_matrices.py:59-72 ; boundary_conditions.py:57-94
    if bc_type == "dirichlet":
        K[0, 1] = 0
        K[-1, -2] = 0
        K[0, 0] = 0
        K[-1, -1] = 0
        return K
    K[0, 1] *= 2
    K[-1, -2] *= 2
    if bc_type == "impedance":
        K[0, 0] += c["beta_x"]
        K[-1, -1] += c["beta_x"]
    elif bc_type == "impedance2":
        K[0, 0] += c["beta_x"] + c["gamma"] / dz**2
        K[0, 1] += -c["gamma"] / dz**2
        K[-1, -1] += c["beta_x"] + c["gamma"] / dz**2
        K[-1, -2] += -c["gamma"] / dz**2
    return K


def step_operators_1d(g: Grid, bc_type, nx: Optional[int] = None):
    """(A0, B0) = (I - K, I + K) for one transverse axis (operator_matrices.py:47-57)."""
    c = bc_coefficients(g, bc_type)
    nx = g.nx if nx is None else nx
    K = _apply_bc_1d(_laplacian_1d(nx, c["mu_x"], c["mu_y"]), bc_type, c, g.dz)
    K = K.tocsr()
    Ix = sp.eye(nx)
    return (Ix - K).tocsr(), (Ix + K).tocsr()


def step_operators_2d(g: Grid, bc_type):
    """(A0, B0) for two transverse axes (operator_matrices.py:75-126).

    Flat index = block * nx + within, with the x-operator (mu_x, beta_x) acting
    inside blocks and mu_y coupling blocks.  Arrays are C-order (ix*ny + iy), so
    the x-operator acts along the array's second axis (quirk Q7)."""
    c = bc_coefficients(g, bc_type)
    nx, ny = g.nx, g.ny
    Ax, Bx = step_operators_1d(g, bc_type, nx)
    Ix = sp.eye(nx)
    Iy = sp.eye(ny)
    T = sp.diags([1, 1], [-1, 1], shape=(ny, ny))
    A = (sp.kron(Iy, Ax) + sp.kron(T, -c["mu_y"] * Ix)).tolil()
    B = (sp.kron(Iy, Bx) + sp.kron(T, c["mu_y"] * Ix)).tolil()
    if bc_type == "dirichlet":
        # boundary_conditions.py:96-109
        Ixd = sp.eye(nx).toarray()
        A[:nx, :nx] = Ixd
        A[-nx:, -nx:] = Ixd
        B[:nx, :nx] = Ixd
        B[-nx:, -nx:] = Ixd
        A[:nx, nx : 2 * nx] = 0
        A[-nx:, -2 * nx : -nx] = 0
        B[:nx, nx : 2 * nx] = 0
        B[-nx:, -2 * nx : -nx] = 0
        return A.tocsr(), B.tocsr()
    # boundary_conditions.py:111-119 (Neumann block couplings x2)
    A[:nx, nx : 2 * nx] = A[:nx, nx : 2 * nx] * 2
    A[-nx:, -2 * nx : -nx] = A[-nx:, -2 * nx : -nx] * 2
    B[:nx, nx : 2 * nx] = B[:nx, nx : 2 * nx] * 2
    B[-nx:, -2 * nx : -nx] = B[-nx:, -2 * nx : -nx] * 2
    if bc_type == "neumann":
        return A.tocsr(), B.tocsr()
    if bc_type == "impedance":
        # boundary_conditions.py:121-130
        by = c["beta_y"] * np.eye(nx)
        A[:nx, :nx] = A[:nx, :nx].toarray() - by
        A[-nx:, -nx:] = A[-nx:, -nx:].toarray() - by
        B[:nx, :nx] = B[:nx, :nx].toarray() + by
        B[-nx:, -nx:] = B[-nx:, -nx:].toarray() + by
        return A.tocsr(), B.tocsr()
    if bc_type == "impedance2":
        # the reference calls a method that does not exist (AttributeError)
        raise AttributeError("impedance2 has no two-axis implementation in the reference")
    raise ValueError(
        "Invalid boundary condition type. Please choose from dirichlet, neumann or dirichlet."
    )


def step_operators(g: Grid, bc_type):
    """Dispatch as ``OperatorMatrices.get_matrix_system`` (operator_matrices.py:16-23)."""
    if g.dimension == 2:
        return step_operators_1d(g, bc_type)
    if g.dimension == 3:
        return step_operators_2d(g, bc_type)
    raise ValueError("Unsupported dimension")


# ----------------------------------------------------------------------------
# Object term  (base_simulation_space.py:315-383)
# ----------------------------------------------------------------------------


def object_contribution(g: Grid, n: np.ndarray, grad: bool = False) -> np.ndarray:
    """C[..., j-1] = (dz/2) * (coef[..., j-1] + coef[..., j]) / 2, coef = (k/i)(n-1)."""
    n = np.asarray(n)
    if grad:
        coef = (g.k / 1j) * np.ones_like(n) if not g.exact_ref_coeff else (g.k / 1j) * n
    else:
        coef = (g.k / 1j) * (n - 1) if not g.exact_ref_coeff else (g.k / 2j) * (n**2 - 1)
    if g.empty_space_px > 0:
        coef = np.pad(
            coef,
            ((g.empty_space_px, g.empty_space_px), (0, 0)),
            mode="constant",
            constant_values=0.0,
        )
    return (g.dz / 2) * (coef[..., :-1] + coef[..., 1:]) / 2


# ----------------------------------------------------------------------------
# Test boundary right-hand sides  (boundary_condition_test.py:35-114)
# ----------------------------------------------------------------------------


def exact_impedance_rhs(g: Grid, x: np.ndarray, z: np.ndarray, y: Optional[np.ndarray] = None):
    """Return f(j) = 0.5 * (b(z[j-1]) + b(z[j])) with the analytic Neumann modes.

    Follows ``BoundaryConditionsTest.test_exact_impedance_rhs_step``."""
    c = bc_coefficients(g, "impedance")
    a = c["a"]

    def u_nm(n, m):
        return lambda xx, yy, zz: (
            np.exp(-a * ((n**2 + m**2) * np.pi**2) * zz)
            * np.cos(n * np.pi * xx)
            * np.cos(m * np.pi * yy)
        )

    if g.dimension == 2:

        def ue(xx, zz):
            return u_nm(1, 0)(xx, 0, zz) + 0.5 * u_nm(2, 0)(xx, 0, zz) + 0.2 * u_nm(5, 0)(xx, 0, zz)

        def b_at(zz):
            ubc = np.zeros(g.nx, dtype=complex)
            ubc[0] -= 2 * c["beta_x"] * ue(x[0], zz)
            ubc[-1] -= 2 * c["beta_x"] * ue(x[-1], zz)
            return ubc

    else:

        def ue(xx, yy, zz):
            return u_nm(1, 1)(xx, yy, zz) + 0.5 * u_nm(2, 2)(xx, yy, zz) + 0.2 * u_nm(5, 5)(xx, yy, zz)

        def b_at(zz):
            ubc = np.zeros((g.nx, g.ny), dtype=complex)
            ubc[0, :] -= 2 * c["beta_x"] * ue(x[0], y, zz)
            ubc[-1, :] -= 2 * c["beta_x"] * ue(x[-1], y, zz)
            ubc[:, 0] -= 2 * c["beta_y"] * ue(x, y[0], zz)
            ubc[:, -1] -= 2 * c["beta_y"] * ue(x, y[-1], zz)
            return ubc.flatten()

    return lambda j: 0.5 * (b_at(z[j - 1]) + b_at(z[j]))


# ----------------------------------------------------------------------------
# Iterative LU slice stepping  (iterative_lu.py:66-197)
# ----------------------------------------------------------------------------


@dataclass
class LUCache:
    A_lu_list: list
    B_list: list
    b: np.ndarray


def build_lu_cache(g: Grid, bc_type, n: np.ndarray, mode: str = "forward", C: Optional[np.ndarray] = None) -> LUCache:
    """``_construct_solve_cache`` + ``_build_lu_factors`` (iterative_lu.py:66-133)."""
    assert mode in MODES, f"Invalid mode: {mode!r}"
    A, B = step_operators(g, bc_type)
    b = np.zeros(g.block_size, dtype=complex)  # operator_matrices.py:143-157 (probe None)
    if mode == "reverse":
        A, B, b = B, A, -b
    if C is None:
        C = object_contribution(g, n)
    C = np.asarray(C).reshape(-1, g.nz - 1)
    A_lu_list, B_list = [], []
    for j in range(1, g.nz):
        if mode == "reverse":
            Cm = -sp.diags(C[:, -j])
        elif mode == "adjoint":
            Cm = sp.diags(C[:, -j])
        else:
            Cm = sp.diags(C[:, j - 1])
        Aw = A - Cm
        Bw = B + Cm
        if mode == "adjoint":
            Aw, Bw = Aw.conj().T, Bw.conj().T
        A_lu_list.append(spla.splu(Aw.tocsc()))
        B_list.append(Bw.tocsr())
    return LUCache(A_lu_list, B_list, b)


def solve_single_probe(
    g: Grid,
    cache: LUCache,
    probe: Optional[np.ndarray],
    mode: str = "forward",
    rhs_block: Optional[np.ndarray] = None,
    test_bcs: Optional[Callable[[int], np.ndarray]] = None,
) -> np.ndarray:
    """``_solve_single_probe_impl`` (iterative_lu.py:138-197). Returns (block_size, nz)."""
    N = g.block_size
    if probe is None:
        probe = np.zeros((N,), dtype=complex)
    u = np.zeros((N, g.nz), dtype=complex)
    u[:, 0] = np.asarray(probe).flatten()
    b = cache.b
    for j in range(1, g.nz):
        if rhs_block is not None:
            b = rhs_block[:, -j] if mode == "adjoint" else rhs_block[:, j - 1]
        if test_bcs is not None:
            b = test_bcs(j)
        rhs = cache.B_list[j - 1] @ u[:, j - 1] + b
        u[:, j] = cache.A_lu_list[j - 1].solve(rhs)
    if mode == "adjoint":
        u = np.flip(u, axis=1)
    return u


def solve_batched(
    g: Grid,
    cache: LUCache,
    probes: np.ndarray,
    mode: str = "forward",
    rhs_block: Optional[np.ndarray] = None,
    per_probe_rhs: bool = False,
    keep: str = "all",
) -> np.ndarray:
    """All probes at once through the same per-slice factors.

    Algorithmically the loop of ``solve_single_probe`` with the probe index
    vectorised (SuperLU multi-RHS solve).  ``rhs_block`` is (N, L) shared by all
    probes (reference ``solve`` behaviour, quirk Q13) or (P, N, L) when
    ``per_probe_rhs``.  ``keep='exit'`` returns only the last plane (P, N)."""
    P = probes.shape[0]
    N = g.block_size
    L = g.nz - 1
    U = np.asarray(probes, dtype=complex).reshape(P, N).T.copy()  # (N, P)
    out = None
    if keep == "all":
        out = np.zeros((P, N, g.nz), dtype=complex)
        out[:, :, 0] = U.T
    for j in range(1, g.nz):
        rhs = cache.B_list[j - 1] @ U
        col = L - j if mode == "adjoint" else j - 1
        if rhs_block is not None:
            if per_probe_rhs:
                rhs = rhs + rhs_block[:, :, col].T
            else:
                rhs = rhs + rhs_block[:, col][:, None]
        else:
            rhs = rhs + cache.b[:, None]
        U = cache.A_lu_list[j - 1].solve(rhs)
        if keep == "all":
            out[:, :, j] = U.T
    if keep == "exit":
        return U.T.copy()
    if mode == "adjoint":
        out = out[:, :, ::-1].copy()
    return out


def solve(g: Grid, bc_type, n: np.ndarray, probes: np.ndarray, mode: str = "forward", rhs_block=None, test_bcs=None):
    """``BaseForwardModelSolver.solve`` for one projection and one angle
    (base/base_solver.py:84-154).  probes: (P, *transverse). Returns (P, N, nz)."""
    cache = build_lu_cache(g, bc_type, n, mode)
    P = probes.shape[0]
    out = np.zeros((P, g.block_size, g.nz), dtype=complex)
    for s in range(P):
        out[s] = solve_single_probe(g, cache, probes[s].reshape(-1), mode, rhs_block, test_bcs)
    return out


# ----------------------------------------------------------------------------
# Whole space-z system  (forward_model.py:108-209, full_lu.py:72-172)
# ----------------------------------------------------------------------------


def full_system_matrix(g: Grid, bc_type, n: np.ndarray) -> sp.csc_matrix:
    """``return_forward_model_matrix``: kron(I,A0) - kron(S,B0) - Ck."""
    A0, B0 = step_operators(g, bc_type)
    L = g.nz - 1
    N = g.block_size
    M = sp.kron(sp.eye(L, format="csr"), A0, format="csr") - sp.kron(
        sp.diags([1], [-1], shape=(L, L), format="csr"), B0, format="csr"
    )
    C = object_contribution(g, n).reshape(N, L)
    Ck = sp.diags(C.T.flatten()) + sp.diags([C[..., 1:].T.flatten()], offsets=[-N])
    return (M - Ck).tocsc()


def full_lu_solve(g: Grid, bc_type, n: np.ndarray, probe: np.ndarray, mode: str = "forward", rhs_block=None) -> np.ndarray:
    """``PWEFullLUSolver._solve_single_probe_impl`` (full_lu.py:114-172).

    Note the missing D_1 u_0 term of the first block (quirk Q4)."""
    A0, B0 = step_operators(g, bc_type)
    N = g.block_size
    L = g.nz - 1
    lu = spla.splu(full_system_matrix(g, bc_type, n))
    if rhs_block is not None:
        b = rhs_block
    else:
        b = np.concatenate([B0 @ probe.reshape(-1), np.zeros(N * (L - 1), dtype=complex)])
    u = lu.solve(b, trans="H") if mode == "adjoint" else lu.solve(b)
    u = u.reshape(L, N).T
    return np.concatenate([probe.reshape(N, 1), u], axis=1)


def pint_matvec_exact(A, B, C, L, v, mode="forward"):
    """``_pintobj_matvec_exact`` (_pint_utils.py:27-54)."""
    Nx = A.shape[0]
    V = np.reshape(v, (Nx, L), order="F")
    U = np.zeros_like(V)
    if mode == "forward":
        U[:, :] += (A @ V) - (C * V)
        U[:, 1:] += -(B @ V[:, :-1] + C[:, 1:] * V[:, :-1])
    elif mode == "adjoint":
        A_H = A.conj().T
        B_H = B.conj().T
        Cc = C.conj()
        U[:, :] += (A_H @ V) - (Cc * V)
        U[:, :-1] += -(B_H @ V[:, 1:] + Cc[:, 1:] * V[:, 1:])
    else:
        raise ValueError("mode must be 'forward' or 'adjoint'")
    return np.reshape(U, (Nx * L,), order="F")
