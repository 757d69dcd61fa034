"""Pin the FP64 oracle's PWE restatement against independent checks (CPU).

Each check is chosen so that a plausible restatement mistake (wrong sign of
the object term, dropped boundary factor, swapped sub/super diagonal, wrong
adjoint slice index) fails at least one of them.
"""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import pwe
from tests.cases import convergence_case_2d, convergence_case_3d, random_object, rel_l2

BCS = ["neumann", "impedance", "dirichlet"]


# --------------------------------------------------------------------------- convergence
@pytest.mark.parametrize("bc", BCS)
def test_convergence_2d(bc):
    """Reference test_boundary_conditions_2d.py:223-275: monotone L-inf error, rate in [1.5, 2.5]."""
    errs = []
    for step in (16, 32, 64):
        g, n, probe, exact, test_bcs, _, _ = convergence_case_2d(step, bc)
        u = pwe.solve(g, bc, n, probe[None], "forward", test_bcs=test_bcs)[0]
        errs.append(np.max(np.abs(u - exact)))
    rates = [np.log2(errs[i - 1] / errs[i]) for i in (1, 2)]
    assert errs[0] > errs[1] > errs[2], errs
    assert 1.5 <= np.mean(rates) <= 2.5, rates


@pytest.mark.parametrize("bc", BCS)
def test_convergence_3d(bc):
    """Reference test_boundary_conditions_3d.py:262-300 (last rate in [1.5, 2.5])."""
    steps = (8, 16, 32) + ((64,) if bc == "dirichlet" else ())
    errs = []
    for step in steps:
        g, n, probe, exact, test_bcs = convergence_case_3d(step, bc)
        u = pwe.solve(g, bc, n, probe.reshape(1, -1), "forward", test_bcs=test_bcs)[0]
        u = u.reshape(g.nx, g.ny, g.nz)
        errs.append(np.max(np.abs(u - exact)))
    assert all(errs[i] < errs[i - 1] for i in range(1, len(errs))), errs
    rate = np.log2(errs[-2] / errs[-1])
    assert 1.5 <= rate <= 2.5, (rate, errs)


# --------------------------------------------------------------------------- operators
def test_operator_rows_against_hand_values():
    """Boundary rows written out by hand from boundary_conditions.py:57-94."""
    g = pwe.Grid(dimension=2, nx=6, nz=4, dx=0.5, dz=0.25, k=3.0)
    mu = (1j / (2 * g.k)) * 0.5 * g.dz / g.dx**2
    beta = mu * 2j * g.k * g.dx
    for bc in ("dirichlet", "neumann", "impedance"):
        A, B = pwe.step_operators(g, bc)
        A = A.toarray()
        B = B.toarray()
        # interior row
        assert np.allclose(A[2, 1:4], [-mu, 1 + 2 * mu, -mu])
        assert np.allclose(B[2, 1:4], [mu, 1 - 2 * mu, mu])
        if bc == "dirichlet":
            assert np.allclose(A[0, :2], [1, 0]) and np.allclose(A[-1, -2:], [0, 1])
        elif bc == "neumann":
            assert np.allclose(A[0, :2], [1 + 2 * mu, -2 * mu]) and np.allclose(A[-1, -2:], [-2 * mu, 1 + 2 * mu])
        else:
            assert np.allclose(A[0, :2], [1 + 2 * mu - beta, -2 * mu])
            assert np.allclose(B[-1, -2:], [2 * mu, 1 - 2 * mu + beta])
    # non-string boundary type behaves as Neumann on one axis (quirk Q1)
    class E:  # stands for an Enum member
        pass

    An, _ = pwe.step_operators(g, "neumann")
    Ae, _ = pwe.step_operators(g, E())
    assert np.allclose(An.toarray(), Ae.toarray())
    g3 = pwe.Grid(dimension=3, nx=4, ny=4, nz=3, dx=0.5, dy=0.5, dz=0.25, k=3.0)
    with pytest.raises(ValueError):
        pwe.step_operators(g3, E())


def test_object_contribution_closed_form():
    g = pwe.Grid(dimension=2, nx=3, nz=4, dx=1.0, dz=0.5, k=2.0)
    n = 1 + np.arange(12).reshape(3, 4) * (1e-3 + 2e-4j)
    C = pwe.object_contribution(g, n)
    coef = -1j * g.k * (n - 1)
    assert np.allclose(C, g.dz / 4 * (coef[:, :-1] + coef[:, 1:]))
    Cg = pwe.object_contribution(g, n, grad=True)
    assert np.allclose(Cg, g.dz * g.k / 2j)


# --------------------------------------------------------------------------- stepping
def _thomas(a, d, c, r):
    """Textbook Thomas algorithm (FP64), sub a[i] (row i, col i-1), sup c[i]."""
    n = len(d)
    cp = np.zeros(n, complex)
    dp = np.zeros(n, complex)
    cp[0] = c[0] / d[0]
    dp[0] = r[0] / d[0]
    for i in range(1, n):
        m = d[i] - a[i] * cp[i - 1]
        cp[i] = c[i] / m if i < n - 1 else 0
        dp[i] = (r[i] - a[i] * dp[i - 1]) / m
    x = np.zeros(n, complex)
    x[-1] = dp[-1]
    for i in range(n - 2, -1, -1):
        x[i] = dp[i] - cp[i] * x[i + 1]
    return x


@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "impedance", "impedance2"])
def test_forward_step_matches_thomas_and_solve_banded(bc, rng):
    g = pwe.Grid(dimension=2, nx=37, nz=9, dx=2.0, dz=4.0, k=2 * np.pi)
    n = random_object(g.nx, g.nz, rng, amp=1e-2)
    probe = rng.standard_normal(g.nx) + 1j * rng.standard_normal(g.nx)
    u = pwe.solve(g, bc, n, probe[None], "forward")[0]
    A0, B0 = pwe.step_operators(g, bc)
    C = pwe.object_contribution(g, n)
    v = probe.copy()
    for j in range(1, g.nz):
        A = (A0 - sp.diags(C[:, j - 1])).toarray()
        B = (B0 + sp.diags(C[:, j - 1])).toarray()
        r = B @ v
        a = np.r_[0, np.diag(A, -1)]
        c = np.r_[np.diag(A, 1), 0]
        v_th = _thomas(a, np.diag(A), c, r)
        ab = np.zeros((3, g.nx), complex)
        ab[0, 1:] = np.diag(A, 1)
        ab[1] = np.diag(A)
        ab[2, :-1] = np.diag(A, -1)
        v_sb = sla.solve_banded((1, 1), ab, r)
        assert rel_l2(v_th, v_sb) < 1e-13
        v = v_sb
        assert rel_l2(u[:, j], v) < 1e-12


def test_cos_mode_eigen_amplification():
    """A Neumann cos mode is an eigenvector of K; amplification (1+l)/(1-l) per free-space step."""
    g = pwe.Grid(dimension=2, nx=33, nz=7, dx=1 / 32, dz=0.05, k=50.0)
    x = np.arange(g.nx)
    for m in (1, 3, 6):
        v = np.cos(np.pi * m * x / (g.nx - 1))
        mu = (1j / (2 * g.k)) * 0.5 * g.dz / g.dx**2
        lam = mu * (2 * np.cos(np.pi * m / (g.nx - 1)) - 2)
        u = pwe.solve(g, "neumann", np.ones((g.nx, g.nz)), v[None], "forward")[0]
        for j in range(g.nz):
            assert np.allclose(u[:, j], v * ((1 + lam) / (1 - lam)) ** j, atol=1e-12)


def test_free_space_norm_conserved():
    g = pwe.Grid(dimension=2, nx=64, nz=20, dx=2.0, dz=4.0, k=2 * np.pi)
    x = np.arange(g.nx)
    probe = np.exp(-((x - 32) ** 2) / 6.0**2).astype(complex)
    u = pwe.solve(g, "dirichlet", np.ones((g.nx, g.nz)), probe[None], "forward")[0]
    norms = np.linalg.norm(u[1:-1], axis=0)
    assert np.allclose(norms, norms[0], rtol=1e-10)


def test_iterative_equals_full_lu_when_first_slice_empty(rng):
    """Reference test_boundary_conditions_3d.py:302-316 pattern; Q4 requires C[:, 0] = 0."""
    g = pwe.Grid(dimension=2, nx=12, nz=8, dx=2.0, dz=4.0, k=2 * np.pi)
    n = random_object(g.nx, g.nz, rng, amp=1e-2)
    n[:, :2] = 1.0  # C[:, 0] = 0
    probe = rng.standard_normal(g.nx) + 0j
    u_it = pwe.solve(g, "impedance", n, probe[None], "forward")[0]
    u_full = pwe.full_lu_solve(g, "impedance", n, probe, "forward")
    assert np.allclose(u_it, u_full, atol=1e-12)


def test_adjoint_modes_against_full_lu_and_identity(rng):
    """The reference adjoint sweep reproduces A_j^{-H} (B_j)^H recursively; the exact
    adjoint of the space-z system is full_lu with trans='H'."""
    g = pwe.Grid(dimension=2, nx=10, nz=7, dx=2.0, dz=4.0, k=2 * np.pi)
    n = random_object(g.nx, g.nz, rng, amp=5e-2)
    N, L = g.nx, g.nz - 1
    M = pwe.full_system_matrix(g, "neumann", n)
    w = rng.standard_normal(N * L) + 1j * rng.standard_normal(N * L)
    v = rng.standard_normal(N * L) + 1j * rng.standard_normal(N * L)
    lhs = np.vdot(v, M @ w)
    rhs = np.vdot(M.conj().T @ v, w)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # exact adjoint (full LU, trans='H') with a source in the last block only
    e = np.zeros(N * L, complex)
    e[-N:] = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    lam_full = pwe.full_lu_solve(g, "neumann", n, np.zeros(N, complex), "adjoint", rhs_block=e)
    from oracle.gradient import exact_adjoint_states

    lam = exact_adjoint_states(g, "neumann", n, e[-N:][None])[0]
    assert np.allclose(lam_full[:, 1:], lam[:, 1:], atol=1e-12)
    # reference-form adjoint: differs from the exact one only through B's slice index (Q3)
    cache = pwe.build_lu_cache(g, "neumann", n, "adjoint")
    rb = np.zeros((N, L), complex)
    rb[:, -1] = e[-N:]
    ref = pwe.solve_single_probe(g, cache, None, "adjoint", rhs_block=rb)
    A0, B0 = pwe.step_operators(g, "neumann")
    C = pwe.object_contribution(g, n)
    vv = np.linalg.solve((A0 - sp.diags(C[:, L - 1])).toarray().conj().T, e[-N:])
    assert np.allclose(ref[:, L - 1], vv)  # first adjoint step lands in column L-1 after the flip
    vv2 = np.linalg.solve((A0 - sp.diags(C[:, L - 2])).toarray().conj().T,
                          (B0 + sp.diags(C[:, L - 2])).toarray().conj().T @ vv)
    assert np.allclose(ref[:, L - 2], vv2)
    assert np.allclose(lam[:, L], vv)


def test_pint_matvec_adjoint_identity(rng):
    """Reference test_pint_utils.py:18-47 (rtol 1e-12)."""
    g = pwe.Grid(dimension=2, nx=9, nz=6, dx=2.0, dz=4.0, k=2 * np.pi)
    n = random_object(g.nx, g.nz, rng, amp=5e-2)
    A0, B0 = pwe.step_operators(g, "impedance")
    C = pwe.object_contribution(g, n)
    L = g.nz - 1
    w = rng.standard_normal(g.nx * L) + 1j * rng.standard_normal(g.nx * L)
    v = rng.standard_normal(g.nx * L) + 1j * rng.standard_normal(g.nx * L)
    Aw = pwe.pint_matvec_exact(A0, B0, C, L, w, "forward")
    Ahv = pwe.pint_matvec_exact(A0, B0, C, L, v, "adjoint")
    assert abs(np.vdot(v, Aw) - np.vdot(Ahv, w)) <= 1e-12 * abs(np.vdot(v, Aw))
    # the mat-vec is the full system matrix up to Q4's first-block D_1 coupling
    M = pwe.full_system_matrix(g, "impedance", n)
    Vf = np.reshape(w, (g.nx, L), order="F").T.ravel()
    Uf = np.reshape(Aw, (g.nx, L), order="F").T.ravel()
    assert np.allclose(M @ Vf, Uf)


def test_reverse_mode_inverts_forward_free_space():
    """With n = 1 the reverse sweep (A and B swapped) undoes the forward one."""
    g = pwe.Grid(dimension=2, nx=40, nz=10, dx=2.0, dz=4.0, k=2 * np.pi)
    x = np.arange(g.nx)
    probe = np.exp(-((x - 20) ** 2) / 25.0).astype(complex)
    n = np.ones((g.nx, g.nz), complex)
    u = pwe.solve(g, "impedance", n, probe[None], "forward")[0]
    back = pwe.solve(g, "impedance", n, u[:, -1][None], "reverse")[0]
    assert np.max(np.abs(back[:, -1] - probe)) < 1e-8 * np.max(np.abs(probe))


def test_reference_scan_windows_example():
    """Reduced windows of the reference's scan-frame rule, worked by hand for
    nx = 20, 3 scan points, step 4, pad 8: W = 12, centres 6, 10, 14 (round half
    to even on 5.5, 9.5, 13.5), windows [0, 11), [4, 15), [8, 19)."""
    from thick_ptycho_amd import SimulationSpace

    assert SimulationSpace.reference_scan_windows(20, 3, 4, 8) == [(0, 11), (4, 15), (8, 19)]
    ss = SimulationSpace(20, 6, 1.0, 1.0, num_probes=3, reduced_windows=[(0, 11), (4, 15), (8, 19)])
    assert ss.solve_reduced_domain and ss.effective_nx == 11 and ss.block_size == 11
    n = np.arange(20 * 6).reshape(20, 6) + 1.0
    C = ss.create_object_contribution(n, scan_index=1)
    assert C.shape == (11, 5)
    coef = (ss.k / 1j) * (n[4:15] - 1)
    assert np.allclose(C, (ss.dz / 2) * (coef[:, :-1] + coef[:, 1:]) / 2)
