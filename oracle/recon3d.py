"""Streaming FP64 gradient and steepest-descent reconstruction for two transverse axes.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The mathematics is that of ``gradient.exact_gradient_3d`` / ``reconstruct_3d``
(the exact adjoint of the CN slice stepping of pwe.py, the per-probe 2D
far-field amplitude misfit, thick_ptycho/reconstruction/pwe_reconstructor.py:100-164
for the structure of the gradient).  Those hold every plane of every probe
and a SuperLU factor per slice; at 256 x 256 x 128 with 400 probes that is
~55 GB of wavefields and hours of single-threaded triangular solves, so this
version changes only how the same linear systems are solved and stored:

* each slice operator A_s (ny blocks of nx x nx on the diagonal, diagonal
  couplings between neighbouring blocks) is factored by the dense block
  Thomas recurrence in FP64, S_0 = A_00, S_j = A_jj - F_{j-1} W_{j-1} E_{j-1},
  W_j = S_j^{-1} (E_j = A_{j,j+1}, F_j = A_{j+1,j}).  A_s^H has blocks
  A_jj^H and couplings conj(F), conj(E), and its recurrence yields W_j^H:
  the adjoint step reuses the forward factors of its slice;
* the forward wavefield is kept only every ``seg`` slices and each segment
  (with its factors) is recomputed during the backward sweep.

Pinned against SuperLU (``scipy.sparse.linalg.splu``) on the assembled
operators of every boundary type and against ``gradient.exact_gradient_3d`` /
``reconstruct_3d`` (themselves pinned by central finite differences) at small
sizes, tests/test_oracle_recon3d.py.
"""

from __future__ import annotations

from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp

from . import gradient as og
from . import pwe


def operator_blocks(M: sp.spmatrix, nb: int):
    """(D (m, nb, nb) dense diagonal blocks, E (m-1, nb) upper couplings, F (m-1, nb)
    lower couplings) of a block-tridiagonal matrix with diagonal couplings."""
    M = M.tocsr()
    m = M.shape[0] // nb
    D = np.empty((m, nb, nb), dtype=complex)
    E = np.empty((m - 1, nb), dtype=complex)
    F = np.empty((m - 1, nb), dtype=complex)
    for j in range(m):
        r = slice(j * nb, (j + 1) * nb)
        D[j] = M[r, r].toarray()
        if j < m - 1:
            rn = slice((j + 1) * nb, (j + 2) * nb)
            Ej, Fj = M[r, rn].tocsr(), M[rn, r].tocsr()
            E[j], F[j] = Ej.diagonal(), Fj.diagonal()
            if (Ej - sp.diags(E[j])).count_nonzero() or (Fj - sp.diags(F[j])).count_nonzero():
                raise ValueError("couplings between blocks are not diagonal")
            if j < m - 2 and M[r, (j + 2) * nb:].count_nonzero():
                raise ValueError("matrix is not block tridiagonal")
    return D, E, F


# Entries of S_j and W_j below FLUSH times the block's largest are set to zero.
# W_j decays geometrically away from the diagonal (by ~1e-9 every 6 entries at
# dx = 2, dz = 4), so far entries reach the subnormal range, where x86 arithmetic
# is ~100x slower; dropping what lies 14 orders below FP64 rounding changes no
# FP64 digit of the result.
FLUSH = 1e-30


def _flush(a: np.ndarray) -> np.ndarray:
    a[np.abs(a) < FLUSH * np.abs(a).max()] = 0
    return a


class BlockTriLU:
    """Dense block Thomas factorisation W_j = S_j^{-1} (FP64); ``solve(R, herm)``
    solves A X = R, or A^H X = R with the same factors."""

    def __init__(self, D: np.ndarray, E: np.ndarray, F: np.ndarray):
        m = D.shape[0]
        self.E, self.F = E, F
        self.W = np.empty_like(D)
        for j in range(m):
            S = D[j].copy() if j == 0 else D[j] - F[j - 1][:, None] * self.W[j - 1] * E[j - 1][None, :]
            self.W[j] = _flush(np.linalg.inv(_flush(S)))

    def solve(self, R: np.ndarray, herm: bool = False) -> np.ndarray:
        m, nb = self.W.shape[:2]
        shape = R.shape
        R = np.asarray(R, dtype=complex).reshape(m, nb, -1)
        Z = np.empty_like(R)
        tmp = np.empty_like(R[0])
        if herm:  # W_j^H x = conj(W_j^T conj(x))
            lo, up = np.conj(self.E), np.conj(self.F)

            def mv(j, x, out):
                np.conjugate(x, out=x)
                np.matmul(self.W[j].T, x, out=out)
                np.conjugate(out, out=out)
        else:
            lo, up = self.F, self.E

            def mv(j, x, out):
                np.matmul(self.W[j], x, out=out)

        tmp[...] = R[0]
        mv(0, tmp, Z[0])
        for j in range(1, m):
            np.multiply(lo[j - 1][:, None], Z[j - 1], out=tmp)
            np.subtract(R[j], tmp, out=tmp)
            mv(j, tmp, Z[j])
        corr = np.empty_like(tmp)
        for j in range(m - 2, -1, -1):
            np.multiply(up[j][:, None], Z[j + 1], out=tmp)
            mv(j, tmp, corr)
            Z[j] -= corr
        return Z.reshape(shape)


def _slice_ops(g: pwe.Grid, bc_type, n: np.ndarray):
    nx, N, L = g.nx, g.block_size, g.nz - 1
    A0, B0 = pwe.step_operators(g, bc_type)
    D0, E0, F0 = operator_blocks(A0, nx)
    C = pwe.object_contribution(g, n).reshape(N, L)
    idx = np.arange(nx)

    def factor(s):  # A_s = A0 - diag(C[:, s-1])
        D = D0.copy()
        D[:, idx, idx] -= C[:, s - 1].reshape(-1, nx)
        return BlockTriLU(D, E0, F0)

    def B(s):  # B_s = B0 + diag(C[:, s-1])
        return (B0 + sp.diags(C[:, s - 1])).tocsr()

    return factor, B


def forward_exit_3d(g: pwe.Grid, bc_type, n: np.ndarray, probes: np.ndarray) -> np.ndarray:
    """Exit planes (P, N) of the forward CN stepping, as pwe.solve_batched(keep='exit')."""
    factor, B = _slice_ops(g, bc_type, n)
    U = np.asarray(probes, dtype=complex).reshape(probes.shape[0], -1).T.copy()
    for s in range(1, g.nz):
        U = factor(s).solve(B(s) @ U)
    return U.T.copy()


def gradient_3d(g: pwe.Grid, bc_type, n: np.ndarray, probes: np.ndarray, I_meas: np.ndarray, seg: int = 16,
                seed: Optional[np.ndarray] = None, log: Optional[Callable[[str], None]] = None):
    """(dJ/dRe n + i dJ/dIm n of shape (nx, ny, nz), J) as ``gradient.exact_gradient_3d``."""
    N, L = g.block_size, g.nz - 1
    P = probes.shape[0]
    factor, B = _slice_ops(g, bc_type, n)
    U = np.asarray(probes, dtype=complex).reshape(P, N).T.copy()  # (N, P)
    ckpt = {0: U}
    for s in range(1, L + 1):
        U = factor(s).solve(B(s) @ U)
        if s % seg == 0 or s == L:
            ckpt[s] = U
        if log and s % 16 == 0:
            log(f"forward slice {s}/{L}")
    exit_w = U.T
    J = og.amplitude_loss_2d(exit_w, I_meas, g.nx, g.ny)
    if seed is None:
        seed = og.amplitude_loss_seed_2d(exit_w, I_meas, g.nx, g.ny)
    G = np.zeros((N, L), dtype=complex)
    lam = None
    for a in range(((L - 1) // seg) * seg, -1, -seg):
        b = min(a + seg, L)
        planes, facs = {a: ckpt[a]}, {}
        for s in range(a + 1, b + 1):
            facs[s] = factor(s)
            planes[s] = facs[s].solve(B(s) @ planes[s - 1])
        for s in range(b, a, -1):
            rhs = np.asarray(seed).T if s == L else B(s + 1).conj().T @ lam
            lam = facs[s].solve(rhs, herm=True)
            G[:, s - 1] = np.einsum("ip,ip->i", np.conj(lam), planes[s] + planes[s - 1])
        if log:
            log(f"adjoint segment {a + 1}..{b}")
    Gn = np.zeros((N, g.nz), dtype=complex)
    Gn[:, :-1] += G
    Gn[:, 1:] += G
    Gn *= (g.dz / 4) * (g.k / 1j)
    return (2 * np.conj(Gn)).reshape(g.nx, g.ny, g.nz), J


def lipschitz_step(g: pwe.Grid, bc_type, n: np.ndarray, probes, I_meas, h: float = 1e-5, seg: int = 16, log=None):
    """0.5 / L with L = ||grad(n - d) - grad(n)|| / ||d||, d = h grad(n) / max|grad(n)|."""
    g0, _ = gradient_3d(g, bc_type, n, probes, I_meas, seg, log=log)
    d = (h / np.max(np.abs(g0))) * g0
    g1, _ = gradient_3d(g, bc_type, n - d, probes, I_meas, seg, log=log)
    return 0.5 / (np.linalg.norm(g1 - g0) / np.linalg.norm(d))


def reconstruct_3d(g: pwe.Grid, bc_type, n0: np.ndarray, probes, I_meas, iters: int, step: float, seg: int = 16,
                   log=None):
    """n <- n - step grad J(n), ``iters`` times; returns (n, [J before each iteration])."""
    n = np.array(n0, dtype=complex)
    hist = []
    for it in range(iters):
        grad, J = gradient_3d(g, bc_type, n, probes, I_meas, seg, log=log)
        hist.append(J)
        if log:
            log(f"iteration {it}: J = {J:.6e}")
        n = n - step * grad
    return n, hist
