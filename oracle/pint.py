"""FP64 restatement of the reference's parallel-in-time (PiT) whole-system solve.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows thick_ptycho/forward_model/pwe/solvers/full_pint.py:71-275 and
.../solvers/utils/_pint_utils.py:27-199:

* the system is the whole space-z one of ``pwe.full_system_matrix``, applied
  matrix-free by ``pwe.pint_matvec_exact`` with (A0, B0, C) -- or, in adjoint
  mode, (A0^H, B0^H, conj(C) with its slices reversed) in the same
  lower-bidiagonal form (full_pint.py:148-160);
* the right-hand side is B0 u_0 in the first block and zeros elsewhere
  (forward_model.py:155-193), or ``rhs_block`` as given;
* the preconditioner is the alpha-circulant block one, P_alpha = I (x) A -
  C_alpha (x) B with C_alpha the down-shift with alpha in its top-right corner,
  applied as Gamma^{-1} F blockdiag(A - omega_k B)^{-1} F^{-1} Gamma with
  gamma_j = alpha^{j/L}, omega_k = alpha^{1/L} exp(2 pi i k / L), F = fft(norm=
  'ortho') along z (_pint_utils.py:81-114, 148-159);
* GMRES with right preconditioning, the solution being R y (full_pint.py:240-262).

The reference builds its preconditioner from the free-space step matrices
whatever the mode (the object-averaged ``update`` is not called,
full_pint.py:143-144); ``blocks_from`` chooses which pair the blocks use here.

Pinned in tests/test_oracle_pint.py: the preconditioner against a sparse solve
of the assembled P_alpha, the GMRES solution against the whole-system SuperLU
solve (``pwe.full_lu_solve``) in both modes.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.fft import fft, ifft

from . import pwe


def shifts(alpha: float, L: int):
    """(gamma (L,), omega (L,)) of the alpha-circulant preconditioner."""
    gamma = alpha ** (np.arange(L) / L)
    omega = np.exp(2j * np.pi * np.arange(L) / L) * alpha ** (1.0 / L)
    return gamma.astype(complex), omega


class AlphaCirculant:
    """P_alpha^{-1} for the blocks (A, B) (_pint_utils.py:81-114)."""

    def __init__(self, A: sp.spmatrix, B: sp.spmatrix, L: int, alpha: float):
        self.N, self.L = A.shape[0], L
        self.gamma, self.omega = shifts(alpha, L)
        self.lus = [spla.splu((A - w * B).astype(complex).tocsc()) for w in self.omega]

    def apply(self, v: np.ndarray) -> np.ndarray:
        X = self.gamma[:, None] * np.asarray(v, dtype=complex).reshape(self.L, self.N)
        Xh = ifft(X, axis=0, norm="ortho")
        for k in range(self.L):
            Xh[k] = self.lus[k].solve(Xh[k])
        return (fft(Xh, axis=0, norm="ortho") / self.gamma[:, None]).ravel()

    @staticmethod
    def assembled(A: sp.spmatrix, B: sp.spmatrix, L: int, alpha: float) -> sp.csc_matrix:
        """P_alpha = I (x) A - C_alpha (x) B, for checks."""
        Ca = sp.diags([np.ones(L - 1)], [-1], shape=(L, L), format="lil").astype(complex)
        Ca[0, L - 1] = alpha
        return (sp.kron(sp.eye(L), A) - sp.kron(Ca.tocsr(), B)).tocsc()


def system(g: pwe.Grid, bc_type, n: np.ndarray, mode: str = "forward"):
    """(A, B, C (N, L)) of the whole-system mat-vec for ``mode``."""
    A0, B0 = pwe.step_operators(g, bc_type)
    C = pwe.object_contribution(g, n).reshape(g.block_size, g.nz - 1)
    if mode == "adjoint":
        return A0.conj().T.tocsr(), B0.conj().T.tocsr(), np.flip(C.conj(), axis=1)
    return A0.tocsr(), B0.tocsr(), C


def pit_solve(g: pwe.Grid, bc_type, n: np.ndarray, probe: Optional[np.ndarray], mode: str = "forward",
              rhs_block: Optional[np.ndarray] = None, alpha: float = 1e-2, atol: float = 1e-8, rtol: float = 1e-5,
              restart: int = 20, maxiter: Optional[int] = None, blocks_from: str = "mode"):
    """(u (N, nz) with the probe as plane 0, GMRES info, iterations).

    ``blocks_from='mode'`` builds P_alpha from the mode's (A, B); 'forward' from
    the free-space forward pair as the reference does in every mode."""
    N, L = g.block_size, g.nz - 1
    A, B, C = system(g, bc_type, n, mode)
    A0, B0 = pwe.step_operators(g, bc_type)
    pre = AlphaCirculant(A if blocks_from == "mode" else A0.tocsr(), B if blocks_from == "mode" else B0.tocsr(),
                         L, alpha)
    if rhs_block is not None:
        b = np.asarray(rhs_block, dtype=complex).reshape(N, L).ravel(order="F")
    else:
        p = np.zeros(N, dtype=complex) if probe is None else np.asarray(probe, dtype=complex).reshape(N)
        b = np.concatenate([B0 @ p, np.zeros(N * (L - 1), dtype=complex)])
    AR = spla.LinearOperator((N * L, N * L), dtype=complex,
                             matvec=lambda y: pwe.pint_matvec_exact(A, B, C, L, pre.apply(y)))
    its = []
    y, info = spla.gmres(AR, b, atol=atol, rtol=rtol, restart=restart, maxiter=maxiter,
                         callback=lambda r: its.append(r), callback_type="pr_norm")
    u = pre.apply(y).reshape(L, N).T
    u0 = np.zeros((N, 1), dtype=complex) if probe is None else np.asarray(probe, dtype=complex).reshape(N, 1)
    return np.concatenate([u0, u], axis=1), info, len(its)
