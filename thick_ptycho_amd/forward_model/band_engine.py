"""Device-side plumbing for CN stepping with two transverse axes (3D objects).

The per-slice operator on N = nx * ny unknowns (flat index block * nx + i, the
reference's ordering, operator_matrices.py:75-126) is block tridiagonal with
tridiagonal diagonal blocks and scaled-identity couplings.  ``Block3D`` holds it
as three block types (first, interior, last) plus the coupling scalars;
``factor`` runs ``tp_band_factor`` (block LU with banded Schur inverses, FP64)
for one (object, mode), widening the band until the dropped entries lie below
single-precision rounding; ``BandScanSpec`` / ``sweep`` step batches of
probes through the factors in one persistent launch.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from .. import _lib
from .operators import bc_scalars, laplacian_1d

# band half-widths the sweep kernel is built for
BAND_WIDTHS = (4, 6, 8, 12, 16, 24, 32)
# LDS of one factor workgroup: the FP64 band of one block
FACTOR_LDS = 160 * 1024
# largest |W_j(i, i +- w)| / max |W_j(i, i)| accepted: the entries beyond the band
# are smaller still, so what is dropped stays below complex64 rounding
EDGE_TOL = 2.0 ** -28


@dataclass
class Block3D:
    """Block-tridiagonal operator: tri[type] = (sub, diag, sup) of the nx x nx
    diagonal blocks (type 0 first block, 1 interior, 2 last), A(j, j-1) = f[j] I,
    A(j, j+1) = e[j] I."""

    tri: np.ndarray  # complex128 (3, 3, nx)
    f: np.ndarray  # complex128 (ny,)
    e: np.ndarray  # complex128 (ny,)

    @property
    def nx(self) -> int:
        return self.tri.shape[2]

    @property
    def ny(self) -> int:
        return self.f.shape[0]

    def conj_transpose(self) -> "Block3D":
        tri = np.zeros_like(self.tri)
        tri[:, 0, 1:] = np.conj(self.tri[:, 2, :-1])
        tri[:, 1] = np.conj(self.tri[:, 1])
        tri[:, 2, :-1] = np.conj(self.tri[:, 0, 1:])
        f = np.zeros_like(self.f)
        e = np.zeros_like(self.e)
        f[1:] = np.conj(self.e[:-1])
        e[:-1] = np.conj(self.f[1:])
        return Block3D(tri, f, e)

    def block_type(self, j: int) -> int:
        return 0 if j == 0 else (2 if j == self.ny - 1 else 1)

    def to_sparse(self):
        """scipy CSR of the full N x N operator (host checks and small systems)."""
        import scipy.sparse as sp

        nx, ny = self.nx, self.ny
        blocks = []
        for t in range(3):
            sub, dia, sup = self.tri[t]
            blocks.append(sp.diags([sub[1:], dia, sup[:-1]], [-1, 0, 1], shape=(nx, nx)))
        rows = [[None] * ny for _ in range(ny)]
        I = sp.eye(nx)
        for j in range(ny):
            rows[j][j] = blocks[self.block_type(j)]
            if j > 0:
                rows[j][j - 1] = self.f[j] * I
            if j < ny - 1:
                rows[j][j + 1] = self.e[j] * I
        return sp.bmat(rows, format="csr")


def step_operators_3d(nx: int, ny: int, k: float, dx: float, dy: float, dz: float, bc_type) -> Tuple[Block3D, Block3D]:
    """(A0, B0) = (I - K, I + K) for two transverse axes.

    Inside a block: the one-axis stencil with diagonal -2(mu_x + mu_y) and the
    x boundary rules; across blocks: mu_y couplings with the block rules of
    boundary_conditions.py:96-130 (dirichlet: first/last diagonal blocks are I
    and their couplings into the domain 0; neumann: those couplings doubled;
    impedance: neumann plus -/+ beta_y I on the first/last diagonal blocks of
    A0/B0).  impedance2 has no two-axis form in the reference (it calls a
    missing method) and any other value is rejected there with ValueError."""
    if bc_type == "impedance2":
        raise AttributeError("impedance2 has no two-axis implementation in the reference")
    if bc_type not in ("dirichlet", "neumann", "impedance"):
        raise ValueError("Invalid boundary condition type. Please choose from dirichlet, neumann or dirichlet.")
    c = bc_scalars(k, dx, dz, bc_type, dy=dy)
    mu_y = c["mu_y"]
    K = laplacian_1d(nx, k, dx, dz, bc_type, mu_y=mu_y)
    ones = np.ones(nx, dtype=complex)
    Ax = np.stack([-K.sub, 1 - K.diag, -K.sup])
    Bx = np.stack([K.sub.copy(), 1 + K.diag, K.sup.copy()])
    triA = np.stack([Ax, Ax, Ax]).astype(complex)
    triB = np.stack([Bx, Bx, Bx]).astype(complex)
    fA = np.full(ny, -mu_y, dtype=complex)
    eA = np.full(ny, -mu_y, dtype=complex)
    fA[0] = 0
    eA[-1] = 0
    fB, eB = -fA, -eA
    if bc_type == "dirichlet":
        ident = np.stack([np.zeros(nx, complex), ones, np.zeros(nx, complex)])
        for t in (0, 2):
            triA[t] = ident
            triB[t] = ident
        eA[0] = eB[0] = 0
        fA[-1] = fB[-1] = 0
    else:
        eA[0] *= 2
        eB[0] *= 2
        fA[-1] *= 2
        fB[-1] *= 2
        if bc_type == "impedance":
            by = c["beta_y"]
            for t in (0, 2):
                triA[t, 1] -= by
                triB[t, 1] += by
    return Block3D(triA, fA, eA), Block3D(triB, fB, eB)


@dataclass
class BandPlan:
    lhs: Block3D
    rhs: Block3D
    lhs_sign: float
    rhs_sign: float
    conj_c: int
    lhs_col0: int
    rhs_col0: int
    col_step: int
    b_sign: float = 1.0
    sigma: bool = False

    def check_complementary(self) -> None:
        """The sweep uses rhs = diag(sigma) - lhs."""
        ok = (np.allclose(self.rhs.tri[:, 0, 1:], -self.lhs.tri[:, 0, 1:], rtol=0, atol=1e-14)
              and np.allclose(self.rhs.tri[:, 2, :-1], -self.lhs.tri[:, 2, :-1], rtol=0, atol=1e-14)
              and np.allclose(self.rhs.tri[:, 1] + self.lhs.tri[:, 1], 2.0, rtol=0, atol=1e-12)
              and np.allclose(self.rhs.f[1:], -self.lhs.f[1:], rtol=0, atol=1e-14)
              and np.allclose(self.rhs.e[:-1], -self.lhs.e[:-1], rtol=0, atol=1e-14))
        if not ok:
            raise ValueError("rhs operator is not 2I minus the lhs operator")


def band_mode_plan(A: Block3D, B: Block3D, mode: str, L: int) -> BandPlan:
    """Per-mode operators (iterative_lu.py:90-128), as for one transverse axis:
    forward A0 - D_s; adjoint (A0 - D)^H with D = C[:, L-s] (the reference form);
    reverse B0 + D with b -> -b; exact_adjoint (A0 - D_{L-s+1})^H with the
    right-hand operator (B0 + D_{L-s+2})^H."""
    if mode == "forward":
        return BandPlan(A, B, -1.0, 1.0, 0, 0, 0, 1)
    if mode == "adjoint":
        return BandPlan(A.conj_transpose(), B.conj_transpose(), -1.0, 1.0, 1, L - 1, L - 1, -1)
    if mode == "reverse":
        return BandPlan(B, A, 1.0, -1.0, 0, L - 1, L - 1, -1, b_sign=-1.0)
    if mode == "exact_adjoint":
        return BandPlan(A.conj_transpose(), B.conj_transpose(), -1.0, 1.0, 1, L - 1, L, -1, sigma=True)
    raise AssertionError(f"Invalid mode: {mode!r}")


@dataclass
class BandFactors:
    W: torch.Tensor
    cpl: torch.Tensor
    sigma: Optional[torch.Tensor]
    nx: int
    ny: int
    steps: int
    w: int
    mode: str
    b_sign: float
    edge: float
    keep: list = field(default_factory=list)

    @property
    def N(self) -> int:
        return self.nx * self.ny


def _factor_once(C, plan: BandPlan, w: int, L: int, nx: int, ny: int):
    dev = C.device
    tri = torch.from_numpy(np.ascontiguousarray(plan.lhs.tri)).to(dev)
    cpl = torch.from_numpy(np.ascontiguousarray(np.stack([plan.lhs.f, plan.lhs.e], axis=1))).to(dev)
    lib = _lib.load()
    W = torch.empty(int(lib.tp_band_coef_bytes(nx, ny, L, w)) // 8, dtype=torch.complex64, device=dev)
    sigma = torch.empty((L, nx * ny), dtype=torch.complex64, device=dev) if plan.sigma else None
    edge = torch.zeros(L, dtype=torch.float32, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(
        lib.tp_band_factor(_lib.ptr(C), nx, ny, L, _lib.ptr(tri), _lib.ptr(cpl), plan.lhs_sign, plan.rhs_sign,
                           plan.conj_c, plan.lhs_col0, plan.rhs_col0, plan.col_step, L, w, int(plan.sigma),
                           _lib.ptr(W), _lib.ptr(sigma), _lib.ptr(edge), _lib.ptr(status), _lib.stream_handle()),
        "tp_band_factor",
    )
    st = int(status.item())
    e = float(edge.max().item())
    if st != 0:
        raise RuntimeError("singular per-slice factor (zero pivot) in tp_band_factor")
    return W, sigma, e, [tri, cpl]


def factor(C: torch.Tensor, A: Block3D, B: Block3D, mode: str, w: Optional[int] = None) -> BandFactors:
    """Factor every slice's LHS for ``mode`` (C: complex128 CUDA (N, L)).

    The band half-width starts at ``w`` (default 6) and grows through
    BAND_WIDTHS until the band edge is below EDGE_TOL (synchronises once per try)."""
    assert C.is_cuda and C.dtype == torch.complex128 and C.dim() == 2
    N, L = C.shape
    nx, ny = A.nx, A.ny
    assert N == nx * ny, (N, nx, ny)
    C = C.contiguous()
    plan = band_mode_plan(A, B, mode, L)
    plan.check_complementary()
    widths = [x for x in BAND_WIDTHS if x >= (w or 6) and x < nx and nx * (2 * x + 4) * 16 <= FACTOR_LDS]
    if not widths:
        raise ValueError(f"no band half-width >= {w or 6} below nx = {nx} fits the factor ({BAND_WIDTHS})")
    e = float("inf")
    for wt in widths:
        W, sigma, e, keep = _factor_once(C, plan, wt, L, nx, ny)
        if e <= EDGE_TOL:
            cpl32 = torch.from_numpy(np.stack([plan.lhs.f, plan.lhs.e], axis=1).astype(np.complex64)).to(C.device)
            return BandFactors(W, cpl32.contiguous(), sigma, nx, ny, L, wt, mode, plan.b_sign, e, keep=keep + [C])
    raise ValueError(
        f"the inverse blocks of the {nx}x{ny} per-slice operator do not decay to {EDGE_TOL:.1e} within a band of "
        f"half-width {widths[-1]} (edge ratio {e:.2e}); wider bands are not built"
    )


class BandFactorizer:
    """``tp_band_factor`` for one operator and mode into preallocated buffers with a
    fixed band half-width: ``run(C)`` launches without host synchronisation (it can be
    captured into a HIP graph); ``check()`` synchronises and verifies the band edge and
    the pivots of the last run."""

    def __init__(self, A: Block3D, B: Block3D, mode: str, L: int, w: int, device,
                 source: Optional["BandFactorizer"] = None):
        nx, ny = A.nx, A.ny
        self.plan = band_mode_plan(A, B, mode, L)
        self.plan.check_complementary()
        # adjoint modes from the forward factors of the same object: their LHS of
        # step s is the conjugate transpose of the forward LHS of slice L - s + 1
        if source is not None:
            assert mode in ("adjoint", "exact_adjoint") and source.plan is not None
            assert source.mode == "forward" and (source.nx, source.ny, source.L, source.w) == (nx, ny, L, w)
        self.source = source
        self.mode = mode
        self.nx, self.ny, self.L, self.w = nx, ny, L, w
        self.tri = torch.from_numpy(np.ascontiguousarray(self.plan.lhs.tri)).to(device)
        self.cpl = torch.from_numpy(np.ascontiguousarray(np.stack([self.plan.lhs.f, self.plan.lhs.e], axis=1))).to(device)
        lib = _lib.load()
        W = torch.empty(int(lib.tp_band_coef_bytes(nx, ny, L, w)) // 8, dtype=torch.complex64, device=device)
        sigma = torch.empty((L, nx * ny), dtype=torch.complex64, device=device) if self.plan.sigma else None
        self.edge = torch.zeros(L, dtype=torch.float32, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)
        cpl32 = torch.from_numpy(np.stack([self.plan.lhs.f, self.plan.lhs.e], axis=1).astype(np.complex64)).to(device)
        self.fac = BandFactors(W, cpl32.contiguous(), sigma, nx, ny, L, w, mode, self.plan.b_sign, float("nan"),
                               keep=[self.tri, self.cpl])

    def run(self, C: torch.Tensor) -> BandFactors:
        pl = self.plan
        if self.source is not None:  # the source must have run on the same C
            lib = _lib.load()
            _lib.check(lib.tp_band_conj_transpose(_lib.ptr(self.source.fac.W), self.nx, self.ny, self.L, self.w, 1,
                                                  _lib.ptr(self.fac.W), _lib.stream_handle()), "tp_band_conj_transpose")
            if pl.sigma:
                _lib.check(lib.tp_band_sigma(_lib.ptr(C), self.nx, self.ny, self.L, pl.lhs_sign, pl.rhs_sign, pl.conj_c,
                                             pl.lhs_col0, pl.rhs_col0, pl.col_step, self.L, _lib.ptr(self.fac.sigma),
                                             _lib.stream_handle()), "tp_band_sigma")
            return self.fac
        _lib.check(
            _lib.load().tp_band_factor(_lib.ptr(C), self.nx, self.ny, self.L, _lib.ptr(self.tri), _lib.ptr(self.cpl),
                                       pl.lhs_sign, pl.rhs_sign, pl.conj_c, pl.lhs_col0, pl.rhs_col0, pl.col_step,
                                       self.L, self.w, int(pl.sigma), _lib.ptr(self.fac.W), _lib.ptr(self.fac.sigma),
                                       _lib.ptr(self.edge), _lib.ptr(self.status), _lib.stream_handle()),
            "tp_band_factor",
        )
        return self.fac

    def check(self) -> float:
        if self.source is not None:
            self.fac.edge = self.source.check()
            return self.fac.edge
        if int(self.status.item()) != 0:
            raise RuntimeError("singular per-slice factor (zero pivot) in tp_band_factor")
        e = float(self.edge.max().item())
        if e > EDGE_TOL:
            raise RuntimeError(f"band half-width {self.w} too narrow for this object (edge ratio {e:.2e})")
        self.fac.edge = e
        return e


@dataclass
class BandScanSpec:
    """One batch of probes stepped through ``fac`` (planes of N = nx*ny values).

    u0: complex64 (P, N); out: (P, L+1, N) or None; exit: (P, N) or None;
    ckpt: (P, L//every + 1, N) or None; rhs: (P or 1, L, N) in step order or None;
    flip: write plane s at index L - s."""

    fac: BandFactors
    u0: torch.Tensor
    out: Optional[torch.Tensor] = None
    exit: Optional[torch.Tensor] = None
    ckpt: Optional[torch.Tensor] = None
    ckpt_every: int = 0
    rhs: Optional[torch.Tensor] = None
    rhs_steps: int = 0
    flip: bool = False
    work: Optional[torch.Tensor] = None  # (P, 2, N) ping-pong planes when out is None (allocated if not given)
    _keep: list = field(default_factory=list)

    def struct(self) -> _lib.BandScan:
        f = self.fac
        N = f.N
        P = self.u0.shape[0]
        s = _lib.BandScan()
        s.W = _lib.ptr(f.W)
        s.cpl = _lib.ptr(f.cpl)
        s.sigma = _lib.ptr(f.sigma)
        assert self.u0.dtype == torch.complex64 and self.u0.shape[-1] == N and self.u0.stride(-1) == 1
        s.u0 = _lib.ptr(self.u0)
        s.u0_pstride = self.u0.stride(0)
        if self.out is not None:
            assert self.out.shape[0] == P and self.out.shape[-1] == N and self.out.stride(-1) == 1
            s.out = _lib.ptr(self.out)
            s.out_pstride = self.out.stride(0)
            s.out_zstride = self.out.stride(1)
            s.out_z0 = f.steps if self.flip else 0
            s.out_zdir = -1 if self.flip else 1
        else:
            work = self.work
            if work is None:
                work = torch.empty((P, 2, N), dtype=torch.complex64, device=self.u0.device)
                self._keep.append(work)
            assert work.shape[0] >= P and work.stride(0) == 2 * N and work.stride(-1) == 1
            s.work = _lib.ptr(work)
            s.work_pstride = 2 * N
        if self.exit is not None:
            s.exitp = _lib.ptr(self.exit)
            s.exit_pstride = self.exit.stride(0)
        if self.ckpt is not None:
            s.ckpt = _lib.ptr(self.ckpt)
            s.ckpt_pstride = self.ckpt.stride(0)
            s.ckpt_zstride = self.ckpt.stride(1)
            s.ckpt_every = self.ckpt_every
        if self.rhs is not None:
            r = self.rhs if f.b_sign == 1.0 else self.rhs * f.b_sign
            if r.stride(-1) != 1:
                r = r.contiguous()
            self._keep.append(r)
            s.rhs = _lib.ptr(r)
            s.rhs_pstride = 0 if r.shape[0] == 1 else r.stride(0)
            s.rhs_sstride = r.stride(1)
            s.rhs_steps = self.rhs_steps
        s.nx, s.ny, s.w, s.steps, s.probes = f.nx, f.ny, f.w, f.steps, P
        return s


def sweep(specs: Sequence[BandScanSpec], probes_per_block: int = 1, stream=None, schedule: int = 0) -> None:
    """All probes of up to eight batches through every slice in one launch;
    ``probes_per_block`` (1, 2, 4) probes share each workgroup and each load of
    the band coefficients; ``schedule`` 1 trades prefetch depth for two
    workgroups per CU (w <= 8, 1 or 2 probes per workgroup)."""
    assert 1 <= len(specs) <= 8
    arr = (_lib.BandScan * len(specs))(*[sp.struct() for sp in specs])
    _lib.check(_lib.load().tp_band_sweep_ex(arr, len(specs), int(probes_per_block), int(schedule),
                                            _lib.stream_handle(stream)), "tp_band_sweep_ex")


def sweep_shapes(w: int):
    """(probes_per_block, schedule) pairs built for band half-width w."""
    shapes = [(1, 0), (2, 0), (4, 0)]
    if w <= 8:
        shapes += [(1, 1), (2, 1)]
    return shapes


def residual(bf: "BandFactorizer", C: torch.Tensor, parts: Sequence[torch.Tensor], r: torch.Tensor,
             stream=None) -> torch.Tensor:
    """r_s = 2 v_{s-1} - A_s (v_{s-1} + v_s) in FP64 (tp_band_residual) for every step
    of the factorizer's operator (its own FP64 blocks and C, an (N, L) complex128
    view of any strides), v the sum of the complex64 plane sets ``parts`` (each
    (P, L+1, N), the same strides); r: complex64 (P, L, N) with unit last stride.
    A complex64 sweep of the same factors on r (rhs_steps = L) gives the correction
    of the planes: refinement of the forward stepping (sigma = 2 modes)."""
    pl = bf.plan
    if bf.source is not None or pl.sigma or pl.b_sign != 1.0:
        raise ValueError("residuals are built for the forward factorizer")
    if not 1 <= len(parts) <= 3:
        raise ValueError("one to three plane sets")
    u = parts[0]
    P, L = u.shape[0], bf.L
    N = bf.nx * bf.ny
    for x in parts:
        assert x.dtype == torch.complex64 and x.shape[0] == P and x.shape[1] >= L + 1 and x.shape[-1] == N
        assert x.stride() == u.stride() and x.stride(-1) == 1
    assert C.dtype == torch.complex128 and tuple(C.shape) == (N, L)
    assert r.dtype == torch.complex64 and r.shape[0] == P and r.shape[1] >= L and r.stride(-1) == 1
    ptrs = [_lib.ptr(x) for x in parts] + [0] * (3 - len(parts))
    _lib.check(_lib.load().tp_band_residual(ptrs[0], ptrs[1], ptrs[2], u.stride(0), u.stride(1), P, bf.nx, bf.ny, L,
                                            _lib.ptr(C), C.stride(0), C.stride(1), L, _lib.ptr(bf.tri),
                                            _lib.ptr(bf.cpl), pl.lhs_sign, pl.conj_c, pl.lhs_col0, pl.col_step, 0,
                                            _lib.ptr(r), r.stride(0), r.stride(1), _lib.stream_handle(stream)),
               "tp_band_residual")
    return r
