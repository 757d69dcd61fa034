"""Device-side plumbing for the 2D (one transverse axis) CN stepping kernels.

``factor`` runs ``tp_tri_factor`` for one (object, mode) and keeps the
per-step coefficients on the GPU; ``TriScanSpec`` describes one batch of
probes stepped through those factors; ``sweep`` launches the persistent
batched kernel over up to eight such batches (scans of different widths in
one launch) and ``step`` the plain one-launch-per-slice form.

Widths above 2048 rows take the wide layout (``operators.layout_for``): each
probe's rows are split over several wavefronts of one workgroup
(``tp_tri_factor_wide`` / ``tp_tri_sweep_wide``); ``TriFactors.wide`` marks
factors in that layout.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from .. import _lib
from .operators import ModePlan, Tridiag, layout_for, mode_plan


@dataclass
class TriFactors:
    coef: torch.Tensor
    n: int
    steps: int
    chunk: int
    mode: str
    b_sign: float
    status: torch.Tensor  # int32 [2]: zero-pivot flag, coupling reach (tp_tri_factor's d_status)
    sigma: bool
    keep: list = field(default_factory=list)
    waves: int = 1  # wavefronts per probe (wide layout only)
    wide: bool = False  # factored for tp_tri_sweep_wide (row i at (i % chunk) * 64 * waves + i / chunk)
    plan: Optional[ModePlan] = None  # the operator it was factored from (for residuals)
    first: int = 0  # steps of the full factor set before this one (step_slice views)

    def check(self) -> None:
        """Raise if a pivot was exactly zero (synchronises)."""
        if int(self.status[0].item()) != 0:
            raise RuntimeError("singular per-slice factor (zero pivot) in tp_tri_factor")


def check_pivots(facs) -> None:
    """Raise RuntimeError if any of the factor sets hit an exactly zero pivot
    (one synchronisation for all of them); the reference's splu raises there."""
    st = [f.status for f in facs if isinstance(f, TriFactors)]
    if st and int(torch.stack([x[0] for x in st]).max().item()) != 0:
        raise RuntimeError("singular per-slice factor (zero pivot) in tp_tri_factor")


# The per-mode plan and its device copy of the base diagonals depend only on the
# step operators (A0, B0), the mode and L -- not on the object -- so repeated
# factorisations (a new object every pass or iteration) reuse them, like the
# reference's cached step matrices (forward_model.py:47-66).  Keyed by the
# operator objects themselves (kept alive by the cache), at most 64 entries.
_PLAN_CACHE: "dict" = {}


def _plan_and_base(A: Tridiag, B: Tridiag, mode: str, L: int, dev):
    key = (id(A), id(B), mode, L, str(dev))
    hit = _PLAN_CACHE.get(key)
    if hit is not None and hit[0] is A and hit[1] is B:
        return hit[2], hit[3]
    plan: ModePlan = mode_plan(A, B, mode, L)
    plan.check_complementary()
    base = torch.from_numpy(plan.base_array().astype(np.complex128)).to(dev)
    if len(_PLAN_CACHE) >= 64:
        _PLAN_CACHE.pop(next(iter(_PLAN_CACHE)))
    _PLAN_CACHE[key] = (A, B, plan, base)
    return plan, base


def factor(C: torch.Tensor, A: Tridiag, B: Tridiag, mode: str, chunk: Optional[int] = None,
           stream=None, waves: Optional[int] = None, out: Optional[torch.Tensor] = None) -> TriFactors:
    """Factor every slice's LHS for ``mode``.

    C: complex128 CUDA tensor (n, L) from create_object_contribution(...).reshape(-1, nz-1).
    chunk / waves: the layout; by default ``layout_for(n)``.  ``waves`` given
    (even 1) selects the wide kernels with that many wavefronts per probe.
    out: optional complex64 buffer of coef_elems(L, chunk, waves) values."""
    assert C.is_cuda and C.dtype == torch.complex128 and C.dim() == 2
    n, L = C.shape
    C = C.contiguous()
    if chunk is None and waves is None:
        chunk, nw = layout_for(n)
        wide = nw > 1
    else:
        chunk = chunk or layout_for(n)[0]
        wide = waves is not None
        nw = waves if wide else 1
    dev = C.device
    plan, base = _plan_and_base(A, B, mode, L, dev)
    elems = coef_elems(L, chunk, nw)
    if out is not None:
        assert out.dtype == torch.complex64 and out.is_contiguous() and out.numel() == elems
        coef = out
    else:
        coef = torch.empty(elems, dtype=torch.complex64, device=dev)
    status = new_status(dev)
    lib = _lib.load()
    sh = _lib.stream_handle(stream)
    if wide:
        _lib.check(
            lib.tp_tri_factor_wide(_lib.ptr(C), n, L, _lib.ptr(base), plan.lhs_sign, plan.rhs_sign, plan.conj_c,
                                   plan.lhs_col0, plan.rhs_col0, plan.col_step, L, chunk, nw, int(plan.sigma),
                                   _lib.ptr(coef), _lib.ptr(status), sh),
            "tp_tri_factor_wide",
        )
    else:
        _lib.check(
            lib.tp_tri_factor(_lib.ptr(C), n, L, _lib.ptr(base), plan.lhs_sign, plan.rhs_sign, plan.conj_c,
                              plan.lhs_col0, plan.rhs_col0, plan.col_step, L, chunk, int(plan.sigma), _lib.ptr(coef),
                              _lib.ptr(status), sh),
            "tp_tri_factor",
        )
    return TriFactors(coef, n, L, chunk, mode, plan.b_sign, status, plan.sigma, keep=[base, C], waves=nw, wide=wide,
                      plan=plan)


def near_offdiagonals(plan: Optional[ModePlan], n: int):
    """tp_tri_scan.offd for the near-carry step: the factored LHS's off-diagonals
    (uniform sub on rows 2..n-2 and sup on rows 1..n-3, and the boundary values
    sub[1], sub[n-1], sup[0], sup[n-2]) as 12 floats, or None when the operator
    is not of that form.  The step multiplies by them as purely imaginary
    numbers, which the CN operators of every boundary condition and mode are
    (the paraxial term i dz / (4 k dx^2) with real k, dx, dz); any real part
    leaves the scan on the full carries."""
    if plan is None or n < 4:
        return None
    sub, sup = plan.lhs.sub, plan.lhs.sup
    s_sub, s_sup = sub[n // 2], sup[n // 2]
    if not (np.all(sub[2:n - 1] == s_sub) and np.all(sup[1:n - 2] == s_sup)):
        return None
    if np.any(sub[1:n].real != 0) or np.any(sup[:n - 1].real != 0):
        return None
    vals = np.array([s_sub, s_sup, sub[1], sub[n - 1], sup[0], sup[n - 2]], dtype=np.complex64)
    return [float(x) for v in vals for x in (v.real, v.imag)]


def new_status(device) -> torch.Tensor:
    """Zeroed factor status words: [0] zero-pivot flag, [1] coupling reach
    (the rows a carry reaches in a chunk; below the chunk the sweep skips the
    64-lane carry scans)."""
    return torch.zeros(2, dtype=torch.int32, device=device)


def coef_elems(L: int, chunk: int, waves: int = 1) -> int:
    """Complex elements of one factor set of L slices (tp_tri_coef_bytes[_wide] / 8)."""
    return L * 6 * 64 * waves * chunk


def factor_many(jobs, chunk: int, stream=None, out: Optional[list] = None, waves: Optional[int] = None,
                status: Optional[torch.Tensor] = None) -> list:
    """``factor`` for up to eight independent (C, A, B, mode) jobs in one launch
    (tp_tri_factor_batch); returns their TriFactors in order.  ``out`` optionally
    gives each job's coefficient buffer (complex64, coef_elems(L, chunk, waves)
    values).  ``waves`` selects the wide layout, factored one job per launch.
    ``status`` (from new_status) is shared by every job when given: per-probe
    factor sets stepped in one sweep need the zero-pivot flag and the coupling
    reach of all of them."""
    assert 1 <= len(jobs) <= _lib.TP_MAX_SCANS
    if waves is not None:
        facs = [factor(C, A, B, mode, chunk=chunk, waves=waves, stream=stream, out=None if out is None else out[k])
                for k, (C, A, B, mode) in enumerate(jobs)]
        if status is not None:  # one status word pair for all of them
            for f in facs:
                status.copy_(torch.maximum(status, f.status))
                f.status = status
        return facs
    facs, arr = [], (_lib.TriFactorJob * len(jobs))()
    for k, (C, A, B, mode) in enumerate(jobs):
        assert C.is_cuda and C.dtype == torch.complex128 and C.dim() == 2
        n, L = C.shape
        C = C.contiguous()
        plan, base = _plan_and_base(A, B, mode, L, C.device)
        if out is not None:
            coef = out[k]
            assert coef.dtype == torch.complex64 and coef.is_contiguous() and coef.numel() == coef_elems(L, chunk)
        else:
            coef = torch.empty(coef_elems(L, chunk), dtype=torch.complex64, device=C.device)
        st = status if status is not None else new_status(C.device)
        j = arr[k]
        j.C, j.base, j.coef, j.status = _lib.ptr(C), _lib.ptr(base), _lib.ptr(coef), _lib.ptr(st)
        j.lhs_sign, j.rhs_sign = plan.lhs_sign, plan.rhs_sign
        j.n, j.ncol, j.steps, j.conj_c = n, L, L, plan.conj_c
        j.lhs_col0, j.rhs_col0, j.col_step, j.with_sigma = plan.lhs_col0, plan.rhs_col0, plan.col_step, int(plan.sigma)
        facs.append(TriFactors(coef, n, L, chunk, mode, plan.b_sign, st, plan.sigma, keep=[base, C], plan=plan))
    _lib.check(_lib.load().tp_tri_factor_batch(arr, len(jobs), chunk, _lib.stream_handle(stream)),
               "tp_tri_factor_batch")
    return facs


@dataclass
class TriScanSpec:
    """One batch of probes stepped through ``fac``.

    u0:   complex64 (P, n) initial planes (contiguous rows).
    out:  complex64 (P, L+1, n) full field, or None.
    exit: complex64 (P, n) last plane, or None.
    ckpt: complex64 (P, L//every + 1, n) checkpoints, or None.
    rhs:  complex64 (P or 1, L, n) extra right-hand side per step (step order), or None.
    flip: write plane s at index L - s (the reference adjoint's output flip).
    """

    fac: TriFactors
    u0: torch.Tensor
    out: Optional[torch.Tensor] = None
    exit: Optional[torch.Tensor] = None
    ckpt: Optional[torch.Tensor] = None
    ckpt_every: int = 0
    rhs: Optional[torch.Tensor] = None
    flip: bool = False
    rhs_steps: int = 0
    coef_pstride: int = 0  # per-probe factor sets (reduced domain), see tp_tri_scan
    gfwd: Optional[torch.Tensor] = None
    gfwd_z0: int = 0
    gfwd_zdir: int = 1
    grad: Optional[torch.Tensor] = None
    grad_z0: int = 0
    grad_zdir: int = 1
    grad_skip_row0: bool = False
    near: bool = True  # False: full 64-lane carries whatever the factors' coupling reach

    def struct(self) -> _lib.TriScan:
        f = self.fac
        P = self.u0.shape[0]
        n, L = f.n, f.steps
        s = _lib.TriScan()
        s.coef = _lib.ptr(f.coef)
        s.coef_pstride = int(self.coef_pstride)
        if self.near and not f.wide and f.status.numel() >= 2:
            s.reach = _lib.ptr(f.status) + f.status.element_size()
            offd = near_offdiagonals(f.plan, f.n)
            if offd is not None:
                s.offd[:] = offd
                s.has_offd = 1
        s.sigma = int(f.sigma)
        assert self.u0.dtype == torch.complex64 and self.u0.shape[-1] == n and self.u0.stride(-1) == 1
        s.u0 = _lib.ptr(self.u0)
        s.u0_pstride = self.u0.stride(0) if P > 1 else n
        if self.rhs is not None:
            nsteps = self.rhs_steps or L
            assert self.rhs.dtype == torch.complex64 and self.rhs.dim() == 3 and self.rhs.stride(-1) == 1
            assert self.rhs.shape[1] >= nsteps and self.rhs.shape[2] == n and self.rhs.shape[0] in (1, P)
            s.rhs = _lib.ptr(self.rhs)
            s.rhs_pstride = self.rhs.stride(0) if self.rhs.shape[0] > 1 else 0
            s.rhs_sstride = self.rhs.stride(1)
            s.rhs_steps = self.rhs_steps
        if self.grad is not None:
            g = self.gfwd
            assert g is not None and g.dtype == torch.complex64 and g.shape[0] == P and g.shape[2] == n
            assert g.stride(-1) == 1 and 0 <= self.gfwd_z0 < g.shape[1] and 0 <= self.gfwd_z0 + self.gfwd_zdir * L < g.shape[1]
            r = self.grad
            assert r.dtype == torch.complex64 and r.dim() == 2 and r.shape[1] == n and r.stride(-1) == 1
            assert 0 <= self.grad_z0 < r.shape[0] and 0 <= self.grad_z0 + self.grad_zdir * (L - 1) < r.shape[0]
            s.gfwd = _lib.ptr(g)
            s.gfwd_pstride = g.stride(0)
            s.gfwd_zstride = g.stride(1)
            s.gfwd_z0 = self.gfwd_z0
            s.gfwd_zdir = self.gfwd_zdir
            s.grad = _lib.ptr(r)
            s.grad_zstride = r.stride(0)
            s.grad_z0 = self.grad_z0
            s.grad_zdir = self.grad_zdir
            s.grad_skip_row0 = int(self.grad_skip_row0)
        if self.out is not None:
            assert self.out.dtype == torch.complex64 and self.out.shape == (P, L + 1, n) and self.out.stride(-1) == 1
            s.out = _lib.ptr(self.out)
            s.out_pstride = self.out.stride(0)
            s.out_zstride = self.out.stride(1)
            s.out_z0 = L if self.flip else 0
            s.out_zdir = -1 if self.flip else 1
        if self.ckpt is not None:
            assert self.ckpt_every >= 1 and self.ckpt.shape == (P, L // self.ckpt_every + 1, n)
            s.ckpt = _lib.ptr(self.ckpt)
            s.ckpt_pstride = self.ckpt.stride(0)
            s.ckpt_zstride = self.ckpt.stride(1)
            s.ckpt_every = self.ckpt_every
        if self.exit is not None:
            assert self.exit.dtype == torch.complex64 and self.exit.shape == (P, n) and self.exit.stride(-1) == 1
            s.exitp = _lib.ptr(self.exit)
            s.exit_pstride = self.exit.stride(0)
        s.n = n
        s.steps = L
        s.probes = P
        return s


def step_slice(fac: TriFactors, first: int, count: int) -> TriFactors:
    """View of ``fac`` restricted to steps first+1 .. first+count (no copy)."""
    assert 0 <= first and first + count <= fac.steps and count >= 1
    per = coef_elems(1, fac.chunk, fac.waves)
    return TriFactors(fac.coef[first * per:(first + count) * per], fac.n, count, fac.chunk, fac.mode, fac.b_sign,
                      fac.status, fac.sigma, keep=[fac], waves=fac.waves, wide=fac.wide, plan=fac.plan,
                      first=fac.first + first)


def residual(fac: TriFactors, planes, stream=None) -> torch.Tensor:
    """r_s = 2 v_{s-1} - A_s (v_{s-1} + v_s) in FP64 (tp_tri_residual) for the
    steps of ``fac`` (a full factor set or a step_slice view): the residual of the
    FP64 recurrence at the iterate v, given as one complex64 (P, steps+1, n)
    plane set or a list of up to three with equal strides whose sum is v (plane
    0 the state before the view's first step).  Returns complex64 (P, steps, n),
    the right-hand side of a refinement sweep of the same factors."""
    if fac.plan is None or fac.sigma:
        raise ValueError("residuals are built for factors of the sigma = 2 modes from tri_engine.factor")
    parts = list(planes) if isinstance(planes, (list, tuple)) else [planes]
    assert 1 <= len(parts) <= 3
    P, S1, n = parts[0].shape
    for q in parts:
        assert q.shape == parts[0].shape and q.stride() == parts[0].stride() and q.dtype == torch.complex64
    assert S1 == fac.steps + 1 and n == fac.n and parts[0].stride(-1) == 1
    root = fac
    while root.keep and isinstance(root.keep[0], TriFactors):  # step_slice views keep their parent
        root = root.keep[0]
    base, C = root.keep[:2]
    plan = fac.plan
    ptrs = [_lib.ptr(q) for q in parts] + [None] * (3 - len(parts))
    r = torch.empty((P, fac.steps, n), dtype=torch.complex64, device=parts[0].device)
    _lib.check(_lib.load().tp_tri_residual(ptrs[0], ptrs[1], ptrs[2], parts[0].stride(0), parts[0].stride(1), P, n,
                                           fac.steps, _lib.ptr(C), C.shape[1], _lib.ptr(base), plan.lhs_sign,
                                           plan.conj_c, plan.lhs_col0, plan.col_step, fac.first, _lib.ptr(r),
                                           r.stride(0), r.stride(1), _lib.stream_handle(stream)), "tp_tri_residual")
    return r


def _structs(specs: Sequence[TriScanSpec], allow_half_chunk: bool = False):
    if not 1 <= len(specs) <= _lib.TP_MAX_SCANS:
        raise ValueError(f"between 1 and {_lib.TP_MAX_SCANS} scans per launch")
    chunk = max(s.fac.chunk for s in specs)
    structs = []
    for s in specs:
        st = s.struct()
        if s.fac.chunk != chunk:
            if not (allow_half_chunk and 2 * s.fac.chunk == chunk and chunk >= 16 and s.grad is None):
                raise ValueError("all scans of one launch must be factored with the same chunk (the persistent "
                                 "sweep also takes half a chunk of 16 or 32 for scans without gradients)")
            st.chunk = s.fac.chunk
        structs.append(st)
    return (_lib.TriScan * len(specs))(*structs), chunk


def batch_for(probes: int, device=None) -> int:
    """Probes per workgroup for a launch of ``probes`` probes: 4 when workgroups
    of 4 still give every compute unit one, else 1 (a few hundred probes in
    workgroups of 4 would leave most units idle)."""
    dev = torch.device("cuda") if device is None else torch.device(device)
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    return 4 if (probes + 3) // 4 >= cus else 1


def sweep(specs: Sequence[TriScanSpec], probes_per_block: Optional[int] = None, stream=None) -> None:
    """Persistent batched sweep over all steps of all scans (one launch).  Scans
    factored with half the widest scan's chunk (16 or 32) step with that many
    rows per lane inside the same launch.  ``probes_per_block`` None: from the
    launch's probe count (``batch_for``).  Wide-layout scans go to
    tp_tri_sweep_wide, one launch per (chunk, waves)."""
    wide = [s for s in specs if s.fac.wide]
    if wide:
        groups = {}
        for spec in wide:
            groups.setdefault((spec.fac.chunk, spec.fac.waves), []).append(spec)
        lib = _lib.load()
        for (chunk, waves), group in groups.items():
            arr = (_lib.TriScan * len(group))(*[g.struct() for g in group])
            _lib.check(lib.tp_tri_sweep_wide(arr, len(group), chunk, waves, _lib.stream_handle(stream)),
                       "tp_tri_sweep_wide")
        specs = [s for s in specs if not s.fac.wide]
        if not specs:
            return
    if probes_per_block is None:
        probes_per_block = batch_for(sum(int(s.u0.shape[0]) for s in specs), specs[0].u0.device)
    arr, chunk = _structs(specs, allow_half_chunk=True)
    if chunk >= 32:
        # LDS of 2048-row blocks: two probes per workgroup, one with gradient accumulation
        probes_per_block = 1 if specs[0].grad is not None else min(probes_per_block, 2)
    lib = _lib.load()
    _lib.check(lib.tp_tri_sweep(arr, len(specs), chunk, probes_per_block, _lib.stream_handle(stream)), "tp_tri_sweep")


def _chunk_groups(specs: Sequence[TriScanSpec]):
    """(struct array, count, chunk) per chunk among ``specs``: the plain form
    launches each chunk separately."""
    groups = {}
    for spec in specs:
        if spec.fac.wide:
            raise ValueError("the plain one-launch-per-slice form is built for the one-wavefront layout (n <= 2048); "
                             "wide grids step in the persistent sweep")
        groups.setdefault(spec.fac.chunk, []).append(spec)
    out = []
    for chunk in sorted(groups, reverse=True):
        arr, _ = _structs(groups[chunk])
        out.append((arr, len(groups[chunk]), chunk))
    return out


def step(specs: Sequence[TriScanSpec], s: int, stream=None) -> None:
    """Plain form: one launch per chunk width advancing every probe by step ``s``."""
    lib = _lib.load()
    sh = _lib.stream_handle(stream)
    for arr, n, chunk in _chunk_groups(specs):
        _lib.check(lib.tp_tri_step(arr, n, chunk, s, sh), "tp_tri_step")


def sweep_plain(specs: Sequence[TriScanSpec], stream=None) -> None:
    """Plain form over all steps: one launch per slice (and chunk width)."""
    L = specs[0].fac.steps
    for spec in specs:
        if spec.fac.steps != L or spec.out is None:
            raise ValueError("plain sweeps need equal step counts and full-field outputs")
    groups = _chunk_groups(specs)
    lib = _lib.load()
    sh = _lib.stream_handle(stream)
    for s in range(1, L + 1):
        for arr, n, chunk in groups:
            _lib.check(lib.tp_tri_step(arr, n, chunk, s, sh), "tp_tri_step")


class PlainGraph:
    """The plain form (one ``tp_tri_step`` launch per slice for every scan) captured
    once into a HIP graph and replayed.

    The scan descriptors, and so every device pointer, are fixed at capture:
    a replay steps the same buffers again.  New probes or a new object go into
    the same tensors in place (``u0.copy_(...)``, a new factorisation written
    into ``fac.coef``) between replays."""

    def __init__(self, specs: Sequence[TriScanSpec]):
        self.specs = list(specs)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):  # first launches outside the capture
            sweep_plain(self.specs)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            sweep_plain(self.specs)

    def replay(self) -> None:
        self.graph.replay()
