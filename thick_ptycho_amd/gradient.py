"""Gradients of the far-field amplitude misfit through CN slice stepping (GPU).

Loss (per scan):  J(n) = sum_p sum_k (|fft(u_L^p)|_k - sqrt(I_k^p))^2, with u_L the
exit plane of the forward sweep.

``amplitude_gradient`` returns the exact gradient dJ/dRe(n) + i dJ/dIm(n): forward
sweep keeping every ``checkpoint``-th plane, the loss seed dJ/d conj(u_L) from
``tp_amplitude_seed``, then the exact adjoint sweep segment by segment in
reverse, each segment's forward planes recomputed from its checkpoint, with
the per-slice accumulation sum_p conj(lambda_s) (u_s + u_{s-1}) fused into the
adjoint kernel.  Memory: (L/checkpoint + 1) planes per probe instead of L + 1.

``reference_gradient`` reproduces ReconstructorPWE.compute_grad_least_squares
(thick_ptycho/reconstruction/pwe_reconstructor.py:100-164): the reference
adjoint (quirk Q3), the exit-wave error of base_reconstructor.py:87-207 (the
seed scaled by 1/N), g = g0 (u_s + u_{s-1}) without the u_0 term, and the
(grad_re, grad_im) split with the reference's signs (its loop passes
grad_Ak = -get_gradient(n), pwe_reconstructor.py:257, 281).

``ExitWaves`` is a torch.autograd.Function mapping n -> exit waves whose
backward runs the same checkpointed exact adjoint.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from .forward_model import ms_engine, operators, tri_engine


def object_contribution(n: torch.Tensor, k: float, dz: float) -> torch.Tensor:
    """C[:, j-1] = (dz/2) (coef[:, j-1] + coef[:, j]) / 2 with coef = (k/i)(n - 1), on the GPU.

    Same formula as base_simulation_space.py:315-383 with the default coefficient."""
    coef = (k / 1j) * (n - 1)
    return ((dz / 2) * (coef[:, :-1] + coef[:, 1:]) / 2).contiguous()


def amplitude_seed(exit_w: torch.Tensor, sqrt_i: torch.Tensor, scale: float = 1.0,
                   exit_lo=None, sqrt_i_lo: torch.Tensor = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """(seed, loss per probe): seed = scale * dJ/d conj(exit).  With ``exit_lo``
    (a correction plane or a list of up to two) the exit plane is exit plus the
    corrections (a refined plane) and the seed is formed in FP64
    (tp_amplitude_seed_refined, widths that are powers of two from 64 to 8192),
    with the amplitudes sqrt_i + sqrt_i_lo when the low part is given; other
    widths take the complex64 seed of the summed plane."""
    if exit_lo is None:
        return ms_engine.amplitude_seed(exit_w.contiguous(), sqrt_i.contiguous(), scale)
    parts = [exit_w.contiguous()] + [q.contiguous() for q in (exit_lo if isinstance(exit_lo, (list, tuple))
                                                               else [exit_lo])]
    P, n = exit_w.shape
    if not refined_seed_width(n):
        return ms_engine.amplitude_seed(sum(parts).contiguous(), sqrt_i.contiguous(), scale)
    assert len(parts) <= 3
    seed = torch.empty_like(exit_w)
    loss = torch.empty(P, dtype=torch.float32, device=exit_w.device)
    ptrs = [_lib.ptr(q) for q in parts] + [None] * (3 - len(parts))
    a_hi = sqrt_i.contiguous()
    a_lo = None if sqrt_i_lo is None else sqrt_i_lo.contiguous()
    _lib.check(_lib.load().tp_amplitude_seed_refined(ptrs[0], ptrs[1], ptrs[2], n, _lib.ptr(a_hi), _lib.ptr(a_lo), n,
                                                     P, float(scale), _lib.ptr(seed), _lib.ptr(loss),
                                                     _lib.stream_handle()), "tp_amplitude_seed_refined")
    return seed, loss


def refined_seed_width(n: int) -> bool:
    """Widths the FP64 seed of a refined exit plane takes (powers of two, 64..8192)."""
    return 64 <= n <= 8192 and (n & (n - 1)) == 0


def amplitude_parts(sqrt_i, device, shape):
    """(hi, lo) float32 of measured amplitudes given in any precision: hi the
    rounded value and lo what rounding left of FP64 data (0 for float32 data)."""
    a64 = _dev_tensor(sqrt_i, torch.float64, device).reshape(shape)
    hi = a64.to(torch.float32)
    return hi.contiguous(), (a64 - hi.to(torch.float64)).to(torch.float32).contiguous()


def _probe_parts(probes, device, N: int):
    """(hi, lo): the probes rounded to complex64 and their rounding error, the
    initial plane of a refinement sweep."""
    p64 = _dev_tensor(probes, torch.complex128, device).reshape(-1, N)
    hi = p64.to(torch.complex64)
    return hi.contiguous(), (p64 - hi.to(torch.complex128)).to(torch.complex64).contiguous()


# Refinement steps of the exit plane behind the amplitude seed: the complex64
# sweep's relative error (~2e-5 after 511 slices) per step, so two reach the
# FP64 level the amplitude loss needs at badly fitting points.
REFINE_STEPS = 2


class _Sweeper:
    """Forward factors + adjoint factors of one object, and the checkpointed sweeps."""

    def __init__(self, C: torch.Tensor, A, B, adjoint_mode: str, checkpoint: int, ppb: Optional[int] = None):
        self.n, self.L = C.shape
        self.fwd = tri_engine.factor(C, A, B, "forward")
        self.adj = tri_engine.factor(C, A, B, adjoint_mode)
        self.K = max(1, min(checkpoint if checkpoint > 0 else self.L, self.L))
        self._ppb = ppb
        self._dev = C.device

    def ppb_for(self, P: int) -> int:
        """Probes per workgroup for a sweep of P probes: ``ppb`` when given, else
        tri_engine.batch_for."""
        return self._ppb if self._ppb is not None else tri_engine.batch_for(P, self._dev)

    def refine_exit(self, ck: torch.Tensor, e0: torch.Tensor, steps: int = REFINE_STEPS) -> list:
        """Corrections [e1_L, ..., e_L] of the exit plane by ``steps`` refinement
        steps, segment by segment in forward order: the planes are recomputed from
        their checkpoint, and for each step the FP64 residual of the recurrence at
        the iterate so far (tri_engine.residual of the plane and the earlier
        corrections) drives a complex64 sweep of the same factors, the first from
        e0 (the probes' own rounding error), the others from zero.  Each step
        multiplies the exit-plane error by about the sweep's relative error."""
        P = e0.shape[0]
        L, K, n = self.L, self.K, self.n
        dev = e0.device
        segs = [torch.empty((P, K + 1, n), dtype=torch.complex64, device=dev) for _ in range(steps)]
        state = [e0.contiguous()] + [torch.zeros_like(e0) for _ in range(steps - 1)]
        for g in range(math.ceil(L / K)):
            s0 = g * K
            ln = min(s0 + K, L) - s0
            fw = tri_engine.step_slice(self.fwd, s0, ln)
            views = [q[:, : ln + 1] for q in segs]
            tri_engine.sweep([tri_engine.TriScanSpec(fw, ck[:, g], out=views[0])], probes_per_block=self.ppb_for(P))
            for k in range(1, steps + 1):
                r = tri_engine.residual(fw, views[:k])
                nxt = torch.empty_like(e0)
                out = views[k] if k < steps else None
                tri_engine.sweep([tri_engine.TriScanSpec(fw, state[k - 1], rhs=r, out=out, exit=nxt)],
                                 probes_per_block=self.ppb_for(P))
                state[k - 1] = nxt
        return state

    def forward(self, u0: torch.Tensor):
        P = u0.shape[0]
        dev = u0.device
        ck = torch.empty((P, self.L // self.K + 1, self.n), dtype=torch.complex64, device=dev)
        ex = torch.empty((P, self.n), dtype=torch.complex64, device=dev)
        tri_engine.sweep([tri_engine.TriScanSpec(self.fwd, u0, ckpt=ck, ckpt_every=self.K, exit=ex)],
                         probes_per_block=self.ppb_for(P))
        return ck, ex

    def adjoint(self, ck: torch.Tensor, seed: torch.Tensor, skip_row0: bool = False) -> torch.Tensor:
        """Accumulate G[s-1] = sum_p conj(lambda_s) (u_s + u_{s-1}) over s = L..1 (complex64 (L, n))."""
        P = seed.shape[0]
        dev = seed.device
        L, K, n = self.L, self.K, self.n
        grad = torch.zeros((L, n), dtype=torch.complex64, device=dev)
        lam = torch.zeros((P, n), dtype=torch.complex64, device=dev)
        seg = torch.empty((P, K + 1, n), dtype=torch.complex64, device=dev)
        seed3 = seed.reshape(P, 1, n)
        nseg = math.ceil(L / K)
        for g in range(nseg - 1, -1, -1):
            s0 = g * K
            s1 = min(s0 + K, L)
            ln = s1 - s0
            segv = seg[:, : ln + 1]
            fw = tri_engine.step_slice(self.fwd, s0, ln)
            tri_engine.sweep([tri_engine.TriScanSpec(fw, ck[:, g], out=segv)], probes_per_block=self.ppb_for(P))
            aw = tri_engine.step_slice(self.adj, L - s1, ln)
            first = s1 == L
            spec = tri_engine.TriScanSpec(
                aw, lam, exit=lam,
                rhs=seed3 if first else None, rhs_steps=1 if first else 0,
                gfwd=segv, gfwd_z0=ln, gfwd_zdir=-1,
                grad=grad, grad_z0=s1 - 1, grad_zdir=-1, grad_skip_row0=skip_row0,
            )
            ppb = self.ppb_for(P)
            tri_engine.sweep([spec], probes_per_block=ppb if ppb in (1, 4) else 4)
        return grad


def _operators(nx: int, k: float, dx: float, dz: float, bc):
    return operators.step_operators_1d(nx, k, dx, dz, bc)


def _dev_tensor(x, dtype, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to(device=device, dtype=dtype)


def amplitude_gradient(n, probes, sqrt_i, k: float, dx: float, dz: float, bc="dirichlet", checkpoint: int = 32,
                       device=None, seed=None, exit_data=None, refine: bool = True):
    """(loss, grad) of J(n) for one scan; grad complex128 (N, nz) on the GPU in the
    dJ/dRe(n) + i dJ/dIm(n) convention.

    ``seed`` (P, N), when given, replaces the loss seed dJ/d conj(u_L) that is
    otherwise computed from the exit waves and ``sqrt_i`` (the loss is still
    evaluated from them).  ``exit_data`` (P, N) switches to the exit-wave loss
    J = sum |u_L - exit_data|^2 (seed u_L - exit_data; ``sqrt_i`` unused).

    ``refine`` (default) corrects the exit plane by two refinement steps
    (``_Sweeper.refine_exit``) and forms the amplitude seed in FP64: the
    amplitude loss amplifies exit-plane errors by orders of magnitude in
    far-field bins far below the peak, so the complex64 plane alone leaves the
    gradient ~4e-3 from the FP64 one at a 1024-wide, 512-slice fit."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    n_t = _dev_tensor(n, torch.complex128, device)
    N, nz = n_t.shape
    A, B = _operators(N, k, dx, dz, bc)
    C = object_contribution(n_t, k, dz)
    sw = _Sweeper(C, A, B, "exact_adjoint", checkpoint)
    u0, u0_lo = _probe_parts(probes, device, N)
    ck, ex = sw.forward(u0)
    ex_lo = sw.refine_exit(ck, u0_lo) if refine and seed is None else None
    if exit_data is not None:
        sd = ex - _dev_tensor(exit_data, torch.complex64, device).reshape(u0.shape[0], N)
        if ex_lo is not None:
            sd = sd + sum(ex_lo)
        loss = (sd.abs() ** 2).sum(dim=1)
    else:
        amp, amp_lo = amplitude_parts(sqrt_i, device, (u0.shape[0], N))
        sd, loss = amplitude_seed(ex, amp, 1.0, exit_lo=ex_lo, sqrt_i_lo=amp_lo)
    if seed is not None:
        sd = _dev_tensor(seed, torch.complex64, device).reshape(u0.shape[0], N)
    G = sw.adjoint(ck, sd)
    Gn = torch.zeros((N, nz), dtype=torch.complex128, device=device)
    Gt = G.to(torch.complex128).transpose(0, 1)
    Gn[:, :-1] += Gt
    Gn[:, 1:] += Gt
    Gn *= (dz / 4) * (k / 1j)
    tri_engine.check_pivots([sw.fwd, sw.adj])  # with the loss read below: one synchronisation point
    return float(loss.double().sum().item()), 2 * torch.conj(Gn)


def amplitude_gradient_reduced(n, probes, sqrt_i, windows, k: float, dx: float, dz: float, bc="dirichlet",
                               device=None, exit_data=None):
    """(loss, grad) for reduced-domain scans: scan s sees only rows [a_s, b_s) of
    n (nx, nz), with probes / sqrt_i (or ``exit_data``) of the window width W;
    J = sum_s J_s and grad complex128 (nx, nz) = the window sum of the per-scan
    exact gradients (the accumulation of pwe_reconstructor.py:144-164 over
    simulation_space_2d.py:80-135 windows).

    Per-scan forward and exact-adjoint factor sets go into two buffers (eight
    factorisations per launch), one persistent sweep steps every scan through
    its own set in each direction keeping all planes, and
    tp_window_grad_accumulate sums the scans' slice products into the full
    width."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    n_t = _dev_tensor(n, torch.complex128, device)
    nx, nz = n_t.shape
    L = nz - 1
    win = [(int(a), int(b)) for a, b in windows]
    S = len(win)
    W = win[0][1] - win[0][0]
    if any(b - a != W or a < 0 or b > nx for a, b in win):
        raise ValueError("windows must all have the same width and lie inside the object")
    A, B = _operators(W, k, dx, dz, bc)
    C = object_contribution(n_t, k, dz)  # (nx, L); a window's rows are its object term
    chunk, waves = operators.layout_for(W)
    per = tri_engine.coef_elems(L, chunk, waves)
    wide = waves > 1
    whole = {}
    for mode in ("forward", "exact_adjoint"):
        coef = torch.empty((S, per), dtype=torch.complex64, device=device)
        status = tri_engine.new_status(device)  # shared: one sweep steps every window's set
        first = None
        for i in range(0, S, _lib.TP_MAX_SCANS):
            jobs = [(C[a:b], A, B, mode) for a, b in win[i:i + _lib.TP_MAX_SCANS]]
            facs = tri_engine.factor_many(jobs, chunk, out=[coef[s] for s in range(i, min(i + _lib.TP_MAX_SCANS, S))],
                                          waves=waves if wide else None, status=status)
            first = first or facs[0]
        whole[mode] = tri_engine.TriFactors(coef.reshape(-1), W, L, chunk, mode, first.b_sign, first.status,
                                            first.sigma, keep=[coef, facs], waves=waves, wide=wide, plan=first.plan)
    u0 = _dev_tensor(probes, torch.complex64, device).reshape(S, W)
    U = torch.empty((S, L + 1, W), dtype=torch.complex64, device=device)
    tri_engine.sweep([tri_engine.TriScanSpec(whole["forward"], u0, out=U, coef_pstride=per)], probes_per_block=1)
    ex = U[:, L].contiguous()
    if exit_data is not None:
        sd = ex - _dev_tensor(exit_data, torch.complex64, device).reshape(S, W)
        loss = (sd.abs() ** 2).sum(dim=1)
    else:
        sd, loss = amplitude_seed(ex, _dev_tensor(sqrt_i, torch.float32, device).reshape(S, W), 1.0)
    Lam = torch.empty_like(U)
    zero = torch.zeros((S, W), dtype=torch.complex64, device=device)
    tri_engine.sweep([tri_engine.TriScanSpec(whole["exact_adjoint"], zero, out=Lam, rhs=sd.reshape(S, 1, W).contiguous(),
                                             rhs_steps=1, coef_pstride=per)], probes_per_block=1)
    wins = torch.tensor([a for a, _ in win], dtype=torch.int32, device=device)
    G = torch.empty((L, nx), dtype=torch.complex64, device=device)
    _lib.check(_lib.load().tp_window_grad_accumulate(_lib.ptr(U), _lib.ptr(Lam), _lib.ptr(wins), S, L, W, nx,
                                                       _lib.ptr(G), _lib.stream_handle()), "tp_window_grad_accumulate")
    Gn = torch.zeros((nx, nz), dtype=torch.complex128, device=device)
    Gt = G.to(torch.complex128).transpose(0, 1)
    Gn[:, :-1] += Gt
    Gn[:, 1:] += Gt
    Gn *= (dz / 4) * (k / 1j)
    return float(loss.double().sum().item()), 2 * torch.conj(Gn)


def reference_gradient(n, probes, sqrt_i, k: float, dx: float, dz: float, bc="dirichlet", checkpoint: int = 32,
                       device=None, exit_error=None, refine: bool = True):
    """(grad_real, grad_imag), each float64 (L*N,) step-major, as
    ReconstructorPWE.compute_grad_least_squares with phase retrieval.

    ``exit_error`` (P, N), when given, is the exit-wave error e = u_L - P_a(u_L)
    supplied by the caller (the reference's ``exit_wave_error`` argument,
    pwe_reconstructor.py:100); otherwise it is computed from the exit waves
    and ``sqrt_i`` (base_reconstructor.py:87-121, 171-207) and ``sqrt_i`` may be None
    only when ``exit_error`` is given."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    n_t = _dev_tensor(n, torch.complex128, device)
    N, nz = n_t.shape
    A, B = _operators(N, k, dx, dz, bc)
    C = object_contribution(n_t, k, dz)
    sw = _Sweeper(C, A, B, "adjoint", checkpoint)
    u0, u0_lo = _probe_parts(probes, device, N)
    ck, ex = sw.forward(u0)
    if exit_error is not None:
        err = _dev_tensor(exit_error, torch.complex64, device).reshape(u0.shape[0], N)
    else:
        amp, amp_lo = amplitude_parts(sqrt_i, device, (u0.shape[0], N))
        ex_lo = sw.refine_exit(ck, u0_lo) if refine else None
        err, _ = amplitude_seed(ex, amp, 1.0 / N, exit_lo=ex_lo, sqrt_i_lo=amp_lo)
    G = sw.adjoint(ck, err, skip_row0=True).to(torch.complex128)
    return tuple(t.cpu().numpy() for t in reference_gradient_parts(G, k, dz))


def reference_gradient_parts(G: torch.Tensor, k: float, dz: float):
    """(grad_real, grad_imag) float64 (L*N,) step-major from the accumulation
    G[s-1] = sum_p conj(lambda_s)(u_s + u_{s-1}) (no u_0 term) of the reference
    adjoint states: with g = grad_Ak u = -g0 (u_s + u_{s-1}), g0 = dz k / 2i,
    grad_re = -sum Re(conj(g) lambda) = Re(g0 G) and grad_im = -sum
    Re(conj(i g) lambda) = -Im(g0 G)."""
    X = ((dz * k / 2j) * G.to(torch.complex128)).reshape(-1)
    return X.real.contiguous(), (-X.imag).contiguous()


class ExitWaves(torch.autograd.Function):
    """n (complex128 (N, nz), CUDA) -> exit planes (complex64 (P, N)) of the CN forward sweep.

    backward runs the checkpointed exact adjoint; the returned gradient follows
    torch's convention for complex inputs (conjugate Wirtinger)."""

    @staticmethod
    def forward(ctx, n, probes, k: float, dx: float, dz: float, bc: str = "dirichlet", checkpoint: int = 32):
        N, nz = n.shape
        A, B = _operators(N, k, dx, dz, bc)
        C = object_contribution(n.detach(), k, dz)
        sw = _Sweeper(C, A, B, "exact_adjoint", checkpoint)
        u0 = probes.detach().to(torch.complex64).reshape(-1, N).contiguous()
        ck, ex = sw.forward(u0)
        ctx.sw = sw
        ctx.k, ctx.dz, ctx.nz = k, dz, nz
        ctx.save_for_backward(ck)
        return ex

    @staticmethod
    def backward(ctx, grad_exit):
        (ck,) = ctx.saved_tensors
        sw = ctx.sw
        G = sw.adjoint(ck, grad_exit.to(torch.complex64).contiguous())
        N = G.shape[1]
        Gn = torch.zeros((N, ctx.nz), dtype=torch.complex128, device=G.device)
        Gt = G.to(torch.complex128).transpose(0, 1)
        Gn[:, :-1] += Gt
        Gn[:, 1:] += Gt
        Gn *= (ctx.dz / 4) * (ctx.k / 1j)
        return torch.conj(Gn), None, None, None, None, None, None


# ---------------------------------------------------------------------------
# Multislice: exact gradient of the far-field amplitude misfit.
# psi_{z+1} = P (T_z psi_z), P = F^-1 H F (ms_solver.py:143-158).  With
# lambda_L = dJ/d conj(psi_L), mu_z = P^H lambda_{z+1} and lambda_z = conj(T_z) mu_z,
# so the adjoint is the same multislice sweep with conj(H), started from the
# seed, whose transmission at step j is conj(T_{L-j}) (1 at step 0): its plane j
# is mu_{L-j}.  Then dJ/dRe n_z + i dJ/dIm n_z = -2 i k dz conj(T_z) sum_p conj(psi_z) mu_z
# (tp_plane_conj_dot), and n's last column has zero gradient.
# ---------------------------------------------------------------------------


def _ms_gradient_core(T, u0, run, seed_fn, k: float, dz: float, seed=None):
    L, N = T.shape
    P = u0.shape[0]
    dev = T.device
    u = torch.empty((P, L + 1, N), dtype=torch.complex64, device=dev)
    run(T, u0, u, False)
    ex = u[:, L].contiguous()
    sd, loss = seed_fn(ex)
    if seed is not None:
        sd = _dev_tensor(seed, torch.complex64, dev).reshape(P, N).contiguous()
    return float(loss.double().sum()), _ms_backward(T, u, sd, run, k, dz)


def _ms_backward(T, u, sd, run, k: float, dz: float):
    """dJ/dRe n + i dJ/dIm n, complex128 (N, L + 1), from all forward planes u
    (P, L + 1, N) and the seed sd = dJ/d conj(psi_L) (P, N)."""
    L, N = T.shape
    P = u.shape[0]
    dev = T.device
    Tadj = torch.empty_like(T)
    Tadj[0] = 1
    if L > 1:
        Tadj[1:] = T.flip(0)[: L - 1].conj()
    lam = torch.empty_like(u)
    run(Tadj, sd, lam, True)
    G = torch.zeros((L, N), dtype=torch.complex64, device=dev)
    _lib.check(_lib.load().tp_plane_conj_dot(_lib.ptr(u), (L + 1) * N, 0, 1, _lib.ptr(lam), (L + 1) * N, L, -1, P, L,
                                             N, _lib.ptr(G), _lib.stream_handle()), "tp_plane_conj_dot")
    grad = torch.zeros((N, L + 1), dtype=torch.complex128, device=dev)
    grad[:, :L] = ((-2j * k * dz) * (T.conj() * G)).transpose(0, 1).to(torch.complex128)
    return grad


class MSExitWaves(torch.autograd.Function):
    """n (complex128 (nx, nz), CUDA) -> exit planes (complex64 (P, nx)) of the
    multislice (ms_solver.py:90-159); backward runs the conj(H) adjoint sweep
    from the incoming gradient (torch's convention: grad_exit = 2 dJ/d conj(exit))
    and returns dJ/dRe n + i dJ/dIm n.  Every forward plane is kept for the
    backward pass (P nz nx complex64)."""

    @staticmethod
    def forward(ctx, n, probes, dx: float, k: float, dz: float, n_medium: float = 1.0):
        from .forward_model import ms_engine

        nt = n.detach().to(torch.complex128).contiguous()
        nx, nz = nt.shape
        T = ms_engine.transmission(nt, k, dz, n_medium)
        H = ms_engine.propagator(nx, dx, k, dz, nt.device)
        u0 = probes.detach().to(nt.device, torch.complex64).reshape(-1, nx).contiguous()
        u = torch.empty((u0.shape[0], nz, nx), dtype=torch.complex64, device=nt.device)
        ms_engine.sweep([ms_engine.MsScanSpec(T, H, u0, out=u)])
        ctx.save_for_backward(T, H, u)
        ctx.k, ctx.dz = k, dz
        return u[:, nz - 1].contiguous()

    @staticmethod
    def backward(ctx, grad_exit):
        from .forward_model import ms_engine

        T, H, u = ctx.saved_tensors

        def run(Tz, x0, out, conj_h):
            ms_engine.sweep([ms_engine.MsScanSpec(Tz, H, x0, out=out, conj_h=conj_h)])

        sd = (0.5 * grad_exit).to(torch.complex64).contiguous()
        return _ms_backward(T, u, sd, run, ctx.k, ctx.dz), None, None, None, None, None


def ms_amplitude_gradient(n, probes, sqrt_i, dx: float, k: float, dz: float, n_medium: float = 1.0, seed=None):
    """(J, grad) for J = sum_p sum (|fft(psi_L)| - sqrt(I))^2 through the multislice,
    grad = dJ/dRe n + i dJ/dIm n, complex128 (nx, nz) on the GPU.  n (nx, nz), probes
    (P, nx), sqrt_i (P, nx); ``seed`` (P, nx) replaces the loss seed when given."""
    from .forward_model import ms_engine

    dev = torch.device("cuda", torch.cuda.current_device())
    n_t = _dev_tensor(n, torch.complex128, dev)
    nx, nz = n_t.shape
    T = ms_engine.transmission(n_t, k, dz, n_medium)
    H = ms_engine.propagator(nx, dx, k, dz, dev)
    u0 = _dev_tensor(probes, torch.complex64, dev).reshape(-1, nx).contiguous()
    a = _dev_tensor(sqrt_i, torch.float32, dev).reshape(-1, nx).contiguous()

    def run(Tz, x0, out, conj_h):
        ms_engine.sweep([ms_engine.MsScanSpec(Tz, H, x0, out=out, conj_h=conj_h)])

    J, grad = _ms_gradient_core(T, u0, run, lambda ex: amplitude_seed(ex, a), k, dz, seed)
    return J, grad


def ms_amplitude_gradient_2d(n, probes, sqrt_i, dx: float, dy: float, k: float, dz: float, n_medium: float = 1.0,
                             seed=None):
    """Two transverse axes: n (nx, ny, nz), probes / sqrt_i (P, nx, ny), per-probe fft2
    far field; grad complex128 (nx, ny, nz)."""
    from .forward_model import ms_engine

    dev = torch.device("cuda", torch.cuda.current_device())
    n_t = _dev_tensor(n, torch.complex128, dev)
    nx, ny, nz = n_t.shape
    N = nx * ny
    T = ms_engine.transmission(n_t.reshape(N, nz), k, dz, n_medium)
    H = ms_engine.propagator_2d(nx, ny, dx, dy, k, dz, dev)
    u0 = _dev_tensor(probes, torch.complex64, dev).reshape(-1, N).contiguous()
    P = u0.shape[0]
    a = _dev_tensor(sqrt_i, torch.float32, dev).reshape(P, N).contiguous()
    lib = _lib.load()
    work = torch.empty((3, P, N), dtype=torch.complex64, device=dev)

    def run(Tz, x0, out, conj_h):
        _lib.check(lib.tp_ms2d_sweep(_lib.ptr(Tz), _lib.ptr(H), _lib.ptr(x0), _lib.ptr(out), None, nx, ny,
                                     Tz.shape[0], P, int(conj_h), _lib.ptr(work), _lib.stream_handle()),
                   "tp_ms2d_sweep")

    def seed_fn(ex):
        sd = torch.empty_like(ex)
        loss = torch.zeros(P, dtype=torch.float32, device=dev)
        _lib.check(lib.tp_amplitude_seed2d(_lib.ptr(ex), _lib.ptr(a), nx, ny, P, 1.0, _lib.ptr(sd), _lib.ptr(loss),
                                           _lib.ptr(work), _lib.stream_handle()), "tp_amplitude_seed2d")
        return sd, loss

    J, grad = _ms_gradient_core(T, u0, run, seed_fn, k, dz, seed)
    return J, grad.reshape(nx, ny, nz)
