"""CPU restatement of the learned preconditioner and of BiCGSTAB (test infrastructure).

Neither exists in the reference; they restate what the HIP library computes
(thick_ptycho_amd/csrc/tp_learn.hip, tp_krylov.hip) so that the GPU kernels are
checked against an independent FP64 computation:

* ``expand`` maps the complex parameters (layer l: [9][OC][IC], OC/IC complex
  channels (16,1), (16,16), (16,16), (1,16)) to real weights [9][2OC][2IC] with
  the block [[a, -b], [b, a]] per complex weight a + ib;
* ``forward`` / ``loss_and_grad`` run the 4-layer 3x3 network (zero padding,
  NHWC planes (P, ny, nx, C)) in FP64; with ``bf16=True`` activations and
  weights are rounded to bfloat16 (round to nearest even) at the points where
  the GPU rounds them, so kernel and oracle differ only by FP32 vs FP64
  accumulation;
* ``bicgstab`` is van der Vorst's BiCGSTAB (SIAM J. Sci. Stat. Comput. 13 (1992)
  631-644) with right preconditioning, the iteration the GPU runs.

Pinning: ``conv3x3`` against ``scipy.signal.correlate2d``; the real network
with structured weights against complex arithmetic on complex planes;
``loss_and_grad`` (without rounding) against central finite differences of the
loss; ``bicgstab`` against ``scipy.sparse.linalg.spsolve`` (solution) and
``scipy.sparse.linalg.bicgstab`` (residual history)."""

from __future__ import annotations

import numpy as np

LAYERS = ((16, 1), (16, 16), (16, 16), (1, 16))  # (OC, IC) complex channels
OFFSETS = np.cumsum([0] + [9 * oc * ic for oc, ic in LAYERS])
NPARAM = int(OFFSETS[-1])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, on):
    """bf16(float32(x)) where the GPU rounds an FP32 result to BF16; exact otherwise."""
    return bf16(np.asarray(x, dtype=np.float32)) if on else np.asarray(x, dtype=np.float64)


def layer_params(params: np.ndarray, l: int) -> np.ndarray:
    oc, ic = LAYERS[l]
    return np.asarray(params[OFFSETS[l]:OFFSETS[l + 1]]).reshape(9, oc, ic)


def expand(params: np.ndarray, l: int) -> np.ndarray:
    """Real weights [9][2 OC][2 IC] of layer l."""
    w = layer_params(params, l)
    a, b = w.real, w.imag
    oc, ic = LAYERS[l]
    W = np.zeros((9, 2 * oc, 2 * ic))
    W[:, 0::2, 0::2] = a
    W[:, 0::2, 1::2] = -b
    W[:, 1::2, 0::2] = b
    W[:, 1::2, 1::2] = a
    return W


def conv3x3(x: np.ndarray, W: np.ndarray) -> np.ndarray:
    """out[p, y, x, o] = sum_tap sum_i x[p, y+dy-1, x+dx-1, i] W[tap, o, i] (zero outside)."""
    P, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((P, H, Wd, W.shape[1]))
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        out += np.einsum("pyxi,oi->pyxo", xp[:, dy:dy + H, dx:dx + Wd, :], W[tap], optimize=True)
    return out


def conv3x3_input_grad(dout: np.ndarray, W: np.ndarray) -> np.ndarray:
    """dL/dx of conv3x3 given dL/dout."""
    P, H, Wd, _ = dout.shape
    dxp = np.zeros((P, H + 2, Wd + 2, W.shape[2]))
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        dxp[:, dy:dy + H, dx:dx + Wd, :] += np.einsum("pyxo,oi->pyxi", dout, W[tap], optimize=True)
    return dxp[:, 1:H + 1, 1:Wd + 1, :]


def conv3x3_weight_grad(x: np.ndarray, dout: np.ndarray) -> np.ndarray:
    """dL/dW[tap, o, i] of conv3x3."""
    P, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    g = np.zeros((9, dout.shape[3], x.shape[3]))
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        g[tap] = np.einsum("pyxo,pyxi->oi", dout, xp[:, dy:dy + H, dx:dx + Wd, :], optimize=True)
    return g


def _planes(v: np.ndarray, nx: int, ny: int) -> np.ndarray:
    v = np.asarray(v).reshape(-1, ny, nx)
    return np.stack([v.real, v.imag], axis=-1)


def forward(params: np.ndarray, v: np.ndarray, nx: int, ny: int, bf16_round: bool = True):
    """M(v) for complex planes v (P, ny*nx); returns (out complex (P, N), activations, scale)."""
    v = np.asarray(v, dtype=np.complex128).reshape(-1, ny * nx)
    P = v.shape[0]
    ms = np.mean(np.abs(v) ** 2, axis=1)
    s = np.where(ms > 0, np.sqrt(ms), 1.0).astype(np.float32)
    inv = (np.float32(1.0) / s).astype(np.float32)
    if bf16_round:
        vf = v.astype(np.complex64)
        x0 = np.stack([(vf.real * inv[:, None]).astype(np.float32), (vf.imag * inv[:, None]).astype(np.float32)], -1)
        x0 = bf16(x0).reshape(P, ny, nx, 2)
    else:
        x0 = _planes(v / s[:, None].astype(np.float64), nx, ny)
    Ws = [_r(expand(params, l), bf16_round) for l in range(4)]
    x1 = _r(conv3x3(x0, Ws[0]), bf16_round)
    x2 = _r(conv3x3(x1, Ws[1]), bf16_round)
    x3 = _r(conv3x3(x2, Ws[2]), bf16_round)
    o = conv3x3(x3, Ws[3])
    if bf16_round:
        o = o.astype(np.float32) * s[:, None, None, None]
    else:
        o = o * s[:, None, None, None].astype(np.float64)
    out = (o[..., 0] + 1j * o[..., 1]).reshape(P, -1)
    return out, (x0, x1, x2, x3, Ws), s


def loss_and_grad(params: np.ndarray, v: np.ndarray, target: np.ndarray, nx: int, ny: int, bf16_round: bool = True):
    """loss = mean |M_n(v/s) - target/s|^2 and its gradient w.r.t. the complex
    parameters, returned as dL/da + i dL/db (complex, [NPARAM])."""
    v = np.asarray(v, dtype=np.complex128).reshape(-1, ny * nx)
    P, N = v.shape
    out, (x0, x1, x2, x3, Ws), s = forward(params, v, nx, ny, bf16_round)
    sc = s.astype(np.float64)[:, None]
    if bf16_round:  # FP32 as on the GPU: d = (out - target) * (1/s)
        inv = (np.float32(1.0) / s).astype(np.float32)[:, None]
        o32 = out.astype(np.complex64)
        t32 = np.asarray(target).astype(np.complex64)
        d = ((o32.real - t32.real) * inv + 1j * ((o32.imag - t32.imag) * inv)).astype(np.complex64)
    else:
        d = (out - np.asarray(target)) / sc
    loss = float(np.sum(np.abs(d.astype(np.complex128)) ** 2) / (P * N))
    norm = 1.0 / (P * N)
    g = (2.0 * norm) * d
    dl = np.stack([g.real, g.imag], -1).reshape(P, ny, nx, 2)
    dl = _r(dl, bf16_round)
    grads = [None] * 4
    grads[3] = conv3x3_weight_grad(x3, dl)
    d3 = _r(conv3x3_input_grad(dl, Ws[3]), bf16_round)
    grads[2] = conv3x3_weight_grad(x2, d3)
    d2 = _r(conv3x3_input_grad(d3, Ws[2]), bf16_round)
    grads[1] = conv3x3_weight_grad(x1, d2)
    d1 = _r(conv3x3_input_grad(d2, Ws[1]), bf16_round)
    grads[0] = conv3x3_weight_grad(x0, d1)
    cg = []
    for l in range(4):
        G = grads[l]
        ga = G[:, 0::2, 0::2] + G[:, 1::2, 1::2]
        gb = G[:, 1::2, 0::2] - G[:, 0::2, 1::2]
        cg.append((ga + 1j * gb).reshape(-1))
    return loss, np.concatenate(cg)


def identity_params(rng=None, noise: float = 0.0) -> np.ndarray:
    """Parameters of the identity map (centre tap, channel 0 through every layer), plus noise."""
    p = np.zeros(NPARAM, dtype=np.complex128)
    for l in range(4):
        w = np.zeros((9,) + LAYERS[l], dtype=np.complex128)
        w[4, 0, 0] = 1.0
        p[OFFSETS[l]:OFFSETS[l + 1]] = w.reshape(-1)
    if noise and rng is not None:
        p = p + noise * (rng.standard_normal(NPARAM) + 1j * rng.standard_normal(NPARAM))
    return p


def bicgstab(A, b: np.ndarray, x0: np.ndarray, M=None, tol: float = 1e-6, maxiter: int = 100):
    """Right-preconditioned BiCGSTAB; returns (x, iterations, residual history ||r_k|| / ||b||)."""
    Mf = (lambda z: z) if M is None else M
    x = np.array(x0, dtype=np.complex128)
    r = b - A @ x
    rh = r.copy()
    p = r.copy()
    rho = np.vdot(rh, r)
    bn = np.linalg.norm(b)
    hist = [np.linalg.norm(r) / bn]
    v = np.zeros_like(r)
    for k in range(1, maxiter + 1):
        ph = Mf(p)
        v = A @ ph
        alpha = rho / np.vdot(rh, v)
        s = r - alpha * v
        if np.linalg.norm(s) <= tol * bn:
            x = x + alpha * ph
            hist.append(np.linalg.norm(s) / bn)
            return x, k, hist
        sh = Mf(s)
        t = A @ sh
        omega = np.vdot(t, s) / np.vdot(t, t)
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        hist.append(np.linalg.norm(r) / bn)
        if np.linalg.norm(r) <= tol * bn:
            return x, k, hist
        rho_new = np.vdot(rh, r)
        beta = (rho_new / rho) * (alpha / omega)
        p = r + beta * (p - omega * v)
        rho = rho_new
    return x, -1, hist
