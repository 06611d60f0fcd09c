"""CPU oracle of the learned preconditioner's network (test infrastructure).

Only the tests import this module; it checks the GPU's BF16 U-Net.  It has no
counterpart in the reference: the network, its training and its use in
flexible CG are this project's own (the learned form of the solver).  It
restates, in numpy, the forward pass, the mean-squared-error loss and the
backward pass of the network built by tealeaf-kokkos_amd/csrc/tl_unet.hip:

* 15 convolutions, 3x3 taps, zero padding, NHWC activations; layer shapes
  (cin -> cout): 8->16, 16->16 | pool | 16->32, 32->32 | pool | 32->64,
  64->64 | pool | 64->128, 128->128 | up+cat(64) | 192->64, 64->64 |
  up+cat(32) | 96->32, 32->32 | up+cat(16) | 48->16, 16->16 | 16->16 (no
  ReLU, channel 0 is the output);
* 2x2 average pooling, nearest 2x upsampling concatenated before the skip;
* loss = mean over pixels of (out[..., 0] - target)^2, or the A-norm of
  d = out[..., 0] - target over the batch, sum(d.A d) / sum(t.A t), A the
  5-point operator of per-pixel face coefficients (the training loss).

With ``bf16=True`` values are rounded to bfloat16 (round to nearest even)
at the points where the GPU stores them (weights, every activation and every
activation gradient), sums are formed in float64; with ``bf16=False``
everything is float64 and the gradients are exact derivatives of the loss,
which the CPU tests check against finite differences.
"""
from __future__ import annotations

import numpy as np

SHAPES = [(8, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True),
          (64, 128, True), (128, 128, True), (192, 64, True), (64, 64, True), (96, 32, True), (32, 32, True),
          (48, 16, True), (16, 16, True), (16, 16, False)]


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(params: np.ndarray):
    """Flat parameters (layer order, W[cout][9][cin] then b[cout]) -> list of (W, b)."""
    out, k = [], 0
    for cin, cout, _ in SHAPES:
        nw = cout * 9 * cin
        w = params[k:k + nw].reshape(cout, 9, cin).astype(np.float64)
        k += nw
        b = params[k:k + cout].astype(np.float64)
        k += cout
        out.append((w, b))
    assert k == len(params)
    return out


def flatten(grads) -> np.ndarray:
    return np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])


def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """x[n, y + dy, x + dx, c] with zeros outside."""
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    ys, yd = slice(max(dy, 0), h + min(dy, 0)), slice(max(-dy, 0), h - max(dy, 0))
    xs, xd = slice(max(dx, 0), w + min(dx, 0)), slice(max(-dx, 0), w - max(dx, 0))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(x, w, b=None):
    """y[p][co] = sum_t sum_ci w[co][t][ci] x[p + off(t)][ci] (+ b)."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dy, dx) in enumerate(TAPS):
        y += _shift(x, dy, dx) @ w[:, t, :].T
    return y if b is None else y + b


def conv_bwd_data(g, w):
    """dx[p][ci] = sum_t sum_co w[co][t][ci] g[p - off(t)][co]."""
    dx = np.zeros(g.shape[:3] + (w.shape[2],))
    for t, (dy, ddx) in enumerate(TAPS):
        dx += _shift(g, -dy, -ddx) @ w[:, t, :]
    return dx


def conv_bwd_weight(g, x):
    """dW[co][t][ci] = sum_p g[p][co] x[p + off(t)][ci]."""
    cout, cin = g.shape[3], x.shape[3]
    dw = np.zeros((cout, 9, cin))
    gf = g.reshape(-1, cout)
    for t, (dy, dx) in enumerate(TAPS):
        dw[:, t, :] = gf.T @ _shift(x, dy, dx).reshape(-1, cin)
    return dw


def pool(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


def pool_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(g):
    return g[:, 0::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 0::2] + g[:, 1::2, 1::2]


def pack_input(r, f):
    n, h, w = r.shape
    x = np.zeros((n, h, w, 8))
    x[..., 0] = r
    x[..., 1] = f
    return x


def forward(params, r, f, bf: bool = True):
    """Activations of every layer (dict) for input planes r, f of shape (n, h, w)."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(params)
    wq = [(q(w), b) for w, b in layers]
    act = {"x0": q(pack_input(r, f))}

    def c(i, x):
        w, b = wq[i]
        y = conv(x, w, b)
        if SHAPES[i][2]:
            y = np.maximum(y, 0.0)
        return q(y)

    a = act
    a["e0a"] = c(0, a["x0"]); a["e0b"] = c(1, a["e0a"]); a["p0"] = q(pool(a["e0b"]))
    a["e1a"] = c(2, a["p0"]); a["e1b"] = c(3, a["e1a"]); a["p1"] = q(pool(a["e1b"]))
    a["e2a"] = c(4, a["p1"]); a["e2b"] = c(5, a["e2a"]); a["p2"] = q(pool(a["e2b"]))
    a["e3a"] = c(6, a["p2"]); a["e3b"] = c(7, a["e3a"])
    a["cat2"] = np.concatenate([up(a["e3b"]), a["e2b"]], axis=3)
    a["d2a"] = c(8, a["cat2"]); a["d2b"] = c(9, a["d2a"])
    a["cat1"] = np.concatenate([up(a["d2b"]), a["e1b"]], axis=3)
    a["d1a"] = c(10, a["cat1"]); a["d1b"] = c(11, a["d1a"])
    a["cat0"] = np.concatenate([up(a["d1b"]), a["e0b"]], axis=3)
    a["d0a"] = c(12, a["cat0"]); a["d0b"] = c(13, a["d0a"])
    a["out"] = c(14, a["d0b"])
    return a, wq


def a_apply(v, kx, ky):
    """The 5-point operator on samples v[n][h][w] with per-pixel west (kx)
    and south (ky) face coefficients; faces on the sample's edge carry no flux."""
    out = v.copy()
    out[:, :, :-1] += kx[:, :, 1:] * (v[:, :, :-1] - v[:, :, 1:])
    out[:, :, 1:] += kx[:, :, 1:] * (v[:, :, 1:] - v[:, :, :-1])
    out[:, :-1, :] += ky[:, 1:, :] * (v[:, :-1, :] - v[:, 1:, :])
    out[:, 1:, :] += ky[:, 1:, :] * (v[:, 1:, :] - v[:, :-1, :])
    return out


def a_diag(kx, ky):
    """Diagonal of a_apply's operator: 1 + the coefficients of the faces
    inside the sample."""
    d = np.ones(kx.shape)
    d[:, :, :-1] += kx[:, :, 1:]
    d[:, :, 1:] += kx[:, :, 1:]
    d[:, :-1, :] += ky[:, 1:, :]
    d[:, 1:, :] += ky[:, 1:, :]
    return d


def smooth(v, kx, ky, sweeps, omega):
    """M v, M = (I - omega D^-1 A)^sweeps: the error propagation of
    `sweeps` weighted Jacobi sweeps on A z = r."""
    D = a_diag(kx, ky)
    for _ in range(sweeps):
        v = v - omega * a_apply(v, kx, ky) / D
    return v


def smooth_t(g, kx, ky, sweeps, omega):
    """M^T g = (I - omega A D^-1)^sweeps g."""
    D = a_diag(kx, ky)
    for _ in range(sweeps):
        g = g - omega * a_apply(g / D, kx, ky)
    return g


def loss_grad(params, r, f, target, bf: bool = True, kx=None, ky=None, sweeps: int = 0, omega: float = 0.8):
    """(loss, flat parameter gradients) of out[..., 0] against target: the
    mean squared error, or with kx, ky the A-norm (d.A d)/(t.A t) of
    d = out - target over the batch; with sweeps > 0 the A-norm after
    `sweeps` weighted Jacobi sweeps, (Md.A Md)/max(Mt.A Mt, batch mean) per
    sample averaged over the samples (M their error propagation)."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    a, wq = forward(params, r, f, bf)
    pixels = target.size
    d = a["out"][..., 0] - target
    g_out = np.zeros_like(a["out"])
    if kx is None:
        loss = float(np.sum(d * d) / pixels)
        g_out[..., 0] = q(2.0 * d / pixels)
    else:
        if sweeps == 0:
            ad = a_apply(d, kx, ky)
            den = float(np.sum(target * a_apply(target, kx, ky)))
            loss = float(np.sum(d * ad) / den)
            g_out[..., 0] = q(2.0 * ad / den)
        else:
            # per sample sum(Md.A Md) / max(sum(Mt.A Mt), its batch mean),
            # averaged over the samples
            n = d.shape[0]
            md = smooth(d, kx, ky, sweeps, omega)
            mt = smooth(target, kx, ky, sweeps, omega)
            ad = a_apply(md, kx, ky)
            den = np.sum(mt * a_apply(mt, kx, ky), axis=(1, 2))
            den = np.maximum(den, den.mean())
            loss = float(np.mean(np.sum(md * ad, axis=(1, 2)) / den))
            g_out[..., 0] = q(smooth_t(2.0 * ad / (n * den[:, None, None]), kx, ky, sweeps, omega))
    grads = [None] * len(SHAPES)

    def back(i, x, y, g, need_dx=True):
        if SHAPES[i][2]:
            g = g * (y > 0)
        grads[i] = (conv_bwd_weight(g, x), g.reshape(-1, g.shape[3]).sum(axis=0))
        return q(conv_bwd_data(g, wq[i][0])) if need_dx else None

    g = back(14, a["d0b"], a["out"], g_out)
    g = back(13, a["d0a"], a["d0b"], g)
    g = back(12, a["cat0"], a["d0a"], g)
    g_d1b, g_e0b = q(up_bwd(g[..., :32])), g[..., 32:]
    g = back(11, a["d1a"], a["d1b"], g_d1b)
    g = back(10, a["cat1"], a["d1a"], g)
    g_d2b, g_e1b = q(up_bwd(g[..., :64])), g[..., 64:]
    g = back(9, a["d2a"], a["d2b"], g_d2b)
    g = back(8, a["cat2"], a["d2a"], g)
    g_e3b, g_e2b = q(up_bwd(g[..., :128])), g[..., 128:]
    g = back(7, a["e3a"], a["e3b"], g_e3b)
    g = back(6, a["p2"], a["e3a"], g)
    g = back(5, a["e2a"], a["e2b"], q(g_e2b + pool_bwd(g)))
    g = back(4, a["p1"], a["e2a"], g)
    g = back(3, a["e1a"], a["e1b"], q(g_e1b + pool_bwd(g)))
    g = back(2, a["p0"], a["e1a"], g)
    g = back(1, a["e0a"], a["e0b"], q(g_e0b + pool_bwd(g)))
    back(0, a["x0"], a["e0a"], g, need_dx=False)
    return loss, flatten(grads)


def init_params(seed: int) -> np.ndarray:
    """He-normal weights (numpy generator), zero biases; the first layer's
    padding channels 2..7 zero."""
    rng = np.random.default_rng(seed)
    out = []
    for i, (cin, cout, relu) in enumerate(SHAPES):
        fan = 9 * (2 if i == 0 else cin)
        w = rng.normal(0.0, np.sqrt(2.0 / fan) * (1.0 if relu else 0.5), size=(cout, 9, cin))
        if i == 0:
            w[:, :, 2:] = 0.0
        out += [w.ravel(), np.zeros(cout)]
    return np.concatenate(out).astype(np.float32)
