"""CPU ORACLE — test infrastructure only (imported by tests/ and nothing that ships).

Reference of the learned preconditioner's network in numpy (float64), with
its exact backward pass, so the card's forward and backward kernels can be
checked. The network has no counterpart in the reference solver: it is a
3-level 3D convolutional encoder-decoder (8/16/32 channels, 3x3x3 kernels,
ReLU, skip connections) applied as the preconditioner of flexible CG.

Layout: activations (nz, ny, nx, C) channels last, cells (k*ny+j)*nx+i.
Weights of a layer: W[co, tap, ci], tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero
padding at the domain boundary; bias b[co].

`round16=True` rounds every stored activation and the weights to binary16 as
the card stores them (accumulation stays float64), so a comparison with the
card sees only accumulation-order differences.

Parity status: the backward pass is pinned by central finite differences of
the loss (tests/test_learned.py); the forward pass by a direct triple-loop
convolution on a tiny volume and by the identity network (a single centre
tap reproduces its input).
"""
from __future__ import annotations

import numpy as np

# (name, level, cin, cout, relu, skip added before the ReLU)
LAYERS = [
    ("e1a", 0, 4, 8, True, None),
    ("e1b", 0, 8, 8, True, None),
    ("e2a", 1, 8, 16, True, None),
    ("e2b", 1, 16, 16, True, None),
    ("e3a", 2, 16, 32, True, None),
    ("e3b", 2, 32, 32, True, None),
    ("d2a", 1, 32, 16, True, "e2b"),
    ("d2b", 1, 16, 16, True, None),
    ("d1a", 0, 16, 8, True, "e1b"),
    ("d1b", 0, 8, 8, True, None),
    ("head", 0, 8, 1, False, None),
]
INPUT_CHANNELS = 4


def param_count():
    return sum(co * 27 * ci + co for _, _, ci, co, _, _ in LAYERS)


def unpack(theta):
    """Flat parameter vector -> [(W, b)] in layer order (W[co, tap, ci])."""
    out, off = [], 0
    for _, _, ci, co, _, _ in LAYERS:
        W = theta[off:off + co * 27 * ci].reshape(co, 27, ci)
        off += co * 27 * ci
        b = theta[off:off + co]
        off += co
        out.append((W, b))
    return out


def init(seed=0):
    """He-style initialisation, biases zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, _, ci, co, relu, _ in LAYERS:
        std = np.sqrt((2.0 if relu else 1.0) / (27 * ci))
        parts.append((rng.standard_normal(co * 27 * ci) * std).ravel())
        parts.append(np.zeros(co))
    return np.concatenate(parts)


def r16(a, on):
    return a.astype(np.float16).astype(np.float64) if on else a


def _shift(x, dz, dy, dx):
    """y[k, j, i] = x[k+dz, j+dy, i+dx], zero outside."""
    nz, ny, nx = x.shape[:3]
    y = np.zeros_like(x)
    zs, ze = max(0, -dz), min(nz, nz - dz)
    ys, ye = max(0, -dy), min(ny, ny - dy)
    xs, xe = max(0, -dx), min(nx, nx - dx)
    y[zs:ze, ys:ye, xs:xe] = x[zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv(x, W, b):
    out = np.zeros(x.shape[:3] + (W.shape[0],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(x, dz, dy, dx) @ W[:, t, :].T
    return out + b


def conv_backward(x, W, g):
    """Gradients of sum(g * conv(x, W, b)) w.r.t. x, W, b."""
    gx = np.zeros_like(x)
    gW = np.zeros_like(W)
    for t, (dz, dy, dx) in enumerate(TAPS):
        xs = _shift(x, dz, dy, dx)
        gW[:, t, :] = np.einsum("zyxo,zyxi->oi", g, xs)
        gx += _shift(g @ W[:, t, :], -dz, -dy, -dx)
    return gx, gW, g.sum(axis=(0, 1, 2))


def pool(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_backward(g):
    nz, ny, nx, c = g.shape
    return g.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def forward(theta, x, round16=False):
    """Returns (y, cache); x (nz, ny, nx, 4), y (nz, ny, nx)."""
    params = unpack(r16(theta, round16))
    acts = {}
    cache = []
    h = r16(x, round16)
    for (name, level, ci, co, relu, skip), (W, b) in zip(LAYERS, params):
        if name == "e2a" or name == "e3a":
            h = r16(pool(h), round16)
        if name == "d2a" or name == "d1a":
            h = up(h)
        pre = conv(h, W, b)
        if skip is not None:
            pre = pre + acts[skip]
        out = np.maximum(pre, 0.0) if relu else pre
        out = r16(out, round16) if relu else out
        cache.append((h, pre))
        acts[name] = out
        h = out
    return h[..., 0], cache


def backward(theta, cache, gy, round16=False):
    """Gradient of sum(gy * y) w.r.t. theta (flat)."""
    params = unpack(r16(theta, round16))
    grads = [None] * len(LAYERS)
    g = gy[..., None]
    skip_grads = {}
    for li in range(len(LAYERS) - 1, -1, -1):
        name, level, ci, co, relu, skip = LAYERS[li]
        h, pre = cache[li]
        if name in skip_grads:
            g = g + skip_grads.pop(name)
        if relu:
            g = g * (pre > 0)
        if skip is not None:
            skip_grads[skip] = g
        gx, gW, gb = conv_backward(h, params[li][0], g)
        grads[li] = (gW, gb)
        if name == "d2a" or name == "d1a":
            gx = up_backward(gx)
        if name == "e2a" or name == "e3a":
            gx = pool_backward(gx)
        g = gx
    return np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def features(r, wet, ghost, solid, sigma):
    """Network input of a residual r: (r * sigma, wet, ghost row, solid)."""
    return np.stack([r * sigma, wet, ghost, solid], axis=-1).astype(np.float64)


def loss_and_grad(theta, x, target, wet, round16=False):
    """Mean squared error over wet cells and its gradient."""
    y, cache = forward(theta, x, round16)
    n = max(int(wet.sum()), 1)
    d = (y - target) * wet
    loss = float((d * d).sum() / n)
    gy = 2.0 * d / n
    return loss, backward(theta, cache, gy, round16), y
