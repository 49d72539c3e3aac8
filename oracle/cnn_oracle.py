"""FP64 CPU oracle of the subgrid-closure network (test infrastructure only: imported by
tests/ as the checker of the GPU kernels, never by the product).

Network (the closure configuration, no reference counterpart): four 3x3 convolutions
with circular padding, 3 -> 32 -> 32 -> 32 -> 1 channels, GELU (erf form) after the
first three, mean-squared-error loss, Adam (Kingma & Ba 2015, bias-corrected moments).

Parameter vector layout (shared with include/dns_closure.h): for each layer, the
weights W[co][ci][ky][kx] (cross-correlation: out[y][x] += W[ky][kx] in[y+ky-1][x+kx-1],
indices mod n) followed by the biases b[co].

Pinned by tests/test_cnn_oracle.py against PyTorch's CPU conv2d(padding_mode="circular"),
gelu and autograd on the same inputs, and by central finite differences of the loss.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

LAYERS = [(3, 32), (32, 32), (32, 32), (32, 1)]  # (Cin, Cout)


def param_count() -> int:
    return sum(co * ci * 9 + co for ci, co in LAYERS)


def unpack(p):
    """Flat parameter vector -> [(W, b)] per layer (views)."""
    out, o = [], 0
    for ci, co in LAYERS:
        W = p[o:o + co * ci * 9].reshape(co, ci, 3, 3)
        o += co * ci * 9
        b = p[o:o + co]
        o += co
        out.append((W, b))
    return out


def _shift(x, dy, dx):
    """xs[..., y, x] = x[..., (y + dy) mod n, (x + dx) mod n]."""
    return np.roll(x, shift=(-dy, -dx), axis=(-2, -1))


def conv(x, W, b):
    B, ci, n, _ = x.shape
    out = np.zeros((B, W.shape[0], n, n))
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("oc,bcyx->boyx", W[:, :, ky, kx], _shift(x, ky - 1, kx - 1))
    return out + b[None, :, None, None]


def conv_backward(x, W, dout):
    """Gradients of conv(x, W, b) given d(out): (dx, dW, db)."""
    dx = np.zeros_like(x)
    dW = np.zeros_like(W)
    for ky in range(3):
        for kx in range(3):
            xs = _shift(x, ky - 1, kx - 1)
            dW[:, :, ky, kx] = np.einsum("boyx,bcyx->oc", dout, xs)
            dx += _shift(np.einsum("oc,boyx->bcyx", W[:, :, ky, kx], dout), -(ky - 1), -(kx - 1))
    return dx, dW, dout.sum(axis=(0, 2, 3))


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, x):
    """x: [B][3][n][n] -> y [B][n][n] and the cache for backward."""
    layers = unpack(np.asarray(p, dtype=np.float64))
    acts, zs = [np.asarray(x, dtype=np.float64)], []
    for i, (W, b) in enumerate(layers):
        z = conv(acts[-1], W, b)
        zs.append(z)
        acts.append(gelu(z) if i < len(layers) - 1 else z)
    return acts[-1][:, 0], (layers, acts, zs)


def loss_and_grad(p, x, t):
    """MSE over all B n^2 outputs and its gradient w.r.t. the flat parameters."""
    y, (layers, acts, zs) = forward(p, x)
    t = np.asarray(t, dtype=np.float64)
    diff = y - t
    loss = float(np.mean(diff * diff))
    d = (2.0 / diff.size) * diff[:, None]
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        if i < len(layers) - 1:
            d = d * gelu_grad(zs[i])
        dx, dW, db = conv_backward(acts[i], layers[i][0], d)
        grads.append((dW, db))
        d = dx
    flat = []
    for dW, db in reversed(grads):
        flat.append(dW.ravel())
        flat.append(db.ravel())
    return loss, np.concatenate(flat)


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
