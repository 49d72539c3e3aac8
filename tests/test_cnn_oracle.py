"""Pins the FP64 closure-network oracle (oracle/cnn_oracle.py): forward, loss gradient
and Adam step against PyTorch's CPU implementation (conv2d with circular padding, exact
GELU, autograd, torch.optim.Adam) and against central finite differences."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cnn_oracle as co  # noqa: E402

torch = pytest.importorskip("torch")


def _torch_net(p):
    import torch.nn.functional as F

    layers = co.unpack(p)
    ts = []
    for W, b in layers:
        ts.append(torch.tensor(W, dtype=torch.float64, requires_grad=True))
        ts.append(torch.tensor(b, dtype=torch.float64, requires_grad=True))

    def fwd(x):
        h = x
        for i in range(4):
            h = F.conv2d(F.pad(h, (1, 1, 1, 1), mode="circular"), ts[2 * i], ts[2 * i + 1])
            if i < 3:
                h = F.gelu(h)
        return h[:, 0]

    return ts, fwd


def _case(seed, B=2, n=8):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(co.param_count()) * 0.2
    x = rng.standard_normal((B, 3, n, n))
    t = rng.standard_normal((B, n, n))
    return p, x, t


def test_param_count():
    assert co.param_count() == 3 * 32 * 9 + 32 + 2 * (32 * 32 * 9 + 32) + 32 * 9 + 1 == 19681


def test_forward_and_gradient_match_torch():
    p, x, t = _case(0)
    ts, fwd = _torch_net(p)
    yt = fwd(torch.tensor(x))
    y, _ = co.forward(p, x)
    assert np.max(np.abs(y - yt.detach().numpy())) < 1e-12 * max(1.0, np.max(np.abs(y)))
    loss_t = torch.mean((yt - torch.tensor(t)) ** 2)
    loss_t.backward()
    g_t = np.concatenate([a.grad.numpy().ravel() for a in ts])
    loss, g = co.loss_and_grad(p, x, t)
    assert abs(loss - loss_t.item()) < 1e-12 * loss
    assert np.max(np.abs(g - g_t)) < 1e-11 * np.max(np.abs(g_t))


def test_gradient_matches_finite_differences():
    p, x, t = _case(1, B=1, n=6)
    _, g = co.loss_and_grad(p, x, t)
    rng = np.random.default_rng(2)
    for i in rng.choice(p.size, 25, replace=False):
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (co.loss_and_grad(pp, x, t)[0] - co.loss_and_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_adam_matches_torch():
    p, x, t = _case(3)
    ts, fwd = _torch_net(p)
    opt = torch.optim.Adam(ts, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    q = p.copy()
    for step in range(1, 4):
        opt.zero_grad()
        loss_t = torch.mean((fwd(torch.tensor(x)) - torch.tensor(t)) ** 2)
        loss_t.backward()
        opt.step()
        _, g = co.loss_and_grad(q, x, t)
        q, m, v = co.adam(q, g, m, v, step)
    q_t = np.concatenate([a.detach().numpy().ravel() for a in ts])
    assert np.max(np.abs(q - q_t)) < 1e-12
