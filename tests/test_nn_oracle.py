"""Pins the CPU oracle of the learned initial guess (oracle/nn_oracle.py).

* conv / backward-data / backward-weight against torch.nn.functional.conv3d in
  float64 on the CPU (an independent implementation: tap order, padding, flips);
* the whole network, loss and parameter gradient (no rounding) against the same
  network assembled from torch float64 ops with autograd;
* the gradient against central finite differences;
* the BF16 rounding against torch's float32 -> bfloat16 conversion (nearest even).
"""
import os
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as F

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import nn_oracle as nno  # noqa: E402


def to_t(x):  # [d0,d1,d2,C] -> [1,C,d0,d1,d2]
    return torch.from_numpy(np.ascontiguousarray(np.moveaxis(x, 3, 0)))[None]


def from_t(t):
    return np.moveaxis(t[0].detach().numpy(), 0, 3)


def w_to_t(w):  # [co,27,ci] -> [co,ci,3,3,3]
    co, _, ci = w.shape
    return torch.from_numpy(np.ascontiguousarray(w.transpose(0, 2, 1).reshape(co, ci, 3, 3, 3)))


def w_from_t(wt):
    co, ci = wt.shape[:2]
    return wt.detach().numpy().reshape(co, ci, 27).transpose(0, 2, 1)


@pytest.mark.parametrize("shape,ci,co", [((6, 5, 7), 3, 4), ((4, 4, 4), 16, 8)])
def test_conv_matches_torch(shape, ci, co):
    rng = np.random.default_rng(1)
    x = rng.normal(size=shape + (ci,))
    w = rng.normal(size=(co, 27, ci))
    b = rng.normal(size=co)
    g = rng.normal(size=shape + (co,))
    xt = to_t(x).requires_grad_(True)
    wt = w_to_t(w).requires_grad_(True)
    bt = torch.from_numpy(b).requires_grad_(True)
    yt = F.conv3d(xt, wt, bt, padding=1)
    (yt * to_t(g)).sum().backward()
    np.testing.assert_allclose(nno.conv(x, w, b), from_t(yt), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(nno.conv_bwd_data(g, w), from_t(xt.grad), rtol=1e-12, atol=1e-12)
    gw, gb = nno.conv_bwd_weight(g, x)
    np.testing.assert_allclose(gw, w_from_t(wt.grad), rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(gb, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


def torch_net(theta, x):
    """the same encoder-decoder from torch float64 ops"""
    L = nno.unflatten(theta)
    ps = []
    for w, b in L:
        ps.append(w_to_t(w).requires_grad_(True))
        ps.append(torch.from_numpy(b.copy()).requires_grad_(True))

    def cv(i, a, relu=True):
        o = F.conv3d(a, ps[2 * i], ps[2 * i + 1], padding=1)
        return F.relu(o) if relu else o

    up = lambda a: F.interpolate(a, scale_factor=2, mode="nearest")
    x0 = to_t(x)
    a0 = cv(0, x0)
    a1 = cv(1, F.avg_pool3d(a0, 2))
    a2 = cv(2, F.avg_pool3d(a1, 2))
    a3 = cv(3, F.avg_pool3d(a2, 2))
    d2 = cv(4, torch.cat([up(a3), a2], 1))
    d1 = cv(5, torch.cat([up(d2), a1], 1))
    d0 = cv(6, torch.cat([up(d1), a0], 1))
    y = cv(7, d0, relu=False)[0, 0]
    return y, ps


def test_network_loss_and_gradient_match_torch():
    rng = np.random.default_rng(2)
    d = (8, 8, 16)
    theta = nno.init_params(3)
    # biases non-zero so every bias path is exercised
    L = nno.unflatten(theta)
    theta = nno.flatten([(w, rng.normal(0, 0.05, b.shape)) for w, b in L])
    x = rng.normal(size=d + (2,))
    t = rng.normal(size=d)
    m = (rng.random(d) < 0.7).astype(np.float64)
    loss, grad, y = nno.loss_grad(theta, x, t, m, rnd=False)
    yt, ps = torch_net(theta, x)
    mt = torch.from_numpy(m)
    lt = (((yt - torch.from_numpy(t)) * mt) ** 2).sum() / mt.sum()
    lt.backward()
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-10, atol=1e-12)
    assert abs(loss - lt.item()) <= 1e-10 * abs(loss)
    gt = nno.flatten([(w_from_t(ps[2 * i].grad), ps[2 * i + 1].grad.numpy()) for i in range(8)])
    np.testing.assert_allclose(grad, gt, rtol=1e-8, atol=1e-12 * np.abs(gt).max())


def test_gradient_finite_differences():
    rng = np.random.default_rng(4)
    d = (8, 8, 8)
    theta = nno.init_params(5)
    x = rng.normal(size=d + (2,))
    t = rng.normal(size=d)
    m = np.ones(d)
    _, grad, _ = nno.loss_grad(theta, x, t, m, rnd=False)
    # coordinates in every layer (weights and biases)
    offs = np.cumsum([0] + [co * 27 * ci + co for ci, co in zip(nno.CIN, nno.COUT)])
    idx = []
    for lo, hi in zip(offs[:-1], offs[1:]):
        idx += list(rng.integers(lo, hi, 3)) + [hi - 1]
    h = 1e-6
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = h
        lp = nno.loss_grad(theta + e, x, t, m, rnd=False)[0]
        lm = nno.loss_grad(theta - e, x, t, m, rnd=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(grad[i])), (i, fd, grad[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(6)
    v = np.concatenate([rng.normal(size=4000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 4000),
                        np.float32([0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.5, 3.0e38, 1e-40])])
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(nno.bf16(v), ref)
    # halfway cases go to even
    assert nno.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert nno.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
