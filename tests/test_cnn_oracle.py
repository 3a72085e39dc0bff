"""Pins of the closure-network oracle (oracle/cnn_oracle.py): its analytic gradients against
central finite differences in float64, the adjoint identity of its convolution, and the bf16
rounding it emulates."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cnn_oracle as co  # noqa: E402


def small_net(seed=0, scale=1.0):
    lay, n = co.layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for l, (ow, cout, cin, ob) in enumerate(lay):
        real = co.NIN if l == 0 else cin
        w = rng.normal(size=(9, cout, cin)) * np.sqrt(2.0 / (9 * real)) * scale
        if l == 0:
            w[..., co.NIN:] = 0.0
        p[ow:ow + w.size] = w.ravel()
        p[ob:ob + cout] = rng.normal(size=cout) * 0.1
    return p


def test_bf16_rounding():
    assert co.bf16(1.0) == 1.0
    assert co.bf16(1.0 + 2 ** -8) == 1.0           # tie -> even
    assert co.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6  # tie -> even (up)
    assert co.bf16(1.0 + 2 ** -7 + 2 ** -12) == 1.0 + 2 ** -7
    x = np.random.default_rng(1).normal(size=1000)
    assert np.max(np.abs(co.bf16(x) - x) / np.abs(x)) <= 2 ** -8


def test_conv_adjoint():
    rng = np.random.default_rng(2)
    x, g = rng.normal(size=(2, 5, 6, 3)), rng.normal(size=(2, 5, 6, 4))
    w = rng.normal(size=(9, 4, 3))
    lhs = np.sum(co.conv(x, w) * g)
    rhs = np.sum(x * co.conv_t(g, w))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert abs(lhs - np.sum(co.conv_wgrad(x, g) * w)) <= 1e-12 * abs(lhs)


def test_gradients_match_finite_differences():
    p = small_net(3)
    rng = np.random.default_rng(4)
    inp = rng.normal(size=(2, 5, 6, co.NIN))
    gout = rng.normal(size=(2, 5, 6))
    out, cache = co.forward(p, inp, emulate=False)
    g = co.backward(p, cache, gout, emulate=False)
    lay, n = co.layout()
    # a few weights and the bias of every layer (the padded input channels excluded)
    idx = []
    for l, (ow, cout, cin, ob) in enumerate(lay):
        for _ in range(4):
            t, o, c = rng.integers(9), rng.integers(cout), rng.integers(co.NIN if l == 0 else cin)
            idx.append(ow + (t * cout + o) * cin + c)
        idx.append(ob)
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (np.sum(co.forward(pp, inp, emulate=False)[0] * gout) - np.sum(co.forward(pm, inp, emulate=False)[0] * gout)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_mse_and_adam():
    out = np.arange(24.0).reshape(1, 4, 6)
    tgt = out + 1.0
    loss, d = co.mse(out, tgt)
    assert loss == 1.0 and np.all(d[:, [0, -1]] == 0) and np.allclose(d[:, 1:-1], -2.0 / 12)
    p, m, v = co.adam(np.ones(3), np.zeros(3), np.zeros(3), np.array([1.0, -2.0, 0.0]), 1, 0.1)
    np.testing.assert_allclose(p, [0.9, 1.1, 1.0], rtol=0, atol=1e-7)  # the first step moves by lr * sign(g)
