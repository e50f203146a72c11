"""Pins for the radiance-cache MLP oracle (oracle/mlp.py), no GPU:
* the encoding's closed form ((sin, cos)(pi 2^b p) at fixed points);
* BF16 rounding (ties to even, exact values kept);
* the backward pass against central finite differences of the float64
  network (a wrong transpose, mask or loss scale fails it);
* Adam against the published update rule on a scalar.
"""
import math

import numpy as np
import pytest

from oracle import mlp


def test_encoding_closed_form():
    f = np.zeros((1, 12))
    f[0, 0], f[0, 1] = 0.25, 0.5
    e = mlp.encode(f)
    assert e.shape == (1, 160) and np.all(e[0, 144:] == 0)
    for b in range(6):
        assert e[0, 2 * b] == pytest.approx(math.sin(math.pi * 0.25 * 2 ** b), abs=1e-12)
        assert e[0, 2 * b + 1] == pytest.approx(math.cos(math.pi * 0.25 * 2 ** b), abs=1e-12)
        assert e[0, 12 + 2 * b] == pytest.approx(math.sin(math.pi * 0.5 * 2 ** b), abs=1e-12)
    assert np.all(e[0, 24:144:2] == 0.0) and np.all(e[0, 25:144:2] == 1.0)  # zero inputs


def test_bf16_rounding():
    x = np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.140625, 1e-40, np.inf])
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[3] == -2.5 and r[4] == 3.140625 and np.isinf(r[6])
    assert r[1] == 1.0  # halfway, ties to even
    assert r[2] == np.float32(1.0 + 2 ** -7)  # 1 + 3/512 rounds up to 1 + 1/128
    assert r[5] > 0  # denormals are kept


def _weights(seed, scale=1.0):
    rs = np.random.default_rng(seed)
    w = np.zeros(mlp.PARAMS)
    for l in range(mlp.LAYERS):
        W = mlp.layer(w, l)
        b = math.sqrt(6.0 / mlp.REAL_IN[l]) * scale
        W[:mlp.REAL_OUT[l], :mlp.REAL_IN[l]] = rs.uniform(-b, b, (mlp.REAL_OUT[l], mlp.REAL_IN[l]))
    return w


def test_backward_matches_finite_differences():
    rs = np.random.default_rng(1)
    w = _weights(2)
    feat = rs.uniform(0, 1, (8, 12))
    tgt = rs.uniform(0, 2, (8, 3))
    loss, g = mlp.Net(w, "f64").loss_and_grads(feat, tgt)
    # the analytic gradient treats the denominator as a constant: compare with
    # finite differences of the same quantity (denominator frozen at w)
    y0, _ = mlp.Net(w, "f64").forward(feat)
    den = y0 * y0 + 0.01

    def frozen(wv):
        y, _ = mlp.Net(wv, "f64").forward(feat)
        return float(((y - tgt) ** 2 / den).sum() / (3 * feat.shape[0]))

    real = [i for l in range(mlp.LAYERS)
            for i in range(mlp.OFFSET[l], mlp.OFFSET[l + 1])
            if (i - mlp.OFFSET[l]) // mlp.IN_DIM[l] < mlp.REAL_OUT[l]
            and (i - mlp.OFFSET[l]) % mlp.IN_DIM[l] < mlp.REAL_IN[l]]
    idx = rs.choice(real, 60, replace=False)
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (frozen(wp) - frozen(wm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    # padding entries carry no gradient
    assert np.all(mlp.mask_padding(g) == g)
    assert loss == pytest.approx(frozen(w))


def test_adam_scalar_rule():
    w = np.zeros(mlp.PARAMS, np.float32)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    g = np.zeros_like(w)
    g[0] = 0.5
    w2, m2, v2 = mlp.adam(w, m, v, g, 1, 1e-3)
    # first step: m_hat = g, v_hat = g^2  ->  w -= lr * g / (|g| + eps)
    assert w2[0] == pytest.approx(-1e-3 * 0.5 / (0.5 + 1e-8), rel=1e-6)
    assert m2[0] == pytest.approx(0.05) and v2[0] == pytest.approx(0.001 * 0.25)
    g[:] = 1.0  # padding stays untouched
    w3, _, _ = mlp.adam(w, m, v, g, 1, 1e-3)
    assert w3[mlp.OFFSET[0] + 150] == 0.0 and w3[mlp.OFFSET[4] + 5 * 64] == 0.0
