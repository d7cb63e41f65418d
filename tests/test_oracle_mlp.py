"""Pins of the network oracle (oracle/mlp.py): finite differences for every
gradient, GELU against mpmath, Adam's first step in closed form, bfloat16
rounding against a brute-force nearest-value search."""

import os
import sys

import mpmath
import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp  # noqa: E402


def _net(widths, seed):
    rng = np.random.default_rng(seed)
    W = [rng.normal(0, 1 / np.sqrt(widths[k]), (widths[k + 1], widths[k]))
         for k in range(len(widths) - 1)]
    b = [rng.normal(0, 0.3, widths[k + 1]) for k in range(len(widths) - 1)]
    return W, b


def test_gradients_match_central_differences():
    widths = [5, 7, 6, 3]
    W, b = _net(widths, 1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4, widths[0]))
    t = rng.normal(size=(4, widths[-1]))
    _, dW, db = mlp.backward(W, b, x, t)
    h = 1e-6
    for k in range(len(W)):
        for idx in np.ndindex(W[k].shape):
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[k][idx] += h
            Wm[k][idx] -= h
            fd = (mlp.loss(Wp, b, x, t) - mlp.loss(Wm, b, x, t)) / (2 * h)
            assert abs(fd - dW[k][idx]) <= 1e-7 + 1e-6 * abs(fd), (k, idx)
        for i in range(b[k].size):
            bp = [v.copy() for v in b]
            bm = [v.copy() for v in b]
            bp[k][i] += h
            bm[k][i] -= h
            fd = (mlp.loss(W, bp, x, t) - mlp.loss(W, bm, x, t)) / (2 * h)
            assert abs(fd - db[k][i]) <= 1e-7 + 1e-6 * abs(fd), (k, i)


def test_gelu_against_mpmath():
    for z in (-6.0, -2.5, -1.0, -0.1, 0.0, 0.3, 1.7, 4.0):
        exact = 0.5 * z * (1 + mpmath.erf(z / mpmath.sqrt(2)))
        assert abs(float(mlp.gelu(np.array([z]))[0]) - float(exact)) < 1e-15
        d = mpmath.diff(lambda s: 0.5 * s * (1 + mpmath.erf(s / mpmath.sqrt(2))), z)
        assert abs(float(mlp.gelu_grad(np.array([z]))[0]) - float(d)) < 1e-14


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    w = rng.normal(size=50)
    g = rng.normal(size=50)
    w1, m1, v1 = mlp.adam(w, g, np.zeros(50), np.zeros(50), 1, lr=1e-3, eps=1e-8)
    np.testing.assert_allclose(w1, w - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-12)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=1e-12)


def test_bf16_rounding_nearest_ties_to_even():
    one = np.float32(1.0)
    cases = {
        1 + 2.0 ** -8: 1.0,                     # tie, even below
        1 + 3 * 2.0 ** -8: 1 + 2.0 ** -6,       # tie, even above
        1 + 2.0 ** -8 + 2.0 ** -20: 1 + 2.0 ** -7,
        -(1 + 2.0 ** -8): -1.0,
    }
    for x, want in cases.items():
        assert mlp.round_bf16(np.array([x], dtype=np.float32))[0] == np.float32(want), x
    assert mlp.round_bf16(np.array([one]))[0] == one
    assert np.isnan(mlp.round_bf16(np.array([np.nan], dtype=np.float32))[0])
    rng = np.random.default_rng(4)
    x = (rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, 2000)).astype(np.float32)
    got = mlp.round_bf16(x)
    u = x.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32).astype(np.float64)
    hi = ((u & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dl, dh = np.abs(xd - lo), np.abs(hi - xd)
    even_lo = ((u >> 16) & 1) == 0
    want = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(even_lo, lo, hi)))
    np.testing.assert_array_equal(got.astype(np.float64), want)
