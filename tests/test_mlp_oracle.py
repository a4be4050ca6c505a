"""Pins the CPU oracle of the learned warm start's network (oracle/mlp_oracle.py)
against what the mathematics fixes: central finite differences of the float64
network, GELU's closed-form values and its erf form, bf16 rounding against an
independent frexp/rint rounding, and Adam's first step (the sign of the
gradient times the learning rate)."""
import numpy as np
import pytest

import mlp_oracle as mo


def rand_params(rng, in_dim, out_dim, scale=0.5):
    shapes = [(in_dim, mo.HIDDEN), (mo.HIDDEN,), (mo.HIDDEN, mo.HIDDEN), (mo.HIDDEN,), (mo.HIDDEN, out_dim),
              (out_dim,)]
    parts = []
    for shp in shapes:
        fan = shp[0] if len(shp) == 2 else mo.HIDDEN
        parts.append(rng.normal(0, scale / np.sqrt(fan), size=shp).ravel())
    return np.concatenate(parts)


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    in_dim, out_dim = 5, 3
    flat = rand_params(rng, in_dim, out_dim, scale=1.5)
    x = rng.normal(size=(7, in_dim))
    t = rng.normal(size=(7, out_dim))
    loss, g = mo.loss_gradients(flat, x, t, in_dim, out_dim, exact=True)
    n = flat.size
    pick = np.concatenate([rng.choice(n, 300, replace=False),
                           np.arange(in_dim * 128, in_dim * 128 + 128),          # b1
                           np.arange(n - out_dim, n)])                          # b3
    h = 1e-6
    for i in pick:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        lp, _ = mo.loss_gradients(fp, x, t, in_dim, out_dim, exact=True)
        lm, _ = mo.loss_gradients(fm, x, t, in_dim, out_dim, exact=True)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)) + 1e-8, (i, num, g[i])


def test_gelu_closed_forms():
    from math import erf, sqrt
    assert mo.gelu(0.0) == 0.0
    assert abs(mo.gelu_grad(0.0) - 0.5) < 1e-15
    xs = np.linspace(-6, 6, 241)
    exact = np.array([0.5 * v * (1 + erf(v / sqrt(2))) for v in xs])
    assert np.max(np.abs(mo.gelu(xs) - exact)) < 1e-3          # tanh form vs erf form
    assert np.allclose(mo.gelu(np.array([20.0, -20.0])), [20.0, 0.0], atol=1e-12)
    h = 1e-6
    num = (mo.gelu(xs + h) - mo.gelu(xs - h)) / (2 * h)
    assert np.max(np.abs(num - mo.gelu_grad(xs))) < 1e-8


def test_bf16_rounding():
    assert mo.bf16(np.float32(1.0)) == 1.0
    assert mo.bf16(np.float32(1 + 2 ** -8)) == 1.0                        # tie -> even (down)
    assert mo.bf16(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)  # tie -> even (up)
    assert mo.bf16(np.float32(-(1 + 2 ** -7))) == np.float32(-(1 + 2 ** -7))
    rng = np.random.default_rng(5)
    x = (rng.normal(size=20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    m, e = np.frexp(x.astype(np.float64))
    ref = np.ldexp(np.rint(m * 256.0) / 256.0, e)   # 8 significant bits, rint = ties to even
    assert np.array_equal(mo.bf16(x).astype(np.float64), ref)


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    w = np.zeros(4)
    w1, m, v = mo.adam(w, g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    assert np.allclose(w1[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)
    assert w1[3] == 0.0


def test_rounded_mode_stays_close_to_exact():
    rng = np.random.default_rng(8)
    in_dim, out_dim = 24, 3
    flat = rand_params(rng, in_dim, out_dim)
    x = rng.normal(size=(64, in_dim))
    y0 = mo.forward(flat, x, in_dim, out_dim, exact=True)
    y1 = mo.forward(flat, x, in_dim, out_dim, exact=False)
    rel = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
    assert 1e-5 < rel < 3e-2   # bf16 rounding is visible but small
