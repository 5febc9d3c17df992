"""Pins of the surrogate MLP's CPU oracle (runs without a GPU)."""
import math

import numpy as np

import mlp_oracle as MO

DIMS = (7, 12, 2, 3)


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even is 1.0
    assert MO.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3 * 2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: even mantissa is 1 + 2^-6
    assert MO.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    # just above the tie rounds up
    assert MO.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)) == 1.0 + 2.0 ** -7
    # exactly representable values are unchanged; sign symmetric
    for x in (0.0, 1.0, -2.5, 3.0 * 2.0 ** 100, 2.0 ** -130):
        assert MO.bf16(np.float32(x)) == np.float32(x)
        assert MO.bf16(np.float32(-x)) == -np.float32(x)
    assert np.isnan(MO.bf16(np.float32("nan")))
    # 8 significant bits: relative rounding error at most 2^-8
    x = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    assert np.max(np.abs(MO.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_gelu_closed_forms():
    assert MO.gelu(np.array(0.0)) == 0.0
    # x * Phi(x): Phi(1) = 0.841344746...
    assert abs(MO.gelu(np.array(1.0)) - 0.8413447460685429) < 1e-15
    x = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (MO.gelu(x + h) - MO.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - MO.gelu_grad(x))) < 1e-8


def test_gradient_matches_finite_differences():
    g = np.random.default_rng(1)
    p = MO.init_params(*DIMS, seed=2).astype(np.float64)
    p += 0.05 * g.standard_normal(p.size)  # nonzero biases too
    x = g.standard_normal((9, DIMS[0]))
    y = g.standard_normal((9, DIMS[3]))
    _, grad = MO.loss_and_grad(p, x, y, DIMS, bf16_ops=False)
    fd = np.zeros_like(p)
    h = 1e-6
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = MO.loss_and_grad(p + e, x, y, DIMS, bf16_ops=False)
        lm, _ = MO.loss_and_grad(p - e, x, y, DIMS, bf16_ops=False)
        fd[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - grad)) < 1e-7 * max(1.0, np.max(np.abs(grad)))


def test_forward_layer_shapes_and_linear_limit():
    # with one hidden unit pushed far into GELU's linear regime the network is affine
    p = np.zeros(sum(k * n + n for k, n in MO.shapes(1, 1, 1, 1)))
    # W0 = 1, b0 = 50 (gelu(x) = x for large x), W1 = 2, b1 = -100
    p[:] = [1.0, 50.0, 2.0, -100.0]
    x = np.array([[1.5], [-3.0]])
    out, _, _ = MO.forward(p, x, (1, 1, 1, 1), bf16_ops=False)
    assert np.allclose(out[:, 0], 2 * x[:, 0], atol=1e-12)
    assert MO.shapes(52, 256, 3, 2) == [(52, 256), (256, 256), (256, 256), (256, 2)]
    n = sum(k * n + n for k, n in MO.shapes(52, 256, 3, 2))
    assert n == 52 * 256 + 256 + 2 * (256 * 256 + 256) + 256 * 2 + 2


def test_bf16_mode_close_to_exact():
    g = np.random.default_rng(3)
    p = MO.init_params(*DIMS, seed=4)
    x = g.standard_normal((16, DIMS[0]))
    a, _, _ = MO.forward(p, x, DIMS, bf16_ops=True)
    b, _, _ = MO.forward(p, x, DIMS, bf16_ops=False)
    assert np.max(np.abs(a - b)) < 0.05 * np.max(np.abs(b))
    assert not np.array_equal(a, b)  # the rounding is really applied


def test_adam_oracle_closed_forms():
    """Adam with bias correction (Kingma & Ba 2015): the first step moves every
    weight by lr * g / (|g| + eps); a constant gradient keeps that step; a gradient
    that flips sign on step 2 gives m_hat = -g / 19 and v_hat = g^2 exactly
    (m2 = 0.09 g - 0.1 g over 1 - 0.81, v2 = 0.001999 g^2 over 1 - 0.998001)."""
    g = np.array([0.3, -2.0, 1e-3, 0.0, 5e-9])
    w0 = np.array([0.1, -0.2, 0.3, 0.4, 0.5])
    lr = 1e-3
    w1, m1, v1 = MO.adam_step(w0, g, 0.0, 0.0, 1, lr)
    np.testing.assert_allclose(w0 - w1, lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-18)
    w2, m2, v2 = MO.adam_step(w1, g, m1, v1, 2, lr)
    np.testing.assert_allclose(w1 - w2, lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-18)
    w2f, _, _ = MO.adam_step(w1, -g, m1, v1, 2, lr)
    np.testing.assert_allclose(w1 - w2f, lr * (-g / 19.0) / (np.abs(g) + 1e-8), rtol=1e-11, atol=1e-18)
    # without the bias correction the first step would be 0.1 / sqrt(0.001) = 3.16 times larger
    assert abs((w0 - w1)[0]) < 1.01 * lr
