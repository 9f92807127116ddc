"""Pins the FP64 MLP oracle: analytic gradients against central finite differences, GELU
against its closed form and symmetry, Adam against a hand-evaluated first step."""
import math

import numpy as np

from oracle import mlp


def test_gelu_closed_form():
    z = np.linspace(-4, 4, 81)
    ref = np.array([0.5 * t * (1 + math.erf(t / math.sqrt(2))) for t in z])
    np.testing.assert_allclose(mlp.gelu(z), ref, rtol=1e-14, atol=1e-15)
    # gelu(z) - gelu(-z) = z
    np.testing.assert_allclose(mlp.gelu(z) - mlp.gelu(-z), z, atol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(mlp.gelu_grad(z), (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h), atol=1e-9)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init_params(7, 9, 5, seed=1)
    for k in p:
        p[k] = p[k] + 0.1 * rng.normal(size=p[k].shape)
    X, T = rng.normal(size=(11, 7)), rng.normal(size=(11, 5))
    _, g = mlp.loss_and_grads(p, X, T)
    h = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = mlp.loss_and_grads(p, X, T)
            flat[idx] = old - h
            lm, _ = mlp.loss_and_grads(p, X, T)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k].reshape(-1)[idx]) <= 1e-7 * max(1.0, abs(fd)), (k, idx)


def test_adam_first_step():
    p = {"w": np.array([1.0, -2.0])}
    g = {"w": np.array([0.5, -0.25])}
    m = {"w": np.zeros(2)}
    v = {"w": np.zeros(2)}
    mlp.adam_step(p, g, m, v, t=1, lr=0.1)
    # first bias-corrected step moves each weight by lr * sign(g) (up to eps)
    np.testing.assert_allclose(p["w"], [1.0 - 0.1, -2.0 + 0.1], atol=1e-7)
