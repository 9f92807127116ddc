"""FP64 CPU oracle of the subgrid-closure MLP (TEST INFRASTRUCTURE ONLY).

The closure has no counterpart in the reference; this module is the checker for the
project's own FP32 HIP training and inference kernels.  Network: x (n_in) -> Linear ->
GELU -> Linear -> GELU -> Linear -> y (n_out), GELU(z) = z/2 (1 + erf(z/sqrt 2)); loss
L = mean over rows and outputs of (y - t)^2 (sum / (rows * n_out)).  Weights are stored
row-major W[in, out] so that a layer is Y = X W + b.  Pinned by central finite differences in
tests/test_oracle_mlp.py.
"""
import math

import numpy as np

SQRT1_2 = 1.0 / math.sqrt(2.0)


def _erf(z):
    from scipy.special import erf
    return erf(z)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z * SQRT1_2)) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def init_params(n_in, n_hidden, n_out, seed=0):
    """He-style uniform initialisation (FP64); the GPU takes the FP32 rounding of these."""
    rng = np.random.default_rng(seed)
    def lin(a, b):
        bound = math.sqrt(6.0 / a) * 0.5
        return rng.uniform(-bound, bound, size=(a, b)), np.zeros(b)
    W1, b1 = lin(n_in, n_hidden)
    W2, b2 = lin(n_hidden, n_hidden)
    W3, b3 = lin(n_hidden, n_out)
    return {"W1": W1, "b1": b1, "W2": W2, "b2": b2, "W3": W3, "b3": b3}


def forward(p, X):
    z1 = X @ p["W1"] + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = gelu(z2)
    y = h2 @ p["W3"] + p["b3"]
    return y, (X, z1, h1, z2, h2)


def loss_and_grads(p, X, T):
    y, (X, z1, h1, z2, h2) = forward(p, X)
    n = y.size
    r = y - T
    loss = float(np.sum(r * r) / n)
    dy = 2.0 * r / n
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    dz2 = (dy @ p["W3"].T) * gelu_grad(z2)
    g["W2"], g["b2"] = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * gelu_grad(z1)
    g["W1"], g["b1"] = X.T @ dz1, dz1.sum(0)
    return loss, g


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected), in place."""
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
