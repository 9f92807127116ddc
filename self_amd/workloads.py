"""Seeded synthetic workloads for the 2D DGSEM configurations.

Each builder takes the model node coordinates x[e, j, i, 2] (SEMQuad%x) and returns numpy
arrays in the package's C-order convention (interior[e, v, j, i]).  Definitions:

  gravity_wave_lsw   linear SW gravity-wave release: u = v = 0, eta = 0.01 exp(-r^2/0.01)
                     around a seeded centre in [0.3, 0.7]^2 (the 0.01 denominator is the
                     reference example's width convention); g = H = 1, f = 0.
  nonlinear_sw       H = 1 + 0.1 exp(-r^2/0.1^2) at a seeded centre; u, v ~ N(0, 0.01^2) per
                     node, smoothed per element by keeping Legendre modes of degree <= 2 in
                     each direction; Hu = H u, Hv = H v.  Topography b = 0.05 * sum_k a_k
                     cos(2 pi (m_k x + n_k y) + phi_k) / sum_k a_k over 16 seeded integer
                     wave vectors (periodic on the unit square).
  euler_hydrostatic  isothermal column rho = exp(-y), p = exp(-y) (R T = 1, gravity
                     potential y), rho multiplied by 1 + 1e-3 exp(-r^2/0.05^2) at a seeded
                     centre with p unchanged, at rest; rho E = p/(gamma - 1).
"""
import numpy as np


def _legendre_vandermonde(nodes, kmax):
    V = np.zeros((len(nodes), kmax + 1))
    V[:, 0] = 1.0
    if kmax >= 1:
        V[:, 1] = nodes
    for k in range(2, kmax + 1):
        V[:, k] = ((2 * k - 1) * nodes * V[:, k - 1] - (k - 1) * V[:, k - 2]) / k
    return V


def modal_filter(u, nodes, weights, keep):
    """Per-element L2 projection of nodal values u[e, j, i] onto tensor Legendre modes of
    degree <= keep (exact with the element's Gauss quadrature)."""
    N = len(nodes) - 1
    V = _legendre_vandermonde(nodes, N)                 # V[i, k] = P_k(x_i)
    norm = 2.0 / (2.0 * np.arange(N + 1) + 1.0)
    A = (V * weights[:, None]).T / norm[:, None]        # modal coefficients c = A u
    mask = (np.arange(N + 1) <= keep).astype(float)
    F = V @ (mask[:, None] * A)                         # nodal filter operator
    return np.einsum("jb,ia,eba->eji", F, F, u)


def gravity_wave_lsw(x, seed=0):
    rng = np.random.default_rng(seed)
    xc, yc = rng.uniform(0.3, 0.7, size=2)
    X, Y = x[..., 0], x[..., 1]
    sol = np.zeros((x.shape[0], 3) + x.shape[1:3])
    sol[:, 2] = 0.01 * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / 0.01)
    return sol


def topography_modes(X, Y, seed, amplitude=0.05, nmodes=16, kmax=4):
    rng = np.random.default_rng(seed + 7919)
    m = rng.integers(-kmax, kmax + 1, size=nmodes)
    n = rng.integers(-kmax, kmax + 1, size=nmodes)
    a = rng.uniform(0.5, 1.0, size=nmodes)
    phi = rng.uniform(0.0, 2 * np.pi, size=nmodes)
    b = np.zeros_like(X)
    for k in range(nmodes):
        b += a[k] * np.cos(2 * np.pi * (m[k] * X + n[k] * Y) + phi[k])
    return amplitude * b / a.sum()


def nonlinear_sw(x, nodes, weights, seed=0, width=0.1, bump=0.1, vel_std=0.01):
    rng = np.random.default_rng(seed)
    xc, yc = rng.uniform(0.25, 0.75, size=2)
    X, Y = x[..., 0], x[..., 1]
    H = 1.0 + bump * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / width ** 2)
    u = modal_filter(rng.normal(0.0, vel_std, size=X.shape), nodes, weights, 2)
    v = modal_filter(rng.normal(0.0, vel_std, size=X.shape), nodes, weights, 2)
    sol = np.stack([H * u, H * v, H], axis=1)
    return np.ascontiguousarray(sol), np.ascontiguousarray(topography_modes(X, Y, seed))


def euler_hydrostatic(x, seed=0, gamma=1.4, amp=1e-3, width=0.05):
    rng = np.random.default_rng(seed)
    xc, yc = rng.uniform(0.3, 0.7, size=2)
    X, Y = x[..., 0], x[..., 1]
    p = np.exp(-Y)
    rho = np.exp(-Y) * (1.0 + amp * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / width ** 2))
    sol = np.zeros((x.shape[0], 4) + x.shape[1:3])
    sol[:, 2] = rho
    sol[:, 3] = p / (gamma - 1.0)
    return sol, np.ascontiguousarray(Y.copy())


def cfl_dt(dx_min, wave_speed, cfl=0.5):
    """dt = CFL * min(dx) / max(wave speed) (pattern of src/models/SELF_CompressibleIdealGas2D.f90:909-918)."""
    return cfl * dx_min / wave_speed
