"""Pins the closure oracle's fine -> coarse L2 projection and mean-free correction (CPU)."""
import numpy as np
import pytest

import oracle
from oracle import closure as oc


def _nodes_1d(n_el, N, x0=0.0, x1=1.0):
    """Physical coordinates [element, node] of a uniform 1D mesh on Gauss nodes (oracle order)."""
    xg, _ = oracle.legendre_gauss(N)
    h = (x1 - x0) / n_el
    return x0 + (np.arange(n_el)[:, None] + (xg[None, :] + 1.0) / 2.0) * h


def _field(f, nx, ny, N):
    """interior[e, 1, j, i] of f(x, y) on a raster-ordered nx x ny structured mesh."""
    X, Y = _nodes_1d(nx, N), _nodes_1d(ny, N)
    out = np.empty((nx * ny, 1, N + 1, N + 1))
    for iy in range(ny):
        for ix in range(nx):
            out[ix + nx * iy, 0] = f(X[ix][None, :], Y[iy][:, None])
    return out


@pytest.mark.parametrize("Nf,Nc,r", [(7, 3, 4), (5, 3, 2), (3, 3, 3), (2, 1, 2)])
def test_projection_reproduces_coarse_polynomials(Nf, Nc, r):
    nxc, nyc = 2, 3
    rng = np.random.default_rng(Nf + 10 * Nc + 100 * r)
    c = rng.standard_normal((Nc + 1, Nc + 1))

    def poly(x, y, deg=Nc):
        return sum(c[m, n] * x ** m * y ** n for m in range(deg + 1) for n in range(deg + 1))

    fine = _field(poly, r * nxc, r * nyc, Nf)
    got = oc.project_l2_structured(fine, Nf, Nc, nxc, nyc, r)
    want = _field(poly, nxc, nyc, Nc)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # degree Nc + 2 is neither in the coarse space nor integrated exactly by the coarse Gauss
    # rule against l_i (degree Nc + 1 would be: its projection is its Gauss-node interpolant)
    fine2 = _field(lambda x, y: x ** (Nc + 2) * y, r * nxc, r * nyc, Nf)
    got2 = oc.project_l2_structured(fine2, Nf, Nc, nxc, nyc, r)
    want2 = _field(lambda x, y: x ** (Nc + 2) * y, nxc, nyc, Nc)
    assert np.max(np.abs(got2 - want2)) > 1e-8


def test_projection_conserves_mass():
    Nf, Nc, r, nxc, nyc = 7, 3, 4, 3, 2
    rng = np.random.default_rng(7)
    fine = rng.standard_normal((r * nxc * r * nyc, 3, Nf + 1, Nf + 1))
    coarse = oc.project_l2_structured(fine, Nf, Nc, nxc, nyc, r)
    _, wf = oracle.legendre_gauss(Nf)
    _, wc = oracle.legendre_gauss(Nc)
    hf, hc = 1.0 / (r * nxc), 1.0 / nxc
    for v in range(3):
        mf = oc.mass(fine, wf, np.full(fine.shape[0], (hf / 2) ** 2), v)
        mc = oc.mass(coarse, wc, np.full(coarse.shape[0], (hc / 2) ** 2), v)
        assert abs(mf - mc) <= 1e-14 * np.sum(np.abs(fine[:, v])) * (hf / 2) ** 2 * 4


def test_projection_matches_mass_matrix_solve():
    """Independent route: coarse mass matrix (not assumed diagonal) and right-hand side by a
    24-point Gauss-Legendre rule (numpy) on every sub-element, then a dense solve."""
    Nf, Nc, r = 2, 2, 3
    rng = np.random.default_rng(3)
    fine = rng.standard_normal((r * r, 1, Nf + 1, Nf + 1))
    xf, _ = oracle.legendre_gauss(Nf)
    xc, _ = oracle.legendre_gauss(Nc)
    g, wg = np.polynomial.legendre.leggauss(24)
    Pc = Nc + 1
    M = np.zeros((Pc * Pc, Pc * Pc))
    b = np.zeros(Pc * Pc)
    for bb in range(r):
        for a in range(r):
            uf = fine[a + r * bb, 0]  # [q, p]
            # sub-element points in coarse coordinates, fine reference coordinates = g
            xi_c = -1.0 + (2 * a + 1 + g) / r
            eta_c = -1.0 + (2 * bb + 1 + g) / r
            lx, ly = oc.lagrange_eval(xc, xi_c), oc.lagrange_eval(xc, eta_c)  # [i, s]
            fx, fy = oc.lagrange_eval(xf, g), oc.lagrange_eval(xf, g)
            u = np.einsum("ps,qt,qp->ts", fx, fy, uf)  # u at (xi=s, eta=t)
            w2 = np.outer(wg, wg) / r ** 2  # [t, s]
            phi = np.einsum("jt,is->jits", ly, lx).reshape(Pc * Pc, g.size, g.size)  # basis (j,i) at (t,s)
            b += np.einsum("kts,ts,ts->k", phi, u, w2)
            M += np.einsum("kts,lts,ts->kl", phi, phi, w2)
    want = np.linalg.solve(M, b).reshape(Pc, Pc)
    got = oc.project_l2_structured(fine, Nf, Nc, 1, 1, r)[0, 0]
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_correction_is_mean_free_in_mass():
    rng = np.random.default_rng(5)
    N, nVar, nEl = 3, 3, 12
    Y = rng.standard_normal((nEl, nVar * 16)).astype(np.float32)
    mu = rng.standard_normal(nVar * 16).astype(np.float32)
    sd = (1 + rng.random(nVar * 16)).astype(np.float32)
    _, qW = oracle.legendre_gauss(N)
    J = 0.01 * (1 + rng.random(nEl))
    c = oc.correction(Y, mu, sd, qW, J, N, nVar)
    raw = (Y.astype(np.float64) * sd + mu).reshape(nEl, nVar, 4, 4)
    assert np.array_equal(c[:, :2], raw[:, :2])
    assert abs(oc.mass(c, qW, J)) <= 1e-15 * oc.mass(np.abs(raw), qW, J)
    d = c[:, 2] - raw[:, 2]
    assert np.ptp(d) <= 4 * np.finfo(float).eps * np.max(np.abs(raw))  # a uniform shift
