"""CPU oracle of the learned-closure data path (TEST INFRASTRUCTURE ONLY).

The closure has no counterpart in the reference.  This module is the FP64 checker for the
library's fine -> coarse L2 projection, the closure's input features and its mean-free
tendency correction.  Arrays follow the package convention interior[e, v, j, i] (C order),
structured meshes are raster ordered (e = ix + nx * iy).

Pinned in tests/test_oracle_closure.py by polynomial exactness, mass conservation and an
independent projection (explicit coarse mass matrix, 24-point quadrature per sub-interval).
"""
import numpy as np

import oracle


def lagrange_eval(nodes, s):
    """l_i(s) for every node i and point s: array [len(nodes), len(s)] (product formula)."""
    nodes = np.asarray(nodes, np.float64)
    s = np.atleast_1d(np.asarray(s, np.float64))
    out = np.ones((nodes.size, s.size))
    for i in range(nodes.size):
        for k in range(nodes.size):
            if k != i:
                out[i] *= (s - nodes[k]) / (nodes[i] - nodes[k])
    return out


def projection_factor(Nf, Nc, r):
    """Px[a, i, p] = w_p^f l_i^c(xi_c(a, p)) / (r w_i^c), xi_c(a, p) = -1 + (2a + 1 + xi_p^f) / r."""
    xf, wf = oracle.legendre_gauss(Nf)
    xc, wc = oracle.legendre_gauss(Nc)
    Px = np.empty((r, Nc + 1, Nf + 1))
    for a in range(r):
        xi = -1.0 + (2.0 * a + 1.0 + xf) / r
        Px[a] = lagrange_eval(xc, xi) * wf[None, :] / (r * wc[:, None])
    return Px


def project_l2_structured(fine, Nf, Nc, nxc, nyc, r):
    """fine[(r nxc)(r nyc), nVar, Nf+1, Nf+1] -> coarse[nxc nyc, nVar, Nc+1, Nc+1]."""
    nVar = fine.shape[1]
    Pf = Nf + 1
    f = fine.reshape(nyc, r, nxc, r, nVar, Pf, Pf)  # [J, b, I, a, v, q, p]
    Px = projection_factor(Nf, Nc, r)
    c = np.einsum("bjq,aip,JbIavqp->JIvji", Px, Px, f, optimize=True)
    return np.ascontiguousarray(c.reshape(nxc * nyc, nVar, Nc + 1, Nc + 1))


def features(S, J, dsdx11, dsdx22, inMean, inScale):
    """Closure inputs: per element the nodal state (v, j, i order) then J, dsdx(1,1),
    dsdx(2,2); x = (raw - mean) / scale in FP64, rounded to FP32."""
    nEl = S.shape[0]
    raw = np.concatenate([S.reshape(nEl, -1), np.stack([J, dsdx11, dsdx22], axis=1)], axis=1)
    return ((raw - inMean.astype(np.float64)) / inScale.astype(np.float64)).astype(np.float32)


def correction(Y, outMean, outScale, qW, J, N, nVar, mass_var=2):
    """dS/dt correction from network outputs Y[nEl, nVar (N+1)^2] (FP32): y * scale + mean in
    FP64, then the mass variable's correction minus its mesh mean (weights w_i w_j J)."""
    nEl = Y.shape[0]
    P = N + 1
    c = (Y.astype(np.float64) * outScale.astype(np.float64) + outMean.astype(np.float64)).reshape(nEl, nVar, P, P)
    w2 = np.outer(qW, qW)
    total = np.sum(J[:, None, None] * w2[None] * c[:, mass_var])
    area = np.sum(J) * np.sum(qW) ** 2
    c[:, mass_var] -= total / area
    return c


def mass(S, qW, J, var=2):
    """sum_e J_e sum_ij w_i w_j S[e, var, j, i] for affine elements (J per element)."""
    return float(np.sum(J[:, None, None] * np.outer(qW, qW)[None] * S[:, var]))
