"""Adjoints for training the closure through the solver, on the GPU: the shallow-water
tendency adjoint against the transposed finite-difference Jacobian of the oracle's tendency
(CPU), the closure-correction adjoint against finite differences of the oracle's correction,
and the a-posteriori rollout gradient against finite differences of the rollout loss."""
import numpy as np
import pytest

import oracle
from oracle import closure as oc
from tests.cases import Case

pytestmark = pytest.mark.gpu


def _oracle_lag(om, og):
    """The engine's lagged face normal velocities after SetSolution: u.n of the current face
    traces with the affine (side node 0) normals, shape (nEl, 4, N+1)."""
    solB = om.field("solB", om.boundary_shape)  # (nEl, 4, nVar, N+1)
    H = solB[:, :, 2]
    nx, ny = og.nHatb[:, :, :, 0], og.nHatb[:, :, :, 1]
    return (solB[:, :, 0] / H) * nx + (solB[:, :, 1] / H) * ny


def _jacobian(f, x0, eps=1e-7):
    """Central-difference Jacobian of f at x0 (flattened)."""
    flat = x0.ravel()
    cols = []
    for k in range(flat.size):
        d = np.zeros(flat.size)
        d[k] = eps
        cols.append((f((flat + d).reshape(x0.shape)) - f((flat - d).reshape(x0.shape))) / (2 * eps))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("N,n", [(3, 4), (2, 3), (5, 3)])
def test_tendency_vjp_matches_oracle_jacobian(N, n):
    """mu = J^T lambda on the GPU against the full Jacobians of the oracle's tendency by central
    differences: with respect to the state S (the lagged velocities held fixed) and with respect
    to the previous stage's state Sp from whose face traces the lagged velocities are formed.
    The step 1e-7 keeps the differences off the kinks of the Rusanov dissipation coefficient
    (a maximum of four wave speeds): at 1e-6 one face of the N = 3 case crosses one."""
    c = Case("sw", N, n, n, periodic=True, seed=7, metrics="affine")
    m, _ = c.gpu_model()
    S0 = c.sol.copy()
    rng = np.random.default_rng(3)
    Sp = S0 * (1.0 + 1e-3 * rng.standard_normal(S0.shape))  # a distinct previous state
    om = c.om
    shp = om.boundary_shape
    om.set_solution(Sp)  # SetSolution's BoundaryInterp: solB = traces of Sp
    solBp = om.field("solB", shp).copy()
    un = _oracle_lag(om, c.geom)

    def f(S):  # tendency at S, lag from Sp
        om.set_solution(S)
        om.field("solB", shp)[...] = solBp
        return om.tendency().ravel()

    def f_lag(P):  # tendency at S0, lag from P
        om.set_solution(P)
        lagB = om.field("solB", shp).copy()
        om.set_solution(S0)
        om.field("solB", shp)[...] = lagB
        return om.tendency().ravel()

    lam = rng.standard_normal(S0.shape)
    want = (_jacobian(f, S0).T @ lam.ravel()).reshape(S0.shape)
    want_lag = (_jacobian(f_lag, Sp).T @ lam.ravel()).reshape(S0.shape)
    import self_amd
    muP = self_amd.DeviceArray(S0.shape)
    got = m.TendencyVJP(S0, un, lam, Sprev=Sp, muPrev=muP).to_host()
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    err_lag = np.max(np.abs(muP.to_host() - want_lag)) / np.max(np.abs(want_lag))
    assert err <= 1e-6, err
    assert err_lag <= 1e-6, err_lag
    # without the previous state the lag is held fixed and mu is the same
    got2 = m.TendencyVJP(S0, un, lam).to_host()
    assert np.array_equal(got, got2)


def test_closure_vjp_matches_finite_differences():
    import self_amd
    from tests.test_gpu_closure import _coarse_model, _mlp, _norm
    m, sol, J, d11, d22, qW = _coarse_model(8, 3, seed=4)
    net, _ = _mlp(13)
    inMean, inScale, outMean, outScale = _norm(14, scale_out=0.05)
    m.SetClosure(net, inMean, inScale, outMean, outScale)
    S = m.solution
    rng = np.random.default_rng(15)
    lam = rng.standard_normal(S.shape)

    def corr(state):
        X = oc.features(state, J, d11, d22, inMean, inScale)
        return oc.correction(net.forward(X), outMean, outScale, qW, J, 3, 3)

    net.zero_grads()
    mu = self_amd.DeviceArray(S.shape, zero=True)
    m.ClosureVJP(S, lam, mu)
    g = net.get_grads()
    # parameters: directional derivative along a random direction
    p0 = net.get_params()
    dirn = {k: rng.standard_normal(v.shape).astype(np.float32) for k, v in p0.items()}
    h = 1e-2
    net.set_params({k: p0[k] + h * dirn[k] for k in p0})
    cp = corr(S)
    net.set_params({k: p0[k] - h * dirn[k] for k in p0})
    cm = corr(S)
    net.set_params(p0)
    fd = np.sum(lam * (cp - cm)) / (2 * h)
    adj = sum(float(np.sum(g[k].astype(np.float64) * dirn[k])) for k in p0)
    assert abs(adj - fd) <= 2e-2 * abs(fd), (adj, fd)
    # state, through the network inputs
    v = rng.standard_normal(S.shape)
    hs = 1e-3
    fd_s = np.sum(lam * (corr(S + hs * v) - corr(S - hs * v))) / (2 * hs)
    adj_s = float(np.sum(mu.to_host() * v))
    assert abs(adj_s - fd_s) <= 2e-2 * abs(fd_s), (adj_s, fd_s)


def test_rollout_gradient_matches_finite_differences():
    """The a-posteriori loss of three corrected coarse steps: its adjoint gradient against a
    central difference of the loss along the gradient's own direction (a random direction's
    derivative is a sum of cancelling terms, below the FP32 network's rounding)."""
    import self_amd
    from self_amd import closure as cl
    from tests.test_gpu_closure import _coarse_model, _mlp, _norm
    m, sol, J, d11, d22, qW = _coarse_model(8, 3, seed=5)
    net, _ = _mlp(16)
    norm_arrays = _norm(17, scale_out=0.2)
    m.SetClosure(net, *norm_arrays)
    dt = 2e-3
    P0 = m.solution
    # targets: the plain run perturbed, so the loss is not at a stationary point
    m2, *_ = _coarse_model(8, 3, seed=5)
    truths = []
    for _ in range(3):
        m2.ForwardStep(1, dt)
        truths.append(m2.solution * 1.001)
    w2 = np.outer(qW, qW)
    weights = (J[:, None, None] * w2[None])[:, None]
    loss = cl.rollout_loss_grad(m, net, P0, truths, dt, weights)
    g = net.get_grads()
    p0 = net.get_params()
    gnorm = np.sqrt(sum(float(np.sum(g[k].astype(np.float64) ** 2)) for k in p0))
    dirn = {k: (g[k] / gnorm).astype(np.float32) for k in p0}
    h = 3e-3
    net.set_params({k: p0[k] + h * dirn[k] for k in p0})
    lp = cl.rollout_loss_grad(m, net, P0, truths, dt, weights, want_grad=False)
    net.set_params({k: p0[k] - h * dirn[k] for k in p0})
    lm = cl.rollout_loss_grad(m, net, P0, truths, dt, weights, want_grad=False)
    net.set_params(p0)
    fd = (lp - lm) / (2 * h)
    adj = sum(float(np.sum(g[k].astype(np.float64) * dirn[k])) for k in p0)
    assert np.isfinite(loss) and loss > 0
    # measured 1.110356e-7 against 1.110369e-7 (h = 3e-3) and 1.110355e-7 (h = 1e-3)
    assert abs(adj - fd) <= 1e-3 * abs(fd), (adj, fd)


def test_aposteriori_training_reduces_the_window_loss():
    """Training through the solver on one window of a small configuration (coarse 16 x 16 N = 3,
    fine 64 x 64 N = 7): from the zero correction (the plain run) the rollout loss of that
    window goes down, and the corrected run conserves mass."""
    from self_amd import closure as cl
    cf = cl.CoarseFine(nxc=16)
    wins = cl.rollout_windows(cf, [0], 4, every=25, n_steps=50)
    w = wins[1]
    mlp, norm, hist = cl.aposteriori_network(cf, [w], 150, lr=1e-3)
    weights = cl.l2_weights(cf)
    C = cf.coarse_model(w[0])
    C.SetClosure(mlp, *norm.arrays(0.0))
    plain = cl.rollout_loss_grad(C, mlp, w[1], w[2], cf.dt_c, weights, want_grad=False)
    C.SetClosure(mlp, *norm.arrays(1.0))
    trained = cl.rollout_loss_grad(C, mlp, w[1], w[2], cf.dt_c, weights, want_grad=False)
    print(f"window loss plain {plain:.6e} trained {trained:.6e} ratio {trained / plain:.4f}", flush=True)
    assert np.isfinite(hist).all()
    assert trained < plain
    m0 = cf.mass(w[1])
    assert abs(cf.mass(C.solution) - m0) <= 1e-13 * abs(m0)
