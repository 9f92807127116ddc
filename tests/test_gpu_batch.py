"""Batched runs: uneven meshes and degrees advanced together must give
each run bit-for-bit the state it reaches when advanced alone, and match the oracle."""
import numpy as np
import pytest

import self_amd
from self_amd import workloads
from tests.cases import Case, rel_err

pytestmark = pytest.mark.gpu


def _run(kind, N, n, seed, periodic=True):
    c = Case(kind, N, n, n, periodic=periodic, bcid=self_amd.BC_NONORMALFLOW, seed=seed, metrics="affine")
    m, g = c.gpu_model()
    return c, m, g


@pytest.mark.parametrize("kind", ["sw", "lsw", "euler"])
def test_batch_equals_separate(kind):
    specs = [(7, 5, 1), (3, 9, 2), (5, 4, 3), (7, 3, 4), (3, 6, 5), (5, 7, 6)]
    periodic = kind == "sw"
    runs = [_run(kind, N, n, s, periodic) for N, n, s in specs]
    alone = [_run(kind, N, n, s, periodic) for N, n, s in specs]
    dts = [c.dt() for c, _, _ in runs]
    b = self_amd.Batch([m for _, m, _ in runs], dts)
    groups, dof = b.info()
    assert groups == 3
    assert dof == sum(n * n * (N + 1) ** 2 * (4 if kind == "euler" else 3) for N, n, _ in specs)
    b.ForwardStep(5)
    b.scatter()
    for (c, m, _), (_, ma, _), dt in zip(runs, alone, dts):
        ma.ForwardStep(5, dt)
        np.testing.assert_array_equal(m.solution, ma.solution)
    # the oracle agrees with the batched result of two of the runs
    for k in (0, 1):
        c, m, _ = runs[k]
        c.om.rk3_steps(5, dts[k])
        assert rel_err(m.solution, c.om.solution).max() <= 1e-12
    # each model continues alone after the batch
    for (c, m, _), (_, ma, _), dt in zip(runs, alone, dts):
        m.ForwardStep(2, dt)
        ma.ForwardStep(2, dt)
        np.testing.assert_array_equal(m.solution, ma.solution)


@pytest.mark.parametrize("kind", ["sw", "lsw", "euler"])
def test_batch_long_calls_equal_separate(kind):
    """Calls of many steps replay graphs of 2 to 16 steps (the largest as often as it fits, each
    smaller one at most once, a one-step graph for an odd count) from either trace buffer; every
    run still equals its separate solve bit for bit, in both launch modes (the persistent one
    for SW)."""
    specs = [(7, 5, 1), (3, 9, 2), (5, 4, 3), (7, 3, 4)]
    periodic = kind == "sw"
    runs = [_run(kind, N, n, s, periodic) for N, n, s in specs]
    alone = [_run(kind, N, n, s, periodic) for N, n, s in specs]
    dts = [c.dt() for c, _, _ in runs]
    b = self_amd.Batch([m for _, m, _ in runs], dts)
    b.ForwardStep(37)  # 16 + 16 + 4 + 1: the other trace buffer is current afterwards
    if kind == "sw":
        b.SetMode(self_amd.Batch.PERSISTENT)
    b.ForwardStep(22)  # 16 + 4 + 2
    b.SetMode(self_amd.Batch.PER_GROUP)
    b.ForwardStep(9)  # 8 + 1
    b.scatter()
    for (c, m, _), (_, ma, _), dt in zip(runs, alone, dts):
        ma.ForwardStep(68, dt)
        np.testing.assert_array_equal(m.solution, ma.solution)
        assert abs(m.t - ma.t) <= 1e-12 * abs(ma.t)


@pytest.mark.parametrize("periodic", [True, False])
def test_batch_persistent_equals_separate(periodic):
    """One persistent launch per stage for all degree groups: every run equals its separate
    solve bit for bit, also when the mode changes between steps."""
    specs = [(7, 5, 1), (3, 9, 2), (5, 4, 3), (7, 3, 4), (3, 6, 5), (5, 7, 6), (3, 2, 7)]
    runs = [_run("sw", N, n, s, periodic) for N, n, s in specs]
    alone = [_run("sw", N, n, s, periodic) for N, n, s in specs]
    dts = [c.dt() for c, _, _ in runs]
    b = self_amd.Batch([m for _, m, _ in runs], dts)
    b.SetMode(self_amd.Batch.PERSISTENT)
    b.ForwardStep(3)
    b.SetMode(self_amd.Batch.PER_GROUP)
    b.ForwardStep(1)
    b.SetMode(self_amd.Batch.PERSISTENT)
    b.ForwardStep(2)
    b.scatter()
    for (c, m, _), (_, ma, _), dt in zip(runs, alone, dts):
        ma.ForwardStep(6, dt)
        np.testing.assert_array_equal(m.solution, ma.solution)
    c, m, _ = runs[1]
    c.om.rk3_steps(6, dts[1])
    assert rel_err(m.solution, c.om.solution).max() <= 1e-12


def test_batch_persistent_rejects_other_models():
    _, m1, _ = _run("lsw", 3, 3, 1, periodic=False)
    b = self_amd.Batch([m1], [1e-3])
    with pytest.raises(self_amd.SelfAmdError):
        b.SetMode(self_amd.Batch.PERSISTENT)
    _, m2, _ = _run("sw", 4, 3, 1)
    b = self_amd.Batch([m2], [1e-3])
    with pytest.raises(self_amd.SelfAmdError):
        b.SetMode(self_amd.Batch.PERSISTENT)
    b.ForwardStep(1)  # the batch keeps its per-group graphs


def test_batch_rejects_mixed_kinds():
    _, m1, _ = _run("sw", 3, 3, 1)
    _, m2, _ = _run("lsw", 3, 3, 1)
    with pytest.raises(self_amd.SelfAmdError):
        self_amd.Batch([m1, m2], [1e-3, 1e-3])


@pytest.mark.parametrize("mode", [self_amd.Batch.PER_GROUP, self_amd.Batch.PERSISTENT])
def test_full_size_ensemble(mode):
    """The batched configuration as benchmarked: 512 SW runs, n x n periodic elements with
    n ~ U{8..64} and N ~ U{3,5,7} (seed 0), in either launch mode.  After 5 batched steps, 16
    runs spread over the batch equal the same runs advanced alone bit for bit, and two of them
    (one per end of the degree range) match the oracle."""
    import oracle
    from tests.cases import KIND, flatten_metrics
    rng = np.random.default_rng(0)
    n_runs, steps = 512, 5
    sizes = rng.integers(8, 65, size=n_runs)
    degrees = rng.choice([3, 5, 7], size=n_runs)
    cache, models, dts, inits = {}, [], [], []
    for r in range(n_runs):
        n, N = int(sizes[r]), int(degrees[r])
        if (n, N) not in cache:
            mesh = self_amd.Mesh2D.structured(n, n, periodic=(True, True))
            interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
            geom = self_amd.SEMQuad(interp, mesh)
            cache[(n, N)] = (mesh, interp, geom, geom.x.to_host(), geom.CovariantArcMin())
        mesh, interp, geom, x, arcmin = cache[(n, N)]
        sol, topo = workloads.nonlinear_sw(x, interp.controlPoints, interp.qWeights, seed=1000 + r)
        m = self_amd.ShallowWater(mesh, geom)
        m.SetTopography(topo)
        m.SetSolution(sol)
        models.append(m)
        dts.append(workloads.cfl_dt(arcmin, 1.2, 0.5))
        inits.append((sol, topo))
    b = self_amd.Batch(models, dts)
    groups, _ = b.info()
    assert groups == 3
    b.SetMode(mode)
    b.ForwardStep(steps)
    b.scatter()
    for r in np.linspace(0, n_runs - 1, 16).astype(int):
        mesh, interp, geom, _, _ = cache[(int(sizes[r]), int(degrees[r]))]
        m = self_amd.ShallowWater(mesh, geom)
        m.SetTopography(inits[r][1])
        m.SetSolution(inits[r][0])
        m.ForwardStep(steps, dts[r])
        np.testing.assert_array_equal(models[r].solution, m.solution, err_msg=f"run {r}")
    for N in (3, 7):
        r = int(np.flatnonzero(degrees == N)[0])
        mesh = cache[(int(sizes[r]), N)][0]
        og = oracle.Geometry(N, mesh.nodeCoords, mesh.nGeo)
        flatten_metrics(og)
        om = oracle.Model(KIND["sw"], N, mesh.sideInfo, og)
        om.set_topography(inits[r][1])
        om.set_solution(inits[r][0])
        om.rk3_steps(steps, dts[r])
        assert rel_err(models[r].solution, om.solution).max() <= 1e-12, f"run {r}"
