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


def test_batch_rejects_mixed_kinds():
    _, m1, _ = _run("sw", 3, 3, 1)
    _, m2, _ = _run("lsw", 3, 3, 1)
    with pytest.raises(self_amd.SelfAmdError):
        self_amd.Batch([m1, m2], [1e-3, 1e-3])
