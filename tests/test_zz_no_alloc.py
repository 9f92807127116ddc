"""The no-allocation integral entry point (selfamd_model2d_integral_noalloc) against
selfamd_model2d_integral, through the C ABI: bit-identical results on two models in a row (the
second smaller, so it reuses the library's fixed partial-sum array after a larger call), and a
mesh beyond that array's capacity rejected with an error."""
import ctypes

import numpy as np
import pytest

import self_amd
from self_amd import workloads
from self_amd._lib import SelfAmdError, call

pytestmark = pytest.mark.gpu


def _sw_model(n, N, seed):
    mesh = self_amd.Mesh2D.structured(n, n, periodic=(True, True))
    interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
    geom = self_amd.SEMQuad(interp, mesh)
    sol, topo = workloads.nonlinear_sw(geom.x.to_host(), interp.controlPoints, interp.qWeights, seed=seed)
    m = self_amd.ShallowWater(mesh, geom)
    m.SetTopography(topo)
    m.SetSolution(sol)
    return m, geom


def test_integral_noalloc_matches_integral_on_two_calls():
    out = self_amd.DeviceArray((1,), zero=True)
    # 96 x 96 elements at N = 7 (589,824 nodes, 2,304 partial sums), then 10 x 10 at N = 3
    for n, N, seed in ((96, 7, 11), (10, 3, 12)):
        m, _geom = _sw_model(n, N, seed)
        m.ForwardStep(2, 1e-4)
        for var in range(3):
            ref = m.Integral(var)
            m.IntegralNoAlloc(var, out)
            self_amd.synchronize()
            got = float(out.to_host()[0])
            assert got == ref, (n, N, var, got, ref)
        # after the larger call the array holds stale partial sums beyond this mesh's; the
        # smaller mesh's total above must not have read them (checked by equality), and a
        # repeated call gives the same bits
        m.IntegralNoAlloc(2, out)
        self_amd.synchronize()
        assert float(out.to_host()[0]) == m.Integral(2)


def test_integral_noalloc_rejects_meshes_beyond_its_scratch():
    # 513 x 512 elements at N = 7: 16,809,984 nodes, above the 16,777,216 the array holds
    mesh = self_amd.Mesh2D.structured(513, 512, periodic=(True, True))
    interp = self_amd.Lagrange(7, self_amd.GAUSS, 7, self_amd.GAUSS)
    geom = self_amd.SEMQuad(interp, mesh)
    m = self_amd.ShallowWater(mesh, geom)
    m.SetSolution(np.ones(m.shape))
    out = self_amd.DeviceArray((1,), zero=True)
    with pytest.raises(SelfAmdError, match="at most 16777216"):
        m.IntegralNoAlloc(2, out)
    # the allocating entry point still takes it: H = 1 over the unit square
    assert abs(m.Integral(2) - 1.0) <= 1e-9
    # a null output is an argument error, not a launch
    with pytest.raises(SelfAmdError, match="null argument"):
        call("selfamd_model2d_integral_noalloc", m._h, 2, ctypes.c_void_p())
