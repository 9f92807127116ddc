"""Product host setup (libself_amd.so, CPU-only entry points) against the oracle: Lagrange
matrices bit-exact, mesh connectivity bit-exact (structured generator and HOPr fixtures)."""
import json
import os

import numpy as np
import pytest

import oracle
import self_amd

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("N,ctrl,M,tgt", [(3, 1, 3, 1), (5, 1, 14, 5), (7, 1, 7, 1), (7, 2, 10, 5), (1, 5, 7, 1),
                                          (6, 5, 3, 1)])
def test_lagrange_bit_exact(N, ctrl, M, tgt):
    p = self_amd.Lagrange(N, ctrl, M, tgt)
    o = oracle.lagrange_init(N, ctrl, M, tgt)
    for k, v in o.items():
        np.testing.assert_array_equal(getattr(p, k), v, err_msg=k)


@pytest.mark.parametrize("nx,ny,periodic,bcid", [(1, 1, False, 102), (4, 3, False, 101), (5, 4, True, 102),
                                                 (2, 2, True, 100), (16, 9, False, 100)])
def test_structured_mesh(nx, ny, periodic, bcid):
    m = self_amd.Mesh2D.structured(nx, ny, -1.0, 3.0, 0.0, 2.0, periodic=(periodic, periodic), bcid=bcid)
    si = m.sideInfo
    # flips recomputed by the oracle (reference RecalculateFlip) are unchanged
    np.testing.assert_array_equal(oracle.recalculate_flip(si, m.globalNodeIDs, 1), si)
    opp = {0: 2, 1: 3, 2: 0, 3: 1}
    for e in range(nx * ny):
        ix, iy = e % nx, e // nx
        for s in range(4):
            e2, code, bc = si[e, s, 2], si[e, s, 3], si[e, s, 4]
            onb = (s == 0 and iy == 0) or (s == 2 and iy == ny - 1) or (s == 1 and ix == nx - 1) or (s == 3 and ix == 0)
            if onb and not periodic:
                assert e2 == 0 and bc == bcid
            else:
                assert bc == 0 and code == 10 * (opp[s] + 1)
                assert si[e2 - 1, opp[s], 2] == e + 1
                assert abs(si[e, s, 1]) == abs(si[e2 - 1, opp[s], 1])
    # element corner coordinates
    dx, dy = 4.0 / nx, 2.0 / ny
    np.testing.assert_allclose(m.nodeCoords[-1, 1, 1], [3.0, 2.0])
    np.testing.assert_allclose(m.nodeCoords[0, 0, 0], [-1.0, 0.0])
    np.testing.assert_allclose(m.nodeCoords[0, 1, 1], [-1.0 + dx, dy])


@pytest.mark.parametrize("name", ["hopr_block2d", "hopr_geophysical_block2d_medium", "hopr_circle"])
def test_hopr_conversion_bit_exact(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        fx = json.load(f)
    si3 = np.array(fx["SideInfo"]["data"], np.int32).reshape(fx["SideInfo"]["shape"])
    nc3 = np.array(fx["NodeCoords"]["data"]).reshape(fx["NodeCoords"]["shape"])
    gn3 = np.array(fx["GlobalNodeIDs"]["data"], np.int32)
    m = self_amd.Mesh2D.from_hopr(fx["Ngeo"], si3, nc3, gn3)
    si, nc, gn = oracle.hopr_to_mesh2d(fx["Ngeo"], si3, nc3, gn3)
    np.testing.assert_array_equal(m.sideInfo, si)
    np.testing.assert_array_equal(m.nodeCoords, nc)
    np.testing.assert_array_equal(m.globalNodeIDs, gn)
    m.ResetBoundaryConditionType(self_amd.BC_NONORMALFLOW)
    assert set(np.unique(m.sideInfo[..., 4])) <= {0, self_amd.BC_NONORMALFLOW}


def test_bad_arguments_raise():
    with pytest.raises(self_amd.SelfAmdError):
        self_amd.Mesh2D.structured(0, 3)
    with pytest.raises(self_amd.SelfAmdError):
        self_amd.Lagrange(0)
