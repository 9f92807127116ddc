"""Operator-level parity on the GPU: every per-operator entry point of the C ABI against the
oracle on the same seeded inputs.  Bar: bit-exact for integer/copy work (side exchange),
relative max-norm <= 1e-12 per variable for floating point."""
import ctypes

import numpy as np
import pytest

import oracle
import self_amd
from self_amd import _lib
from tests.cases import Case, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def dev(a, dtype=np.float64):
    assert self_amd.device_count() >= 1, "GPU tests need a ROCm device"
    return self_amd.DeviceArray.from_host(np.ascontiguousarray(a, dtype=dtype))


def zeros(shape):
    return self_amd.DeviceArray(shape, zero=True)


def p(t):
    """Argument marker: the DeviceArray itself travels to run(), which keeps it referenced
    (and so allocated) until the call has completed."""
    assert isinstance(t, self_amd.DeviceArray)
    return t


def run(name, *args):
    held = args  # device arrays stay alive for the whole call
    _lib.call(name, *[a.ptr if isinstance(a, self_amd.DeviceArray) else a for a in held], None)
    self_amd.synchronize()
    del held


def host(t):
    return t.to_host()


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(11)


@pytest.mark.parametrize("N", [3, 5, 7])
def test_boundary_interp(N, rng):
    L = oracle.lagrange_init(N, oracle.GAUSS, N, oracle.GAUSS)
    nEl, nVar = 37, 3
    f = rng.normal(size=(nEl, nVar, N + 1, N + 1))
    bM = dev(L["bMatrix"])
    out = zeros((nEl, 4, nVar, N + 1))
    run("selfamd_scalar_boundary_interp_2d", p(bM), p(dev(f)), p(out), N, nVar, nEl)
    assert rel_err(host(out), oracle.scalar_boundary_interp(N, L["bMatrix"], f), 2).max() <= TOL
    fv = rng.normal(size=(nEl, nVar, N + 1, N + 1, 2))
    outv = zeros((nEl, 4, nVar, N + 1, 2))
    run("selfamd_vector_boundary_interp_2d", p(bM), p(dev(fv)), p(outv), N, nVar, nEl)
    assert rel_err(host(outv), oracle.vector_boundary_interp(N, L["bMatrix"], fv), 2).max() <= TOL
    ft = rng.normal(size=(nEl, 1, N + 1, N + 1, 2, 2))
    outt = zeros((nEl, 4, 1, N + 1, 2, 2))
    run("selfamd_tensor_boundary_interp_2d", p(bM), p(dev(ft)), p(outt), N, 1, nEl)
    assert rel_err(host(outt), oracle.tensor_boundary_interp(N, L["bMatrix"], ft), 2).max() <= TOL


@pytest.mark.parametrize("N,M", [(7, 14), (3, 7), (1, 7)])
def test_grid_interp(N, M, rng):
    L = oracle.lagrange_init(N, oracle.GAUSS if N > 1 else oracle.UNIFORM, M, oracle.UNIFORM)
    nEl, nVar = 13, 2
    f = rng.normal(size=(nEl, nVar, N + 1, N + 1))
    out = zeros((nEl, nVar, M + 1, M + 1))
    iM = dev(L["iMatrix"])
    run("selfamd_scalar_grid_interp_2d", p(iM), p(dev(f)), p(out), N, M, nVar, nEl)
    assert rel_err(host(out), oracle.scalar_grid_interp(N, M, L["iMatrix"], f)).max() <= TOL
    fv = rng.normal(size=(nEl, nVar, N + 1, N + 1, 2))
    outv = zeros((nEl, nVar, M + 1, M + 1, 2))
    run("selfamd_vector_grid_interp_2d", p(iM), p(dev(fv)), p(outv), N, M, nVar, nEl)
    assert rel_err(host(outv), oracle.vector_grid_interp(N, M, L["iMatrix"], fv)).max() <= TOL


@pytest.mark.parametrize("N", [3, 7])
def test_derivatives(N, rng):
    L = oracle.lagrange_init(N, oracle.GAUSS, N, oracle.GAUSS)
    nEl, nVar = 21, 3
    f = rng.normal(size=(nEl, nVar, N + 1, N + 1, 2))
    bF = rng.normal(size=(nEl, 4, nVar, N + 1))
    g = zeros((nEl, nVar, N + 1, N + 1, 2, 2))
    run("selfamd_vector_gradient_2d", p(dev(L["dMatrix"])), p(dev(f)), p(g), N, nVar, nEl)
    assert rel_err(host(g), oracle.vector_gradient(N, L["dMatrix"], f)).max() <= TOL
    d = zeros((nEl, nVar, N + 1, N + 1))
    run("selfamd_vector_divergence_2d", p(dev(L["dMatrix"])), p(dev(f)), p(d), N, nVar, nEl)
    assert rel_err(host(d), oracle.vector_divergence(N, L["dMatrix"], f)).max() <= TOL
    run("selfamd_vector_dg_divergence_2d", p(dev(L["dgMatrix"])), p(dev(L["bMatrix"])), p(dev(L["qWeights"])),
        p(dev(f)), p(dev(bF)), p(d), N, nVar, nEl)
    ref = oracle.vector_dg_divergence(N, L["dgMatrix"], L["bMatrix"], L["qWeights"], f, bF)
    assert rel_err(host(d), ref).max() <= TOL
    t = rng.normal(size=(nEl, 1, N + 1, N + 1, 2, 2))
    det = zeros((nEl, 1, N + 1, N + 1))
    run("selfamd_determinant_tensor_2d", p(dev(t)), p(det), N, 1, nEl)
    assert rel_err(host(det), oracle.determinant(N, t)).max() <= TOL


@pytest.mark.parametrize("periodic", [False, True])
def test_side_exchange_bit_exact(periodic, rng):
    N, nVar = 7, 3
    mesh = self_amd.Mesh2D.structured(6, 5, periodic=(periodic, periodic))
    nEl = mesh.nElem
    b = rng.normal(size=(nEl, 4, nVar, N + 1))
    ext = zeros(b.shape)
    si = dev(mesh.sideInfo, dtype=np.int32)
    run("selfamd_side_exchange_scalar_2d", p(dev(b)), p(ext), p(si), N, nVar, nEl)
    np.testing.assert_array_equal(host(ext), oracle.side_exchange_scalar(N, mesh.sideInfo, b))
    bv = rng.normal(size=(nEl, 4, 1, N + 1, 2))
    extv = zeros(bv.shape)
    run("selfamd_side_exchange_vector_2d", p(dev(bv)), p(extv), p(si), N, 1, nEl)
    np.testing.assert_array_equal(host(extv), oracle.side_exchange_vector(N, mesh.sideInfo, bv))


@pytest.mark.parametrize("source", ["structured", "hopr_circle"])
def test_geometry_generation(source):
    import json
    import os
    if source == "structured":
        mesh = self_amd.Mesh2D.structured(7, 5, 0.0, 2.0, -1.0, 1.0)
    else:
        with open(os.path.join(os.path.dirname(__file__), "golden", source + ".json")) as f:
            fx = json.load(f)
        mesh = self_amd.Mesh2D.from_hopr(
            fx["Ngeo"], np.array(fx["SideInfo"]["data"], np.int32).reshape(fx["SideInfo"]["shape"]),
            np.array(fx["NodeCoords"]["data"]).reshape(fx["NodeCoords"]["shape"]),
            np.array(fx["GlobalNodeIDs"]["data"], np.int32))
    for N in (3, 7):
        g = self_amd.SEMQuad(self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS), mesh)
        og = oracle.Geometry(N, mesh.nodeCoords, mesh.nGeo)
        for k in ("x", "xb", "dxds", "dxdsb", "dsdx", "dsdxb", "nHatb", "nScaleb", "J", "Jb"):
            a, b = getattr(g, k).to_host(), getattr(og, k)
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
            assert err <= TOL, (k, N, err)
        assert abs(g.CovariantArcMin() - og.covariant_arc_min()) <= TOL * og.covariant_arc_min()


def test_mapped_operators(rng):
    N = 7
    mesh = self_amd.Mesh2D.structured(4, 4)
    og = oracle.Geometry(N, mesh.nodeCoords, 1)
    L = oracle.lagrange_init(N, oracle.GAUSS, N, oracle.GAUSS)
    nEl, nVar = mesh.nElem, 2
    f = rng.normal(size=(nEl, nVar, N + 1, N + 1))
    dsdx, J = dev(og.dsdx), dev(og.J)
    gr = zeros((nEl, nVar, N + 1, N + 1, 2))
    run("selfamd_gradient_sf_2d", p(dev(L["dMatrix"])), p(dsdx), p(J), p(dev(f)), p(gr), N, nVar, nEl)
    assert rel_err(host(gr), oracle.gradient_sf(N, L["dMatrix"], og.dsdx, og.J, f)).max() <= TOL
    fw = dev(f)
    run("selfamd_jacobian_weight_2d", p(J), p(fw), N, nVar, nEl)
    assert rel_err(host(fw), oracle.jacobian_weight(N, og.J, f)).max() <= TOL
    fv = rng.normal(size=(nEl, nVar, N + 1, N + 1, 2))
    fvd = dev(fv)
    run("selfamd_contravariant_projection_2d", p(dsdx), p(fvd), N, nVar, nEl)
    assert rel_err(host(fvd), oracle.contravariant_projection(N, og.dsdx, fv)).max() <= TOL
    G, S, D = (rng.normal(size=(nEl, 3, N + 1, N + 1)) for _ in range(3))
    Gd, Sd = dev(G), dev(S)
    run("selfamd_update_grk", p(Gd), p(Sd), p(dev(D)), -5.0 / 9.0, 15.0 / 16.0, 1e-3, N, 3, nEl)
    Gr = -5.0 / 9.0 * G + D
    np.testing.assert_allclose(host(Gd), Gr, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(host(Sd), S + (15.0 / 16.0 * 1e-3) * Gr, rtol=1e-15, atol=1e-15)
    out = zeros(G.shape)
    run("selfamd_calculate_dsdt", p(dev(G)), p(dev(S)), p(out), N, 3, nEl)
    np.testing.assert_array_equal(host(out), S - G)


def _physics_inputs(kind, bcid, periodic=False):
    """Oracle model after one CalculateTendency, with the physical flux recomputed (the
    tendency projects it in place): every register of the stage is populated."""
    c = Case(kind, 5, 4, 4, periodic=periodic, bcid=bcid, seed=5)
    oracle.lib().or_calculate_tendency(c.om._p)
    oracle.lib().or_flux(c.om._p)
    return c


@pytest.mark.parametrize("bcid", [self_amd.BC_NONORMALFLOW, self_amd.BC_RADIATION])
def test_sw_physics_operators(bcid):
    c = _physics_inputs("sw", bcid)
    om, N, nEl = c.om, c.N, c.mesh.nElem
    P = N + 1
    F = lambda n, s: om.field(n, s).copy()
    ish, bsh = om.interior_shape, om.boundary_shape
    sol, solB, solE = dev(F("sol", ish)), dev(F("solB", bsh)), dev(F("solE", bsh))
    vel, velB = zeros((nEl, P, P, 2)), zeros((nEl, 4, P, 2))
    # velocity from the boundary values as they stood before this stage's interpolation:
    # feed the oracle's velB source (solB of the set-solution call is the same data here)
    run("selfamd_sw2d_velocity", p(sol), p(solB), p(vel), p(velB), N, nEl)
    assert rel_err(host(vel)[None], F("vel", (nEl, P, P, 2))[None], 0).max() <= TOL
    si = dev(c.mesh.sideInfo, dtype=np.int32)
    velE = zeros((nEl, 4, P, 2))
    run("selfamd_side_exchange_vector_2d", p(velB), p(velE), p(si), N, 1, nEl)
    np.testing.assert_array_equal(host(velE), F("velE", (nEl, 4, P, 2)))
    ext = dev(F("solE", bsh) * (c.mesh.sideInfo[..., 2][:, :, None, None] != 0))
    g = dev(c.geom.nHatb)
    run("selfamd_sw2d_set_boundary_condition", p(solB), p(ext), p(dev(F("topoB", (nEl, 4, P)))), p(g), p(si), N, nEl)
    assert rel_err(host(ext), F("solE", bsh), 2).max() <= TOL
    src = zeros(ish)
    run("selfamd_sw2d_source", p(src), p(sol), p(dev(F("gradTopo", (nEl, P, P, 2)))), 1.0, 0.0, N, nEl)
    assert rel_err(host(src), F("source", ish)).max() <= TOL
    fl = zeros(ish + (2,))
    run("selfamd_sw2d_flux", p(fl), p(sol), p(vel), 1.0, N, nEl)
    assert rel_err(host(fl), F("flux", ish + (2,))).max() <= TOL
    fbn = zeros(bsh)
    run("selfamd_sw2d_riemann", p(fbn), p(solB), p(solE), p(dev(F("velB", (nEl, 4, P, 2)))),
        p(dev(F("velE", (nEl, 4, P, 2)))), p(g), p(dev(c.geom.nScaleb)), 1.0, N, nEl)
    assert rel_err(host(fbn), F("fluxBN", bsh), 2).max() <= TOL


def test_lsw_physics_operators():
    c = _physics_inputs("lsw", self_amd.BC_NONORMALFLOW)
    om, N, nEl = c.om, c.N, c.mesh.nElem
    P = N + 1
    F = lambda n, s: om.field(n, s).copy()
    ish, bsh = om.interior_shape, om.boundary_shape
    sol, solB = dev(F("sol", ish)), dev(F("solB", bsh))
    si = dev(c.mesh.sideInfo, dtype=np.int32)
    ext = dev(F("solE", bsh) * (c.mesh.sideInfo[..., 2][:, :, None, None] != 0))
    run("selfamd_lsw2d_set_boundary_condition", p(solB), p(ext), p(dev(c.geom.nHatb)), p(si), N, nEl)
    assert rel_err(host(ext), F("solE", bsh), 2).max() <= TOL
    fc = np.full((nEl, P, P), 0.7)
    om.set_coriolis(fc)
    oracle.lib().or_source(om._p)
    src = zeros(ish)
    run("selfamd_lsw2d_source", p(src), p(sol), p(dev(fc)), N, nEl)
    assert rel_err(host(src), F("source", ish)).max() <= TOL
    fl = zeros(ish + (2,))
    run("selfamd_lsw2d_flux", p(fl), p(sol), p(dev(F("topo", (nEl, P, P)))), 1.0, N, nEl)
    assert rel_err(host(fl), F("flux", ish + (2,))).max() <= TOL
    fbn = zeros(bsh)
    run("selfamd_lsw2d_riemann", p(fbn), p(solB), p(dev(F("solE", bsh))), p(dev(F("topoB", (nEl, 4, P)))),
        p(dev(c.geom.nHatb)), p(dev(c.geom.nScaleb)), 1.0, N, nEl)
    assert rel_err(host(fbn), F("fluxBN", bsh), 2).max() <= TOL


@pytest.mark.parametrize("bcid", [self_amd.BC_NONORMALFLOW, self_amd.BC_PRESCRIBED])
def test_euler_physics_operators(bcid):
    c = _physics_inputs("euler", bcid)
    om, N, nEl = c.om, c.N, c.mesh.nElem
    P = N + 1
    F = lambda n, s: om.field(n, s).copy()
    ish, bsh = om.interior_shape, om.boundary_shape
    sol = dev(F("sol", ish))
    for name, fld, args in (("selfamd_euler2d_diagnostics", "diag", (1.4, 1.0)),
                            ("selfamd_euler2d_primitive", "prim", (1.4,)),
                            ("selfamd_euler2d_entropy_vars", "ent", (1.4,))):
        out = zeros(ish)
        run(name, p(sol), p(out), *args, N, nEl)
        assert rel_err(host(out), F(fld, ish)).max() <= TOL, name
    si = dev(c.mesh.sideInfo, dtype=np.int32)
    mask = (c.mesh.sideInfo[..., 2] != 0)[:, :, None, None]
    solE, primE, diagE = dev(F("solE", bsh) * mask), dev(F("primE", bsh) * mask), dev(F("diagE", bsh) * mask)
    run("selfamd_euler2d_set_boundary_condition", p(dev(F("solB", bsh))), p(solE), p(dev(F("presSolB", bsh))),
        p(dev(F("primB", bsh))), p(primE), p(dev(F("presPrimB", bsh))), p(dev(F("diagB", bsh))), p(diagE),
        p(dev(F("presDiagB", bsh))), p(dev(c.geom.nHatb)), p(si), N, nEl)
    for a, k in ((solE, "solE"), (primE, "primE"), (diagE, "diagE")):
        assert rel_err(host(a), F(k, bsh), 2).max() <= TOL, k
    src = zeros(ish)
    run("selfamd_euler2d_source", p(src), p(sol), p(dev(F("env", (nEl, 2, P, P)))),
        p(dev(F("envGrad", (nEl, 2, P, P, 2)))), N, nEl)
    assert rel_err(host(src), F("source", ish)).max() <= TOL
    fl = zeros(ish + (2,))
    run("selfamd_euler2d_flux", p(fl), p(sol), p(dev(F("prim", ish))), p(dev(F("diag", ish))), N, nEl)
    assert rel_err(host(fl), F("flux", ish + (2,))).max() <= TOL
    fbn = zeros(bsh)
    run("selfamd_euler2d_riemann", p(fbn), p(dev(F("solB", bsh))), p(dev(F("solE", bsh))), p(dev(F("primB", bsh))),
        p(dev(F("primE", bsh))), p(dev(F("diagB", bsh))), p(dev(F("diagE", bsh))), p(dev(c.geom.nHatb)),
        p(dev(c.geom.nScaleb)), N, nEl)
    assert rel_err(host(fbn), F("fluxBN", bsh), 2).max() <= TOL
