"""Pins the oracle's Riemann solvers and wall boundary condition with what the mathematics
fixes, independently of the formulas restated in the oracle:

* consistency: F*(q, q, n) = F(q).n nScale, with F the physical flux written from the PDE;
* conservation: F*(qL, qR, n) = -F*(qR, qL, -n);
* dissipation: the coefficient alpha recovered from F* - (F(qL) + F(qR)).n nScale / 2 is
  compared with the eigenvalues of finite-difference Jacobians d(F.n)/dq of the physical
  flux: the SW Rusanov alpha (src/models/SELF_ShallowWater.f90:512-513) is the larger of the
  two sides' spectral radii, the Euler NaiveLLF alpha
  (src/models/SELF_CompressibleIdealGas2D.f90:1531) is max|u.n| + max c with u.n and c read
  off the eigenvalues (u.n - c, u.n, u.n, u.n + c); states are chosen so that the side with
  the larger |u.n| has the smaller c, where other combinations differ;
* the LSW flux (src/models/SELF_LinearShallowWater.f90:543-549) is the exact Godunov flux of
  the linear system, A+ qL + A- qR from the eigen-decomposition of A = d(F.n)/dq;
* no-normal-flow walls (src/models/SELF_ShallowWater.f90:324-325,
  src/models/SELF_LinearShallowWater.f90:336-337,
  src/models/SELF_CompressibleIdealGas2D.f90:1260-1274): the exterior state is the interior
  state with the velocity reflected, u - 2 (u.n) n, for arbitrary unit normals, including
  the curved Circle mesh's;
* walled meshes conserve mass to round-off for all three models;
* the whole tendency of a smooth state converges to the PDE's -div F + S (volume fluxes,
  sources and their assembly).
"""
import json
import os

import numpy as np
import pytest

import oracle
import self_amd
from tests.cases import Case

GAMMA = 1.4
G = 1.0


# ------------------------------------------------------------------ physical fluxes F(q).n
def flux_sw(q, n):
    Hu, Hv, H = q
    un = (Hu * n[0] + Hv * n[1]) / H
    p = 0.5 * G * H * H
    return np.array([Hu * un + p * n[0], Hv * un + p * n[1], H * un])


def flux_euler(q, n):
    ru, rv, r, rE = q
    un = (ru * n[0] + rv * n[1]) / r
    p = (GAMMA - 1.0) * (rE - 0.5 * (ru * ru + rv * rv) / r)
    return np.array([ru * un + p * n[0], rv * un + p * n[1], r * un, (rE + p) * un])


def flux_lsw(q, n, H):
    u, v, eta = q
    return np.array([G * eta * n[0], G * eta * n[1], H * (u * n[0] + v * n[1])])


def jacobian(f, q, step=1e-6):
    """Central finite-difference Jacobian df/dq (exact for a linear f with step 1)."""
    q = np.asarray(q, float)
    k = len(q)
    A = np.zeros((k, k))
    for j in range(k):
        h = step * max(1.0, abs(q[j]))
        e = np.zeros(k)
        e[j] = h
        A[:, j] = (f(q + e) - f(q - e)) / (2 * h)
    return A


def eig_real(A):
    lam = np.linalg.eigvals(A)
    assert np.abs(lam.imag).max() < 1e-6 * max(1.0, np.abs(lam.real).max())
    return np.sort(lam.real)


# ------------------------------------------------------------------ face fixtures
class Faces:
    """A small oracle model whose face arrays (boundary, extBoundary, normals) are set
    directly; every face node is an independent Riemann problem."""

    def __init__(self, kind, N=3, n=2, seed=0):
        self.kind = kind
        self.case = Case(kind, N, n, n, periodic=True, generic=False)
        om = self.case.om
        self.om, self.N, self.nEl, self.nVar = om, N, om.nElem, om.nVar
        self.P = N + 1
        self.nface = self.nEl * 4 * self.P
        self.rng = np.random.default_rng(seed)

    def boundary(self, name, nvar):
        return self.om.field(name, (self.nEl, 4, nvar, self.P))

    def set_normals(self, nx, ny, ns):
        g = self.case.geom
        g.nHatb[..., 0] = nx.reshape(self.nEl, 4, self.P)
        g.nHatb[..., 1] = ny.reshape(self.nEl, 4, self.P)
        g.nScaleb[...] = ns.reshape(self.nEl, 4, self.P)

    def put(self, name, vals):
        """vals: (nface, k) in face-node order (e, s, i) -> boundary array (e, s, v, i)."""
        k = vals.shape[1]
        self.boundary(name, k)[...] = vals.reshape(self.nEl, 4, self.P, k).transpose(0, 1, 3, 2)

    def get(self, name, k):
        return self.boundary(name, k).transpose(0, 1, 3, 2).reshape(self.nface, k).copy()

    def set_states(self, qL, qR, extra_H=None):
        """qL, qR: (nface, nVar) conservative states (LSW: u, v, eta) on the two sides."""
        self.put("solB", qL)
        self.put("solE", qR)
        if self.kind == "sw":
            vel = lambda q: q[:, :2] / q[:, 2:3]
            self.om.field("velB", (self.nEl, 4, self.P, 2))[...] = vel(qL).reshape(self.nEl, 4, self.P, 2)
            self.om.field("velE", (self.nEl, 4, self.P, 2))[...] = vel(qR).reshape(self.nEl, 4, self.P, 2)
        elif self.kind == "lsw":
            self.om.field("topoB", (self.nEl, 4, self.P))[...] = extra_H.reshape(self.nEl, 4, self.P)
        else:
            for q, sfx in ((qL, "B"), (qR, "E")):
                ru, rv, r, rE = q.T
                u, v = ru / r, rv / r
                ke = 0.5 * (ru * u + rv * v)
                p = (GAMMA - 1.0) * (rE - ke)
                self.put("prim" + sfx, np.stack([u, v, r, p], 1))
                self.put("diag" + sfx, np.stack([ke, rE + p, np.sqrt(GAMMA * p / r), p / r], 1))

    def riemann(self):
        oracle.lib().or_riemann(self.om._p)
        return self.get("fluxBN", self.nVar)


def random_normals(rng, n):
    th = rng.uniform(0, 2 * np.pi, n)
    return np.cos(th), np.sin(th), rng.uniform(0.5, 2.0, n)


def sw_states(rng, n):
    """Left sides: large |u.n|, shallow (small c); right sides: slow, deep (large c); the
    spectral radii |u.n| + c of the two sides then differ from max|u.n| + max c."""
    HL, HR = rng.uniform(0.3, 0.6, n), rng.uniform(1.5, 2.5, n)
    uL, vL = rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n)
    uR, vR = rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)
    qL = np.stack([HL * uL, HL * vL, HL], 1)
    qR = np.stack([HR * uR, HR * vR, HR], 1)
    return qL, qR


def euler_states(rng, n):
    rL, rR = rng.uniform(0.8, 1.2, n), rng.uniform(0.8, 1.2, n)
    pL, pR = rng.uniform(0.2, 0.4, n), rng.uniform(2.0, 3.0, n)
    uL, vL = rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n)
    uR, vR = rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)
    mk = lambda r, u, v, p: np.stack([r * u, r * v, r, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)], 1)
    return mk(rL, uL, vL, pL), mk(rR, uR, vR, pR)


def lsw_states(rng, n):
    qL = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-0.5, 0.5, n)], 1)
    qR = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-0.5, 0.5, n)], 1)
    return qL, qR, rng.uniform(0.5, 2.0, n)


def physical_flux(kind, q, n, H=None):
    if kind == "sw":
        return flux_sw(q, n)
    if kind == "euler":
        return flux_euler(q, n)
    return flux_lsw(q, n, H)


def make(kind, seed=0):
    f = Faces(kind, seed=seed)
    nx, ny, ns = random_normals(f.rng, f.nface)
    H = None
    if kind == "sw":
        qL, qR = sw_states(f.rng, f.nface)
    elif kind == "euler":
        qL, qR = euler_states(f.rng, f.nface)
    else:
        qL, qR, H = lsw_states(f.rng, f.nface)
    return f, (nx, ny, ns), qL, qR, H


# ------------------------------------------------------------------ Riemann solvers
@pytest.mark.parametrize("kind", ["sw", "lsw", "euler"])
def test_riemann_consistency(kind):
    """F*(q, q, n) = F(q).n nScale."""
    f, (nx, ny, ns), qL, _, H = make(kind, seed=1)
    f.set_normals(nx, ny, ns)
    f.set_states(qL, qL, H)
    Fs = f.riemann()
    for k in range(f.nface):
        want = physical_flux(kind, qL[k], (nx[k], ny[k]), None if H is None else H[k]) * ns[k]
        np.testing.assert_allclose(Fs[k], want, rtol=1e-13, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("kind", ["sw", "lsw", "euler"])
def test_riemann_conservation(kind):
    """F*(qL, qR, n) = -F*(qR, qL, -n): the two sides of a face see opposite fluxes."""
    f, (nx, ny, ns), qL, qR, H = make(kind, seed=2)
    f.set_normals(nx, ny, ns)
    f.set_states(qL, qR, H)
    a = f.riemann()
    f.set_normals(-nx, -ny, ns)
    f.set_states(qR, qL, H)
    b = f.riemann()
    np.testing.assert_allclose(a, -b, rtol=1e-14, atol=1e-15 * np.abs(a).max())


def _alpha(f, kind, Fs, qL, qR, nx, ny, ns, H=None):
    """alpha per face node from F* = (F(qL).n + F(qR).n + alpha (qL - qR)) nScale / 2,
    checked to be the same in every component."""
    out = np.zeros(f.nface)
    for k in range(f.nface):
        n = (nx[k], ny[k])
        avg = 0.5 * (physical_flux(kind, qL[k], n) + physical_flux(kind, qR[k], n))
        a = 2.0 * (Fs[k] / ns[k] - avg) / (qL[k] - qR[k])
        np.testing.assert_allclose(a, a[0], rtol=1e-9)
        out[k] = a[0]
    return out


def test_sw_rusanov_alpha_is_max_spectral_radius():
    f, (nx, ny, ns), qL, qR, _ = make("sw", seed=3)
    f.set_normals(nx, ny, ns)
    f.set_states(qL, qR)
    alpha = _alpha(f, "sw", f.riemann(), qL, qR, nx, ny, ns)
    differs = 0
    for k in range(f.nface):
        n = (nx[k], ny[k])
        lamL = eig_real(jacobian(lambda q: flux_sw(q, n), qL[k]))
        lamR = eig_real(jacobian(lambda q: flux_sw(q, n), qR[k]))
        rho = max(np.abs(lamL).max(), np.abs(lamR).max())
        assert abs(alpha[k] - rho) <= 1e-7 * rho, (k, alpha[k], rho)
        # the same states with the larger |u.n| and the larger c taken from different sides
        unc = [((l.max() + l.min()) / 2, (l.max() - l.min()) / 2) for l in (lamL, lamR)]
        mixed = max(abs(unc[0][0]), abs(unc[1][0])) + max(unc[0][1], unc[1][1])
        differs += mixed > rho * (1 + 1e-3)
    assert differs > f.nface // 2  # the fixture separates the two candidate formulas


def test_euler_llf_alpha_is_max_speed_plus_max_sound_speed():
    f, (nx, ny, ns), qL, qR, _ = make("euler", seed=4)
    f.set_normals(nx, ny, ns)
    f.set_states(qL, qR)
    alpha = _alpha(f, "euler", f.riemann(), qL, qR, nx, ny, ns)
    for k in range(f.nface):
        n = (nx[k], ny[k])
        sides = []
        for q in (qL[k], qR[k]):
            lam = eig_real(jacobian(lambda s: flux_euler(s, n), q))
            # eigenvalues u.n - c, u.n, u.n, u.n + c
            assert abs(lam[1] - lam[2]) <= 1e-6 * max(1.0, np.abs(lam).max())
            sides.append(((lam[-1] + lam[0]) / 2, (lam[-1] - lam[0]) / 2, np.abs(lam).max()))
        want = max(abs(sides[0][0]), abs(sides[1][0])) + max(sides[0][1], sides[1][1])
        assert abs(alpha[k] - want) <= 1e-7 * want, (k, alpha[k], want)
        assert alpha[k] >= max(sides[0][2], sides[1][2]) * (1 - 1e-9)


def test_euler_sound_speed_matches_flux_jacobian():
    """CalculateDiagnostics' sound speed (src/models/SELF_CompressibleIdealGas2D.f90:984-1127)
    is half the spread of the flux Jacobian's eigenvalues."""
    c = Case("euler", 3, 2, 2, periodic=True, generic=True, seed=7)
    om = c.om
    oracle.lib().or_euler_diagnostics(om._p)
    sol = om.solution
    diag = om.field("diag", om.interior_shape)
    rng = np.random.default_rng(0)
    for _ in range(20):
        e, i, j = rng.integers(om.nElem), rng.integers(4), rng.integers(4)
        q = sol[e, :, j, i]
        th = rng.uniform(0, 2 * np.pi)
        lam = eig_real(jacobian(lambda s: flux_euler(s, (np.cos(th), np.sin(th))), q))
        cs = (lam[-1] - lam[0]) / 2
        assert abs(diag[e, 2, j, i] - cs) <= 1e-7 * cs


def test_lsw_flux_is_exact_godunov():
    f, (nx, ny, ns), qL, qR, H = make("lsw", seed=5)
    f.set_normals(nx, ny, ns)
    f.set_states(qL, qR, H)
    Fs = f.riemann()
    for k in range(f.nface):
        n = (nx[k], ny[k])
        A = jacobian(lambda q: flux_lsw(q, n, H[k]), np.zeros(3), step=1.0)
        lam, R = np.linalg.eig(A)
        lam, R = lam.real, R.real
        Ri = np.linalg.inv(R)
        Ap = R @ np.diag(np.maximum(lam, 0)) @ Ri
        Am = R @ np.diag(np.minimum(lam, 0)) @ Ri
        want = (Ap @ qL[k] + Am @ qR[k]) * ns[k]
        np.testing.assert_allclose(Fs[k], want, rtol=1e-13, atol=1e-14 * np.abs(want).max())


# ------------------------------------------------------------------ no-normal-flow walls
def reflect(u, v, nx, ny):
    un = u * nx + v * ny
    return u - 2 * un * nx, v - 2 * un * ny


def _circle_mesh():
    with open(os.path.join(os.path.dirname(__file__), "golden", "hopr_circle.json")) as fh:
        fx = json.load(fh)
    mesh = self_amd.Mesh2D.from_hopr(
        fx["Ngeo"], np.array(fx["SideInfo"]["data"], np.int32).reshape(fx["SideInfo"]["shape"]),
        np.array(fx["NodeCoords"]["data"]).reshape(fx["NodeCoords"]["shape"]),
        np.array(fx["GlobalNodeIDs"]["data"], np.int32))
    mesh.ResetBoundaryConditionType(self_amd.BC_NONORMALFLOW)
    return mesh


@pytest.mark.parametrize("kind", ["sw", "lsw", "euler"])
@pytest.mark.parametrize("mesh", ["square-random-normals", "circle"])
def test_wall_exterior_state_is_reflection(kind, mesh):
    if mesh == "circle":
        c = Case(kind, 5, 0, 0, hopr=_circle_mesh(), seed=3)
    else:
        c = Case(kind, 4, 2, 2, bcid=oracle.BC_NONORMALFLOW, seed=3)
        rng = np.random.default_rng(11)
        nx, ny, _ = random_normals(rng, c.geom.nScaleb.size)
        c.geom.nHatb[..., 0] = nx.reshape(c.geom.nScaleb.shape)
        c.geom.nHatb[..., 1] = ny.reshape(c.geom.nScaleb.shape)
    om = c.om
    oracle.lib().or_pre_tendency(om._p)
    oracle.lib().or_set_boundary_condition(om._p)
    wall = (c.mesh.sideInfo[:, :, 2] == 0)
    assert wall.any() and (c.mesh.sideInfo[:, :, 4][wall] == oracle.BC_NONORMALFLOW).all()
    nh = c.geom.nHatb
    bshape = om.boundary_shape
    B, E = om.field("solB", bshape), om.field("solE", bshape)
    nx, ny = nh[..., 0][wall], nh[..., 1][wall]
    # the circle's wall normals are not axis-aligned
    assert np.abs(nx * ny).max() > 0.2
    ru, rv = reflect(B[:, :, 0][wall], B[:, :, 1][wall], nx, ny)
    tol = 1e-15 * max(1.0, np.abs(B).max())
    assert np.abs(E[:, :, 0][wall] - ru).max() <= 4 * tol
    assert np.abs(E[:, :, 1][wall] - rv).max() <= 4 * tol
    for v in range(2, om.nVar):
        assert np.array_equal(E[:, :, v][wall], B[:, :, v][wall])
    if kind == "euler":
        PB, PE = om.field("primB", bshape), om.field("primE", bshape)
        pu, pv = reflect(PB[:, :, 0][wall], PB[:, :, 1][wall], nx, ny)
        assert np.abs(PE[:, :, 0][wall] - pu).max() <= 4e-15 * max(1.0, np.abs(PB).max())
        assert np.abs(PE[:, :, 1][wall] - pv).max() <= 4e-15 * max(1.0, np.abs(PB).max())
        for v in (2, 3):
            assert np.array_equal(PE[:, :, v][wall], PB[:, :, v][wall])
        DB, DE = om.field("diagB", bshape), om.field("diagE", bshape)
        assert np.array_equal(DE[wall], DB[wall])


@pytest.mark.parametrize("kind,mesh", [("sw", "square"), ("lsw", "square"), ("euler", "square"),
                                       ("sw", "circle"), ("lsw", "circle"), ("euler", "circle")])
def test_mass_conserved_with_walls(kind, mesh):
    """Mass (H, eta, rho) of a walled domain is conserved: the semi-discrete rate is zero to
    round-off and five RK3 steps keep it.  The wall's mass flux vanishes only if the exterior
    normal velocity is minus the interior one and the exterior thickness/density equals the
    interior one."""
    if mesh == "circle":
        # N > Ngeo = 6: both sides of a curved face interpolate the same geometry exactly, so
        # their normals and scale factors agree to round-off
        c = Case(kind, 7, 0, 0, hopr=_circle_mesh(), seed=4)
    else:
        c = Case(kind, 5, 4, 4, bcid=oracle.BC_NONORMALFLOW, seed=4)
    if kind == "lsw":
        # a continuous bathymetry: the two sides of a face then see the same wave speed
        # (each side evaluates the flux with its own H trace)
        c.fields["H"] = np.full_like(c.fields["H"], 1.3)
        c.apply_oracle(c.om)
    var = 2
    w = c.weights
    d = c.om.tendency()
    rate = np.sum(d[:, var] * c.geom.J * w[None, :, None] * w[None, None, :])
    scale = np.sum(np.abs(d[:, var]) * np.abs(c.geom.J) * w[None, :, None] * w[None, None, :])
    assert abs(rate) <= 1e-13 * scale, (rate, scale)
    m0 = c.om.mass(var)
    c.om.rk3_steps(5, c.dt())
    assert abs(c.om.mass(var) - m0) <= 1e-13 * abs(m0)


# ------------------------------------------------------------------ whole tendency vs the PDE
TWO_PI = 2 * np.pi


def _smooth_fields(kind, X, Y):
    """Smooth periodic states and static fields on the unit square."""
    sx, cx, sy, cy = np.sin(TWO_PI * X), np.cos(TWO_PI * X), np.sin(TWO_PI * Y), np.cos(TWO_PI * Y)
    if kind == "sw":
        H = 1.0 + 0.1 * sx * cy
        u, v = 0.3 + 0.1 * cy, -0.2 + 0.1 * sx * sy
        return np.stack([H * u, H * v, H]), {"topo": 0.05 * cx * sy}
    if kind == "lsw":
        q = np.stack([0.1 * sy, 0.2 * cx * sy, 0.05 * sx * cy])
        return q, {"H": 1.0 + 0.2 * cx * cy, "f": 0.5 + 0.3 * sx}
    r = 1.0 + 0.1 * sx * sy
    u, v, p = 0.2 + 0.1 * cy, -0.1 + 0.1 * sx, 1.0 + 0.1 * cx * cy
    q = np.stack([r * u, r * v, r, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)])
    return q, {"phi": 0.3 * sx + 0.2 * cy, "drag": 0.1 + 0.05 * sy}


def _pde_flux(kind, q, fields):
    """Physical flux tensor (Fx, Fy), each (nVar, ...)."""
    if kind == "sw":
        Hu, Hv, H = q
        p = 0.5 * G * H * H
        return (np.stack([Hu * Hu / H + p, Hv * Hu / H, Hu]), np.stack([Hu * Hv / H, Hv * Hv / H + p, Hv]))
    if kind == "lsw":
        u, v, eta = q
        H = fields["H"]
        z = np.zeros_like(eta)
        return np.stack([G * eta, z, H * u]), np.stack([z, G * eta, H * v])
    ru, rv, r, rE = q
    u, v = ru / r, rv / r
    p = (GAMMA - 1) * (rE - 0.5 * r * (u * u + v * v))
    return (np.stack([ru * u + p, rv * u, ru, (rE + p) * u]), np.stack([ru * v, rv * v + p, rv, (rE + p) * v]))


@pytest.mark.parametrize("kind", ["sw", "lsw", "euler"])
def test_tendency_converges_to_the_pde(kind):
    """The weak-form DG tendency of a smooth periodic state at N = 7 equals -div F(q) + S(q)
    of the PDE to discretisation accuracy, with F the physical flux and div F from central
    differences of F(q(x, y)).  S is the model's source: SW (-f Hv + g H db/dx,
    f Hu + g H db/dy, 0) (src/models/SELF_ShallowWater.f90:345-384), LSW (f v, -f u, 0)
    (src/models/SELF_LinearShallowWater.f90:413-446), Euler (-rho dphi/dx - Cd rho u,
    -rho dphi/dy - Cd rho v, 0, -rho u dphi/dx - rho u dphi/dy)
    (src/models/SELF_CompressibleIdealGas2D.f90:1325-1366).  This pins the volume fluxes, the
    sources and their assembly; a wrong velocity, pressure or enthalpy term moves the result
    by O(1)."""
    N, n = 7, 6
    c = Case(kind, N, n, n, periodic=True, generic=False)
    X, Y = c.geom.x[..., 0], c.geom.x[..., 1]
    q, fields = _smooth_fields(kind, X, Y)
    c.sol = np.ascontiguousarray(q.transpose(1, 0, 2, 3))
    c.fields.update(fields)
    c.apply_oracle(c.om)
    got = c.om.tendency()

    h = 1e-5
    F = lambda x, y: _pde_flux(kind, *[_smooth_fields(kind, x, y)[k] for k in (0, 1)])
    dFx = (F(X + h, Y)[0] - F(X - h, Y)[0]) / (2 * h)
    dFy = (F(X, Y + h)[1] - F(X, Y - h)[1]) / (2 * h)
    want = -(dFx + dFy)
    if kind == "sw":
        b = lambda x, y: _smooth_fields("sw", x, y)[1]["topo"]
        bx, by = (b(X + h, Y) - b(X - h, Y)) / (2 * h), (b(X, Y + h) - b(X, Y - h)) / (2 * h)
        want[0] += G * q[2] * bx
        want[1] += G * q[2] * by
    elif kind == "lsw":
        want[0] += fields["f"] * q[1]
        want[1] -= fields["f"] * q[0]
    else:
        ph = lambda x, y: _smooth_fields("euler", x, y)[1]["phi"]
        gx, gy = (ph(X + h, Y) - ph(X - h, Y)) / (2 * h), (ph(X, Y + h) - ph(X, Y - h)) / (2 * h)
        cd = fields["drag"]
        want[0] += -q[2] * gx - cd * q[0]
        want[1] += -q[2] * gy - cd * q[1]
        want[3] += -q[0] * gx - q[0] * gy
    want = want.transpose(1, 0, 2, 3)
    for v in range(c.om.nVar):
        err = np.abs(got[:, v] - want[:, v]).max() / np.abs(want[:, v]).max()
        assert err < 1e-5, (kind, v, err)
