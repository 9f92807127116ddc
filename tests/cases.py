"""Shared test cases: the same mesh, geometry and seeded state on the oracle and on the GPU."""
import numpy as np

import oracle
import self_amd
from self_amd import workloads

KIND = {"lsw": oracle.LSW, "sw": oracle.SW, "euler": oracle.EULER}


def rel_err(a, b, axis_var=1):
    """Per-variable relative max-norm max|a_v - b_v| / max|b_v| (variables along axis_var).
    Every variable is judged against its own magnitude; a variable whose reference is
    identically zero must match exactly (its entry is 0 if a_v is zero too, else inf)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for v in range(b.shape[axis_var]):
        bv, av = np.take(b, v, axis=axis_var), np.take(a, v, axis=axis_var)
        ref = np.abs(bv).max()
        d = np.abs(av - bv).max()
        out.append(d / ref if ref > 0 else (0.0 if d == 0 else np.inf))
    return np.array(out)


class DeviceGeometry:
    """The oracle's geometry arrays uploaded to the device, in SEMQuad's shapes: the GPU engine
    then sees exactly the metric values the oracle uses (identical inputs)."""

    def __init__(self, og, N):
        self.interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
        for k in ("x", "dsdx", "J", "nHatb", "nScaleb"):
            setattr(self, k, self_amd.DeviceArray.from_host(np.ascontiguousarray(getattr(og, k))))


def flatten_metrics(og):
    """Metric terms made element-constant (node (0,0) and side node 0), in place: the affine
    representation the fused engine uses by default on affine meshes."""
    og.dsdx[...] = og.dsdx[:, :1, :1]
    og.J[...] = og.J[:, :1, :1]
    og.nHatb[...] = og.nHatb[:, :, :1]
    og.nScaleb[...] = og.nScaleb[:, :, :1]


class Case:
    """Mesh + oracle geometry + oracle model with a seeded state for one model kind.

    metrics: "nodal" keeps the per-node metric terms; "affine" flattens them to one value per
    element on both sides.  generic: adds seeded smooth velocity perturbations so that no
    variable's tendency vanishes by symmetry (a relative error of a round-off-sized
    reference value measures nothing)."""

    def __init__(self, kind, N, nx, ny, periodic=False, bcid=self_amd.BC_NONORMALFLOW, seed=1, hopr=None,
                 metrics="nodal", generic=True):
        self.kind, self.N, self.metrics = kind, N, metrics
        if hopr is not None:
            self.mesh = hopr
        else:
            self.mesh = self_amd.Mesh2D.structured(nx, ny, periodic=(periodic, periodic), bcid=bcid)
        lag = oracle.lagrange_init(N, oracle.GAUSS, N, oracle.GAUSS)
        self.nodes, self.weights = lag["controlPoints"], lag["qWeights"]
        self.geom = oracle.Geometry(N, self.mesh.nodeCoords, self.mesh.nGeo)
        if metrics == "affine":
            flatten_metrics(self.geom)
        x = self.geom.x
        rng = np.random.default_rng(seed + 101)
        smooth = lambda amp: amp * workloads.modal_filter(rng.normal(size=x.shape[:3]), self.nodes, self.weights, 2)
        self.fields = {}
        if kind == "sw":
            self.sol, self.fields["topo"] = workloads.nonlinear_sw(x, self.nodes, self.weights, seed=seed)
        elif kind == "lsw":
            self.sol = workloads.gravity_wave_lsw(x, seed=seed)
            if generic:
                self.sol[:, 0] += smooth(0.01)
                self.sol[:, 1] += smooth(0.01)
            self.fields["H"] = 1.0 + smooth(0.05) if generic else np.ones(x.shape[:3])
            self.fields["f"] = smooth(1.0) if generic else np.zeros(x.shape[:3])
        else:
            self.sol, self.fields["phi"] = workloads.euler_hydrostatic(x, seed=seed)
            if generic:
                u, v = smooth(0.01), smooth(0.01)
                rho = self.sol[:, 2]
                self.sol[:, 0] = rho * u
                self.sol[:, 1] = rho * v
                self.sol[:, 3] += 0.5 * rho * (u * u + v * v)
            self.fields["drag"] = 0.1 + smooth(0.05) if generic else np.zeros(x.shape[:3])
        self.sol = np.ascontiguousarray(self.sol)
        self.om = oracle.Model(KIND[kind], N, self.mesh.sideInfo, self.geom)
        self.apply_oracle(self.om)

    def apply_oracle(self, om):
        if self.kind == "sw":
            om.set_topography(self.fields["topo"])
        elif self.kind == "lsw":
            om.set_bathymetry(self.fields["H"])
            om.set_coriolis(self.fields["f"])
        else:
            om.set_gravity(self.fields["phi"])
            om.set_drag(self.fields["drag"])
        om.set_solution(self.sol)
        if self.kind == "euler":
            om.set_prescribed()

    def gpu_model(self, geometry="oracle"):
        """GPU model on the same state.  geometry="oracle": the oracle's metric arrays uploaded
        (identical inputs); "device": metrics generated on the GPU by SEMQuad."""
        if geometry == "oracle":
            geom = DeviceGeometry(self.geom, self.N)
        else:
            geom = self_amd.SEMQuad(self_amd.Lagrange(self.N, self_amd.GAUSS, self.N, self_amd.GAUSS), self.mesh)
        if self.kind == "sw":
            m = self_amd.ShallowWater(self.mesh, geom)
            m.SetTopography(self.fields["topo"])
        elif self.kind == "lsw":
            m = self_amd.LinearShallowWater(self.mesh, geom)
            m.SetBathymetry(self.fields["H"])
            m.SetCoriolis(self.fields["f"])
        else:
            m = self_amd.CompressibleIdealGas2D(self.mesh, geom)
            m.SetGravity(self.fields["phi"])
            m.SetDrag(self.fields["drag"])
        m.SetNodalMetrics(self.metrics == "nodal")
        m.SetSolution(self.sol)
        return m, geom

    def dt(self, cfl=0.5):
        dx = self.geom.covariant_arc_min()
        if self.kind == "euler":
            rho = self.sol[:, 2]
            ke = 0.5 * (self.sol[:, 0] ** 2 + self.sol[:, 1] ** 2) / rho
            p = 0.4 * (self.sol[:, 3] - ke)
            c = (np.sqrt(1.4 * p / rho) + np.sqrt(2 * ke / rho)).max()
            return workloads.cfl_dt(dx, c, cfl)
        if self.kind == "sw":
            H = self.sol[:, 2]
            speed = (np.sqrt(self.sol[:, 0] ** 2 + self.sol[:, 1] ** 2) / H + np.sqrt(H)).max()
            return workloads.cfl_dt(dx, speed, cfl)
        return workloads.cfl_dt(dx, np.sqrt(self.fields["H"].max()), cfl)
