"""Shared test cases: the same mesh, geometry and seeded state on the oracle and on the GPU."""
import numpy as np

import oracle
import self_amd
from self_amd import workloads

KIND = {"lsw": oracle.LSW, "sw": oracle.SW, "euler": oracle.EULER}


def rel_err(a, b, axis_var=1):
    """Per-variable relative max-norm max|a-b|/max|b| (variables along axis 1).  A variable
    whose reference magnitude is below 1e-10 of the largest one is measured against the
    largest one instead (e.g. a tendency that is zero up to round-off)."""
    a = np.asarray(a)
    b = np.asarray(b)
    nv = b.shape[axis_var]
    ref = [np.abs(np.take(b, v, axis=axis_var)).max() for v in range(nv)]
    big = max(ref) if max(ref) > 0 else 1.0
    out = []
    for v in range(nv):
        d = np.abs(np.take(a, v, axis=axis_var) - np.take(b, v, axis=axis_var)).max()
        out.append(d / (ref[v] if ref[v] > 1e-10 * big else big))
    return np.array(out)


class Case:
    """Mesh + oracle geometry + oracle model with a seeded state for one model kind."""

    def __init__(self, kind, N, nx, ny, periodic=False, bcid=self_amd.BC_NONORMALFLOW, seed=1, hopr=None):
        self.kind, self.N = kind, N
        if hopr is not None:
            self.mesh = hopr
        else:
            self.mesh = self_amd.Mesh2D.structured(nx, ny, periodic=(periodic, periodic), bcid=bcid)
        lag = oracle.lagrange_init(N, oracle.GAUSS, N, oracle.GAUSS)
        self.nodes, self.weights = lag["controlPoints"], lag["qWeights"]
        self.geom = oracle.Geometry(N, self.mesh.nodeCoords, self.mesh.nGeo)
        x = self.geom.x
        self.fields = {}
        if kind == "sw":
            self.sol, self.fields["topo"] = workloads.nonlinear_sw(x, self.nodes, self.weights, seed=seed)
        elif kind == "lsw":
            self.sol = workloads.gravity_wave_lsw(x, seed=seed)
            self.fields["H"] = np.ones(x.shape[:3])
            self.fields["f"] = np.zeros(x.shape[:3])
        else:
            self.sol, self.fields["phi"] = workloads.euler_hydrostatic(x, seed=seed)
            self.fields["drag"] = np.zeros(x.shape[:3])
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

    def gpu_model(self):
        interp = self_amd.Lagrange(self.N, self_amd.GAUSS, self.N, self_amd.GAUSS)
        geom = self_amd.SEMQuad(interp, self.mesh)
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
        m.SetSolution(self.sol)
        return m, geom

    def dt(self, cfl=0.5):
        dx = self.geom.covariant_arc_min()
        if self.kind == "euler":
            c = np.sqrt(1.4 * (self.sol[:, 3] * 0.4) / self.sol[:, 2]).max()
            return workloads.cfl_dt(dx, c, cfl)
        if self.kind == "sw":
            H = self.sol[:, 2]
            speed = (np.sqrt(self.sol[:, 0] ** 2 + self.sol[:, 1] ** 2) / H + np.sqrt(H)).max()
            return workloads.cfl_dt(dx, speed, cfl)
        return workloads.cfl_dt(dx, 1.0, cfl)
