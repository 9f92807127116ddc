"""The Fortran host layer (SELF-style modules compiled with amdflang over the C ABI): the
dg2d_run driver builds Lagrange, Mesh2D, SEMQuad and the model through the Fortran API,
computes a tendency and RK3 steps on the GPU; both must match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from tests.cases import Case, rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "self_amd", "fortran", "bin", "dg2d_run")
TOL = 1e-12


@pytest.mark.parametrize("kind,model,N,n,periodic,bcid", [("sw", 2, 7, 6, True, 102), ("lsw", 1, 5, 5, False, 102),
                                                        ("euler", 3, 4, 4, False, 102)])
def test_fortran_driver_parity(kind, model, N, n, periodic, bcid, tmp_path):
    assert os.path.exists(DRIVER), "build the Fortran layer first (make -C self_amd/fortran)"
    c = Case(kind, N, n, n, periodic=periodic, bcid=bcid, seed=8, metrics="affine")
    prefix = str(tmp_path / kind)
    # C-order [e, v, j, i] is the Fortran interior(0:N,0:N,nVar,nEl) memory order
    c.sol.astype(np.float64).tofile(prefix + "_sol.bin")
    f1, f2 = {"sw": ("topo", None), "lsw": ("H", "f"), "euler": ("phi", "drag")}[kind]
    np.ascontiguousarray(c.fields[f1]).tofile(prefix + "_field1.bin")
    (np.ascontiguousarray(c.fields[f2]) if f2 else np.zeros_like(c.fields[f1])).tofile(prefix + "_field2.bin")
    dt, steps = c.dt(), 6
    par = tmp_path / "run.nml"
    par.write_text(f"&run model={model}, N={N}, nx={n}, ny={n}, periodic={'.true.' if periodic else '.false.'}, "
                   f"bcid={bcid}, nSteps={steps}, dt={float(dt)!r}, nodal=.false., prefix='{prefix}' /\n")
    r = subprocess.run([DRIVER, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape = c.sol.shape
    tend = np.fromfile(prefix + "_tend.bin").reshape(shape)
    out = np.fromfile(prefix + "_out.bin").reshape(shape)
    assert rel_err(tend, c.om.tendency()).max() <= TOL
    # CalculateEntropy / Integral of the initial state through the Fortran layer (the oracle's
    # Euler entropy takes the pressure of its last tendency evaluation: the same state here)
    entropy, mass = np.loadtxt(prefix + "_diag.txt")
    assert abs(entropy - c.om.entropy()) <= TOL * abs(c.om.entropy())
    assert abs(mass - c.om.mass(2)) <= TOL * abs(c.om.mass(2))
    c.om.rk3_steps(steps, dt)
    assert rel_err(out, c.om.solution).max() <= TOL
    # WriteModel / ReadModel through SELF_IO: the pickup holds the final state exactly, with N
    reread = np.fromfile(prefix + "_reread.bin").reshape(shape)
    assert np.array_equal(reread, out)
    from self_amd import _lib
    if _lib.io_available():
        sol = np.empty(shape)
        _lib.call_io("selfamd_pickup_read", os.fsencode(prefix + "_pickup.h5"), N, shape[1], shape[0], sol.ctypes.data)
        assert np.array_equal(sol, out)


def test_fortran_driver_on_hopr_mesh(tmp_path):
    """Read_HOPr of the Fortran SELF_IO module: the driver reads the Block2D HOPr mesh file
    (committed copy of the reference's etc/mesh/Block2D/Block2D_mesh.h5), resets its boundary
    conditions to no-normal-flow and runs SW at N = 5; tendency and RK3 steps must match the
    oracle on the mesh that the Python reader builds from the same file."""
    import self_amd
    assert os.path.exists(DRIVER), "build the Fortran layer first (make -C self_amd/fortran)"
    mesh_file = os.path.join(ROOT, "tests", "golden", "hopr_Block2D_mesh.h5")
    mesh = self_amd.Mesh2D.read_hopr(mesh_file)
    mesh.ResetBoundaryConditionType(self_amd.BC_NONORMALFLOW)
    N = 5
    c = Case("sw", N, 0, 0, hopr=mesh, seed=4, metrics="affine")
    prefix = str(tmp_path / "hopr")
    c.sol.astype(np.float64).tofile(prefix + "_sol.bin")
    np.ascontiguousarray(c.fields["topo"]).tofile(prefix + "_field1.bin")
    np.zeros_like(c.fields["topo"]).tofile(prefix + "_field2.bin")
    dt, steps = c.dt(), 6
    par = tmp_path / "run.nml"
    par.write_text(f"&run model=2, N={N}, bcid={self_amd.BC_NONORMALFLOW}, nSteps={steps}, dt={float(dt)!r}, "
                   f"nodal=.false., prefix='{prefix}', meshFile='{mesh_file}' /\n")
    r = subprocess.run([DRIVER, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape = c.sol.shape
    assert shape[0] == 25
    tend = np.fromfile(prefix + "_tend.bin").reshape(shape)
    out = np.fromfile(prefix + "_out.bin").reshape(shape)
    assert rel_err(tend, c.om.tendency()).max() <= TOL
    c.om.rk3_steps(steps, dt)
    assert rel_err(out, c.om.solution).max() <= TOL
