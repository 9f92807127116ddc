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


def _driver_inputs(c, kind, tmp_path):
    prefix = str(tmp_path / kind)
    c.sol.astype(np.float64).tofile(prefix + "_sol.bin")
    f1, f2 = {"sw": ("topo", None), "lsw": ("H", "f"), "euler": ("phi", "drag")}[kind]
    np.ascontiguousarray(c.fields[f1]).tofile(prefix + "_field1.bin")
    (np.ascontiguousarray(c.fields[f2]) if f2 else np.zeros_like(c.fields[f1])).tofile(prefix + "_field2.bin")
    return prefix


@pytest.mark.parametrize("kind,model,N,n,periodic,bcid", [("sw", 2, 7, 5, True, 102), ("sw", 2, 5, 4, False, 102),
                                                        ("lsw", 1, 5, 5, False, 102), ("euler", 3, 4, 4, False, 102),
                                                        ("euler", 3, 3, 4, False, 100)])
def test_fortran_plugin_path_parity(kind, model, N, n, periodic, bcid, tmp_path):
    """A user type extension of each physics (overriding nothing) runs the per-operator path:
    CalculateTendency_Model2D's sequence of PreTendency, MappedScalar2D%BoundaryInterp /
    SideExchange, SetBoundaryCondition, SourceMethod, RiemannSolver, FluxMethod,
    MappedVector2D%ContravariantProjection / Divergence (VectorDGDivergence_2D,
    JacobianWeight), and UpdateGRK3 in the reference's RK3 loop, with the SEMQuad per-node
    metric terms.  Tendency and RK3 steps match the unmodified oracle."""
    assert os.path.exists(DRIVER), "build the Fortran layer first (make -C self_amd/fortran)"
    c = Case(kind, N, n, n, periodic=periodic, bcid=bcid, seed=12, metrics="nodal")
    prefix = _driver_inputs(c, kind, tmp_path)
    dt, steps = c.dt(), 5
    par = tmp_path / "run.nml"
    par.write_text(f"&run model={model}, N={N}, nx={n}, ny={n}, periodic={'.true.' if periodic else '.false.'}, "
                   f"bcid={bcid}, nSteps={steps}, dt={float(dt)!r}, plugins=.true., prefix='{prefix}' /\n")
    r = subprocess.run([DRIVER, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "per-operator path" in r.stdout
    shape = c.sol.shape
    tend = np.fromfile(prefix + "_tend.bin").reshape(shape)
    out = np.fromfile(prefix + "_out.bin").reshape(shape)
    assert rel_err(tend, c.om.tendency()).max() <= TOL
    entropy, mass = np.loadtxt(prefix + "_diag.txt")
    assert abs(mass - c.om.mass(2)) <= TOL * abs(c.om.mass(2))
    c.om.rk3_steps(steps, dt)
    assert rel_err(out, c.om.solution).max() <= TOL


def test_fortran_user_source_method(tmp_path):
    """A user's ShallowWater extension with its own SourceMethod (the parent's source, then a
    constant mass source added on host arrays): its tendency is the oracle's plus that
    constant in the thickness equation, and over RK3 steps the mass grows by exactly
    massSource * area * t (the periodic mesh conserves the rest)."""
    assert os.path.exists(DRIVER), "build the Fortran layer first (make -C self_amd/fortran)"
    N, n, s0 = 5, 6, 0.01
    c = Case("sw", N, n, n, periodic=True, seed=13, metrics="nodal")
    prefix = _driver_inputs(c, "sw", tmp_path)
    dt, steps = c.dt(), 6
    par = tmp_path / "run.nml"
    par.write_text(f"&run model=2, N={N}, nx={n}, ny={n}, periodic=.true., bcid=0, nSteps={steps}, "
                   f"dt={float(dt)!r}, plugins=.true., massSource={s0!r}, prefix='{prefix}' /\n")
    r = subprocess.run([DRIVER, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape = c.sol.shape
    tend = np.fromfile(prefix + "_tend.bin").reshape(shape)
    want = c.om.tendency()
    want[:, 2] += s0
    assert rel_err(tend, want).max() <= TOL
    out = np.fromfile(prefix + "_out.bin").reshape(shape)
    w = c.weights
    mass = lambda s: float(np.sum(s[:, 2] * c.geom.J * w[None, :, None] * w[None, None, :]))
    m0 = mass(c.sol)
    assert abs(mass(out) - (m0 + s0 * steps * dt)) <= 1e-13 * m0


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


def _read_tecplot(path):
    """Header, zone lines and the data rows (float array per zone) of a WriteTecplot file."""
    with open(path) as f:
        lines = f.read().splitlines()
    header, zones, rows = lines[0], [], []
    for line in lines[1:]:
        if line.lstrip().startswith("ZONE"):
            zones.append(line)
            rows.append([])
        else:
            rows[-1].append([float(v) for v in line.split()])
    return header, zones, [np.array(r) for r in rows]


def test_write_tecplot_fortran_and_python(tmp_path):
    """WriteTecplot (src/SELF_Model2D.f90:1101-1207) through the Fortran layer and through
    Python on the same state: the files are identical byte for byte, and the values are the
    element positions and the solution interpolated to the uniform (M+1)^2 plot grid."""
    import self_amd
    N, M, n, steps = 5, 8, 3, 4
    c = Case("sw", N, n, n, periodic=True, bcid=0, seed=8, metrics="affine")
    prefix = str(tmp_path / "tec")
    c.sol.astype(np.float64).tofile(prefix + "_sol.bin")
    np.ascontiguousarray(c.fields["topo"]).tofile(prefix + "_field1.bin")
    np.zeros_like(c.fields["topo"]).tofile(prefix + "_field2.bin")
    par = tmp_path / "run.nml"
    par.write_text(f"&run model=2, N={N}, Mtarget={M}, nx={n}, ny={n}, periodic=.true., bcid=0, nSteps={steps}, "
                   f"dt={float(c.dt())!r}, nodal=.false., prefix='{prefix}' /\n")
    r = subprocess.run([DRIVER, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(prefix + "_out.bin").reshape(c.sol.shape)
    # the same state written from Python, on the geometry SEMQuad generates for the same mesh
    mesh = self_amd.Mesh2D.structured(n, n, periodic=(True, True))
    interp = self_amd.Lagrange(N, self_amd.GAUSS, M, self_amd.UNIFORM)
    geom = self_amd.SEMQuad(interp, mesh)
    m = self_amd.ShallowWater(mesh, geom)
    m.SetSolution(out)
    py = m.WriteTecplot(str(tmp_path / "py.tec"))
    with open(prefix + ".tec") as a, open(py) as b:
        assert a.read() == b.read()
    header, zones, rows = _read_tecplot(py)
    assert header == ' VARIABLES = "X", "Y", "Hu", "Hv", "H", "d/dx(Hu)", "d/dx(Hv)", "d/dx(H)", ' \
                     '"d/dy(Hu)", "d/dy(Hv)", "d/dy(H)"'
    assert zones == [f' ZONE T="el{e + 1:08d}", I={M + 1}, J={M + 1}' for e in range(n * n)]
    iM = interp.iMatrix  # [M+1, N+1]: value at target point a = sum_i iM[a, i] f(i)
    x = geom.x.to_host()
    for e in range(n * n):
        d = rows[e]
        assert d.shape == ((M + 1) ** 2, 11)
        xt = np.einsum("bj,ai,jic->bac", iM, iM, x[e]).reshape(-1, 2)
        st = np.einsum("bj,ai,vji->bav", iM, iM, out[e]).reshape(-1, 3)
        assert np.allclose(d[:, :2], xt, rtol=1e-7, atol=1e-7)
        assert np.allclose(d[:, 2:5], st, rtol=1e-7, atol=1e-7 * np.abs(out).max())
        assert not d[:, 5:].any()
