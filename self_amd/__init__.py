"""self_amd -- MI355X-native DGSEM tendency + low-storage RK3 engine.

Host-side mirror of the SELF model API for the 2D DGSEM path (Lagrange, Mesh2D, SEMQuad,
Model2D and the LinearShallowWater / ShallowWater / CompressibleIdealGas2D plug-ins).  All
numerics run in libself_amd.so (hand-written HIP kernels for gfx950) through its C ABI.
Device memory is allocated through the library's own C ABI (DeviceArray), so the process
loads only the HIP runtime libself_amd.so links.  There is no CPU fallback: without the
native library or a GPU the compute calls raise.

Array convention: a Fortran array a(0:N,0:N,nVar,nEl) is held in C order with the shape
reversed, a[e, v, j, i]; boundary(0:N,nVar,4,nEl) is b[e, s, v, i]; sideInfo(5,4,nEl) is
si[e, s, k].
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import SelfAmdError, call, load

GAUSS, GAUSS_LOBATTO, UNIFORM = 1, 2, 5
BC_PRESCRIBED, BC_RADIATION, BC_NONORMALFLOW = 100, 101, 102
MODEL_LSW, MODEL_SW, MODEL_EULER = 1, 2, 3
PARAM_G, PARAM_FCORI, PARAM_GAMMA, PARAM_RGAS, PARAM_NODAL_METRICS = 1, 2, 3, 4, 5
FIELD_TOPOGRAPHY, FIELD_BATHYMETRY, FIELD_CORIOLIS, FIELD_POTENTIAL, FIELD_DRAG = 1, 2, 3, 4, 5

__all__ = ["DeviceArray", "device_count", "device_name", "synchronize", "Lagrange", "Mesh2D", "SEMQuad",
           "Model2D", "ShallowWater", "LinearShallowWater", "CompressibleIdealGas2D", "Batch", "SelfAmdError", "load", "GAUSS", "GAUSS_LOBATTO", "UNIFORM",
           "BC_PRESCRIBED", "BC_RADIATION", "BC_NONORMALFLOW"]


class DeviceArray:
    """A device allocation owned by Python (hipMalloc through the library's C ABI).

    Holds `shape`/`dtype` metadata for a C-order array; `ptr` is the device address passed
    to entry points.  Copies are synchronous."""

    def __init__(self, shape, dtype=np.float64, zero=False):
        self.shape = tuple(int(n) for n in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        _require_device()
        call("selfamd_malloc", ctypes.byref(self.ptr), self.nbytes)
        if zero:
            call("selfamd_memset", self.ptr, 0, self.nbytes)

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or (a.dtype if isinstance(a, np.ndarray) else np.float64))
        d = cls(a.shape, a.dtype)
        d.copy_from_host(a)
        return d

    def copy_from_host(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        call("selfamd_memcpy", self.ptr, ctypes.c_void_p(a.ctypes.data), self.nbytes, 1)

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        call("selfamd_memcpy", ctypes.c_void_p(out.ctypes.data), self.ptr, self.nbytes, 2)
        return out

    def reshape(self, *shape):
        """View with another shape (same allocation)."""
        shape = shape[0] if len(shape) == 1 and isinstance(shape[0], tuple) else shape
        v = object.__new__(DeviceArray)
        v.shape, v.dtype, v.nbytes, v.ptr = tuple(shape), self.dtype, self.nbytes, self.ptr
        v._base = self
        assert int(np.prod(v.shape)) * v.dtype.itemsize == self.nbytes
        return v

    def __del__(self):
        if getattr(self, "_base", None) is None and getattr(self, "ptr", None) is not None and self.ptr.value:
            try:
                load().selfamd_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


def device_count():
    n = ctypes.c_int(0)
    call("selfamd_device_count", ctypes.byref(n))
    return n.value


def device_name(device=0):
    buf = ctypes.create_string_buffer(256)
    call("selfamd_device_name", device, buf, 256)
    return buf.value.decode()


def time_stamp(t):
    """13-digit time stamp dddd hh mm ss mmm of t seconds (TimeStamp with units 's',
    src/SELF_SupportRoutines.f90:856-924): day, hour, minute, second, millisecond."""
    day = int(t / 86400.0)
    hour = int((t - 86400.0 * day) / 3600.0)
    minute = int((t - 3600.0 * hour - 86400.0 * day) / 60.0)
    rest = t - 60.0 * minute - 3600.0 * hour - 86400.0 * day
    second = int(rest)
    ms = int(np.floor((rest - second) * 1000.0 + 0.5))  # nearest, halves away from zero
    if ms >= 1000:
        ms, second = ms - 1000, second + 1
    if second >= 60:
        second, minute = second - 60, minute + 1
    if minute >= 60:
        minute, hour = minute - 60, hour + 1
    if hour >= 24:
        hour, day = hour - 24, day + 1
    return f"{day:04d}{hour:02d}{minute:02d}{second:02d}{ms:03d}"


def _es16(v):
    """v in the Fortran edit descriptor ES16.7E3: one digit before the point, seven after,
    a three-digit exponent, right-justified in 16 columns."""
    m, e = f"{v:.7E}".split("E")
    return f"{m}E{int(e):+04d}".rjust(16)


def synchronize():
    call("selfamd_device_synchronize")


def _require_device():
    if device_count() < 1:
        raise SelfAmdError("self_amd compute calls need a ROCm GPU (no HIP device visible)")


def _ptr(a):
    """Address for the C ABI: numpy arrays (host) or DeviceArray (device)."""
    if isinstance(a, DeviceArray):
        return a.ptr
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def _dev(x, shape=None):
    """DeviceArray from numpy input (copied) or a DeviceArray (returned as is)."""
    if isinstance(x, DeviceArray):
        return x if shape is None else x.reshape(shape)
    a = np.ascontiguousarray(np.asarray(x, np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return DeviceArray.from_host(a)


class Lagrange:
    """Lagrange interpolant on N+1 control points (Init_Lagrange, src/SELF_Lagrange.f90:521-615).

    Host numpy arrays: controlPoints, targetPoints, bWeights, qWeights, iMatrix[M+1, N+1]
    (= Fortran iMatrix(0:N,0:M)), dMatrix, dgMatrix [N+1, N+1] (Fortran (0:N,0:N)) and
    bMatrix[2, N+1] (Fortran (0:N,0:1)).
    """

    def __init__(self, N, controlNodeType=GAUSS, M=None, targetNodeType=UNIFORM):
        M = N if M is None else M
        self.N, self.M = N, M
        self.controlNodeType, self.targetNodeType = controlNodeType, targetNodeType
        P, Q = N + 1, M + 1
        self.controlPoints = np.zeros(P)
        self.targetPoints = np.zeros(Q)
        self.bWeights = np.zeros(P)
        self.qWeights = np.zeros(P)
        self.iMatrix = np.zeros((Q, P))
        self.dMatrix = np.zeros((P, P))
        self.dgMatrix = np.zeros((P, P))
        self.bMatrix = np.zeros((2, P))
        call("selfamd_lagrange_init", N, controlNodeType, M, targetNodeType,
             *[_ptr(a) for a in (self.controlPoints, self.targetPoints, self.bWeights, self.qWeights,
                                  self.iMatrix, self.dMatrix, self.dgMatrix, self.bMatrix)])
        self._dev = None

    def device(self):
        """Device copies (iMatrix, dMatrix, dgMatrix, bMatrix, qWeights) as DeviceArrays."""
        if self._dev is None:
            self._dev = {k: DeviceArray.from_host(getattr(self, k))
                         for k in ("iMatrix", "dMatrix", "dgMatrix", "bMatrix", "qWeights")}
        return self._dev


class Mesh2D:
    """2-D quad mesh connectivity in the reference convention (src/SELF_Mesh.f90:27-91)."""

    def __init__(self, sideInfo, nodeCoords, globalNodeIDs, nGeo, quadrature=UNIFORM):
        self.sideInfo = np.ascontiguousarray(sideInfo, dtype=np.int32)        # [nEl, 4, 5]
        self.nodeCoords = np.ascontiguousarray(nodeCoords, dtype=np.float64)  # [nEl, G, G, 2]
        self.globalNodeIDs = np.ascontiguousarray(globalNodeIDs, dtype=np.int32)  # [nEl, G, G]
        self.nGeo = nGeo
        self.nElem = self.sideInfo.shape[0]
        self.quadrature = quadrature

    @classmethod
    def structured(cls, nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0, periodic=(False, False), bcid=BC_RADIATION):
        n = nx * ny
        si = np.zeros((n, 4, 5), np.int32)
        nc = np.zeros((n, 2, 2, 2))
        gn = np.zeros((n, 2, 2), np.int32)
        call("selfamd_mesh2d_structured", nx, ny, x0, x1, y0, y1, int(periodic[0]), int(periodic[1]), bcid,
             _ptr(si), _ptr(nc), _ptr(gn))
        m = cls(si, nc, gn, 1)
        m.nx, m.ny = nx, ny
        return m

    @classmethod
    def read_hopr(cls, path):
        """Read_HOPr_Mesh2D (src/SELF_Mesh.f90:852-1016) from a HOPr mesh file: the attribute
        Ngeo and the datasets SideInfo, NodeCoords and GlobalNodeIDs are read with the HDF5 C
        library (libself_amd_io.so), then converted by from_hopr."""
        from ._lib import call_io
        p = os.fsencode(os.path.abspath(path))
        ngeo = ctypes.c_int32()
        call_io("selfamd_hopr_attribute_int", p, b"Ngeo", ctypes.byref(ngeo))

        def read(name, dtype, fn):
            rank, dims = ctypes.c_int32(), (ctypes.c_int64 * 5)()
            call_io("selfamd_hopr_dataset_dims", p, name, ctypes.byref(rank), dims)
            shape = tuple(dims[k] for k in range(rank.value))
            a = np.empty(shape, dtype)
            call_io(fn, p, name, a.ctypes.data, a.size)
            return a

        sideInfo = read(b"SideInfo", np.int32, "selfamd_hopr_read_int")
        nodeCoords = read(b"NodeCoords", np.float64, "selfamd_hopr_read_double")
        globalNodeIDs = read(b"GlobalNodeIDs", np.int32, "selfamd_hopr_read_int").reshape(-1)
        return cls.from_hopr(ngeo.value, sideInfo, nodeCoords, globalNodeIDs)

    @classmethod
    def from_hopr(cls, nGeo, sideInfo3D, nodeCoords3D, globalNodeIDs3D):
        """Read_HOPr_Mesh2D's 3-D -> 2-D conversion on arrays read from a HOPr file."""
        si3 = np.ascontiguousarray(sideInfo3D, np.int32)
        nEl = si3.shape[0] // 6
        G = nGeo + 1
        si = np.zeros((nEl, 4, 5), np.int32)
        nc = np.zeros((nEl, G, G, 2))
        gn = np.zeros((nEl, G, G), np.int32)
        nc3 = np.ascontiguousarray(nodeCoords3D, np.float64)  # named: must outlive the call
        gn3 = np.ascontiguousarray(globalNodeIDs3D, np.int32)
        call("selfamd_mesh2d_from_hopr", nEl, nGeo, _ptr(si3), _ptr(nc3), _ptr(gn3), _ptr(si), _ptr(nc), _ptr(gn))
        return cls(si, nc, gn, nGeo)

    def RecalculateFlip(self):
        call("selfamd_mesh2d_recalculate_flip", self.nElem, self.nGeo, _ptr(self.sideInfo), _ptr(self.globalNodeIDs))

    def ResetBoundaryConditionType(self, bcid):
        call("selfamd_mesh2d_reset_boundary_condition_type", self.nElem, _ptr(self.sideInfo), bcid)


class SEMQuad:
    """Geometry (metric terms) generated on the device (GenerateFromMesh_SEMQuad,
    src/SELF_Geometry.f90:393-552).  Attributes are DeviceArrays in the C-order views:
    x[e,j,i,2], xb[e,s,i,2], dxds/dsdx[e,j,i,2(c),2(r)], dsdxb[e,s,i,2,2], nHatb[e,s,i,2],
    nScaleb[e,s,i], J[e,j,i], Jb[e,s,i]."""

    def __init__(self, interp, mesh):
        N, nEl, P = interp.N, mesh.nElem, interp.N + 1
        self.interp, self.nElem = interp, nEl
        self.x = DeviceArray((nEl, P, P, 2))
        self.xb = DeviceArray((nEl, 4, P, 2))
        self.dxds = DeviceArray((nEl, P, P, 2, 2))
        self.dxdsb = DeviceArray((nEl, 4, P, 2, 2))
        self.dsdx = DeviceArray((nEl, P, P, 2, 2))
        self.dsdxb = DeviceArray((nEl, 4, P, 2, 2))
        self.nHatb = DeviceArray((nEl, 4, P, 2))
        self.nScaleb = DeviceArray((nEl, 4, P))
        self.J = DeviceArray((nEl, P, P))
        self.Jb = DeviceArray((nEl, 4, P))
        nc = DeviceArray.from_host(mesh.nodeCoords)
        call("selfamd_semquad_generate", N, interp.controlNodeType, mesh.nGeo, nEl, _ptr(nc),
             *[_ptr(t) for t in (self.x, self.xb, self.dxds, self.dxdsb, self.dsdx, self.dsdxb, self.nHatb,
                                 self.nScaleb, self.J, self.Jb)], None)

    def CovariantArcMin(self):
        """CovariantArcMin_SEMQuad (src/SELF_Geometry.f90:554-582)."""
        d = self.dxds.to_host()  # [e, j, i, c, r] = dxds(r, c)
        dx = np.sqrt(d[..., 0, 0] ** 2 + d[..., 1, 0] ** 2)  # sqrt(dxds(1,1)^2 + dxds(1,2)^2)
        dy = np.sqrt(d[..., 0, 1] ** 2 + d[..., 1, 1] ** 2)  # sqrt(dxds(2,1)^2 + dxds(2,2)^2)
        N = self.interp.N
        return float(np.min(2.0 * np.minimum(dx, dy) / (N * N)))


class Model2D:
    """Model2D + physics plug-in on one GPU (src/SELF_Model2D.f90:18-71).  The state stays
    resident in the engine; CalculateTendency and ForwardStep run the fused HIP kernels."""

    kind = None

    def __init__(self, mesh, geometry):
        self.mesh, self.geometry = mesh, geometry
        self.N = geometry.interp.N
        self.nElem = mesh.nElem
        self.nVar = load().selfamd_model2d_nvar(self.kind)
        self._h = ctypes.c_void_p()
        call("selfamd_model2d_create", ctypes.byref(self._h), self.kind, self.N, self.nElem, _ptr(mesh.sideInfo),
             _ptr(geometry.dsdx), _ptr(geometry.J), _ptr(geometry.nHatb), _ptr(geometry.nScaleb), None)
        self._buf = DeviceArray(self.shape)
        self.dt = 0.0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().selfamd_model2d_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def shape(self):
        P = self.N + 1
        return (self.nElem, self.nVar, P, P)

    def _param(self, which, value):
        call("selfamd_model2d_set_param", self._h, which, float(value))

    def _field(self, which, values):
        f = _dev(values, (self.nElem, self.N + 1, self.N + 1))
        call("selfamd_model2d_set_field", self._h, which, _ptr(f))

    def SetNodalMetrics(self, on=True):
        """Use per-node metric terms (the reference's representation; required on curved
        meshes) instead of one affine record per element."""
        self._param(PARAM_NODAL_METRICS, 1.0 if on else 0.0)

    def SetSolution(self, interior):
        """SetSolution (src/SELF_Model2D.f90:313-332) from interior[e, v, j, i] (numpy or DeviceArray)."""
        s = _dev(interior, self.shape)
        call("selfamd_model2d_set_solution", self._h, _ptr(s))

    def CalculateEntropy(self):
        """CalculateEntropy (SW src/models/SELF_ShallowWater.f90:149-188, LSW
        src/models/SELF_LinearShallowWater.f90:171-216, Euler
        src/models/SELF_CompressibleIdealGas2D.f90:782-823) of the current state, computed on
        the device; also stored as `self.entropy`.  Euler uses the pressure of the current
        state (the reference: that of its last tendency evaluation)."""
        v = ctypes.c_double()
        call("selfamd_model2d_entropy", self._h, ctypes.byref(v))
        self.entropy = v.value
        return v.value

    def Integral(self, var):
        """sum over elements and nodes of w_i w_j J S(i, j, var, e), on the device (the mass of
        variable `var`; no reference routine, the pattern of Integral_MappedScalar2D,
        src/SELF_MappedData.f90:1270-1311)."""
        v = ctypes.c_double()
        call("selfamd_model2d_integral", self._h, int(var), ctypes.byref(v))
        return v.value

    def Stage(self, stage, dt, Ssave=None, unSave=None):
        """One low-storage RK3 stage (0, 1, 2 in turn), closure included; the stage's input
        state and lagged face normal velocities are copied to the given DeviceArrays first."""
        call("selfamd_model2d_stage", self._h, int(stage), float(dt), Ssave.ptr if Ssave is not None else None,
             unSave.ptr if unSave is not None else None)

    @property
    def lag_shape(self):
        """Shape of the lagged SW face normal velocities of the affine engine, (nElem, 4, N+1)."""
        return (self.nElem, 4, self.N + 1)

    def TendencyVJP(self, S, un, lam, mu=None, Sprev=None, muPrev=None):
        """mu = (d dSdt / dS)^T lam at state S with the lagged normal velocities un (SW, affine,
        periodic); with Sprev (the previous stage's input, from which un was formed) also the
        lag path into the DeviceArray muPrev.  Arrays host or DeviceArrays; returns mu."""
        Sd, ud, ld = _dev(S, self.shape), _dev(un, self.lag_shape), _dev(lam, self.shape)
        mu = mu if mu is not None else DeviceArray(self.shape)
        Spd = _dev(Sprev, self.shape) if Sprev is not None else None
        call("selfamd_model2d_tendency_vjp", self._h, _ptr(Sd), _ptr(ud), _ptr(ld), _ptr(mu),
             _ptr(Spd) if Spd is not None else None, _ptr(muPrev) if muPrev is not None else None)
        return mu

    def ClosureVJP(self, S, lam, mu):
        """Adjoint of the attached closure's correction at state S: adds the network's parameter
        gradients to its gradient buffer and the state cotangent to the DeviceArray mu."""
        Sd, ld = _dev(S, self.shape), _dev(lam, self.shape)
        call("selfamd_model2d_closure_vjp", self._h, _ptr(Sd), _ptr(ld), _ptr(mu))
        return mu

    def IntegralNoAlloc(self, var, out):
        """Integral(var) written to the device double `out` (a DeviceArray) on the model's
        stream by the library's no-allocation entry point; does not synchronise."""
        call("selfamd_model2d_integral_noalloc", self._h, int(var), out.ptr)

    def WriteModel(self, path):
        """WriteModel (Write_Model2D, src/SELF_Model2D.f90:861-1059) to an HDF5 pickup file:
        quadrature, interior and boundary solution, interior and boundary mesh positions, and
        the attribute N that ReadModel needs (libself_amd_io.so).  The geometry must carry x
        and xb (a SEMQuad)."""
        from ._lib import call_io
        interp, g = self.geometry.interp, self.geometry
        sol = self.solution
        dev = DeviceArray.from_host(sol)
        bnd = DeviceArray((self.nElem, 4, self.nVar, self.N + 1))
        bm = DeviceArray.from_host(interp.bMatrix)
        call("selfamd_scalar_boundary_interp_2d", bm.ptr, dev.ptr, bnd.ptr, self.N, self.nVar, self.nElem, None)
        synchronize()
        x, xb = g.x.to_host(), g.xb.to_host()
        host = [np.ascontiguousarray(a, np.float64) for a in (interp.controlPoints, interp.qWeights, interp.dgMatrix,
                                                               interp.dMatrix, interp.bMatrix, sol, bnd.to_host(), x, xb)]
        call_io("selfamd_pickup_write", os.fsencode(os.path.abspath(path)), self.N, self.nVar, self.nElem,
                *[a.ctypes.data for a in host])
        # the flux and fluxDivergence registers of a tendency evaluation of the current state;
        # the flux traces and the solution gradient, which the reference's tendency path never
        # fills, are written as zeros
        fdiv, fl = self.flux_registers()
        call_io("selfamd_pickup_write_registers", os.fsencode(os.path.abspath(path)), self.N, self.nVar, self.nElem,
                fdiv.ctypes.data, fl.ctypes.data, None, None, None)

    def flux_registers(self):
        """(fluxDivergence[e, v, j, i], flux[e, v, j, i, 2]): the registers CalculateTendency
        (src/SELF_Model2D.f90:816-859) leaves for the current state, the contravariant flux and
        its weak divergence divided by J.  The state and the SW lag state are unchanged."""
        P = self.N + 1
        fdiv = DeviceArray(self.shape)
        fl = DeviceArray((self.nElem, self.nVar, P, P, 2))
        call("selfamd_model2d_flux_registers", self._h, _ptr(fdiv), _ptr(fl))
        return fdiv.to_host(), fl.to_host()

    def WriteTecplot(self, filename=None):
        """WriteTecplot (WriteTecplot_Model2D, src/SELF_Model2D.f90:1101-1207): the element
        positions and the solution interpolated on the device to the interpolant's target grid
        ((M+1)^2 points per element, GridInterp), written as ASCII Tecplot, one zone per
        element, X, Y, the variables, then d/dx and d/dy of each variable.  The gradient
        columns are zero: no tendency on this path forms the solution gradient (the
        reference's 2-D models allocate that register and never fill it).  Default file name:
        solution.<time stamp>.tec.  Returns the file name."""
        interp, g = self.geometry.interp, self.geometry
        N, M, nEl, nVar = self.N, interp.M, self.nElem, self.nVar
        if filename is None:
            filename = f"solution.{time_stamp(self.t)}.tec"
        iM = DeviceArray.from_host(np.ascontiguousarray(interp.iMatrix))
        s = DeviceArray(self.shape)
        call("selfamd_model2d_get_solution", self._h, _ptr(s))
        sT = DeviceArray((nEl, nVar, M + 1, M + 1))
        xT = DeviceArray((nEl, M + 1, M + 1, 2))
        call("selfamd_scalar_grid_interp_2d", iM.ptr, s.ptr, sT.ptr, N, M, nVar, nEl, None)
        call("selfamd_vector_grid_interp_2d", iM.ptr, g.x.ptr, xT.ptr, N, M, 1, nEl, None)
        sol, x = sT.to_host(), xT.to_host()
        names = self.var_names
        header = 'VARIABLES = "X", "Y"' + "".join(f', "{n}"' for n in names) + "".join(
            f', "d/dx({n})"' for n in names) + "".join(f', "d/dy({n})"' for n in names)
        # the edit descriptors (ES16.7E3, 1X) repeated: one blank between values and none at the
        # end of a record (a trailing X edit writes nothing)
        zeros = "".join(" " + _es16(0.0) for _ in range(2 * nVar))
        with open(filename, "w") as f:
            f.write(" " + header + "\n")
            for e in range(nEl):
                f.write(f' ZONE T="el{e + 1:08d}", I={M + 1}, J={M + 1}\n')
                for j in range(M + 1):
                    for i in range(M + 1):
                        vals = [x[e, j, i, 0], x[e, j, i, 1]] + [sol[e, v, j, i] for v in range(nVar)]
                        f.write(" ".join(_es16(v) for v in vals) + zeros + "\n")
        return filename

    def ReadModel(self, path):
        """ReadModel (Read_Model2D, src/SELF_Model2D.f90:1061-1099): the solution of a pickup
        file (its N must match), set with SetSolution.  The SW velocity lag restarts from the
        read state, as in the reference, whose pickup holds no lag state either."""
        from ._lib import call_io
        sol = np.empty(self.shape)
        call_io("selfamd_pickup_read", os.fsencode(os.path.abspath(path)), self.N, self.nVar, self.nElem,
                sol.ctypes.data)
        self.SetSolution(sol)
        return sol

    @property
    def solution(self):
        """Host copy of the solution interior[e, v, j, i]."""
        call("selfamd_model2d_get_solution", self._h, _ptr(self._buf))
        return self._buf.to_host()

    def CalculateTendency(self):
        """CalculateTendency (src/SELF_Model2D.f90:816-859); returns a host copy of dSdt[e, v, j, i]."""
        call("selfamd_model2d_tendency", self._h, _ptr(self._buf))
        return self._buf.to_host()

    def ForwardStep(self, nSteps=1, dt=None):
        """nSteps low-storage RK3 steps (LowStorageRK3_timeIntegrator, src/SELF_Model.f90:704-732)."""
        if dt is not None:
            self.dt = dt
        call("selfamd_model2d_forward_step", self._h, int(nSteps), float(self.dt))
        call("selfamd_model2d_synchronize", self._h)

    def time_steps(self, nSteps, dt):
        """Run nSteps and return the hipEvent-measured milliseconds on the model's stream."""
        ms = ctypes.c_double()
        call("selfamd_model2d_time_steps", self._h, int(nSteps), float(dt), ctypes.byref(ms))
        return ms.value

    @property
    def stream(self):
        """The model's HIP stream handle (an integer address), for ordering other work with
        the model's (e.g. the library's copy kernel)."""
        st = ctypes.c_void_p()
        call("selfamd_model2d_stream", self._h, ctypes.byref(st))
        return st.value

    def solution_device(self):
        """Device copy of the solution (a buffer owned by the model, overwritten by the next
        solution/tendency call)."""
        call("selfamd_model2d_get_solution", self._h, _ptr(self._buf))
        return self._buf

    def SetBands(self, nBands):
        """Banded stepping over nBands bands of the element graph (0: off); bit-identical
        results, better Infinity-Cache reuse across the three RK stages."""
        call("selfamd_model2d_set_bands", self._h, int(nBands))

    def SetClosure(self, mlp, inMean=None, inScale=None, outMean=None, outScale=None):
        """Attach an MLP tendency correction (None detaches); normalisation arrays are host
        float32 of the network's input/output widths."""
        if mlp is None:
            call("selfamd_model2d_set_closure", self._h, None, None, None, None, None)
            self._closure = None
            return
        arrs = [np.ascontiguousarray(a, np.float32) for a in (inMean, inScale, outMean, outScale)]
        call("selfamd_model2d_set_closure", self._h, mlp._h, *[_ptr(a) for a in arrs])
        self._closure = (mlp, arrs)

    def closure_features(self):
        """Host copy of the attached closure's normalised features of the current state."""
        mlp = self._closure[0]
        X = DeviceArray((self.nElem, mlp.nIn), np.float32)
        call("selfamd_model2d_closure_features", self._h, _ptr(X))
        return X.to_host()

    @property
    def t(self):
        v = ctypes.c_double()
        call("selfamd_model2d_get_time", self._h, ctypes.byref(v))
        return v.value

    @t.setter
    def t(self, value):
        call("selfamd_model2d_set_time", self._h, float(value))


class ShallowWater(Model2D):
    """Nonlinear shallow water (src/models/SELF_ShallowWater.f90): state (Hu, Hv, H)."""
    kind = MODEL_SW
    var_names = ("Hu", "Hv", "H")  # src/models/SELF_ShallowWater.f90:135-143

    def __init__(self, mesh, geometry, g=1.0, fCori=0.0):
        super().__init__(mesh, geometry)
        self._param(PARAM_G, g)
        self._param(PARAM_FCORI, fCori)

    def SetTopography(self, topo):
        self._field(FIELD_TOPOGRAPHY, topo)


class LinearShallowWater(Model2D):
    """Linear shallow water (src/models/SELF_LinearShallowWater.f90): state (u, v, eta)."""
    kind = MODEL_LSW
    var_names = ("u", "v", "eta")  # src/models/SELF_LinearShallowWater.f90:139-147

    def __init__(self, mesh, geometry, g=1.0):
        super().__init__(mesh, geometry)
        self._param(PARAM_G, g)

    def SetBathymetry(self, H):
        if np.isscalar(H):
            H = np.full((self.nElem, self.N + 1, self.N + 1), float(H))
        self._field(FIELD_BATHYMETRY, H)

    def SetCoriolis(self, f):
        if np.isscalar(f):
            f = np.full((self.nElem, self.N + 1, self.N + 1), float(f))
        self._field(FIELD_CORIOLIS, f)


class CompressibleIdealGas2D(Model2D):
    """Compressible ideal gas (src/models/SELF_CompressibleIdealGas2D.f90): state
    (rho u, rho v, rho, rho E).  Boundary sides with bcid 100/101 use the prescribed state
    taken from the last SetSolution (SetPrescribedSolutionFromSolution)."""
    kind = MODEL_EULER
    var_names = ("rho*u", "rho*v", "rho", "rho*E")  # src/models/SELF_CompressibleIdealGas2D.f90:306-318

    def __init__(self, mesh, geometry, gamma=1.4, R=1.0):
        super().__init__(mesh, geometry)
        self._param(PARAM_GAMMA, gamma)
        self._param(PARAM_RGAS, R)

    def SetGravity(self, phi):
        self._field(FIELD_POTENTIAL, phi)

    def SetDrag(self, cd):
        if np.isscalar(cd):
            cd = np.full((self.nElem, self.N + 1, self.N + 1), float(cd))
        self._field(FIELD_DRAG, cd)


class Batch:
    """Many independent models advanced together (no reference counterpart).

    Runs of one polynomial degree share one fused launch per RK stage and every degree group
    of a step sits in one hipGraph.  `SetMode(Batch.PERSISTENT)` instead advances all groups
    (nonlinear SW, degrees 3, 5, 7) in one persistent launch per stage; both modes give each
    run its separate solve's results bit for bit.  `scatter()` hands each run's state back to
    its model."""

    PER_GROUP = 0
    PERSISTENT = 1

    def __init__(self, models, dts):
        assert len(models) == len(dts) and models
        self.models = list(models)
        arr = (ctypes.c_void_p * len(models))(*[m._h.value for m in models])
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        self._h = ctypes.c_void_p()
        call("selfamd_batch_create", ctypes.byref(self._h), len(models), arr, _ptr(dt))

    def SetMode(self, mode):
        call("selfamd_batch_set_mode", self._h, int(mode))

    def ForwardStep(self, nSteps):
        call("selfamd_batch_forward_step", self._h, int(nSteps))

    def time_steps(self, nSteps):
        ms = ctypes.c_double()
        call("selfamd_batch_time_steps", self._h, int(nSteps), ctypes.byref(ms))
        return ms.value

    def scatter(self):
        call("selfamd_batch_scatter", self._h)

    def info(self):
        g, dof = ctypes.c_int(), ctypes.c_double()
        call("selfamd_batch_info", self._h, ctypes.byref(g), ctypes.byref(dof))
        return g.value, dof.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().selfamd_batch_destroy(h)
            except Exception:
                pass
            self._h = None


# ----------------------------------------------------------------------------------------
# Learned subgrid closure (no reference counterpart)
# ----------------------------------------------------------------------------------------
MLP_KEYS = ("W1", "b1", "W2", "b2", "W3", "b3")


class MLP:
    """FP32 MLP nIn -> nHidden -> nHidden -> nOut (GELU), trained by the library's kernels.

    Parameters are exchanged as a dict of float32 arrays W1[nIn, H], b1[H], W2[H, H], b2[H],
    W3[H, nOut], b3[nOut] (a layer is Y = X W + b)."""

    def __init__(self, nIn, nHidden, nOut, maxBatch):
        self.nIn, self.nHidden, self.nOut, self.maxBatch = nIn, nHidden, nOut, maxBatch
        _require_device()
        self._h = ctypes.c_void_p()
        call("selfamd_mlp_create", ctypes.byref(self._h), nIn, nHidden, nOut, maxBatch)
        n = ctypes.c_size_t()
        call("selfamd_mlp_num_params", self._h, ctypes.byref(n))
        self.nParams = n.value

    def _shapes(self):
        i, h, o = self.nIn, self.nHidden, self.nOut
        return {"W1": (i, h), "b1": (h,), "W2": (h, h), "b2": (h,), "W3": (h, o), "b3": (o,)}

    def pack(self, params):
        return np.concatenate([np.asarray(params[k], np.float32).ravel() for k in MLP_KEYS])

    def unpack(self, flat):
        out, o = {}, 0
        for k, shp in self._shapes().items():
            n = int(np.prod(shp))
            out[k] = flat[o:o + n].reshape(shp)
            o += n
        return out

    def set_params(self, params):
        """Load parameters (dict) and reset the optimiser moments."""
        flat = self.pack(params)
        assert flat.size == self.nParams
        call("selfamd_mlp_set_params", self._h, _ptr(flat))

    def get_params(self):
        flat = np.empty(self.nParams, np.float32)
        call("selfamd_mlp_get_params", self._h, _ptr(flat))
        return self.unpack(flat)

    def get_grads(self):
        flat = np.empty(self.nParams, np.float32)
        call("selfamd_mlp_get_grads", self._h, _ptr(flat))
        return self.unpack(flat)

    def forward(self, X):
        """Y = net(X) for X (rows, nIn) host or DeviceArray; returns a host array."""
        Xd = X if isinstance(X, DeviceArray) else DeviceArray.from_host(np.asarray(X, np.float32))
        rows = Xd.shape[0]
        Y = DeviceArray((rows, self.nOut), np.float32)
        call("selfamd_mlp_forward", self._h, Xd.ptr, Y.ptr, rows, None)
        return Y.to_host()

    def loss_grad(self, X, T, rows=None, x_offset=0, want_loss=True):
        """Forward + backward on rows [x_offset, x_offset + rows) of device arrays X, T."""
        rows = X.shape[0] if rows is None else rows
        xp = ctypes.c_void_p(X.ptr.value + 4 * self.nIn * x_offset)
        tp = ctypes.c_void_p(T.ptr.value + 4 * self.nOut * x_offset)
        loss = ctypes.c_float()
        call("selfamd_mlp_loss_grad", self._h, xp, tp, int(rows), ctypes.byref(loss) if want_loss else None, None)
        return loss.value if want_loss else None

    def adam_step(self, lr, step):
        call("selfamd_mlp_adam_step", self._h, float(lr), int(step), None)

    def backward(self, X, dY, accumulate=False, want_dX=True):
        """Parameter gradients (written, or added with accumulate) and the input cotangent
        (host array, or None) of the output cotangent dY for inputs X (host or DeviceArray)."""
        Xd = X if isinstance(X, DeviceArray) else DeviceArray.from_host(np.asarray(X, np.float32))
        Yd = dY if isinstance(dY, DeviceArray) else DeviceArray.from_host(np.asarray(dY, np.float32))
        rows = Xd.shape[0]
        dX = DeviceArray((rows, self.nIn), np.float32) if want_dX else None
        call("selfamd_mlp_backward", self._h, Xd.ptr, Yd.ptr, rows, int(bool(accumulate)), dX.ptr if dX else None, None)
        return dX.to_host() if dX else None

    def zero_grads(self):
        call("selfamd_mlp_zero_grads", self._h, None)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().selfamd_mlp_destroy(h)
            except Exception:
                pass
            self._h = None


def project_l2_structured(fine, Nf, Nc, nxc, nyc, r, nVar):
    """L2 projection of a structured fine-mesh interior (DeviceArray or numpy, shape
    (nElem_f, nVar, Nf+1, Nf+1)) onto the coarse mesh; returns a DeviceArray."""
    f = _dev(fine)
    out = DeviceArray((nxc * nyc, nVar, Nc + 1, Nc + 1))
    call("selfamd_project_l2_structured_2d", f.ptr, out.ptr, Nf, Nc, nxc, nyc, r, nVar, None)
    return out


# Tuned tiles of the SW stage kernel (self_amd/tuning/sw_tiles.json, written by
# `python -m self_amd.autotune`); without the table the default tiles are used.
from . import autotune as _autotune  # noqa: E402

_autotune.apply()
