"""Parity oracle: ctypes access to liboracle.so, the CPU restatement of the SELF 2D DGSEM
tendency + RK3 path (self_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package, and only as the checker or the timed CPU baseline.
The product (self_amd) never imports it.

Array views follow the same convention as self_amd: a Fortran array a(0:N,0:N,nVar,nEl)
appears as a C-order numpy array a[e, v, j, i].
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

GAUSS, GAUSS_LOBATTO, UNIFORM = 1, 2, 5
LSW, SW, EULER = 1, 2, 3
BC_PRESCRIBED, BC_RADIATION, BC_NONORMALFLOW = 100, 101, 102

# builds of the same source (oracle/Makefile): "parity" is the oracle proper; the "fast"
# builds allow FMA contraction, as an optimising compiler would, and serve only as the timed
# CPU baseline of bench.py
BUILDS = {"parity": LIB_PATH, "fast_v3": os.path.join(_HERE, "liboracle_fast_v3.so"),
          "fast_v4": os.path.join(_HERE, "liboracle_fast_v4.so")}
_libs = {}
_build = "parity"
_lib = None
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_build(name):
    """Selects the library that objects created from now on use (existing objects keep
    theirs)."""
    global _build, _lib
    if name not in BUILDS:
        raise ValueError(f"unknown oracle build {name!r}")
    _build = name
    _lib = None


def lib():
    global _lib
    if _lib is None and _build in _libs:
        _lib = _libs[_build]
    if _lib is None:
        path = BUILDS[_build]
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_geometry_from_mesh.restype = _vp
        L.or_model_create.restype = _vp
        L.or_model_field.restype = ctypes.POINTER(ctypes.c_double)
        L.or_model_field.argtypes = [_vp, ctypes.c_char_p]
        L.or_geometry_field.restype = ctypes.POINTER(ctypes.c_double)
        L.or_geometry_field.argtypes = [_vp, ctypes.c_char_p]
        L.or_model_set_scalar.argtypes = [_vp, ctypes.c_char_p, ctypes.c_double]
        L.or_model_get_scalar.argtypes = [_vp, ctypes.c_char_p]
        L.or_model_get_scalar.restype = ctypes.c_double
        L.or_covariant_arc_min.restype = ctypes.c_double
        L.or_covariant_arc_min.argtypes = [_vp]
        L.or_entropy.restype = ctypes.c_double
        L.or_entropy.argtypes = [_vp]
        L.or_mass.restype = ctypes.c_double
        L.or_mass.argtypes = [_vp, ctypes.c_int]
        L.or_rk3_steps.argtypes = [_vp, ctypes.c_int, ctypes.c_double]
        for f in ("or_model_free", "or_pre_tendency", "or_set_boundary_condition", "or_source", "or_riemann",
                  "or_flux", "or_calculate_tendency", "or_euler_diagnostics", "or_euler_primitive",
                  "or_euler_entropy_vars", "or_euler_set_prescribed", "or_geometry_free"):
            getattr(L, f).argtypes = [_vp]
        L.or_update_grk3.argtypes = [_vp, ctypes.c_int]
        for f in ("or_model_set_solution", "or_sw_set_topography", "or_lsw_set_bathymetry", "or_lsw_set_coriolis",
                  "or_euler_set_gravity", "or_euler_set_drag"):
            getattr(L, f).argtypes = [_vp, _vp]
        _libs[_build] = L
        _lib = L
    return _lib


def P(a):
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], "oracle takes C-contiguous numpy arrays"
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n):
    lib().or_set_threads(int(n))


# ---- a1 ----
def legendre_gauss(N):
    x, w = np.zeros(N + 1), np.zeros(N + 1)
    lib().or_legendre_gauss(N, P(x), P(w))
    return x, w


def legendre_gauss_lobatto(N):
    x, w = np.zeros(N + 1), np.zeros(N + 1)
    lib().or_legendre_gauss_lobatto(N, P(x), P(w))
    return x, w


def lagrange_init(N, ctrl=GAUSS, M=None, tgt=UNIFORM):
    M = N if M is None else M
    p, q = N + 1, M + 1
    out = dict(controlPoints=np.zeros(p), targetPoints=np.zeros(q), bWeights=np.zeros(p), qWeights=np.zeros(p),
               iMatrix=np.zeros((q, p)), dMatrix=np.zeros((p, p)), dgMatrix=np.zeros((p, p)), bMatrix=np.zeros((2, p)))
    lib().or_lagrange_init(N, ctrl, M, tgt, *[P(out[k]) for k in ("controlPoints", "targetPoints", "bWeights",
                                                                  "qWeights", "iMatrix", "dMatrix", "dgMatrix",
                                                                  "bMatrix")])
    return out


# ---- operators (numpy in, numpy out) ----
def _shape_b(f, N, nVar, ncomp):
    nEl = f.shape[0]
    return (nEl, 4, nVar, N + 1) + ((ncomp,) if ncomp == 2 else (2, 2) if ncomp == 4 else ())


def scalar_boundary_interp(N, bM, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, 4, nVar, N + 1))
    lib().or_scalar_boundary_interp_2d(N, P(bM), P(f), P(out), nVar, nEl)
    return out


def vector_boundary_interp(N, bM, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, 4, nVar, N + 1, 2))
    lib().or_vector_boundary_interp_2d(N, P(bM), P(f), P(out), nVar, nEl)
    return out


def tensor_boundary_interp(N, bM, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, 4, nVar, N + 1, 2, 2))
    lib().or_tensor_boundary_interp_2d(N, P(bM), P(f), P(out), nVar, nEl)
    return out


def scalar_grid_interp(N, M, iM, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, nVar, M + 1, M + 1))
    lib().or_scalar_grid_interp_2d(N, M, P(iM), P(f), P(out), nVar, nEl)
    return out


def vector_grid_interp(N, M, iM, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, nVar, M + 1, M + 1, 2))
    lib().or_vector_grid_interp_2d(N, M, P(iM), P(f), P(out), nVar, nEl)
    return out


def vector_gradient(N, dM, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, nVar, N + 1, N + 1, 2, 2))
    lib().or_vector_gradient_2d(N, P(dM), P(f), P(out), nVar, nEl)
    return out


def vector_divergence(N, dM, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, nVar, N + 1, N + 1))
    lib().or_vector_divergence_2d(N, P(dM), P(f), P(out), nVar, nEl)
    return out


def vector_dg_divergence(N, dgM, bM, qW, f, bF):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, nVar, N + 1, N + 1))
    lib().or_vector_dg_divergence_2d(N, P(dgM), P(bM), P(qW), P(f), P(bF), P(out), nVar, nEl)
    return out


def determinant(N, t):
    nEl, nVar = t.shape[0], t.shape[1]
    out = np.zeros((nEl, nVar, N + 1, N + 1))
    lib().or_determinant_tensor_2d(N, P(t), P(out), nVar, nEl)
    return out


def side_exchange_scalar(N, sideInfo, b):
    nEl, nVar = b.shape[0], b.shape[2]
    ext = np.zeros_like(b)
    lib().or_side_exchange_scalar_2d(N, nVar, nEl, P(sideInfo), P(b), P(ext))
    return ext


def side_exchange_vector(N, sideInfo, b):
    nEl, nVar = b.shape[0], b.shape[2]
    ext = np.zeros_like(b)
    lib().or_side_exchange_vector_2d(N, nVar, nEl, P(sideInfo), P(b), P(ext))
    return ext


def gradient_sf(N, dM, dsdx, J, f):
    nEl, nVar = f.shape[0], f.shape[1]
    out = np.zeros((nEl, nVar, N + 1, N + 1, 2))
    lib().or_gradient_sf_2d(N, P(dM), P(dsdx), P(J), P(f), P(out), nVar, nEl)
    return out


def jacobian_weight(N, J, f):
    f = f.copy()
    lib().or_jacobian_weight_2d(N, P(J), P(f), f.shape[1], f.shape[0])
    return f


def contravariant_projection(N, dsdx, f):
    f = f.copy()
    lib().or_contravariant_projection_2d(N, P(dsdx), P(f), f.shape[1], f.shape[0])
    return f


# ---- mesh ----
def recalculate_flip(sideInfo, globalNodeIDs, nGeo):
    si = np.ascontiguousarray(sideInfo, np.int32).copy()
    gn = np.ascontiguousarray(globalNodeIDs, np.int32)  # named: must outlive the call
    lib().or_recalculate_flip(si.shape[0], nGeo, P(si), P(gn))
    return si


def hopr_to_mesh2d(nGeo, sideInfo3D, nodeCoords3D, globalNodeIDs3D):
    si3 = np.ascontiguousarray(sideInfo3D, np.int32)
    nEl, G = si3.shape[0] // 6, nGeo + 1
    si = np.zeros((nEl, 4, 5), np.int32)
    nc = np.zeros((nEl, G, G, 2))
    gn = np.zeros((nEl, G, G), np.int32)
    nc3 = np.ascontiguousarray(nodeCoords3D, np.float64)  # named: must outlive the call
    gn3 = np.ascontiguousarray(globalNodeIDs3D, np.int32)
    lib().or_hopr_to_mesh2d(nEl, nGeo, P(si3), P(nc3), P(gn3), P(si), P(nc), P(gn))
    return si, nc, gn


# ---- geometry ----
class Geometry:
    """or_geometry (SEMQuad) with numpy views of its arrays."""

    def __init__(self, N, nodeCoords, nGeo, ctrl=GAUSS):
        nc = np.ascontiguousarray(nodeCoords, np.float64)
        self.N, self.nElem = N, nc.shape[0]
        self._L = lib()
        self._p = ctypes.c_void_p(self._L.or_geometry_from_mesh(N, ctrl, nGeo, self.nElem, P(nc)))
        p, n = N + 1, self.nElem
        shapes = dict(x=(n, p, p, 2), xb=(n, 4, p, 2), dxds=(n, p, p, 2, 2), dxdsb=(n, 4, p, 2, 2),
                      dsdx=(n, p, p, 2, 2), dsdxb=(n, 4, p, 2, 2), nHatb=(n, 4, p, 2), nScaleb=(n, 4, p),
                      J=(n, p, p), Jb=(n, 4, p))
        for k, shp in shapes.items():
            ptr = self._L.or_geometry_field(self._p, k.encode())
            setattr(self, k, np.ctypeslib.as_array(ptr, shape=shp))

    def covariant_arc_min(self):
        return self._L.or_covariant_arc_min(self._p)

    def __del__(self):
        if getattr(self, "_p", None) is not None:
            self._L.or_geometry_free(self._p)
            self._p = None


class Model:
    """or_model: reference-ordered CPU Model2D with the LSW / SW / Euler plug-ins."""

    def __init__(self, kind, N, sideInfo, geom):
        self.kind, self.N, self.geom = kind, N, geom
        self.sideInfo = np.ascontiguousarray(sideInfo, np.int32)
        self.nElem = self.sideInfo.shape[0]
        self.nVar = 4 if kind == EULER else 3
        self._L = geom._L
        self._p = ctypes.c_void_p(self._L.or_model_create(kind, N, self.nElem, P(self.sideInfo), geom._p))

    def __del__(self):
        if getattr(self, "_p", None) is not None:
            self._L.or_model_free(self._p)
            self._p = None

    def field(self, name, shape):
        return np.ctypeslib.as_array(self._L.or_model_field(self._p, name.encode()), shape=shape)

    def _scal(self, x, nvar=1):
        p, n = self.N + 1, self.nElem
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), (n, nvar, p, p)) if np.ndim(x) <= 1
                                 else np.asarray(x, np.float64).reshape(n, nvar, p, p))
        return a

    @property
    def interior_shape(self):
        p = self.N + 1
        return (self.nElem, self.nVar, p, p)

    @property
    def boundary_shape(self):
        return (self.nElem, 4, self.nVar, self.N + 1)

    def set_scalar(self, name, v):
        self._L.or_model_set_scalar(self._p, name.encode(), float(v))

    def get_scalar(self, name):
        return self._L.or_model_get_scalar(self._p, name.encode())

    def set_solution(self, interior):
        a = np.ascontiguousarray(np.asarray(interior, np.float64).reshape(self.interior_shape))
        self._L.or_model_set_solution(self._p, P(a))

    def set_topography(self, b):
        a = self._scal(b)  # named: must outlive the call
        self._L.or_sw_set_topography(self._p, P(a))

    def set_bathymetry(self, H):
        a = self._scal(H)  # named: must outlive the call
        self._L.or_lsw_set_bathymetry(self._p, P(a))

    def set_coriolis(self, f):
        a = self._scal(f)  # named: must outlive the call
        self._L.or_lsw_set_coriolis(self._p, P(a))

    def set_gravity(self, phi):
        a = self._scal(phi)  # named: must outlive the call
        self._L.or_euler_set_gravity(self._p, P(a))

    def set_drag(self, cd):
        a = self._scal(cd)  # named: must outlive the call
        self._L.or_euler_set_drag(self._p, P(a))

    def set_prescribed(self):
        self._L.or_euler_set_prescribed(self._p)

    @property
    def solution(self):
        return self.field("sol", self.interior_shape).copy()

    def tendency(self):
        self._L.or_calculate_tendency(self._p)
        return self.field("dSdt", self.interior_shape).copy()

    def rk3_steps(self, n, dt):
        self._L.or_rk3_steps(self._p, int(n), float(dt))

    def entropy(self):
        return self._L.or_entropy(self._p)

    def mass(self, var):
        return self._L.or_mass(self._p, int(var))
