"""Model pickup files (WriteModel / ReadModel, src/SELF_Model2D.f90:861-1099) through
libself_amd_io.so on the host: the reference's dataset layout, the attribute N, and an exact
round trip of the solution."""
import ctypes
import os

import numpy as np
import pytest

from self_amd import _lib

pytestmark = pytest.mark.skipif(not _lib.io_available(), reason="libself_amd_io.so not built or HDF5 missing")


def dims(path, name):
    rank, d = ctypes.c_int32(), (ctypes.c_int64 * 5)()
    _lib.call_io("selfamd_hopr_dataset_dims", path, name.encode(), ctypes.byref(rank), d)
    return tuple(d[k] for k in range(rank.value))


def write(path, N, nVar, nEl, rng):
    P = N + 1
    arrays = [rng.normal(size=s) for s in ((P,), (P,), (P, P), (P, P), (2, P), (nEl, nVar, P, P),
                                           (nEl, 4, nVar, P), (nEl, P, P, 2), (nEl, 4, P, 2))]
    _lib.call_io("selfamd_pickup_write", path, N, nVar, nEl, *[a.ctypes.data for a in arrays])
    return arrays


def test_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    N, nVar, nEl = 5, 4, 7
    P = N + 1
    path = os.fsencode(str(tmp_path / "solution.pickup.h5"))
    arrays = write(path, N, nVar, nEl, rng)
    v = ctypes.c_int32()
    _lib.call_io("selfamd_hopr_attribute_int", path, b"N", ctypes.byref(v))
    assert v.value == N
    # Fortran shapes with the dimensions reversed, as the reference's HDF5 writes store them
    assert dims(path, "/quadrature/xi") == (P,)
    assert dims(path, "/quadrature/dgmatrix") == (P, P)
    assert dims(path, "/quadrature/bmatrix") == (2, P)
    assert dims(path, "/state/interior/solution") == (nEl, nVar, P, P)
    assert dims(path, "/state/boundary/solution") == (nEl, 4, nVar, P)
    assert dims(path, "/mesh/interior/x") == (nEl, 1, P, P, 2)
    assert dims(path, "/mesh/boundary/x") == (nEl, 4, 1, P, 2)
    sol = np.empty((nEl, nVar, P, P))
    _lib.call_io("selfamd_pickup_read", path, N, nVar, nEl, sol.ctypes.data)
    assert np.array_equal(sol, arrays[5])
    xb = np.empty((nEl, 4, P, 2))
    _lib.call_io("selfamd_hopr_read_double", path, b"/mesh/boundary/x", xb.ctypes.data, xb.size)
    assert np.array_equal(xb, arrays[8])


def test_register_datasets(tmp_path):
    """The rest of Write_Model2D's datasets (src/SELF_Model2D.f90:981-990): fluxDivergence,
    flux and solutionGradient (interior and boundary), in the reference's layout; a null array
    is written as zeros."""
    rng = np.random.default_rng(7)
    N, nVar, nEl = 4, 3, 6
    P = N + 1
    path = os.fsencode(str(tmp_path / "regs.h5"))
    write(path, N, nVar, nEl, rng)
    fdiv = rng.normal(size=(nEl, nVar, P, P))
    fl = rng.normal(size=(nEl, nVar, P, P, 2))
    flb = rng.normal(size=(nEl, 4, nVar, P, 2))
    _lib.call_io("selfamd_pickup_write_registers", path, N, nVar, nEl, fdiv.ctypes.data, fl.ctypes.data,
                 flb.ctypes.data, None, None)
    assert dims(path, "/state/interior/fluxDivergence") == (nEl, nVar, P, P)
    assert dims(path, "/state/interior/flux") == (nEl, nVar, P, P, 2)
    assert dims(path, "/state/boundary/flux") == (nEl, 4, nVar, P, 2)
    assert dims(path, "/state/interior/solutionGradient") == (nEl, nVar, P, P, 2)
    assert dims(path, "/state/boundary/solutionGradient") == (nEl, 4, nVar, P, 2)
    for name, want in (("/state/interior/fluxDivergence", fdiv), ("/state/interior/flux", fl),
                       ("/state/boundary/flux", flb), ("/state/interior/solutionGradient", np.zeros_like(fl))):
        got = np.empty(want.shape)
        _lib.call_io("selfamd_hopr_read_double", path, name.encode(), got.ctypes.data, got.size)
        assert np.array_equal(got, want), name
    # the solution is still there, unchanged
    sol = np.empty((nEl, nVar, P, P))
    _lib.call_io("selfamd_pickup_read", path, N, nVar, nEl, sol.ctypes.data)


def test_read_rejects_mismatches(tmp_path):
    rng = np.random.default_rng(6)
    path = os.fsencode(str(tmp_path / "p.h5"))
    write(path, 3, 3, 4, rng)
    sol = np.empty(3 * 4 * 4 * 5 * 5)
    with pytest.raises(_lib.SelfAmdError, match="does not match"):
        _lib.call_io("selfamd_pickup_read", path, 4, 3, 4, sol.ctypes.data)
    with pytest.raises(_lib.SelfAmdError, match="expected"):
        _lib.call_io("selfamd_pickup_read", path, 3, 3, 5, sol.ctypes.data)
    with pytest.raises(_lib.SelfAmdError, match="cannot open"):
        _lib.call_io("selfamd_pickup_read", os.fsencode(str(tmp_path / "none.h5")), 3, 3, 4, sol.ctypes.data)
