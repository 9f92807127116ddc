"""HOPr mesh files read through libself_amd_io.so (HDF5 C library) against the committed
fixtures, which tests/golden/make_hopr_fixtures.py wrote from the same files with h5dump
(doubles printed with 17 digits, so the values compare bit-exactly).  Runs where the
reference checkout with its mesh files is present (the build container); the Block2D cases
also run elsewhere on the committed copy of that mesh file (tests/golden/hopr_Block2D_mesh.h5)."""
import ctypes
import json
import os

import numpy as np
import pytest

import self_amd
from self_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
FIXTURES = ["hopr_block2d", "hopr_geophysical_block2d_medium", "hopr_circle"]

pytestmark = pytest.mark.skipif(not _lib.io_available(), reason="libself_amd_io.so not built or HDF5 missing")


def fixture(name):
    with open(os.path.join(HERE, "golden", name + ".json")) as f:
        return json.load(f)


BLOCK2D_COPY = os.path.join(HERE, "golden", "hopr_Block2D_mesh.h5")


def mesh_path(fx):
    p = os.path.join(REF, fx["source"])
    if not os.path.exists(p) and fx["source"] == "etc/mesh/Block2D/Block2D_mesh.h5":
        p = BLOCK2D_COPY
    if not os.path.exists(p):
        pytest.skip(f"mesh file {fx['source']} not present (reference checkout absent)")
    return os.fsencode(p)


def read(path, name, dtype):
    rank, dims = ctypes.c_int32(), (ctypes.c_int64 * 5)()
    _lib.call_io("selfamd_hopr_dataset_dims", path, name.encode(), ctypes.byref(rank), dims)
    a = np.empty(tuple(dims[k] for k in range(rank.value)), dtype)
    fn = "selfamd_hopr_read_int" if dtype == np.int32 else "selfamd_hopr_read_double"
    _lib.call_io(fn, path, name.encode(), a.ctypes.data, a.size)
    return a


@pytest.mark.parametrize("name", FIXTURES)
def test_datasets_and_attributes_match_fixtures(name):
    fx = fixture(name)
    p = mesh_path(fx)
    for attr in ("Ngeo", "nElems", "nBCs", "nUniqueSides"):
        v = ctypes.c_int32()
        _lib.call_io("selfamd_hopr_attribute_int", p, attr.encode(), ctypes.byref(v))
        assert v.value == fx[attr], attr
    for ds, dtype in (("ElemInfo", np.int32), ("SideInfo", np.int32), ("GlobalNodeIDs", np.int32),
                      ("BCType", np.int32), ("NodeCoords", np.float64)):
        a = read(p, ds, dtype)
        assert list(a.shape) == fx[ds]["shape"], ds
        assert np.array_equal(a.ravel(), np.array(fx[ds]["data"], dtype)), ds


@pytest.mark.parametrize("name", FIXTURES)
def test_read_hopr_equals_conversion_of_fixture(name):
    fx = fixture(name)
    p = mesh_path(fx)
    a = self_amd.Mesh2D.read_hopr(os.fsdecode(p))
    b = self_amd.Mesh2D.from_hopr(fx["Ngeo"], np.array(fx["SideInfo"]["data"], np.int32).reshape(fx["SideInfo"]["shape"]),
                                  np.array(fx["NodeCoords"]["data"]).reshape(fx["NodeCoords"]["shape"]),
                                  np.array(fx["GlobalNodeIDs"]["data"], np.int32))
    assert a.nElem == fx["nElems"] and a.nGeo == fx["Ngeo"]
    assert np.array_equal(a.sideInfo, b.sideInfo)
    assert np.array_equal(a.nodeCoords, b.nodeCoords)
    assert np.array_equal(a.globalNodeIDs, b.globalNodeIDs)


def test_errors_are_reported(tmp_path):
    v = ctypes.c_int32()
    with pytest.raises(_lib.SelfAmdError, match="cannot open"):
        _lib.call_io("selfamd_hopr_attribute_int", os.fsencode(str(tmp_path / "none.h5")), b"Ngeo", ctypes.byref(v))
    fx = fixture("hopr_block2d")
    p = mesh_path(fx)
    with pytest.raises(_lib.SelfAmdError, match="no dataset"):
        read(p, "NoSuchDataset", np.int32)
    buf = np.empty(3, np.int32)
    with pytest.raises(_lib.SelfAmdError, match="elements, buffer"):
        _lib.call_io("selfamd_hopr_read_int", p, b"SideInfo", buf.ctypes.data, buf.size)


def test_committed_block2d_copy_matches_fixture():
    """The committed copy of the Block2D mesh file reads as its JSON fixture (no reference needed)."""
    fx = fixture("hopr_block2d")
    p = os.fsencode(BLOCK2D_COPY)
    for ds, dtype in (("SideInfo", np.int32), ("GlobalNodeIDs", np.int32), ("NodeCoords", np.float64)):
        assert np.array_equal(read(p, ds, dtype).ravel(), np.array(fx[ds]["data"], dtype)), ds
    m = self_amd.Mesh2D.read_hopr(BLOCK2D_COPY)
    assert m.nElem == fx["nElems"] == 25 and m.nGeo == fx["Ngeo"] == 1
