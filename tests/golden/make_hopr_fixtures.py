"""Write the HOPr mesh fixtures used by the connectivity tests.

Reads the reference's shipped HOPr mesh files (etc/mesh/*/..._mesh.h5, data the
reference's own example programs load) with the h5dump tool and stores the datasets that
Read_HOPr_Mesh2D consumes (src/SELF_Mesh.f90:852-1016) as JSON: ElemInfo, SideInfo,
NodeCoords, GlobalNodeIDs, BCType and the Ngeo/nElems attributes.  The Block2D mesh file
itself (18 KB) is also copied unchanged to hopr_Block2D_mesh.h5, so that the HDF5 readers can
be run where the reference checkout is absent.  Only this script reads the reference
checkout, and only in the build container; the fixtures travel.

Usage: python tests/golden/make_hopr_fixtures.py [/root/reference]
"""
import json
import os
import re
import shutil
import subprocess
import sys

H5DUMP = "/opt/conda/bin/h5dump"
MESHES = {
    "hopr_block2d": "etc/mesh/Block2D/Block2D_mesh.h5",
    "hopr_geophysical_block2d_medium": "etc/mesh/GeophysicalBlock2DMedium/Block2D_mesh.h5",
    "hopr_circle": "etc/mesh/Circle/Circle_mesh.h5",
}


def _dump(path, kind, name):
    # %.17g: every double printed with enough digits to read back bit-exactly
    out = subprocess.run([H5DUMP, "-y", "-w", "0", "-m", "%.17g", kind, name, path], check=True,
                         capture_output=True, text=True).stdout
    dims = re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", out).group(1)
    shape = [int(d) for d in dims.split(",")]
    body = out[out.index("DATA {") + 6: out.rindex("}", 0, out.rindex("}"))]
    body = body[: body.rindex("}")]
    vals = [v for v in re.split(r"[,\s]+", body.strip()) if v]
    return shape, vals


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for key, rel in MESHES.items():
        path = os.path.join(ref, rel)
        fx = {"source": rel}
        for attr in ("Ngeo", "nElems", "nBCs", "nUniqueSides"):
            _, v = _dump(path, "-a", "/" + attr)
            fx[attr] = int(v[0])
        for ds, conv in (("ElemInfo", int), ("SideInfo", int), ("GlobalNodeIDs", int), ("NodeCoords", float),
                         ("BCType", int)):
            shape, v = _dump(path, "-d", "/" + ds)
            fx[ds] = {"shape": shape, "data": [conv(x) for x in v]}
        with open(os.path.join(here, key + ".json"), "w") as f:
            json.dump(fx, f, separators=(",", ":"))
        print(key, {k: (fx[k]["shape"] if isinstance(fx[k], dict) else fx[k]) for k in fx})
    shutil.copyfile(os.path.join(ref, MESHES["hopr_block2d"]), os.path.join(here, "hopr_Block2D_mesh.h5"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
