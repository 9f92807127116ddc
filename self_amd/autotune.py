"""Tile autotuner for the nonlinear-SW stage kernel (no reference counterpart).

The stage kernel at N = 3, 5, 7 exists in tile variants: work-items per block (64, 128 or
256), an element order that alternates between stages (each stage starts on the data the
previous one left in the Infinity Cache) and an XCD-contiguous element order.  `tune` times
every variant on a periodic SW mesh of the given size, checks that the variants agree to
round-off and returns the fastest.
`python -m self_amd.autotune` writes the table `self_amd/tuning/sw_tiles.json`, which
`apply` (called when the package is imported) loads into the library.
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

from ._lib import call

MODEL_SW = 2
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "sw_tiles.json")


def count(N):
    c = ctypes.c_int()
    call("selfamd_tile_count", MODEL_SW, N, ctypes.byref(c))
    return c.value


def info(N, v):
    t, w, x = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("selfamd_tile_info", MODEL_SW, N, v, ctypes.byref(t), ctypes.byref(w), ctypes.byref(x))
    return {"threads": t.value, "alternate": bool(w.value), "xcd_order": bool(x.value)}


def set_tile(N, v):
    call("selfamd_tile_set", MODEL_SW, N, int(v))


def get_tile(N):
    v = ctypes.c_int()
    call("selfamd_tile_get", MODEL_SW, N, ctypes.byref(v))
    return v.value


def tune(N, nx, steps=10, warmup=3, repeats=3, log=None):
    """Time every variant (and the default tile, -1) of the SW stage kernel on an nx x nx
    periodic mesh at degree N; returns {"best": v, "timings": {v: us per launch}}."""
    import self_amd
    from self_amd import workloads
    mesh = self_amd.Mesh2D.structured(nx, nx, periodic=(True, True))
    interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
    geom = self_amd.SEMQuad(interp, mesh)
    sol, topo = workloads.nonlinear_sw(geom.x.to_host(), interp.controlPoints, interp.qWeights, seed=0)
    dt = workloads.cfl_dt(geom.CovariantArcMin(), 1.1, 0.5)
    before = get_tile(N)
    timings, ref = {}, None
    try:
        for v in [-1] + list(range(count(N))):
            set_tile(N, v)
            m = self_amd.ShallowWater(mesh, geom)
            m.SetTopography(topo)
            m.SetSolution(sol)
            m.ForwardStep(warmup, dt)
            best = min(m.time_steps(steps, dt) for _ in range(repeats))
            s = m.solution
            if ref is None:
                ref = s
            else:
                # variants agree to round-off (face-line sums may be split across lane pairs)
                err = np.max(np.abs(s - ref)) / np.max(np.abs(ref))
                if not err <= 1e-12:
                    raise RuntimeError(f"tile variant {v} at N={N} changed the result (rel {err:.2e})")
            timings[v] = best / (3 * steps) * 1e3
            if log:
                log(f"N={N} {nx}x{nx} variant {v:2d} {info(N, v) if v >= 0 else 'default'}: "
                    f"{timings[v]:.2f} us/launch")
            del m
    finally:
        set_tile(N, before)
    return {"best": min(timings, key=timings.get), "timings": timings}


def apply(path=TABLE):
    """Load a tile table into the library (missing file: keep the defaults)."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        table = json.load(f)
    for key, ent in table.get("entries", {}).items():
        N, v = int(key), int(ent["variant"])
        if count(N) > v:
            set_tile(N, v)
    return table


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--N", type=int, nargs="+", default=[3, 5, 7])
    ap.add_argument("--elements", type=int, default=65536, help="elements of the tuning mesh (square)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args(argv)
    nx = int(round(np.sqrt(a.elements)))
    table = {"kind": "sw", "mesh": f"{nx}x{nx} periodic", "entries": {}}
    for N in a.N:
        r = tune(N, nx, steps=a.steps, log=lambda m: print(m, flush=True))
        v = r["best"]
        ent = {"variant": v, "us_per_launch": r["timings"][v], "default_us_per_launch": r["timings"][-1]}
        if v >= 0:
            ent.update(info(N, v))
        table["entries"][str(N)] = ent
        print(f"N={N}: best variant {v} ({ent['us_per_launch']:.2f} us vs default "
              f"{ent['default_us_per_launch']:.2f} us)", flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
