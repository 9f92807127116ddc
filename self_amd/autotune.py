"""Tile autotuner for the fused stage kernel (no reference counterpart).

The stage kernel (affine metrics, RK update) of nonlinear SW at N = 3, 5, 7 and of linear SW
and Euler at N = 7 exists in tile variants: work-items per block (64, 128 or 256), an element
order that alternates between stages (each stage starts on the data the previous one left in
the Infinity Cache), an XCD-contiguous element order and an occupancy target for the
register allocator.  `tune` times every variant on a mesh of the given size, checks that the
variants agree to round-off and returns the fastest.
`python -m self_amd.autotune` writes the table `self_amd/tuning/sw_tiles.json`, which `apply`
(called when the package is imported) loads into the library.
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

from ._lib import call

KINDS = {"lsw": 1, "sw": 2, "euler": 3}
MODEL_SW = KINDS["sw"]
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "sw_tiles.json")


def count(N, kind=MODEL_SW):
    c = ctypes.c_int()
    call("selfamd_tile_count", kind, N, ctypes.byref(c))
    return c.value


def info(N, v, kind=MODEL_SW):
    t, w, x, o = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("selfamd_tile_info", kind, N, v, ctypes.byref(t), ctypes.byref(w), ctypes.byref(x), ctypes.byref(o))
    return {"threads": t.value, "alternate": bool(w.value), "xcd_order": bool(x.value), "min_waves": o.value}


def set_tile(N, v, kind=MODEL_SW):
    call("selfamd_tile_set", kind, N, int(v))


def get_tile(N, kind=MODEL_SW):
    v = ctypes.c_int()
    call("selfamd_tile_get", kind, N, ctypes.byref(v))
    return v.value


def _model(kind, N, nx):
    """A factory for a model of the given kind on an nx x nx mesh with its seeded workload,
    and a stable dt."""
    import self_amd
    from self_amd import workloads
    if kind == KINDS["euler"]:
        mesh = self_amd.Mesh2D.structured(nx, nx, bcid=self_amd.BC_NONORMALFLOW)
    else:
        mesh = self_amd.Mesh2D.structured(nx, nx, periodic=(True, True))
    interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
    geom = self_amd.SEMQuad(interp, mesh)
    x = geom.x.to_host()
    arc = geom.CovariantArcMin()
    if kind == KINDS["sw"]:
        sol, topo = workloads.nonlinear_sw(x, interp.controlPoints, interp.qWeights, seed=0)

        def make():
            m = self_amd.ShallowWater(mesh, geom)
            m.SetTopography(topo)
            m.SetSolution(sol)
            return m
        return make, workloads.cfl_dt(arc, 1.1, 0.5)
    if kind == KINDS["lsw"]:
        sol = workloads.gravity_wave_lsw(x, seed=0)
        H = np.ones(x.shape[:3])

        def make():
            m = self_amd.LinearShallowWater(mesh, geom)
            m.SetBathymetry(H)
            m.SetSolution(sol)
            return m
        return make, workloads.cfl_dt(arc, 1.0, 0.5)
    sol, phi = workloads.euler_hydrostatic(x, seed=0)

    def make():
        m = self_amd.CompressibleIdealGas2D(mesh, geom)
        m.SetGravity(phi)
        m.SetSolution(sol)
        return m
    return make, workloads.cfl_dt(arc, 1.2, 0.5)


def tune(N, nx, kind=MODEL_SW, steps=10, warmup=3, repeats=3, log=None):
    """Time every variant (and the default tile, -1) of the stage kernel of `kind` on an
    nx x nx mesh at degree N; returns {"best": v, "timings": {v: us per launch}}."""
    make, dt = _model(kind, N, nx)
    before = get_tile(N, kind)
    timings, ref = {}, None
    # bring the GPU to its steady clock first (the first ~100 steps run slower), so that the
    # variants timed first are not penalised
    m = make()
    m.ForwardStep(100, dt)
    del m
    try:
        for v in [-1] + list(range(count(N, kind))):
            set_tile(N, v, kind)
            m = make()
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
                log(f"kind {kind} N={N} {nx}x{nx} variant {v:2d} {info(N, v, kind) if v >= 0 else 'default'}: "
                    f"{timings[v]:.2f} us/launch")
            del m
    finally:
        set_tile(N, before, kind)
    return {"best": min(timings, key=timings.get), "timings": timings}


def apply(path=TABLE):
    """Load a tile table into the library (missing file: keep the defaults).  Keys are
    "<kind>:<N>" (a bare "<N>" means nonlinear SW)."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        table = json.load(f)
    for key, ent in table.get("entries", {}).items():
        kname, _, n = key.rpartition(":")
        kind, N, v = KINDS.get(kname or "sw"), int(n), int(ent["variant"])
        if kind and count(N, kind) > v:
            set_tile(N, v, kind)
    return table


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--kinds", nargs="+", default=["sw", "euler", "lsw"], choices=sorted(KINDS))
    ap.add_argument("--elements", type=int, default=65536, help="elements of the tuning mesh (square)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args(argv)
    nx = int(round(np.sqrt(a.elements)))
    table = {"mesh": f"{nx}x{nx}", "entries": {}}
    for kname in a.kinds:
        kind = KINDS[kname]
        for N in (3, 5, 7):
            if not count(N, kind):
                continue
            r = tune(N, nx, kind, steps=a.steps, log=lambda m: print(m, flush=True))
            v = r["best"]
            ent = {"variant": v, "us_per_launch": r["timings"][v], "default_us_per_launch": r["timings"][-1]}
            if v >= 0:
                ent.update(info(N, v, kind))
            table["entries"][f"{kname}:{N}"] = ent
            print(f"{kname} N={N}: best variant {v} ({ent['us_per_launch']:.2f} us vs default "
                  f"{ent['default_us_per_launch']:.2f} us)", flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
