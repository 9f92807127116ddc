"""Benchmark: nonlinear shallow water 2D, 256x256 elements, N = 7, FP64, periodic, Rusanov
flux, low-storage RK3 -- the fused DGSEM tendency + RK3 engine on one MI355X.

One step = one full RK3 step (3 fused stage launches) over the whole mesh with the state
resident in HBM.  value = DOF-updates/s = DOF * steps / seconds, timed with hipEvents on the
model's stream around graph-replayed steps (warm-up excluded).

roofline: the fused stage kernel is HBM-bound; its algorithmic traffic is 120 bytes per
DOF-update (FP64 state + RK register read/write, one face-trace write and one neighbour-trace
read per stage, one static topography field; affine metrics), i.e. 40 bytes per DOF per stage
launch.  achieved = 40 B * DOF / average launch duration; peak = 8 TB/s (MI355X HBM3E).

cpu_baseline: the oracle (faithful C restatement of the reference Fortran CPU path) timed on
the host for one RK3 step of the same workload with all available host threads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = ("DOF-updates/sec + % HBM roofline, SW2D N=7 256² elems FP64; "
          "closure speedup at equal L2 error")
HBM_PEAK_GBS = 8000.0
BYTES_PER_DOF_UPDATE = 120.0  # SW, N = 7, one static field


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--N", type=int, default=7)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=100, help="RK3 steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for HBM counter runs: 3 dispatches of a 1 GiB FP64 copy (calibration), "
                         "then 2 RK3 steps (6 stage launches) of the workload; prints no JSON")
    return ap.parse_args()


def traffic_from_profiles(root):
    """Measured HBM bytes per stage launch, if a counter summary has been committed."""
    path = os.path.join(root, "profiles", "sw2d_stage_256x256_N7_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def traffic_probe(self_amd, model, dt):
    """Known-traffic copy kernel (1 GiB read + 1 GiB written per dispatch), then the stage
    kernel, so FETCH_SIZE / WRITE_SIZE of the stage launches can be scaled by what the same
    counters read for the copy."""
    from self_amd._lib import call
    n = 1 << 27
    print(f"probe: copy {8 * n} bytes x3", flush=True)
    a = self_amd.DeviceArray((n,), zero=True)
    b = self_amd.DeviceArray((n,), zero=True)
    for _ in range(3):
        call("selfamd_copy_f64", a.ptr, b.ptr, n, None)
    self_amd.synchronize()
    del a, b
    print("probe: 2 RK3 steps", flush=True)
    model.ForwardStep(2, dt)
    self_amd.synchronize()
    print("probe: done", flush=True)
    return 0


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        return 2
    import self_amd
    from self_amd import workloads

    root = os.path.dirname(os.path.abspath(__file__))
    N, nx = a.N, a.nx
    mesh = self_amd.Mesh2D.structured(nx, nx, periodic=(True, True))
    interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
    geom = self_amd.SEMQuad(interp, mesh)
    x = geom.x.to_host()
    sol, topo = workloads.nonlinear_sw(x, interp.controlPoints, interp.qWeights, seed=a.seed)
    model = self_amd.ShallowWater(mesh, geom)
    model.SetTopography(topo)
    model.SetSolution(sol)
    H = sol[:, 2]
    speed = float((np.sqrt(sol[:, 0] ** 2 + sol[:, 1] ** 2) / H + np.sqrt(H)).max())
    dt = workloads.cfl_dt(geom.CovariantArcMin(), speed, 0.5)
    dof = mesh.nElem * (N + 1) ** 2 * 3

    if a.traffic_probe:
        return traffic_probe(self_amd, model, dt)

    model.ForwardStep(a.warmup, dt)
    self_amd.synchronize()
    t0 = time.perf_counter()
    ms = model.time_steps(a.steps, dt)
    self_amd.synchronize()
    wall = time.perf_counter() - t0
    s = model.solution
    if not np.isfinite(s).all():
        print("non-finite state after the timed steps", file=sys.stderr)
        return 1
    sec = ms / 1e3
    value = dof * a.steps / sec
    launch_s = sec / (3 * a.steps)
    achieved = BYTES_PER_DOF_UPDATE / 3.0 * dof / launch_s / 1e9

    cpu = None
    if not a.no_cpu_baseline:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        og = oracle.Geometry(N, mesh.nodeCoords, 1)
        om = oracle.Model(oracle.SW, N, mesh.sideInfo, og)
        om.set_topography(topo)
        om.set_solution(sol)
        oracle.set_threads(threads)
        c0 = time.perf_counter()
        om.rk3_steps(a.cpu_steps, dt)
        csec = time.perf_counter() - c0
        oracle.set_threads(1)
        cpu = {"value": dof * a.cpu_steps / csec, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
               "sample": f"{a.cpu_steps} RK3 step(s) of the same {nx}x{nx} N={N} workload, "
                         f"OpenMP over elements, {csec:.2f} s"}

    out = {
        "metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"nonlinear shallow water 2D, {nx}x{nx} quad elements periodic, N={N} Gauss, "
                               f"FP64, Rusanov flux, low-storage RK3", "elements": mesh.nElem, "N": N,
                   "dof": dof, "dt": dt, "wall_s": wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(root),
                     "kernel": "k_fused<PhysSW, 7, 0, false, 64>", "avg_launch_us": launch_s * 1e6,
                     "bytes_per_launch_alg": BYTES_PER_DOF_UPDATE / 3.0 * dof},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
