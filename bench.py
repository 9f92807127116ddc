"""Benchmark of the fused DGSEM tendency + low-storage RK3 engine on one MI355X.

Default workload: nonlinear shallow water 2D, 256x256 periodic elements, N = 7 Gauss, FP64,
Rusanov flux, RK3, 100 steps.  One step = one full RK3 step (3 fused stage launches) over
the whole mesh with the state resident in HBM.  value = DOF-updates/s = DOF * steps /
seconds, timed with hipEvents on the model's stream around graph-replayed steps (warm-up
excluded).  Other workloads (--workload): euler512 (compressible Euler, 512x512 walls,
N = 7, gravity, 50 steps), lsw10 (linear SW gravity wave, 10x10 walls, N = 7, 200 steps) and
ensemble (512 batched SW runs, 8x8..64x64 elements, N in {3, 5, 7}).

roofline: the stage kernel is HBM-bound.  Its algorithmic traffic per DOF and RK step is
3 (32 + 64 NT / ((N+1) nVar) + 8 p / nVar) - 8 bytes: state and RK register read and written
each stage (no register read at the first stage), NT face-trace values per face node written
once and read once per stage, p static per-node fields read once per stage.  For SW at N = 7
(NT = nVar = 3, p = 1) that is 120 B per DOF-update, 40 B per DOF per stage launch.
achieved = algorithmic bytes per launch / average launch duration; peak = 8 TB/s.

cpu_baseline: the oracle (C restatement of the reference Fortran CPU path, OpenMP over
elements) timed on this host for a bounded number of RK3 steps of the same workload.
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


def alg_bytes_per_dof_step(N, nVar, NT, p):
    """Algorithmic HBM bytes per DOF and RK3 step (three stage launches)."""
    return 3.0 * (32.0 + 64.0 * NT / ((N + 1) * nVar) + 8.0 * p / nVar) - 8.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed RK3 steps (default per workload)")
    # untimed steps before the timed region (they run after the clock-settling copies)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="sw256", choices=["sw256", "euler512", "lsw10", "ensemble", "closure"])
    ap.add_argument("--closure-runs", type=int, default=8, help="fine training runs of the closure workload")
    ap.add_argument("--closure-horizon", type=int, default=None,
                    help="coarse steps of each closure training run (default: the judged horizon)")
    ap.add_argument("--closure-threads", type=int, default=0,
                    help="host threads building the closure training runs' initial states")
    ap.add_argument("--closure-aposteriori", type=int, default=1000,
                    help="steps of training through the solver (closure workload; 0: a-priori only)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--settle-steps", type=int, default=None,
                    help="untimed steps of the workload after the clock-settling copies, before the "
                         "warm-up steps (default: 100 for single runs, 10 for the ensemble)")
    ap.add_argument("--batch-mode", default="per-group", choices=["per-group", "persistent"],
                    help="ensemble: one launch per degree group and stage, or one persistent launch per stage")
    ap.add_argument("--cpu-steps", type=int, default=None, help="RK3 steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nodal-metrics", action="store_true",
                    help="run the engine on the per-node metric terms (the reference's representation) "
                         "instead of one affine record per element")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for HBM counter runs: 3 dispatches of a 1 GiB FP64 copy (calibration), "
                         "then 2 RK3 steps (6 stage launches) of the workload; prints no JSON")
    return ap.parse_args()


def traffic_from_profiles(root, name):
    """Measured HBM bytes per stage launch, if a counter summary has been committed."""
    path = os.path.join(root, "profiles", name)
    if not name or not os.path.isfile(path):
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


def copy_bandwidth(self_amd, nbytes=1 << 30, reps=20):
    """Achievable streaming rate on this GPU: read + write bytes per second of the library's
    copy kernel over `nbytes` (past the 256 MB Infinity Cache), timed with hipEvents, in GB/s."""
    import ctypes
    from self_amd._lib import call
    n = nbytes // 8
    a = self_amd.DeviceArray((n,), zero=True)
    b = self_amd.DeviceArray((n,), zero=True)
    ms = ctypes.c_double()
    call("selfamd_time_copy_f64", a.ptr, b.ptr, n, reps, ctypes.byref(ms))
    del a, b
    return 2 * nbytes * reps / (ms.value / 1e3) / 1e9


# GPU clocks drop after the GPU idles for about 10 ms or more and take about 25 ms of load to
# come back: the first 20 SW steps after such a gap run at 131-137 us per stage launch, the
# next 20 at ~124, then ~113 (one box, 256^2 N = 7).  Setup leaves the GPU idle, so the
# warm-up steps are queued behind this much streaming work on the model's stream.
SETTLE_COPIES = 250  # 1 GiB copies, ~0.1 s at 5 TB/s
# The streaming copies bring the first 20 steps to ~107-109 us per launch; only the workload
# itself takes them to the sustained ~100-103 us (no other settling work measured did: FP64
# per-operator kernels, copies and kernels mixed).  So after the copies the bench runs this many
# untimed steps of the workload itself, then the caller's warm-up steps, then the timed steps.
SETTLE_STEPS = 100
ENSEMBLE_SETTLE_STEPS = 10  # ~30 ms of the 512-run batch, like 100 steps of the 256^2 run
# After the reported measurement the same steps are timed again this many steps later, as a
# check that the reported window ran at the sustained rate.
SUSTAIN_STEPS = 100


def clock_settle(self_amd, stream, nbytes=1 << 30, reps=SETTLE_COPIES):
    """Queues `reps` copies of `nbytes` on `stream` without waiting; returns the buffers, which
    must stay allocated until the stream has drained."""
    from self_amd._lib import call
    n = nbytes // 8
    a = self_amd.DeviceArray((n,), zero=True)
    b = self_amd.DeviceArray((n,), zero=True)
    for _ in range(reps):
        call("selfamd_copy_f64", a.ptr, b.ptr, n, stream)
    return a, b


def tile_name(self_amd, N, phys="PhysSW", kind=2):
    from self_amd import autotune
    v = autotune.get_tile(N, kind) if autotune.count(N, kind) else -1
    if v < 0:
        t = 64 if N >= 6 else (128 if N >= 4 else 256)
        return f"k_fused<{phys}, {N}, 0, false, {t}, 1, false>"
    i = autotune.info(N, v, kind)
    return f"k_fused<{phys}, {N}, 0, false, {i['threads']}, {i['min_waves']}, {str(i['xcd_order']).lower()}>"


# ---------------------------------------------------------------------------------------
# single-model workloads
# ---------------------------------------------------------------------------------------
def setup_single(self_amd, workloads, name, seed):
    if name == "sw256":
        N, nx = 7, 256
        mesh = self_amd.Mesh2D.structured(nx, nx, periodic=(True, True))
    elif name == "euler512":
        N, nx = 7, 512
        mesh = self_amd.Mesh2D.structured(nx, nx, bcid=self_amd.BC_NONORMALFLOW)
    else:
        N, nx = 7, 10
        mesh = self_amd.Mesh2D.structured(nx, nx, bcid=self_amd.BC_NONORMALFLOW)
    interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
    geom = self_amd.SEMQuad(interp, mesh)
    x = geom.x.to_host()
    arcmin = geom.CovariantArcMin()
    if name == "sw256":
        sol, topo = workloads.nonlinear_sw(x, interp.controlPoints, interp.qWeights, seed=seed)
        model = self_amd.ShallowWater(mesh, geom)
        model.SetTopography(topo)
        H = sol[:, 2]
        speed = float((np.sqrt(sol[:, 0] ** 2 + sol[:, 1] ** 2) / H + np.sqrt(H)).max())
        fields, desc = {"topography": topo}, (f"nonlinear shallow water 2D, {nx}x{nx} quad elements periodic, "
                                              f"N={N} Gauss, FP64, Rusanov flux, low-storage RK3")
        nVar, NT, p, steps = 3, 3, 1, 100
    elif name == "euler512":
        sol, phi = workloads.euler_hydrostatic(x, seed=seed)
        model = self_amd.CompressibleIdealGas2D(mesh, geom)
        model.SetGravity(phi)
        rho, rhoE = sol[:, 2], sol[:, 3]
        pr = 0.4 * (rhoE - 0.5 * (sol[:, 0] ** 2 + sol[:, 1] ** 2) / rho)
        speed = float((np.sqrt(sol[:, 0] ** 2 + sol[:, 1] ** 2) / rho + np.sqrt(1.4 * pr / rho)).max())
        fields, desc = {"potential": phi}, (f"compressible Euler 2D, {nx}x{nx} quad elements, no-normal-flow "
                                            f"walls, N={N} Gauss, FP64, LLF flux, gravity, low-storage RK3")
        nVar, NT, p, steps = 4, 9, 1, 50
    else:
        sol = workloads.gravity_wave_lsw(x, seed=seed)
        model = self_amd.LinearShallowWater(mesh, geom)
        Hb = np.ones((mesh.nElem, N + 1, N + 1))
        model.SetBathymetry(Hb)
        speed = 1.0
        fields, desc = {"bathymetry": Hb}, (f"linear shallow water 2D gravity wave, {nx}x{nx} quad elements, "
                                            f"reflecting walls, N={N} Gauss, FP64, low-storage RK3")
        nVar, NT, p, steps = 3, 3, 1, 200
    model.SetSolution(sol)
    dt = workloads.cfl_dt(arcmin, speed, 0.5)
    return dict(N=N, nx=nx, mesh=mesh, geom=geom, model=model, sol=sol, fields=fields, dt=dt, desc=desc,
                nVar=nVar, NT=NT, p=p, steps=steps)


def host_info():
    """The host's CPU model, core counts and the threads the CPU baseline may use: all cores
    this process may run on, capped by OMP_NUM_THREADS where the machine sets it (the GPU
    hosts give one GPU's job 16 cores that way)."""
    model, flags = "unknown", ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = line.split(":", 1)[1]
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(affinity, omp) if omp > 0 else affinity
    return {"cpu_model": model, "nproc": nproc, "affinity_cores": affinity, "omp_num_threads": omp or None,
            "threads": threads, "avx512": " avx512f " in f" {flags} "}


def cpu_single(oracle, name, w, steps_all, steps_one):
    """The oracle's C restatement of the reference CPU path, built as an optimising compiler
    would (-O3, FMA contraction, AVX-512 where the host has it), timed on this host: once on
    one core (the reference has no threading) and once with OpenMP over elements on all the
    cores this job may use (the single-node stand-in for the reference's MPI decomposition)."""
    info = host_info()
    build = "fast_v4" if info["avx512"] else "fast_v3"
    oracle.use_build(build)
    try:
        og = oracle.Geometry(w["N"], w["mesh"].nodeCoords, 1)
        kind = {"sw256": oracle.SW, "euler512": oracle.EULER, "lsw10": oracle.LSW}[name]
        om = oracle.Model(kind, w["N"], w["mesh"].sideInfo, og)
        for k, v in w["fields"].items():
            {"topography": om.set_topography, "potential": om.set_gravity, "bathymetry": om.set_bathymetry}[k](v)
        om.set_solution(w["sol"])
        if name == "euler512":
            om.set_prescribed()
        dof = w["mesh"].nElem * (w["N"] + 1) ** 2 * w["nVar"]
        rates = {}
        for threads, steps in ((1, steps_one), (info["threads"], steps_all)):
            oracle.set_threads(threads)
            c0 = time.perf_counter()
            om.rk3_steps(steps, w["dt"])
            rates[threads] = (dof * steps / (time.perf_counter() - c0), steps, time.perf_counter() - c0)
        oracle.set_threads(1)
        del om, og
    finally:
        oracle.use_build("parity")
    allr, one = rates[info["threads"]], rates[1]
    return {"value": allr[0], "unit": "DOF-updates/s", "cores": info["threads"], "kind": "port",
            "value_1core": one[0], "build": f"oracle/{oracle.BUILDS[build].rsplit('/', 1)[-1]}",
            "cpu_model": info["cpu_model"], "nproc": info["nproc"], "affinity_cores": info["affinity_cores"],
            "omp_num_threads": info["omp_num_threads"],
            "sample": f"RK3 steps of the same {w['nx']}x{w['nx']} N={w['N']} workload: {allr[1]} on "
                      f"{info['threads']} threads (OpenMP over elements) in {allr[2]:.2f} s, {one[1]} on 1 core "
                      f"in {one[2]:.2f} s"}


def run_single(a, root):
    import self_amd
    from self_amd import workloads
    w = setup_single(self_amd, workloads, a.workload, a.seed)
    model, dt, N = w["model"], w["dt"], w["N"]
    # the benchmark meshes are affine: one metric record per element (raises if a mesh is not),
    # unless the reference's per-node metric terms are asked for
    model.SetNodalMetrics(bool(a.nodal_metrics))
    if a.traffic_probe:
        return traffic_probe(self_amd, model, dt)
    steps = a.steps or w["steps"]
    dof = w["mesh"].nElem * (N + 1) ** 2 * w["nVar"]
    settle_steps = SETTLE_STEPS if a.settle_steps is None else a.settle_steps
    settle = clock_settle(self_amd, model.stream)
    model.ForwardStep(a.warmup, dt)
    # then the settling steps, as calls of the timed call's size: the timed call then starts
    # from a trace buffer whose step graphs have been captured and replayed already (a capture,
    # and a graph's first replay, carry host work that leaves the GPU idle)
    done = 0
    while done < settle_steps:
        model.ForwardStep(steps, dt)
        done += steps
    settle_steps = done
    self_amd.synchronize()
    t0 = time.perf_counter()
    ms = model.time_steps(steps, dt)
    self_amd.synchronize()
    wall = time.perf_counter() - t0
    del settle
    if not np.isfinite(model.solution).all():
        print("non-finite state after the timed steps", file=sys.stderr)
        return 1
    sec = ms / 1e3
    launch_s = sec / (3 * steps)
    bpd = alg_bytes_per_dof_step(N, w["nVar"], w["NT"], w["p"])
    achieved = bpd / 3.0 * dof / launch_s / 1e9
    # after the reported measurement: the same steps timed again SUSTAIN_STEPS steps later,
    # reported beside the value
    model.ForwardStep(SUSTAIN_STEPS, dt)
    self_amd.synchronize()
    sus_launch_s = model.time_steps(steps, dt) / 1e3 / (3 * steps)
    sustained = {"after_steps": settle_steps + a.warmup + steps + SUSTAIN_STEPS, "timed_steps": steps,
                 "avg_launch_us": sus_launch_s * 1e6,
                 "frac": bpd / 3.0 * dof / sus_launch_s / 1e9 / HBM_PEAK_GBS,
                 "value": dof / (3 * sus_launch_s)}
    cpu = None
    if not a.no_cpu_baseline:
        import oracle
        steps_all = a.cpu_steps or {"sw256": 100, "euler512": 5, "lsw10": 200}[a.workload]
        steps_one = max(1, steps_all // {"sw256": 10, "euler512": 5, "lsw10": 2}[a.workload])
        cpu = cpu_single(oracle, a.workload, w, steps_all, steps_one)
    phys, kind = {"sw256": ("PhysSW", 2), "euler512": ("PhysEuler", 3), "lsw10": ("PhysLSW", 1)}[a.workload]
    kernel = tile_name(self_amd, N, phys, kind)
    if a.nodal_metrics:
        t = {7: 64, 6: 64, 5: 128, 4: 128}.get(N, 256)
        kernel = f"k_fused<{phys}, {N}, 0, true, {t}, 1, false>"
    traffic = None if a.nodal_metrics else traffic_from_profiles(
        root, {"sw256": "sw2d_stage_256x256_N7_traffic.json",
               "euler512": "euler2d_stage_512x512_N7_traffic.json"}.get(a.workload, ""))
    copy_gbs = copy_bandwidth(self_amd)
    out = {
        "metric": METRIC if a.workload == "sw256" else f"DOF-updates/sec + % HBM roofline, {a.workload}",
        "value": dof * steps / sec, "unit": "DOF-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": ms / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["desc"], "elements": w["mesh"].nElem, "N": N, "dof": dof, "dt": dt, "wall_s": wall,
                   "metrics": ("per-node metric terms (SEMQuad)" if a.nodal_metrics
                               else "affine, one record per element"),
                   "clock_settle": f"{SETTLE_COPIES} streaming copies of 1 GiB, then the warm-up steps, then "
                                   f"{settle_steps} untimed steps of the workload in calls of the timed call's size"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                     "avg_launch_us": launch_s * 1e6, "bytes_per_launch_alg": bpd / 3.0 * dof,
                     "copy_GBs": copy_gbs, "frac_of_copy": achieved / copy_gbs, "sustained": sustained},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


# ---------------------------------------------------------------------------------------
# batched ensemble
# ---------------------------------------------------------------------------------------
def run_ensemble(a, root, n_runs=512, n_check=16):
    import self_amd
    from self_amd import workloads
    rng = np.random.default_rng(a.seed)
    sizes = rng.integers(8, 65, size=n_runs)
    degrees = rng.choice([3, 5, 7], size=n_runs)
    cache = {}
    models, dts, inits = [], [], []
    t0 = time.perf_counter()
    for r in range(n_runs):
        n, N = int(sizes[r]), int(degrees[r])
        if (n, N) not in cache:
            mesh = self_amd.Mesh2D.structured(n, n, periodic=(True, True))
            interp = self_amd.Lagrange(N, self_amd.GAUSS, N, self_amd.GAUSS)
            geom = self_amd.SEMQuad(interp, mesh)
            cache[(n, N)] = (mesh, interp, geom, geom.x.to_host(), geom.CovariantArcMin())
        mesh, interp, geom, x, arcmin = cache[(n, N)]
        sol, topo = workloads.nonlinear_sw(x, interp.controlPoints, interp.qWeights, seed=1000 + r)
        m = self_amd.ShallowWater(mesh, geom)
        m.SetTopography(topo)
        m.SetSolution(sol)
        models.append(m)
        dts.append(workloads.cfl_dt(arcmin, 1.2, 0.5))
        inits.append((sol, topo))
        if (r + 1) % 64 == 0:
            print(f"ensemble: {r + 1}/{n_runs} runs set up", file=sys.stderr, flush=True)
    setup_s = time.perf_counter() - t0
    batch = self_amd.Batch(models, dts)
    persistent = a.batch_mode == "persistent"
    if persistent:
        batch.SetMode(self_amd.Batch.PERSISTENT)
    groups, dof = batch.info()
    steps = a.steps or 20
    # the batch runs on its own stream: the settling copies run beside the warm-up steps on the
    # default stream, so the GPU stays busy up to the timed steps
    settle_steps = ENSEMBLE_SETTLE_STEPS if a.settle_steps is None else a.settle_steps
    settle = clock_settle(self_amd, None)
    batch.ForwardStep(settle_steps + a.warmup)
    self_amd.synchronize()
    ms = batch.time_steps(steps)
    del settle
    batch.scatter()
    # per-run error against solving the run alone (same initial state, same steps)
    errs = []
    for r in np.linspace(0, n_runs - 1, n_check).astype(int):
        n, N = int(sizes[r]), int(degrees[r])
        mesh, interp, geom, x, arcmin = cache[(n, N)]
        m = self_amd.ShallowWater(mesh, geom)
        m.SetTopography(inits[r][1])
        m.SetSolution(inits[r][0])
        m.ForwardStep(settle_steps + a.warmup + steps, dts[r])
        ref, got = m.solution, models[r].solution
        errs.append(float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))))
    bytes_step = 0.0
    for r in range(n_runs):
        N = int(degrees[r])
        bytes_step += int(sizes[r]) ** 2 * (N + 1) ** 2 * 3 * alg_bytes_per_dof_step(N, 3, 3, 1)
    sec = ms / 1e3
    achieved = bytes_step * steps / sec / 1e9
    out = {
        "metric": "aggregate DOF-updates/s, 512 batched SW2D runs", "value": dof * steps / sec,
        "unit": "DOF-updates/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": ms / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n_runs} nonlinear SW 2D runs, n x n periodic elements with n ~ U{{8..64}}, "
                               f"N ~ U{{3,5,7}}, "
                               + ("one persistent launch per RK stage for all degree groups"
                                  if persistent else "one fused launch per degree group and RK stage")
                               + ", one hipGraph per step", "batch_mode": a.batch_mode, "runs": n_runs, "groups": groups, "dof": dof, "setup_s": setup_s,
                   "max_rel_err_vs_separate": max(errs), "runs_checked": n_check,
                   "clock_settle": f"{SETTLE_COPIES} streaming copies of 1 GiB beside {settle_steps} untimed "
                                   f"steps of the batch and the warm-up steps"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     # measured HBM bytes of one RK stage of all degree groups (per-group mode)
                     "traffic": None if persistent else traffic_from_profiles(root, "sw2d_ensemble512_stage_traffic.json"),
                     "traffic_scope": None if persistent else "one RK stage, all degree groups (three launches)",
                     "kernel": "k_batch_persistent<PhysSW>" if persistent
                     else "k_fused<PhysSW, N, 0, false, T> (one per degree group)"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


# ---------------------------------------------------------------------------------------
# learned closure
# ---------------------------------------------------------------------------------------
def run_closure(a, root):
    from self_amd import closure

    def log(m):
        print(m, file=sys.stderr, flush=True)

    r = closure.benchmark(n_train=a.closure_runs, n_steps=a.steps or 500, log=log,
                          aposteriori_iters=a.closure_aposteriori, train_horizon=a.closure_horizon,
                          threads=a.closure_threads)
    runs = r["runs"]
    value = float(np.mean([x["speedup_at_equal_error"] for x in runs]))
    out = {
        "metric": "closure speedup at equal L2 error (coarse 64x64 N=3 SW + MLP vs plain runs, fine 256x256 N=7 truth)",
        "value": value, "unit": "x", "n_gpus": 1, "steps": a.steps or 500, "warmup": 0,
        "ms_per_step": float(np.mean([x["closure_ms"] for x in runs])) / (a.steps or 500),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "nonlinear SW 2D closure: coarse 64x64 periodic N=3 FP64 + FP32 MLP 51-128-128-48 "
                               "GELU, trained on fine 256x256 N=7 runs projected onto the coarse space",
                   "train_runs": a.closure_runs, "train_horizon": a.closure_horizon or (a.steps or 500),
                   "train_samples": r["train_samples"], "alpha": r["alpha"],
                   "skill_train": r["skill_train"], "skill_validation": r["skill_validation"],
                   "data_s": r["data_s"], "train_s": r["train_s"], "runs": runs,
                   "aposteriori": r.get("aposteriori")},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out, default=float), flush=True)
    return 0


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: the engine has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        return 2
    root = os.path.dirname(os.path.abspath(__file__))
    if a.workload == "ensemble":
        return run_ensemble(a, root)
    if a.workload == "closure":
        return run_closure(a, root)
    return run_single(a, root)


if __name__ == "__main__":
    sys.exit(main())
