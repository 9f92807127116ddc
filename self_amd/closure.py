"""Learned subgrid closure for nonlinear shallow water 2D (no reference counterpart).

A coarse run (nxc x nxc periodic elements, degree Nc) is corrected by an FP32 MLP that
maps each element's nodal state plus its metric scalars (J, dsdx(1,1), dsdx(2,2)) to a
tendency correction.  Training data come from fine runs of the same solver (r nxc x r nxc
elements, degree Nf, the seeded nonlinear-SW distribution): for a sampled coarse step
starting at the projected fine state P S_f(t),

    target = (P S_f(t + w dt_c) - Phi_c^w(P S_f(t))) / (w dt_c),

where Phi_c^w is w plain coarse RK3 steps and dt_c = k dt_f (k whole fine steps).  Because an
RK3 step adds exactly dt_c times a constant forcing, a constant correction equal to the
target would make the coarse run land on the projected fine state after the w steps.  The
correction can be scaled by a factor alpha chosen on validation runs (alpha = 0 is the plain
run, the fallback).  The library's kernels do all of it:
the fine and coarse solvers, the L2 projection, the MLP forward/backward and Adam; this
module drives them.  The mass (H) correction is made mean-free over the mesh by the engine,
so the corrected run conserves mass to round-off.
"""
import numpy as np

from . import (GAUSS, DeviceArray, Lagrange, MLP, Mesh2D, SEMQuad, ShallowWater, project_l2_structured, workloads)

NVAR = 3


class CoarseFine:
    """The fine and coarse discretisations of the unit periodic square and their time steps.

    dt_f = cfl * arcmin_f / speed and dt_c = k dt_f with k = floor(cfl * arcmin_c / speed / dt_f),
    with one wave-speed bound for all runs so that every run shares the step sizes."""

    def __init__(self, nxc=64, Nc=3, r=4, Nf=7, cfl=0.5, speed=1.1):
        self.nxc, self.Nc, self.r, self.Nf = nxc, Nc, r, Nf
        self.nxf = nxc * r
        self.mesh_f = Mesh2D.structured(self.nxf, self.nxf, periodic=(True, True))
        self.interp_f = Lagrange(Nf, GAUSS, Nf, GAUSS)
        self.geom_f = SEMQuad(self.interp_f, self.mesh_f)
        self.mesh_c = Mesh2D.structured(nxc, nxc, periodic=(True, True))
        self.interp_c = Lagrange(Nc, GAUSS, Nc, GAUSS)
        self.geom_c = SEMQuad(self.interp_c, self.mesh_c)
        self.xf = self.geom_f.x.to_host()
        self.dt_f = workloads.cfl_dt(self.geom_f.CovariantArcMin(), speed, cfl)
        self.k = max(1, int(workloads.cfl_dt(self.geom_c.CovariantArcMin(), speed, cfl) / self.dt_f))
        self.dt_c = self.k * self.dt_f
        J = self.geom_c.J.to_host()
        d = self.geom_c.dsdx.to_host()
        self.J_c = J[:, 0, 0].copy()
        self.metrics = np.stack([self.J_c, d[:, 0, 0, 0, 0], d[:, 0, 0, 1, 1]], axis=1)
        self.qW_c = np.asarray(self.interp_c.qWeights, np.float64)
        self.nIn = NVAR * (Nc + 1) ** 2 + 3
        self.nOut = NVAR * (Nc + 1) ** 2

    # --- models --------------------------------------------------------------------------
    def initial_state(self, seed):
        """(solution, topography) of the fine run of this seed (host arrays)."""
        return workloads.nonlinear_sw(self.xf, self.interp_f.controlPoints, self.interp_f.qWeights, seed=seed)

    def fine_model(self, seed, state=None, model=None):
        """Fine model started from the seed's initial state (or `state`); `model` is reused
        when given (topography and solution set anew)."""
        sol, topo = state if state is not None else self.initial_state(seed)
        F = model if model is not None else ShallowWater(self.mesh_f, self.geom_f)
        F.SetTopography(topo)
        F.SetSolution(sol)
        return F, topo

    def project(self, fine, nVar=NVAR):
        """DeviceArray (coarse) of the L2 projection of a fine interior."""
        return project_l2_structured(fine, self.Nf, self.Nc, self.nxc, self.nxc, self.r, nVar)

    def coarse_model(self, topo_f, model=None):
        topo_c = self.project(topo_f.reshape(topo_f.shape[0], 1, self.Nf + 1, self.Nf + 1), 1).to_host()
        C = model if model is not None else ShallowWater(self.mesh_c, self.geom_c)
        C.SetTopography(topo_c[:, 0])
        return C

    def raw_features(self, Sc):
        return np.concatenate([Sc.reshape(Sc.shape[0], -1), self.metrics], axis=1)

    def l2(self, S):
        """L2 norm over the coarse mesh, per variable: sqrt(sum_e J_e sum_ij w_i w_j S^2)."""
        w2 = np.outer(self.qW_c, self.qW_c)
        return np.sqrt(np.einsum("e,ji,evji->v", self.J_c, w2, S * S))

    def mass(self, S):
        w2 = np.outer(self.qW_c, self.qW_c)
        return float(np.einsum("e,ji,eji->", self.J_c, w2, S[:, 2]))


def generate(cf, seeds, n_steps=500, stride=20, window=1, log=None, threads=0):
    """Training pairs (raw features, raw targets) from fine runs of the given seeds: one
    sample per element at every `stride`-th coarse step of an n_steps horizon.  The target
    is the mean tendency defect over `window` coarse steps,
    (P S_f(t + w dt_c) - Phi_c^w(P S_f(t))) / (w dt_c): longer windows average out the
    part of the defect carried by fast unresolved waves.  With threads > 0 the initial states
    of the next runs are built on that many host threads while the GPU advances the current
    one, and one fine and one coarse model serve all runs."""
    assert stride >= window, "samples must not overlap (stride >= window)"
    import collections
    from concurrent.futures import ThreadPoolExecutor
    seeds = list(seeds)
    X, T = [], []
    pool = ThreadPoolExecutor(threads) if threads else None
    pending = collections.deque()
    nxt = 0
    Fm = Cm = None
    for i, seed in enumerate(seeds):
        if pool is not None:
            while nxt < len(seeds) and len(pending) < 2 * threads:
                pending.append(pool.submit(cf.initial_state, seeds[nxt]))
                nxt += 1
            F, topo = cf.fine_model(seed, state=pending.popleft().result(), model=Fm)
            C = cf.coarse_model(topo, model=Cm)
            Fm, Cm = F, C
        else:
            F, topo = cf.fine_model(seed)
            C = cf.coarse_model(topo)
        P0, due = None, None
        for step in range(n_steps):
            if step % stride == 0 and step + window <= n_steps:
                P0, due = cf.project(F.solution_device()).to_host(), step + window
            F.ForwardStep(cf.k, cf.dt_f)
            if due is not None and step + 1 == due:
                P1 = cf.project(F.solution_device()).to_host()
                C.SetSolution(P0)
                C.ForwardStep(window, cf.dt_c)
                Sc = C.solution
                due = None
                if not (np.isfinite(P1).all() and np.isfinite(Sc).all()):
                    if log:
                        log(f"closure data: seed {seed} non-finite state at coarse step {step + 1}; run dropped there")
                    break
                X.append(cf.raw_features(P0))
                T.append(((P1 - Sc) / (window * cf.dt_c)).reshape(P0.shape[0], -1))
        if log and (pool is None or (i + 1) % 100 == 0 or i + 1 == len(seeds)):
            log(f"closure data: seed {seed} done ({len(X)} snapshots)")
        del F, C
    if pool is not None:
        pool.shutdown()
    return np.concatenate(X), np.concatenate(T)


class Normalisation:
    """Per-feature / per-output mean and scale (FP32), scale 1 where a column is constant."""

    def __init__(self, X, T):
        self.inMean = X.mean(0).astype(np.float32)
        s = X.std(0)
        self.inScale = np.where(s > 1e-12 * np.maximum(1.0, np.abs(X.mean(0))), s, 1.0).astype(np.float32)
        self.outMean = T.mean(0).astype(np.float32)
        s = T.std(0)
        self.outScale = np.where(s > 0, s, 1.0).astype(np.float32)

    def x(self, X):
        return ((X - self.inMean.astype(np.float64)) / self.inScale.astype(np.float64)).astype(np.float32)

    def t(self, T):
        return ((T - self.outMean.astype(np.float64)) / self.outScale.astype(np.float64)).astype(np.float32)

    def arrays(self, alpha=1.0):
        """Normalisation arrays for the engine; alpha scales the correction (0 = plain run)."""
        return (self.inMean, self.inScale, (alpha * self.outMean).astype(np.float32),
                (alpha * self.outScale).astype(np.float32))

    def save(self, path, params):
        np.savez(path, inMean=self.inMean, inScale=self.inScale, outMean=self.outMean, outScale=self.outScale,
                 **{k: np.asarray(v, np.float32) for k, v in params.items()})

    @classmethod
    def load(cls, path):
        z = np.load(path)  # plain arrays, no pickled objects
        n = object.__new__(cls)
        n.inMean, n.inScale, n.outMean, n.outScale = (z[k] for k in ("inMean", "inScale", "outMean", "outScale"))
        params = {k: z[k] for k in ("W1", "b1", "W2", "b2", "W3", "b3")}
        return n, params


def init_params(nIn, nHidden, nOut, seed=0):
    """He-style uniform initialisation, FP32."""
    rng = np.random.default_rng(seed)

    def lin(a, b):
        bound = np.sqrt(6.0 / a) * 0.5
        return rng.uniform(-bound, bound, size=(a, b)).astype(np.float32), np.zeros(b, np.float32)

    W1, b1 = lin(nIn, nHidden)
    W2, b2 = lin(nHidden, nHidden)
    W3, b3 = lin(nHidden, nOut)
    return {"W1": W1, "b1": b1, "W2": W2, "b2": b2, "W3": W3, "b3": b3}


def train(mlp, Xn, Tn, epochs=20, batch=4096, lr=1e-3, lr_final=1e-4, seed=0, log=None):
    """Adam on mean-squared error over normalised (Xn, Tn) (FP32), with a geometric learning
    rate decay from lr to lr_final; one shuffle up front, chunk order shuffled per epoch.
    Returns the per-epoch mean training loss."""
    rng = np.random.default_rng(seed)
    n = Xn.shape[0]
    perm = rng.permutation(n)
    Xd = DeviceArray.from_host(Xn[perm])
    Td = DeviceArray.from_host(Tn[perm])
    offs = np.arange(0, n - batch + 1, batch)
    total = epochs * len(offs)
    step, hist = 0, []
    for ep in range(epochs):
        losses = []
        for o in rng.permutation(offs):
            step += 1
            rate = lr * (lr_final / lr) ** ((step - 1) / max(1, total - 1))
            want = len(losses) < 8  # a few synchronised loss reads per epoch
            loss = mlp.loss_grad(Xd, Td, batch, int(o), want_loss=want)
            if want:
                losses.append(loss)
            mlp.adam_step(rate, step)
        hist.append(float(np.mean(losses)))
        if log:
            log(f"closure train: epoch {ep + 1}/{epochs} loss {hist[-1]:.4e}")
    return hist


# Williamson low-storage RK3 coefficients (src/SELF_Model.f90:25-27)
RK3_A = (0.0, -5.0 / 9.0, -153.0 / 128.0)
RK3_G = (1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0)


def l2_weights(cf):
    """Quadrature weights of the coarse L2 norm per node, shape (nEl, 1, Nc+1, Nc+1)."""
    w2 = np.outer(cf.qW_c, cf.qW_c)
    return (cf.J_c[:, None, None] * w2[None])[:, None]


def rollout_loss_grad(C, mlp, P0, truths, dt, weights, want_grad=True):
    """A-posteriori loss of K = len(truths) closure-corrected coarse RK3 steps from P0,
    L = (1/K) sum_k sum_v ||S_k - T_k||^2 (quadrature-weighted), and with want_grad its
    gradient with respect to the network parameters, left in the network's gradient buffer.

    The gradient is the adjoint of the K steps on the GPU: per stage (last to first) the
    transposed RK3 update, the tendency adjoint (selfamd_model2d_tendency_vjp, the lagged
    normal velocities held at their forward values) and the closure adjoint
    (selfamd_model2d_closure_vjp: parameter gradients and the state cotangent through the
    network inputs).  Low-storage RK3 stage m: G_m = a_m G_(m-1) + f(S_(m-1)),
    S_m = S_(m-1) + g_m dt G_m; backwards, with lG~ = lG_m + g_m dt lS_m:
    lS_(m-1) = lS_m + f'(S_(m-1))^T lG~ and lG_(m-1) = a_m lG~."""
    K = len(truths)
    shape = C.shape
    C.SetSolution(P0)
    Ss = [[DeviceArray(shape) for _ in range(3)] for _ in range(K)] if want_grad else None
    Us = [[DeviceArray(C.lag_shape) for _ in range(3)] for _ in range(K)] if want_grad else None
    loss, deltas = 0.0, []
    for k in range(K):
        for st in range(3):
            C.Stage(st, dt, Ss[k][st] if want_grad else None, Us[k][st] if want_grad else None)
        r = C.solution - truths[k]
        loss += float(np.sum(weights * r * r)) / K
        deltas.append(2.0 * weights * r / K)
    if not want_grad:
        return loss
    mlp.zero_grads()
    lS, lG = deltas[-1].copy(), np.zeros(shape)
    mu, muP = DeviceArray(shape), DeviceArray(shape)
    lGt_d = DeviceArray(shape)
    pending = 0.0  # lag-path cotangent of the previous stage's input, from the stage after it
    for k in reversed(range(K)):
        for st in (2, 1, 0):
            lGt = lG + RK3_G[st] * dt * lS
            lGt_d.copy_from_host(lGt)
            # the lagged velocities of stage (k, st) were formed from the previous stage's input
            # (none before the first stage: SetSolution forms them from P0, independent of the
            # network)
            prev = Ss[k][st - 1] if st > 0 else (Ss[k - 1][2] if k > 0 else None)
            C.TendencyVJP(Ss[k][st], Us[k][st], lGt_d, mu, Sprev=prev, muPrev=muP if prev is not None else None)
            C.ClosureVJP(Ss[k][st], lGt_d, mu)
            lS = lS + mu.to_host() + pending
            pending = muP.to_host() if prev is not None else 0.0
            lG = RK3_A[st] * lGt
        if k > 0:
            lS += deltas[k - 1]
    return loss


def rollout_windows(cf, seeds, K, every, n_steps=500, log=None):
    """Training windows from fine runs: the projected fine state at every `every`-th coarse step
    and the projected states of the K coarse steps after it.  Returns a list of
    (topography of the run, P0, [T_1..T_K])."""
    out = []
    for seed in seeds:
        F, topo = cf.fine_model(seed)
        states = [cf.project(F.solution_device()).to_host()]
        for step in range(1, n_steps + 1):
            F.ForwardStep(cf.k, cf.dt_f)
            if step % every <= K:
                states.append(cf.project(F.solution_device()).to_host())
            else:
                states.append(None)
        for t0 in range(0, n_steps - K + 1, every):
            win = states[t0:t0 + K + 1]
            if all(w is not None and np.isfinite(w).all() for w in win):
                out.append((topo, win[0], win[1:]))
        if log:
            log(f"closure windows: seed {seed} done ({len(out)} windows)")
        del F
    return out


def train_aposteriori(cf, mlp, norm, windows, iters, lr=1e-4, lr_final=None, seed=0, log=None, log_every=50):
    """Adam on the a-posteriori rollout loss over randomly drawn windows (one window per
    step); the network corrects the coarse run through norm.arrays() at alpha = 1.  Returns
    the per-step losses."""
    rng = np.random.default_rng(seed)
    weights = l2_weights(cf)
    K = len(windows[0][2])
    lr_final = lr if lr_final is None else lr_final
    models = {}
    hist = []
    for it in range(1, iters + 1):
        topo, P0, truths = windows[int(rng.integers(len(windows)))]
        key = id(topo)
        if key not in models:
            C = cf.coarse_model(topo)
            C.SetClosure(mlp, *norm.arrays(1.0))
            models[key] = C
        C = models[key]
        loss = rollout_loss_grad(C, mlp, P0, truths, cf.dt_c, weights)
        rate = lr * (lr_final / lr) ** ((it - 1) / max(1, iters - 1))
        mlp.adam_step(rate, it)
        hist.append(loss)
        if log and it % log_every == 0:
            log(f"closure a-posteriori: step {it}/{iters} loss {np.mean(hist[-log_every:]):.4e}")
    return hist


def skill(mlp, Xn, Tn, rows=65536):
    """A priori skill on normalised data: 1 - sum (y - t)^2 / sum (t - mean t)^2 (R^2 over all
    outputs) on up to `rows` samples."""
    n = min(rows, Xn.shape[0])
    y = np.concatenate([mlp.forward(Xn[o:o + mlp.maxBatch]) for o in range(0, n, mlp.maxBatch)])[:n]
    t = Tn[:n].astype(np.float64)
    return float(1.0 - np.sum((y - t) ** 2) / np.sum((t - t.mean(0)) ** 2))


def run_coarse(cf, C, P0, n_steps, checkpoints):
    """Advance coarse model C from P0; states at the checkpoints and the timed milliseconds
    (the step graph is captured by an untimed step first)."""
    C.SetSolution(P0)
    C.ForwardStep(1, cf.dt_c)
    C.SetSolution(P0)
    C.t = 0.0
    states, ms, done = {}, 0.0, 0
    for c in checkpoints:
        ms += C.time_steps(c - done, cf.dt_c)
        done = c
        states[c] = C.solution
    return states, ms


def evaluate(cf, seed, mlp=None, norm=None, n_steps=500, checkpoints=None, alphas=(1.0,), log=None):
    """Fine truth, plain coarse and (if given) closure-corrected coarse runs of one seed.
    Returns per-checkpoint L2 errors (summed over variables) against the projected fine
    state, mass drifts and timed milliseconds."""
    checkpoints = checkpoints or [n_steps // 5 * (i + 1) for i in range(5)]
    F, topo = cf.fine_model(seed)
    C = cf.coarse_model(topo)
    P0 = cf.project(F.solution_device()).to_host()
    truth, fine_ms, done = {}, 0.0, 0
    for c in checkpoints:
        fine_ms += F.time_steps((c - done) * cf.k, cf.dt_f)
        done = c
        truth[c] = cf.project(F.solution_device()).to_host()
    out = {"checkpoints": checkpoints, "fine_ms": fine_ms, "mass0": cf.mass(P0)}
    plain, out["plain_ms"] = run_coarse(cf, C, P0, n_steps, checkpoints)
    out["plain_err"] = [float(np.sum(cf.l2(plain[c] - truth[c]))) for c in checkpoints]
    out["plain_err_var"] = [float(x) for x in cf.l2(plain[checkpoints[-1]] - truth[checkpoints[-1]])]
    out["plain_mass_drift"] = (cf.mass(plain[checkpoints[-1]]) - out["mass0"]) / out["mass0"]
    if mlp is not None:
        out["closure"] = {}
        for alpha in alphas:
            # alpha = 0 is the fallback: the plain run, without evaluating the network
            if alpha != 0.0:
                C.SetClosure(mlp, *norm.arrays(alpha))
            corr, ms = run_coarse(cf, C, P0, n_steps, checkpoints)
            C.SetClosure(None)
            last = corr[checkpoints[-1]]
            out["closure"][alpha] = {
                "ms": ms, "err": [float(np.sum(cf.l2(corr[c] - truth[c]))) for c in checkpoints],
                "err_var": [float(x) for x in cf.l2(corr[checkpoints[-1]] - truth[checkpoints[-1]])],
                "mass_drift": (cf.mass(last) - out["mass0"]) / out["mass0"] if np.isfinite(last).all() else None}
    if log:
        log(f"closure eval seed {seed}: plain {out['plain_err'][-1]:.4e}" + "".join(
            f", alpha {a}: {r['err'][-1]:.4e}" for a, r in out.get("closure", {}).items()))
    return out


def select_alpha(cf, mlp, norm, seeds, n_steps, alphas=(0.0, 0.25, 0.5, 0.75, 1.0), log=None):
    """Correction scale minimising the mean final L2 error over validation seeds.  alpha = 0
    is the plain run, so the chosen closure is never worse than plain on the validation set."""
    errs = {a: [] for a in alphas}
    for seed in seeds:
        r = evaluate(cf, seed, mlp, norm, n_steps=n_steps, alphas=alphas, log=log)
        for a in alphas:
            e = r["closure"][a]["err"][-1]
            errs[a].append(e if np.isfinite(e) else np.inf)
    mean = {a: float(np.mean(v)) for a, v in errs.items()}
    return min(mean, key=mean.get), mean


def plain_ladder(cf, seed, n_steps, factors=(1, 2, 4), log=None):
    """Plain runs at (q nxc)^2 elements, degree Nc, q in factors, from the projection of the
    same fine initial state, advanced to the same final time (q times more steps of dt_c / q)
    and projected onto the coarse mesh.  Returns {q: (L2 error vs the projected fine state,
    timed ms)} and the fine run's timed ms."""
    sol_f, topo_f = workloads.nonlinear_sw(cf.xf, cf.interp_f.controlPoints, cf.interp_f.qWeights, seed=seed)
    F, _ = cf.fine_model(seed)
    fine_ms = F.time_steps(n_steps * cf.k, cf.dt_f)
    truth = cf.project(F.solution_device()).to_host()
    del F
    sol0 = None
    out = {}
    for q in factors:
        n = cf.nxc * q
        if cf.nxf % n:
            continue
        r = cf.nxf // n
        mesh = Mesh2D.structured(n, n, periodic=(True, True))
        geom = SEMQuad(cf.interp_c, mesh)
        s0 = project_l2_structured(sol_f, cf.Nf, cf.Nc, n, n, r, NVAR).to_host()
        t0 = project_l2_structured(topo_f.reshape(topo_f.shape[0], 1, cf.Nf + 1, cf.Nf + 1), cf.Nf, cf.Nc, n, n, r,
                                   1).to_host()
        m = ShallowWater(mesh, geom)
        m.SetTopography(t0[:, 0])
        m.SetSolution(s0)
        m.ForwardStep(1, cf.dt_c / q)  # capture the step graph outside the timed region
        m.SetSolution(s0)
        ms = m.time_steps(n_steps * q, cf.dt_c / q)
        s = m.solution
        back = project_l2_structured(s, cf.Nc, cf.Nc, cf.nxc, cf.nxc, q, NVAR).to_host() if q > 1 else s
        out[q] = (float(np.sum(cf.l2(back - truth))), ms)
        if log:
            log(f"plain {n}x{n} N={cf.Nc}: error {out[q][0]:.4e}, {ms:.1f} ms")
        if q == 1:
            sol0 = s0
    return out, fine_ms, sol0


def aposteriori_network(cf, windows, iters, lr=1e-3, K=None, seed=0, log=None):
    """A network trained through the solver from a zero correction: normalisation of the
    inputs from the window states and of the outputs from the one-step defects of the windows,
    last layer zero (the untrained network is the plain run), then train_aposteriori."""
    X = np.concatenate([cf.raw_features(w[1]) for w in windows])
    T = []
    models = {}
    for w in windows:
        key = id(w[0])
        if key not in models:
            models[key] = cf.coarse_model(w[0])
        C = models[key]
        C.SetSolution(w[1])
        C.ForwardStep(1, cf.dt_c)
        T.append(((w[2][0] - C.solution) / cf.dt_c).reshape(w[1].shape[0], -1))
    norm = Normalisation(X, np.concatenate(T))
    norm.outMean[:] = 0.0
    mlp = MLP(cf.nIn, 128, cf.nOut, max(4096, cf.nxc * cf.nxc))
    p = init_params(cf.nIn, 128, cf.nOut, seed=seed)
    p["W3"][:] = 0.0
    p["b3"][:] = 0.0
    mlp.set_params(p)
    hist = train_aposteriori(cf, mlp, norm, windows, iters, lr=lr, lr_final=lr / 10, seed=seed, log=log)
    return mlp, norm, hist


def benchmark(n_train=8, n_steps=500, stride=10, window=5, epochs=10, nxc=64, seeds_eval=(1000, 1001), log=None,
              aposteriori_iters=0, aposteriori_K=8, train_horizon=None, threads=0, aposteriori_runs=8):
    """Closure workload: train on fine runs, pick the correction scale on a validation run,
    then compare on held-out runs the corrected coarse run with plain runs of increasing
    resolution.  'speedup_at_equal_error' is the wall-time of the cheapest plain run whose
    error is at most the corrected run's, divided by the corrected run's wall-time (the fine
    run itself when no coarser plain run reaches that error).  Each training run covers
    train_horizon coarse steps (default n_steps); the training through the solver takes its
    windows from the first aposteriori_runs training runs over n_steps."""
    import time
    t0 = time.perf_counter()
    cf = CoarseFine(nxc=nxc)
    # short training runs are sampled in back-to-back windows
    horizon, stride_t = (n_steps, stride) if train_horizon is None else (train_horizon, window)
    X, T = generate(cf, range(n_train), n_steps=horizon, stride=stride_t, window=window, log=log, threads=threads)
    Xv, Tv = generate(cf, [500], n_steps=n_steps, stride=stride, window=window)
    gen_s = time.perf_counter() - t0
    norm = Normalisation(X, T)
    mlp = MLP(cf.nIn, 128, cf.nOut, 8192)
    mlp.set_params(init_params(cf.nIn, 128, cf.nOut, seed=0))
    t1 = time.perf_counter()
    hist = train(mlp, norm.x(X), norm.t(T), epochs=epochs, batch=4096, log=log)
    train_s = time.perf_counter() - t1
    res = {"train_samples": int(X.shape[0]), "data_s": gen_s, "train_s": train_s, "loss": hist,
           "skill_train": skill(mlp, norm.x(X), norm.t(T)), "skill_validation": skill(mlp, norm.x(Xv), norm.t(Tv))}
    alpha, mean = select_alpha(cf, mlp, norm, [600], n_steps, alphas=(0.0, 0.02, 0.05, 0.1, 0.2), log=log)
    res["alpha"], res["alpha_validation_err"] = alpha, mean
    runs = []
    for seed in seeds_eval:
        r = evaluate(cf, seed, mlp, norm, n_steps=n_steps, alphas=(alpha,), log=log)
        ladder, fine_ms, _ = plain_ladder(cf, seed, n_steps, log=log)
        c = r["closure"][alpha]
        err, ms = c["err"][-1], c["ms"]
        cheaper = [(v[1], q) for q, v in ladder.items() if v[0] <= err]
        ref_ms = min(cheaper)[0] if cheaper else fine_ms
        runs.append({"seed": seed, "closure_err": err, "closure_ms": ms, "plain_err": ladder[1][0],
                     "plain_ms": ladder[1][1], "ladder": {str(q): v for q, v in ladder.items()}, "fine_ms": fine_ms,
                     "closure_mass_drift": c["mass_drift"], "speedup_at_equal_error": ref_ms / ms,
                     "speedup_vs_fine": fine_ms / ms})
    res["runs"] = runs
    if aposteriori_iters:
        # training through the solver: windows of the training runs, network from a zero correction
        t2 = time.perf_counter()
        wins = rollout_windows(cf, range(min(n_train, aposteriori_runs)), aposteriori_K, every=25, n_steps=n_steps,
                               log=log)
        mlp_a, norm_a, hist_a = aposteriori_network(cf, wins, aposteriori_iters, K=aposteriori_K, log=log)
        weights = l2_weights(cf)
        val = rollout_windows(cf, [600], aposteriori_K, every=50, n_steps=n_steps)
        plain_l, corr_l = [], []
        for w in val:
            C = cf.coarse_model(w[0])
            C.SetClosure(mlp_a, *norm_a.arrays(0.0))
            plain_l.append(rollout_loss_grad(C, mlp_a, w[1], w[2], cf.dt_c, weights, want_grad=False))
            C.SetClosure(mlp_a, *norm_a.arrays(1.0))
            corr_l.append(rollout_loss_grad(C, mlp_a, w[1], w[2], cf.dt_c, weights, want_grad=False))
        ap = {"windows": len(wins), "K": aposteriori_K, "iters": aposteriori_iters, "train_s": time.perf_counter() - t2,
              "loss_first": float(np.mean(hist_a[:50])), "loss_last": float(np.mean(hist_a[-50:])),
              "validation_window_loss_plain": float(np.mean(plain_l)),
              "validation_window_loss_closure": float(np.mean(corr_l)), "runs": []}
        for seed in seeds_eval:
            r = evaluate(cf, seed, mlp_a, norm_a, n_steps=n_steps, alphas=(1.0,), log=log)
            ap["runs"].append({"seed": seed, "plain_err": r["plain_err"][-1], "closure_err": r["closure"][1.0]["err"][-1],
                               "closure_mass_drift": r["closure"][1.0]["mass_drift"],
                               "error_ratio": r["closure"][1.0]["err"][-1] / r["plain_err"][-1]})
        res["aposteriori"] = ap
    return res
