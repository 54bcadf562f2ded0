"""Learned warm start for time stepping to steady state.

A network trained on steady states that the time stepper itself produced
predicts the steady densities of a new (applied bias, generation amplitude,
donor doping, cone half-angle) draw; the stepper then starts from the
prediction instead of the equilibrium state and still decides
convergence (the same relative steady residual).  A prediction whose residual
after the first block of steps is above `fallback_residual` is dropped and the
instance restarts from the equilibrium state (its steps so far still count).
(After the first 25 steps at refinement 5 the equilibrium start's residual is
about 1.5 and a prediction with 0.4 % error about 0.3.)

The network (pecs_amd.WarmStartNet: hand-written HIP kernels, FP32 training,
BF16 inference) maps 4 scaled parameters to the density coefficients of the
four carriers on the logically Cartesian cell grid of one refinement
(4 x 9 N^2 values at degree 2).  Because each region's cells are numbered by
their (i, j) position, the same output vector serves every cone angle: it is
the field resampled on the region's cell grid, applied to each instance's own
mesh.  Currents are not predicted: the step recomputes them from the
densities (they carry no time derivative), and the Poisson state is
recomputed from the densities when the state is set.

Draws come in groups that share one cone angle (one mesh per group); each
group is one batched context (pecs_create_batch), and the groups are stepped
together (pecs_sweep_*).
"""
import time

import numpy as np

from . import SolarCellGPU, SolarCellSweep, WarmStartNet, workload_params

CARRIERS = "npro"


def draw_groups(rng, ngroups, group):
    """ngroups angles U[0, 15] degrees; per instance applied bias U[-0.5, 0.5],
    generation amplitude logU[0.1, 10], donor doping logU[0.5, 5]."""
    out = []
    for _ in range(ngroups):
        ang = float(rng.uniform(0.0, 15.0))
        inst = [(float(rng.uniform(-0.5, 0.5)), float(10 ** rng.uniform(-1, 1)),
                 float(10 ** rng.uniform(np.log10(0.5), np.log10(5.0)))) for _ in range(group)]
        out.append((ang, inst))
    return out


def features(ang, app, g0, ne):
    """Inputs scaled to about [-1, 1]."""
    return [app / 0.5, np.log10(g0), (np.log10(ne) - 0.2) / 0.5, ang / 7.5 - 1.0]


def instance_params(refine, degree, dt, ang, app, g0, ne):
    p = workload_params(refine, degree, dt)
    p.update(r1=0.5, r2=0.5 + float(np.tan(np.radians(ang))), L=1.5, phi_bi=0.4, phi_app=app, G0=g0, ne=ne)
    return p


def density_slices(g):
    """Slices of the density blocks [u] in each carrier vector of context g."""
    out = {}
    for c, n in zip(CARRIERS, (g.n_semi, g.n_semi, g.n_elec, g.n_elec)):
        out[c] = slice(2 * n // 3, n)
    return out


def step_groups(ctxs, tol, check_every, max_steps, log=None, fallback=None):
    """Step batched contexts together until every instance is steady.
    Returns the steps per (context, instance), the wall seconds and the
    residuals after the first block of steps.  fallback(ci, i, residual) is
    called once per instance after the first block; returning True means it
    was restarted."""
    sw = SolarCellSweep(ctxs, min(4, len(ctxs)))
    done = [np.full(g.batch, -1) for g in ctxs]
    first_res = []
    active = np.ones(len(ctxs), bool)
    t0 = time.perf_counter()
    n = 0
    first = True
    while active.any() and n < max_steps:
        k = min(check_every, max_steps - n)
        sw.step(k)
        n += k
        for ci, g in enumerate(ctxs):
            if not active[ci]:
                continue
            r = g.residuals()
            if first:
                first_res.extend(r.tolist())
                if fallback is not None:
                    for i in range(g.batch):
                        if fallback(ci, i, r[i]):
                            r[i] = np.inf
            d = done[ci]
            d[(d < 0) & (r <= tol)] = n
            if np.all(d >= 0):
                active[ci] = False
        first = False
        sw.set_active(active)
        if log and n % 500 == 0:
            log(f"step {n}: {int(active.sum())} of {len(ctxs)} groups still stepping")
    wall = time.perf_counter() - t0
    sw.close()
    return done, wall, np.array(first_res)


def generate(groups, refine, degree, dt, tol, check_every, max_steps, log=None):
    """Steady states of every draw, cold-started from the equilibrium state.
    Returns (X features, Y densities [n_draws, 4 nu], steps, wall, setup)."""
    t0 = time.perf_counter()
    ctxs = [SolarCellGPU([instance_params(refine, degree, dt, ang, *d) for d in inst]) for ang, inst in groups]
    setup = time.perf_counter() - t0
    for g in ctxs:
        g.initial_state()
    done, wall, first_res = step_groups(ctxs, tol, check_every, max_steps, log)
    X, Y, steps = [], [], []
    for (ang, inst), g, d in zip(groups, ctxs, done):
        sl = density_slices(g)
        for i, dr in enumerate(inst):
            st = g.get_state(i)
            X.append(features(ang, *dr))
            Y.append(np.concatenate([st[c][sl[c]] for c in CARRIERS]))
            steps.append(int(d[i]))
        g.close()
    return np.array(X, np.float32), np.array(Y), np.array(steps), wall, setup, first_res


class Normaliser:
    def __init__(self, Y):
        self.mu = Y.mean(axis=0)
        sd = Y.std(axis=0)
        self.sd = np.maximum(sd, 1e-6 * max(np.abs(Y).max(), 1e-30))

    def fwd(self, Y):
        return ((Y - self.mu) / self.sd).astype(np.float32)

    def back(self, Z):
        return Z.astype(np.float64) * self.sd + self.mu


def train(X, Y, hidden=512, layers=3, epochs=600, batch=64, lr=2e-3, seed=0, log=None):
    """FP32 training with Adam (learning rate decayed by cosine); returns the
    network, the output normaliser and the loss history (per epoch)."""
    norm = Normaliser(Y)
    Z = norm.fwd(Y)
    net = WarmStartNet(X.shape[1], hidden, layers, Y.shape[1], seed=seed)
    rng = np.random.default_rng(seed)
    nb = max(1, len(X) // batch)
    total = epochs * nb
    hist = []
    it = 0
    for ep in range(epochs):
        perm = rng.permutation(len(X))
        acc = 0.0
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            a = lr * 0.5 * (1.0 + np.cos(np.pi * it / total))
            acc += net.train_step(X[idx], Z[idx], max(a, 1e-6))
            it += 1
        hist.append(acc / nb)
        if log and (ep % 50 == 0 or ep == epochs - 1):
            log(f"epoch {ep}: loss {hist[-1]:.3e}")
    return net, norm, hist


def evaluate(net, norm, groups, refine, degree, dt, tol, check_every, max_steps, fallback_residual=10.0,
             bf16=True, log=None):
    """Warm-started time to steady state for the draws of `groups`.
    Returns per-instance steps, fallbacks, prediction errors and wall time."""
    ctxs = [SolarCellGPU([instance_params(refine, degree, dt, ang, *d) for d in inst]) for ang, inst in groups]
    X = np.array([features(ang, *d) for ang, inst in groups for d in inst], np.float32)
    t0 = time.perf_counter()
    pred = norm.back(net.predict(X, bf16=bf16))
    t_pred = time.perf_counter() - t0
    cold_state = {}
    row = 0
    for ci, g in enumerate(ctxs):
        g.initial_state()
        sl = density_slices(g)
        nu = [sl[c].stop - sl[c].start for c in CARRIERS]
        for i in range(g.batch):
            st = g.get_state(i)
            cold_state[(ci, i)] = st
            new = {c: np.zeros_like(st[c]) for c in CARRIERS}
            off = 0
            for c, m in zip(CARRIERS, nu):
                new[c][sl[c]] = pred[row, off:off + m]
                off += m
            new["P"] = None
            g.set_state(new, i)
            row += 1
    fell = set()

    def fallback(ci, i, res):
        if res > fallback_residual:
            ctxs[ci].set_state(cold_state[(ci, i)], i)
            fell.add((ci, i))
            return True
        return False

    done, wall, first_res = step_groups(ctxs, tol, check_every, max_steps, log, fallback)
    steps = np.concatenate(done)
    for g in ctxs:
        g.close()
    return {"steps": steps, "fallbacks": len(fell), "wall": wall, "predict_s": t_pred, "pred": pred,
            "first_block_residual": first_res}
