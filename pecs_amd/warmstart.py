"""Learned warm start for time stepping to steady state.

A network trained on steady states that the time stepper itself produced
predicts the steady densities of a new (applied bias, generation amplitude,
donor doping, cone half-angle) draw; the stepper then starts from the
prediction instead of the equilibrium state and still decides convergence
(the same relative steady residual).

Output space: the densities of the four carriers resampled on a G x G grid
per region (G = 64): the logical cell centres of each region's coarse
quadrilateral, the same points for every refinement and cone angle
(pecs_sample_grid).  A prediction is put on any instance's own DG mesh by
pecs_set_state_from_grid (nodal values of the grid's bilinear interpolant;
currents zero; the potential solved from the densities), so one network
serves refinements 5-8 and every cone angle.

Network (pecs_amd.WarmStartNet, hand-written HIP kernels on the matrix
cores, FP32 training, BF16 inference): a 4-layer MLP encoder, 4 -> 3 x 256
(ReLU) -> 64 x 8 x 8 feature map, and a convolutional decoder: three stages
of nearest x2 upsampling + 3x3 convolution (64, 32, 16 channels, ReLU) and a
final 3x3 convolution to the 4 density channels on the 64 x 64 grid.

Fallback: the predicted residual is the steady residual of the predicted
state preconditioned by one implicit step, relative to the state: the
change of the four densities over the first step from the prediction,
||u^1 - u_pred|| / ||u_pred|| (pecs_last_density_change).  Above
`fallback_residual` (1e-2: the prediction would move by more than 1 % in one
step) the instance restarts from the equilibrium state; the step it took
still counts.  (Measured on refinements 3-4 at dt = 1e-3: 2e-3 for the
equilibrium start, 1e-3 for the exact steady densities resampled through
the 64 x 64 grid, 6e-3 - 8e-3 for predictions with 1.3 - 1.6 % error, 0.29
for the steady state with log-normal noise of width 0.7 on every grid value.)

Draws come in groups that share one mesh (refinement and cone angle); each
group is one batched context (pecs_create_batch), and the groups are stepped
together (pecs_sweep_*).
"""
import time

import numpy as np

from . import SolarCellGPU, SolarCellSweep, WarmStartNet, workload_params

CARRIERS = "npro"
GRID = 64
DECODER = dict(hidden=256, layers=3, c0=64, g0=8, stages=[(64, 1), (32, 1), (16, 1), (4, 0)])


def draw_groups(rng, ngroups, group, refines=(5,)):
    """ngroups meshes: refinement drawn uniformly from `refines`, cone angle
    U[0, 15] degrees; per instance applied bias U[-0.5, 0.5], generation
    amplitude logU[0.1, 10], donor doping logU[0.5, 5].
    Returns [(refine, angle, [(bias, amplitude, doping), ...]), ...]."""
    out = []
    for _ in range(ngroups):
        ref = int(rng.choice(np.asarray(refines)))
        ang = float(rng.uniform(0.0, 15.0))
        inst = [(float(rng.uniform(-0.5, 0.5)), float(10 ** rng.uniform(-1, 1)),
                 float(10 ** rng.uniform(np.log10(0.5), np.log10(5.0)))) for _ in range(group)]
        out.append((ref, ang, inst))
    return out


def features(ang, app, g0, ne):
    """Inputs scaled to about [-1, 1]."""
    return [app / 0.5, np.log10(g0), (np.log10(ne) - 0.2) / 0.5, ang / 7.5 - 1.0]


def instance_params(refine, degree, dt, ang, app, g0, ne):
    p = workload_params(refine, degree, dt)
    p.update(r1=0.5, r2=0.5 + float(np.tan(np.radians(ang))), L=1.5, phi_bi=0.4, phi_app=app, G0=g0, ne=ne)
    return p


def make_contexts(groups, degree, dt):
    return [SolarCellGPU([instance_params(ref, degree, dt, ang, *d) for d in inst]) for ref, ang, inst in groups]


def step_groups(ctxs, tol, check_every, max_steps, log=None, fallback=None):
    """Step batched contexts together until every instance is steady.
    Returns the steps per (context, instance), the wall seconds and the
    predicted residuals (empty without a fallback).

    With fallback(ci, i, rho): one step first; rho is the relative change of
    the densities over it; fallback returns True when it restarted the
    instance (its count then goes on from 1)."""
    sw = SolarCellSweep(ctxs, min(4, len(ctxs)))
    done = [np.full(g.batch, -1) for g in ctxs]
    rho = []
    active = np.ones(len(ctxs), bool)
    t0 = time.perf_counter()
    n = 0
    if fallback is not None:
        sw.step(1)
        n = 1
        for ci, g in enumerate(ctxs):
            r = g.density_change()
            rho.extend(r.tolist())
            for i in range(g.batch):
                fallback(ci, i, r[i])
    while active.any() and n < max_steps:
        k = min(check_every, max_steps - n)
        sw.step(k)
        n += k
        for ci, g in enumerate(ctxs):
            if not active[ci]:
                continue
            r = g.residuals()
            d = done[ci]
            d[(d < 0) & (r <= tol)] = n
            if np.all(d >= 0):
                active[ci] = False
        sw.set_active(active)
        if log and n % 500 == 0:
            log(f"step {n}: {int(active.sum())} of {len(ctxs)} groups still stepping")
    wall = time.perf_counter() - t0
    sw.close()
    return done, wall, np.array(rho)


def generate(groups, degree, dt, tol, check_every, max_steps, G=GRID, log=None):
    """Steady states of every draw, cold-started from the equilibrium state.
    Returns (X features, Y grids [n_draws, 4 G^2], steps, wall, setup)."""
    t0 = time.perf_counter()
    ctxs = make_contexts(groups, degree, dt)
    setup = time.perf_counter() - t0
    for g in ctxs:
        g.initial_state()
    done, wall, _ = step_groups(ctxs, tol, check_every, max_steps, log)
    X, Y, steps = [], [], []
    for (ref, ang, inst), g, d in zip(groups, ctxs, done):
        for i, dr in enumerate(inst):
            X.append(features(ang, *dr))
            Y.append(g.sample_grid(G, i).ravel())
            steps.append(int(d[i]))
        g.close()
    return np.array(X, np.float32), np.array(Y), np.array(steps), wall, setup


class Normaliser:
    """Per-output standardisation, of log(y) when log=True (the densities
    span orders of magnitude across the draws: the network then works in
    relative errors)."""

    FLOOR = 1e-12

    def __init__(self, Y, log=False):
        self.log = log
        T = np.log(np.maximum(Y, self.FLOOR)) if log else Y
        self.mu = T.mean(axis=0)
        sd = T.std(axis=0)
        self.sd = np.maximum(sd, 1e-6 * max(np.abs(T).max(), 1e-30))

    def fwd(self, Y):
        T = np.log(np.maximum(Y, self.FLOOR)) if self.log else Y
        return ((T - self.mu) / self.sd).astype(np.float32)

    def back(self, Z):
        T = Z.astype(np.float64) * self.sd + self.mu
        return np.exp(T) if self.log else T


def make_net(spec=None, seed=0):
    s = dict(DECODER, **(spec or {}))
    return WarmStartNet(4, s["hidden"], s["layers"], 0, seed=seed, decoder=(s["c0"], s["g0"], s["stages"]))


def train(X, Y, spec=None, epochs=300, batch=32, lr=2e-3, seed=0, log=None, log_targets=True):
    """FP32 training with Adam (learning rate decayed by cosine) on the
    standardised log-densities; returns the network, the output normaliser
    and the loss history (per epoch)."""
    norm = Normaliser(Y, log=log_targets)
    Z = norm.fwd(Y)
    net = make_net(spec, seed)
    assert net.nout == Y.shape[1], (net.nout, Y.shape)
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


def predict(net, norm, groups, bf16=True, G=GRID):
    X = np.array([features(ang, *d) for _, ang, inst in groups for d in inst], np.float32)
    return norm.back(net.predict(X, bf16=bf16)).reshape(len(X), 4, G, G)


def evaluate(net, norm, groups, degree, dt, tol, check_every, max_steps, fallback_residual=1e-2, bf16=True,
             log=None, G=GRID, perturb=None):
    """Warm-started time to steady state for the draws of `groups`: predict
    the grids, project them onto every instance's mesh, step with the
    predicted-residual fallback.  perturb(grids) may alter the predictions
    (tests of the fallback).  Returns per-instance steps, fallbacks,
    predicted residuals and the times (predict + projection, stepping)."""
    ctxs = make_contexts(groups, degree, dt)
    cold = {}
    for ci, g in enumerate(ctxs):
        g.initial_state()
        for i in range(g.batch):
            cold[(ci, i)] = g.get_state(i)
    t0 = time.perf_counter()
    pred = predict(net, norm, groups, bf16, G)
    if perturb is not None:
        pred = perturb(pred)
    row = 0
    for g in ctxs:
        for i in range(g.batch):
            g.set_state_from_grid(pred[row], i)
            row += 1
    t_pred = time.perf_counter() - t0
    fell = set()

    def fallback(ci, i, rho):
        if not rho <= fallback_residual:  # NaN falls back too
            ctxs[ci].set_state(cold[(ci, i)], i)
            fell.add((ci, i))
            return True
        return False

    done, wall, rho = step_groups(ctxs, tol, check_every, max_steps, log, fallback)
    for g in ctxs:
        g.close()
    return {"steps": np.concatenate(done), "fallbacks": len(fell), "fell": sorted(fell), "wall": wall,
            "predict_s": t_pred, "pred": pred, "predicted_residual": rho}


def cold_start(groups, degree, dt, tol, check_every, max_steps, log=None, G=GRID):
    """The same draws from the equilibrium state (the method without the
    network): steps, wall seconds and the steady densities on the grid."""
    ctxs = make_contexts(groups, degree, dt)
    for g in ctxs:
        g.initial_state()
    done, wall, _ = step_groups(ctxs, tol, check_every, max_steps, log)
    grids = np.array([g.sample_grid(G, i) for g in ctxs for i in range(g.batch)])
    for g in ctxs:
        g.close()
    return np.concatenate(done), wall, grids
