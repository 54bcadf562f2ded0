"""GPU tests of the graph-replayed time step: pecs_step captures one step as
a HIP graph at its first call and replays it.  On one context (one set of
factors and tuned tile shapes), replays interleaved with eager work (state
reads and writes) and synchronisations must give bitwise the same states and
residuals as launching the kernels one by one from the same state, for the
scalar sweeps (single instance, shared and separate factors) and the MFMA
sweeps (batched instances)."""
import numpy as np
import pytest

import pecs_amd

pytestmark = pytest.mark.gpu

CHUNKS = (3, 1, 4, 2)


def _say(msg):
    print(msg, flush=True)


def _params(B, seed, mu_p=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(B):
        p = pecs_amd.workload_params(4, 2, 5e-4)
        p.update(phi_bi=0.4, phi_app=float(rng.uniform(-0.5, 0.5)), mu_p=mu_p)
        out.append(p)
    return out if B > 1 else out[0]


def _states(g):
    return [g.get_state(i) for i in range(g.batch)]


def _reset(g, states):
    for i, st in enumerate(states):
        g.set_state(st, i)


def _same(a, b, what):
    for x, y in zip(a, b):
        for c in "nproP":
            assert np.array_equal(x[c], y[c]), (what, c, np.abs(x[c] - y[c]).max())


@pytest.mark.parametrize("B,mu_p", [(1, 1.0), (1, 0.5), (3, 1.0)])
def test_graph_replay_matches_eager(B, mu_p, seed):
    prm = _params(B, seed, mu_p)
    _say(f"[gpu] graph replay vs eager, batch {B}, mu_p={mu_p}")
    g = pecs_amd.SolarCellGPU(prm)
    g.initial_state()
    s0 = _states(g)
    g.use_graph(False)
    ref = []
    for n in CHUNKS:
        g.step(n)
        ref.append((_states(g), g.residuals()))
    _reset(g, s0)
    g.use_graph(True)
    for (st, res), n in zip(ref, CHUNKS):
        g.step(n)
        _same(_states(g), st, n)
        assert np.array_equal(g.residuals(), res)
    # a second pass of replays from the same state, after the eager writes
    _reset(g, s0)
    g.step(sum(CHUNKS))
    _same(_states(g), ref[-1][0], "second pass")
    g.close()


def test_graph_and_timing_mix(seed):
    """Per-phase timing launches eagerly; switching back to the graph goes on
    from the same state."""
    prm = _params(1, seed)
    _say("[gpu] graph replay with timing switched on and off")
    g = pecs_amd.SolarCellGPU(prm)
    g.initial_state()
    s0 = _states(g)
    g.use_graph(False)
    g.step(6)
    ref = _states(g)
    _reset(g, s0)
    g.use_graph(True)
    g.step(2)
    g.enable_timing(True)
    g.step(2)
    g.enable_timing(False)
    g.step(2)
    _same(_states(g), ref, "mixed")
    g.close()
