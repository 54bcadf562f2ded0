"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bars (FP64): right-hand sides 1e-13 relative (same quadrature, different
summation order); single solves 1e-11 relative (the oracle solves the full
[J; u] / mixed system with SuperLU, the GPU the SPD Schur / hybridised
systems with a supernodal Cholesky); full trajectories 1e-10 relative in
densities and potential (the trajectory bar).
"""
import numpy as np
import pytest

from oracle.solarcell import OracleSolarCell, default_params

pytestmark = pytest.mark.gpu


def _gpu(prm):
    import pecs_amd
    return pecs_amd.SolarCellGPU(prm)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _say(msg):
    print(msg, flush=True)


def _random_state(o, rng, scale=1e-3):
    st = o.initial_state()
    out = {}
    for k in "npro":
        u = st[k].copy()
        u += scale * rng.standard_normal(u.shape) * (1.0 + np.abs(u))
        out[k] = u
    out["P"] = st["P"] + scale * rng.standard_normal(st["P"].shape)
    return out


CASES = {
    "C1": dict(),
    "C1-config-physics": dict(gen_mode="config", srh=True, ni=0.1),
    "k2-r4-conic-insulated-schottky": dict(k=2, r=4, r2=0.5 + np.tan(np.radians(10.0)), L=1.5, insulated=True,
                                           schottky=True, v_rec_n=0.3, v_rec_p=0.2, mu_n=1.3, mu_p=0.7,
                                           mu_r=2.0, mu_o=0.5, eps_s=2.0, eps_e=3.0, lam2=0.5, phi_bi=0.4,
                                           phi_app=0.1, dt=5e-4),
    "k2-r3-bias": dict(k=2, r=3, phi_bi=1.0, phi_app=0.25, ne=2.0, rinf=30.0, oinf=29.0, dt=5e-4),
}


def _prm(name):
    p = default_params()
    p.update(CASES[name])
    return p


@pytest.mark.parametrize("name", list(CASES))
def test_rhs_parity(name, seed):
    _say(f"[gpu] rhs parity case={name} seed={seed}")
    p = _prm(name)
    o = OracleSolarCell(p)
    g = _gpu(p)
    rng = np.random.default_rng(seed)
    st = _random_state(o, rng)
    bn, bp = o.rhs_semi(st)
    gn, gp = g.rhs_semiconductor(st["n"], st["p"], st["P"], st["r"], st["o"])
    assert _rel(gn, bn) < 1e-13 and _rel(gp, bp) < 1e-13, (_rel(gn, bn), _rel(gp, bp))
    br, bo = o.rhs_elec(st)
    gr, go = g.rhs_electrolyte(st["r"], st["o"], st["P"], st["n"], st["p"])
    assert _rel(gr, br) < 1e-13 and _rel(go, bo) < 1e-13, (_rel(gr, br), _rel(go, bo))
    bP = o.rhs_poisson(st)
    gP = g.rhs_poisson(st["n"], st["p"], st["r"], st["o"])
    assert np.linalg.norm(gP - bP) <= 1e-13 * max(np.linalg.norm(bP), 1.0), np.linalg.norm(gP - bP)
    g.close()


@pytest.mark.parametrize("name", list(CASES))
def test_solve_parity(name, seed):
    _say(f"[gpu] solve parity case={name} seed={seed}")
    import pecs_amd
    p = _prm(name)
    o = OracleSolarCell(p)
    g = _gpu(p)
    rng = np.random.default_rng(seed + 1)
    for which, key in ((pecs_amd.ELECTRONS, "n"), (pecs_amd.HOLES, "p"), (pecs_amd.REDUCTANTS, "r"),
                       (pecs_amd.OXIDANTS, "o"), (pecs_amd.POISSON, "P")):
        n = {"n": g.n_semi, "p": g.n_semi, "r": g.n_elec, "o": g.n_elec, "P": g.n_poisson}[key]
        b = rng.standard_normal(n)
        x_ref = o.solve(key, b)
        x = g.solve(which, b)
        assert _rel(x, x_ref) < 1e-11, (key, _rel(x, x_ref))
    g.close()


@pytest.mark.parametrize("name,nsteps", [("C1", 200), ("C1-config-physics", 50),
                                         ("k2-r4-conic-insulated-schottky", 20), ("k2-r3-bias", 40)])
def test_trajectory_parity(name, nsteps):
    _say(f"[gpu] trajectory case={name} steps={nsteps}")
    p = _prm(name)
    o = OracleSolarCell(p)
    g = _gpu(p)
    st = o.initial_state()
    g.initial_state()
    gs = g.get_state()
    for k in "npro":
        assert np.array_equal(gs[k], st[k])
    assert np.linalg.norm(gs["P"] - st["P"]) <= 1e-12 * max(1.0, np.linalg.norm(st["P"]))
    ref = o.run(nsteps, st)
    g.step(nsteps)
    gs = g.get_state()
    for k in "npro":
        nb = o.cfe.nb
        nc = len(ref[k]) // (3 * nb)
        d_ref = ref[k][2 * nb * nc:]
        d_gpu = gs[k][2 * nb * nc:]
        assert _rel(d_gpu, d_ref) < 1e-10, (k, _rel(d_gpu, d_ref))
        assert _rel(gs[k], ref[k]) < 1e-10, (k, _rel(gs[k], ref[k]))
    assert np.linalg.norm(gs["P"] - ref["P"]) <= 1e-10 * max(np.linalg.norm(ref["P"]), 1e-3), \
        np.linalg.norm(gs["P"] - ref["P"])
    g.close()
