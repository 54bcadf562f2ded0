"""GPU parity on boundary-layer locally refined meshes (local_refine = 1).

The HIP path applies the couplings across hanging faces in separate kernels
after the Schur and multiplier kernels (the faces are absent from the
neighbour lists those read) and factors the lattice-embedded Schur and
multiplier matrices with adaptively sized separators.  Checked against the
oracle (LocalRefinedMesh, hanging-face LDG fluxes, RT hanging-node
constraints) with the uniform-mesh bars: right-hand sides 1e-13, single
solves 1e-11, trajectories 1e-10 relative (potential: 1e-10 plus the part
the density difference induces).  The "reference-input" case is the
reference's input_file.prm mesh (global refinements 4, local refinements 1,
boundary layer 0.1, radii 0.3 / 0.6).
"""
import numpy as np
import pytest

from oracle.solarcell import OracleSolarCell, default_params

pytestmark = pytest.mark.gpu

CASES = {
    "lr-k1-r3": dict(k=1, r=3, local_refine=1, bl=0.05),
    "lr-k2-r3-conic-insulated-schottky": dict(k=2, r=3, r1=0.3, r2=0.6, L=1.2, local_refine=1, bl=0.1,
                                              insulated=True, schottky=True, v_rec_n=0.3, v_rec_p=0.2,
                                              mu_n=1.3, mu_p=0.7, mu_r=2.0, mu_o=0.5, eps_s=2.0, eps_e=3.0,
                                              lam2=0.5, phi_bi=0.4, phi_app=0.1, dt=5e-4),
    "lr-k3-r2-config-physics": dict(k=3, r=2, local_refine=1, bl=0.1, dt=5e-4, gen_mode="config", srh=True,
                                    ni=0.1),
    "reference-input": dict(k=1, r=4, r1=0.3, r2=0.6, L=1.0, H=1.0, local_refine=1, bl=0.1, phi_bi=0.4,
                            phi_app=0.1),
}


def _prm(name):
    p = default_params()
    p.update(CASES[name])
    return p


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


@pytest.mark.parametrize("name", list(CASES))
def test_rhs_and_solve_parity(name, seed):
    import pecs_amd
    _say(f"[gpu] local refinement rhs/solve parity case={name} seed={seed}")
    p = _prm(name)
    o = OracleSolarCell(p)
    g = pecs_amd.SolarCellGPU(p)
    try:
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
        assert np.linalg.norm(gP - bP) <= 1e-13 * max(np.linalg.norm(bP), 1.0)
        for which, key in ((pecs_amd.ELECTRONS, "n"), (pecs_amd.HOLES, "p"), (pecs_amd.REDUCTANTS, "r"),
                           (pecs_amd.OXIDANTS, "o")):
            n = g.n_semi if key in "np" else g.n_elec
            b = rng.standard_normal(n)
            x = g.solve(which, b)
            assert _rel(x, o.solve(key, b)) < 1e-11, (key, _rel(x, o.solve(key, b)))
        # the Poisson solve of an assembled right-hand side (constrained
        # fine-face entries distributed to their coarse faces)
        x = g.solve(pecs_amd.POISSON, bP)
        assert _rel(x, o.solve("P", bP)) < 1e-11, _rel(x, o.solve("P", bP))
    finally:
        g.close()


@pytest.mark.parametrize("name,nsteps", [("lr-k1-r3", 40), ("lr-k2-r3-conic-insulated-schottky", 20),
                                         ("lr-k3-r2-config-physics", 20), ("reference-input", 20)])
def test_trajectory_parity(name, nsteps):
    import pecs_amd
    _say(f"[gpu] local refinement trajectory case={name} steps={nsteps}")
    p = _prm(name)
    o = OracleSolarCell(p)
    g = pecs_amd.SolarCellGPU(p)
    try:
        st = o.initial_state()
        g.initial_state()
        ref = o.run(nsteps, st)
        g.step(nsteps)
        gs = g.get_state()
        errs = {}
        for k in "npro":
            nb = o.cfe.nb
            nc = len(ref[k]) // (3 * nb)
            errs[k + ".density"] = _rel(gs[k][2 * nb * nc:], ref[k][2 * nb * nc:])
            errs[k] = _rel(gs[k], ref[k])
        P_own = o.solve("P", o.rhs_poisson(gs))
        errs["P.step"] = _rel(gs["P"], P_own)
        induced = np.linalg.norm(P_own - ref["P"])
        dP = np.linalg.norm(gs["P"] - ref["P"])
        nP = np.linalg.norm(ref["P"])
        _say(f"[gpu] trajectory {name}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()) +
             f", P {dP / max(nP, 1e-300):.2e}")
        for k, v in errs.items():
            assert v < 1e-10, (k, v)
        assert dP <= 1e-10 * nP + induced, (dP, nP, induced)
    finally:
        g.close()


def test_batch_on_locally_refined_mesh():
    """Three instances (different biases) share the lattice factors; each
    matches its oracle trajectory."""
    import pecs_amd
    _say("[gpu] local refinement batch B=3 k=2 r=3")
    base = _prm("lr-k2-r3-conic-insulated-schottky")
    prms = [dict(base, phi_app=a) for a in (0.0, 0.1, 0.25)]
    g = pecs_amd.SolarCellGPU(prms)
    try:
        g.initial_state()
        g.step(10)
        for i, p in enumerate(prms):
            ref = OracleSolarCell(p).run(10)
            gs = g.get_state(i)
            for k in "npro":
                assert _rel(gs[k], ref[k]) < 1e-10, (i, k, _rel(gs[k], ref[k]))
    finally:
        g.close()


def test_mixed_precision_on_locally_refined_mesh():
    """FP32 cell contractions with the hanging-face corrections in FP64:
    within 1e-5 of the FP64 oracle after 20 steps.  The mode needs the
    wave-uniform cell lists, i.e. cells that are exact translates of each
    other: a planar mesh whose layer width is dyadic (1/8), so that the
    midpoint subdivision is exact -- and enough cells (refinement 5) that
    padding the 18 cell types to whole waves stays within the lists' limit."""
    import pecs_amd
    _say("[gpu] local refinement mixed precision k=1 r=5")
    p = _prm("lr-k1-r3")
    p.update(r=5, bl=0.125)
    g = pecs_amd.SolarCellGPU(p)
    try:
        g.set_mixed_precision(True)
        g.initial_state()
        g.step(20)
        gs = g.get_state()
    finally:
        g.close()
    ref = OracleSolarCell(p).run(20)
    for k in "npro":
        n = len(ref[k]) // 3
        assert _rel(gs[k][2 * n:], ref[k][2 * n:]) < 1e-5, k


def test_unsupported_features_refused():
    import pecs_amd
    g = pecs_amd.SolarCellGPU(_prm("lr-k1-r3"))
    try:
        with pytest.raises(RuntimeError, match="uniformly refined"):
            g.sample_grid(16)
        with pytest.raises(RuntimeError, match="locally refined"):
            g.set_dof_layout(pecs_amd.LAYOUT_DEALII)
    finally:
        g.close()
