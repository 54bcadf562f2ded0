"""GPU tests of batched instances (pecs_create_batch): B bias / illumination /
doping points on one mesh share every factor; each instance must follow its
own oracle trajectory to the trajectory bar (1e-10 relative) and match a
single-instance GPU run to rounding.  B = 3 (6 carrier right-hand sides per
region factor), 9 (18), 20 (40) and 40 (80) exercise the FP64-MFMA sweep
with 16, 32, 64 and 128 padded right-hand sides; the multiplier solve runs
with B right-hand sides.
"""
import numpy as np
import pytest

from oracle.solarcell import OracleSolarCell

pytestmark = pytest.mark.gpu

STEPS = 6


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _batch(B, seed, r=3, k=2):
    import pecs_amd
    rng = np.random.default_rng(seed)
    prms = []
    for _ in range(B):
        p = pecs_amd.workload_params(r, k, 5e-4)
        p.update(phi_bi=0.4, phi_app=float(rng.uniform(-0.5, 0.5)), G0=float(10 ** rng.uniform(-1, 1)),
                 ne=float(10 ** rng.uniform(np.log10(0.5), np.log10(5.0))))
        prms.append(p)
    return prms


@pytest.mark.parametrize("B", [3, 9, 20, 40])
def test_batch_matches_oracle(B, seed):
    import pecs_amd
    prms = _batch(B, seed)
    _say(f"[gpu] batch B={B} r=3 k=2 seed={seed}: create + {STEPS} steps")
    g = pecs_amd.SolarCellGPU(prms)
    assert g.batch == B
    g.initial_state()
    g.step(STEPS)
    res = g.residuals()
    assert res.shape == (B,) and np.all(np.isfinite(res)) and np.all(res > 0)
    check = range(B) if B <= 20 else range(0, B, 5)  # every instance up to 20, then every fifth
    for i in check:
        st = g.get_state(i)
        ref = OracleSolarCell(prms[i]).run(STEPS)
        for c in "nproP":
            assert _rel(st[c], ref[c]) < 1e-10, (B, i, c, _rel(st[c], ref[c]))
    g.close()


def test_batch_size_limit():
    import pecs_amd
    with pytest.raises(RuntimeError, match="batch size"):
        pecs_amd.SolarCellGPU(_batch(65, 3, r=1))


def test_batch_matches_single_instance(seed):
    import pecs_amd
    prms = _batch(3, seed + 1)
    _say("[gpu] batch B=3 vs three single-instance contexts")
    g = pecs_amd.SolarCellGPU(prms)
    g.initial_state()
    g.step(STEPS)
    for i, p in enumerate(prms):
        one = pecs_amd.SolarCellGPU(p)
        one.initial_state()
        one.step(STEPS)
        a, b = g.get_state(i), one.get_state()
        for c in "nproP":
            assert _rel(a[c], b[c]) < 1e-12, (i, c, _rel(a[c], b[c]))
        assert abs(g.residuals()[i] - one.residual()) <= 1e-9 * one.residual()
        one.close()
    # set_state on one instance leaves the others untouched
    st1 = g.get_state(1)
    st0 = g.get_state(0)
    g.set_state(st0, 2)
    assert _rel(g.get_state(2)["n"], st0["n"]) == 0.0
    assert _rel(g.get_state(1)["n"], st1["n"]) == 0.0
    g.close()


def test_batch_rejects_mismatched_matrix_parameters(seed):
    import pecs_amd
    prms = _batch(2, seed)
    prms[1]["mu_n"] = 2.0 * prms[0]["mu_n"]
    with pytest.raises(RuntimeError, match="enters the matrices"):
        pecs_amd.SolarCellGPU(prms)


# PECS_SWEEP_FORCE=m,rows,ks: one MFMA sweep shape on every level (m blocks of
# 16 rows per wave; rows per tile; column segments of ks, 0 unsplit), so the
# shapes the tuner picks only on large meshes (several blocks per wave, split
# tiles and their reduction) run here against the oracle.  Split tiles add
# their segments in a fixed order: they agree with unsplit ones to rounding.
SHAPES = ["1,64,0", "1,32,0", "2,128,0", "4,256,0", "4,256,128", "1,64,128", "2,128,256", "1,32,512",
          "1,32,0,1", "2,64,128,1", "4,128,0,1",  # split 1: right-hand-side groups over waves
          "1,64,0,2", "1,64,256,2",  # split 2: the vector operand staged in LDS (more than 16 right-hand sides)
          "2,128,0,2", "2,64,0,2", "2,128,256,2"]  # ... two blocks per wave (a 64-row tile leaves 2 waves idle)


# m = 4 is built for 128 padded right-hand sides only with the groups split
# over waves (unsplit, the accumulators and operands of 4 blocks would not fit
# the register file)
WIDE = ["1,64,0", "2,128,0", "2,128,256", "1,32,512", "1,16,0,1", "2,32,256,1", "4,64,0,1", "4,64,256,1",
        "1,64,0,2", "1,64,128,2", "2,128,0,2", "2,128,128,2"]


@pytest.mark.parametrize("B,r", [(3, 4), (9, 3), (40, 3)])
def test_forced_sweep_shapes(B, r, seed):
    import os
    import pecs_amd
    prms = _batch(B, seed + 7, r=r)
    shapes = SHAPES if B <= 32 else WIDE
    refs = [OracleSolarCell(p).run(STEPS) for p in prms]
    out = {}
    for shape in shapes:
        _say(f"[gpu] forced MFMA sweep shape {shape}: B={B} r={r} k=2, {STEPS} steps")
        os.environ["PECS_SWEEP_FORCE"] = shape
        try:
            g = pecs_amd.SolarCellGPU(prms)
        finally:
            del os.environ["PECS_SWEEP_FORCE"]
        f = [int(x) for x in shape.split(",")]
        m, rows = f[0], f[1]
        w = -1 if len(f) > 3 and f[3] == 1 and 2 * B > 16 else 0
        if len(f) > 3 and f[3] == 2 and 2 * B > 16:
            w = -2
        cfg = g.sweep_config(0)
        assert all(t[0] == w and t[1] == m and t[2] == rows and t[3] == w and t[4] == m and t[5] == rows
                   for t in cfg), cfg
        g.initial_state()
        g.step(STEPS)
        out[shape] = [g.get_state(i) for i in range(B)]
        g.close()
        for i in range(B):
            for c in "nproP":
                assert _rel(out[shape][i][c], refs[i][c]) < 1e-10, (shape, i, c)
    for shape in shapes[1:]:
        for i in range(B):
            for c in "nproP":
                assert _rel(out[shape][i][c], out[shapes[0]][i][c]) < 1e-12, (shape, i, c)


def test_forced_sweep_shape_rejected():
    import os
    import pecs_amd
    # split covers 32 rows at 32 columns; the LDS-staged sweep takes m 1 or 2 and at most 64 m rows
    for B, shape in ((3, "3,64,0"), (40, "4,256,0"), (9, "1,64,0,1"), (9, "4,128,0,2"), (9, "1,128,0,2"),
                     (9, "2,144,0,2")):
        os.environ["PECS_SWEEP_FORCE"] = shape
        try:
            with pytest.raises(RuntimeError, match="PECS_SWEEP_FORCE"):
                pecs_amd.SolarCellGPU(_batch(B, 5, r=2))
        finally:
            del os.environ["PECS_SWEEP_FORCE"]


@pytest.mark.parametrize("shape", ["1,64,0,2", "2,128,256,2"])
def test_single_solves_on_lds_shaped_factors(shape, seed):
    """Single solves (one right-hand side) on a batched context whose factors
    were shaped for the LDS-staged sweep (17 instances: 34 carrier and 17
    multiplier right-hand sides per factor, so both take the shape): the
    narrow solves run the plain MFMA sweep on the same tiles and match the
    oracle's solves."""
    import os
    import pecs_amd
    _say(f"[gpu] single solves on LDS-shaped factors {shape}: B=17 r=3 k=2")
    prms = _batch(17, seed + 11, r=3)
    os.environ["PECS_SWEEP_FORCE"] = shape
    try:
        g = pecs_amd.SolarCellGPU(prms)
    finally:
        del os.environ["PECS_SWEEP_FORCE"]
    try:
        for which in (0, 1, 2):
            assert all(t[0] == -2 and t[3] == -2 for t in g.sweep_config(which)), (which, g.sweep_config(which))
        o = OracleSolarCell(prms[0])
        rng = np.random.default_rng(seed)
        for which, key in ((pecs_amd.ELECTRONS, "n"), (pecs_amd.HOLES, "p"), (pecs_amd.REDUCTANTS, "r"),
                           (pecs_amd.OXIDANTS, "o"), (pecs_amd.POISSON, "P")):
            n = {"n": g.n_semi, "p": g.n_semi, "r": g.n_elec, "o": g.n_elec, "P": g.n_poisson}[key]
            b = rng.standard_normal(n)
            x = g.solve(which, b)
            assert _rel(x, o.solve(key, b)) < 1e-11, (key, _rel(x, o.solve(key, b)))
    finally:
        g.close()
