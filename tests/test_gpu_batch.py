"""GPU tests of batched instances (pecs_create_batch): B bias / illumination /
doping points on one mesh share every factor; each instance must follow its
own oracle trajectory to the trajectory bar (1e-10 relative) and match a
single-instance GPU run to rounding.  B = 3 (6 carrier right-hand sides per
region factor) and B = 9 (18) exercise the FP64-MFMA sweep with 16 and 32
padded right-hand sides; the multiplier solve runs with B right-hand sides.
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


@pytest.mark.parametrize("B", [3, 9])
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
    for i, p in enumerate(prms):
        st = g.get_state(i)
        ref = OracleSolarCell(p).run(STEPS)
        for c in "nproP":
            assert _rel(st[c], ref[c]) < 1e-10, (B, i, c, _rel(st[c], ref[c]))
    g.close()


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
