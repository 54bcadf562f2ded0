"""GPU tests of the mixed-precision mode (pecs_set_mixed_precision): FP32
contractions with the cell-type tables in the Schur and Poisson cell kernels,
FP64 state, assembly and solves.  Against the FP64 path on the same inputs
the trajectory must stay within FP32 rounding of the contractions (relative
1e-4 after 20 steps; measured ≈ 1e-7) and must differ from it (the FP32
kernels ran); switching back restores the FP64 path exactly; a conic mesh
(one cell type per cell, no wave-uniform lists) is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 20


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _params(B, seed, r, k=2):
    import pecs_amd
    rng = np.random.default_rng(seed)
    prms = []
    for _ in range(B):
        p = pecs_amd.workload_params(r, k, 5e-4)
        p.update(phi_bi=0.4, phi_app=float(rng.uniform(-0.5, 0.5)), G0=float(10 ** rng.uniform(-1, 1)))
        prms.append(p)
    return prms


def _states(g, B):
    return [g.get_state(i) for i in range(B)]


@pytest.mark.parametrize("B,r", [(1, 4), (4, 4)])
def test_mixed_precision_against_fp64(B, r, seed):
    import pecs_amd
    prms = _params(B, seed, r)
    _say(f"[gpu] mixed precision vs FP64: B={B} r={r} k=2 seed={seed}, {STEPS} steps")
    g = pecs_amd.SolarCellGPU(prms if B > 1 else prms[0])
    g.initial_state()
    g.step(STEPS)
    ref = _states(g, B)
    g.set_mixed_precision(True)
    g.initial_state()
    g.step(STEPS)
    mix = _states(g, B)
    worst = 0.0
    for i in range(B):
        for c in "nproP":
            assert np.all(np.isfinite(mix[i][c]))
            d = _rel(mix[i][c], ref[i][c])
            assert d < 1e-4, (B, r, i, c, d)
            worst = max(worst, d)
    assert worst > 1e-12, worst  # the FP32 kernels ran
    g.set_mixed_precision(False)
    g.initial_state()
    g.step(STEPS)
    back = _states(g, B)
    for i in range(B):
        for c in "nproP":
            assert np.array_equal(back[i][c], ref[i][c]), (i, c)
    g.close()
    _say(f"[gpu] mixed precision: worst relative difference {worst:.2e}")


def test_mixed_precision_refused_on_conic_mesh(seed):
    import pecs_amd
    p = pecs_amd.workload_params(3, 2, 5e-4)
    p.update(r1=0.5, r2=0.6, L=1.5, phi_bi=0.4)
    _say("[gpu] mixed precision refused on a conic mesh (r=3)")
    g = pecs_amd.SolarCellGPU(p)
    assert g.uniform_lists() == 0
    with pytest.raises(Exception):
        g.set_mixed_precision(True)
    g.close()
