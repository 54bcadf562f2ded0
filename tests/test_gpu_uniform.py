"""GPU tests of the wave-uniform cell lists of the cell-local kernels
(Schur transform and recovery, Poisson condensation and recovery): on a planar
mesh the lists are on by default (PECS_SCHUR_UNIFORM unset) and the kernels
read their cell-type tables as scalar operands; with PECS_SCHUR_UNIFORM=0
they take the cells in order with per-lane table reads.  Both orders do the
same arithmetic per cell, up to the compiler's choice of fused multiply-adds
in each variant, so the trajectories must agree to rounding: 1e-11 relative
after 6 steps (measured ≈ 1e-13), for a single instance and for a batch of
4 (which also takes the Schur right-hand side's one-pass-per-right-hand-side
form).  The list-based run is also held
to the oracle at the trajectory bar (1e-10 relative).
"""
import os

import numpy as np
import pytest

from oracle.solarcell import OracleSolarCell

pytestmark = pytest.mark.gpu

STEPS = 6


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


def _run(prms, uniform):
    import pecs_amd
    old = os.environ.get("PECS_SCHUR_UNIFORM")
    os.environ["PECS_SCHUR_UNIFORM"] = "1" if uniform else "0"
    try:
        g = pecs_amd.SolarCellGPU(prms if len(prms) > 1 else prms[0])
    finally:
        if old is None:
            del os.environ["PECS_SCHUR_UNIFORM"]
        else:
            os.environ["PECS_SCHUR_UNIFORM"] = old
    flags = g.uniform_lists()
    assert flags == (31 if uniform else 0), flags
    g.initial_state()
    g.step(STEPS)
    out = [g.get_state(i) for i in range(len(prms))]
    g.close()
    return out


@pytest.mark.parametrize("B,r", [(1, 3), (1, 5), (4, 4)])
def test_uniform_lists_match_per_lane_types(B, r, seed):
    prms = _params(B, seed, r)
    _say(f"[gpu] uniform cell lists vs per-lane types: B={B} r={r} k=2 seed={seed}, {STEPS} steps")
    a = _run(prms, True)
    b = _run(prms, False)
    for i in range(B):
        for c in "nproP":
            assert _rel(a[i][c], b[i][c]) < 1e-11, (B, r, i, c, _rel(a[i][c], b[i][c]))
    if r == 3:
        for i, p in enumerate(prms):
            ref = OracleSolarCell(p).run(STEPS)
            for c in "nproP":
                assert _rel(a[i][c], ref[c]) < 1e-10, (i, c, _rel(a[i][c], ref[c]))
