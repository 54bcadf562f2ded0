"""GPU tests of uneven-mesh sweeps (pecs_sweep_*): contexts with different
refinements, geometries and degrees stepped together (their captured step
graphs launched on several streams, or as child graphs of one graph launched
once per step).  Each member must follow exactly the trajectory it follows
when stepped alone (same context: bitwise), the oracle to the trajectory bar,
and inactive members must stay untouched."""
import numpy as np
import pytest

import pecs_amd
from oracle.solarcell import OracleSolarCell

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _members():
    out = []
    for r, k, angle, app in ((3, 2, 0.0, 0.1), (4, 2, 10.0, -0.3), (3, 1, 5.0, 0.4)):
        p = pecs_amd.workload_params(r, k, 5e-4 if k > 1 else 1e-3)
        p.update(phi_bi=0.4, phi_app=app)
        if angle:
            p.update(r1=0.5, r2=0.5 + float(np.tan(np.radians(angle))), L=1.5)
        out.append(p)
    return out


def _state(g):
    return g.get_state(0)


@pytest.mark.parametrize("nstreams", [0, 1, 2])  # 0: one graph launch per step for all members
def test_sweep_matches_members_alone(nstreams):
    prms = _members()
    _say(f"[gpu] sweep of 3 uneven meshes, {nstreams} stream(s)")
    gs = [pecs_amd.SolarCellGPU(p) for p in prms]
    s0 = []
    for g in gs:
        g.initial_state()
        s0.append(_state(g))
    ref = []
    for g in gs:
        g.step(5)
        ref.append((_state(g), g.residual()))
    sw = pecs_amd.SolarCellSweep(gs, nstreams)
    # several rounds: the residual reads the state saved before the last step,
    # which must queue behind the previous step's graph on every stream
    for rnd in range(3):
        for g, st in zip(gs, s0):
            g.set_state(st)
        sw.step(2)
        sw.step(3)
        for i, g in enumerate(gs):
            st = _state(g)
            for c in "nproP":
                assert np.array_equal(st[c], ref[i][0][c]), (rnd, i, c)
            assert g.residual() == ref[i][1], (rnd, i, g.residual(), ref[i][1])
    # the oracle, for the conic member
    o = OracleSolarCell(prms[1]).run(5)
    st = _state(gs[1])
    for c in "nproP":
        rel = np.linalg.norm(st[c] - o[c]) / np.linalg.norm(o[c])
        assert rel < 1e-10, (c, rel)
    # an inactive member keeps its state; the others go on
    sw.set_active([1, 0, 1])
    before = _state(gs[1])
    sw.step(3)
    after = _state(gs[1])
    for c in "nproP":
        assert np.array_equal(before[c], after[c]), c
    sw.close()
    for g in gs:
        g.close()


def test_sweep_rejects_duplicates():
    p = pecs_amd.workload_params(3, 1, 1e-3)
    g = pecs_amd.SolarCellGPU(p)
    with pytest.raises(RuntimeError, match="twice"):
        pecs_amd.SolarCellSweep([g, g])
    with pytest.raises(RuntimeError, match="nstreams"):
        pecs_amd.SolarCellSweep([g], 9)
    g.close()
