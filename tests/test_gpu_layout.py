"""pecs_set_dof_layout(PECS_LAYOUT_DEALII) through the host-pointer calls.

With deal.II's layout switched on, every host-pointer call that carries a
Poisson vector takes and returns it in the reference's layout: the state
(get/set), the Poisson input of the carrier right-hand sides, the Poisson
right-hand side and both vectors of the Poisson solve.  Each is checked
against the same call in the library's own layout followed (or preceded) by
the stateless conversion, which tests/test_dealii_layout.py pins.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("conic", [False, True])
def test_dealii_layout_host_calls(conic, seed):
    import pecs_amd
    p = pecs_amd.workload_params(3, 2, 5e-4)
    p.update(phi_bi=0.4, phi_app=0.1)
    if conic:
        p.update(r2=0.62, L=1.4, insulated=True, schottky=True)
    _say(f"[gpu] deal.II Poisson layout through the host calls (refinement 3, k=2, conic={conic})")
    conv = lambda v, rhs=False: pecs_amd.poisson_layout_convert(p, v, to_dealii=True, rhs=rhs)
    back = lambda v, rhs=False: pecs_amd.poisson_layout_convert(p, v, to_dealii=False, rhs=rhs)
    g = pecs_amd.SolarCellGPU(p)
    g.initial_state()
    g.step(5)
    own = g.get_state()
    b_own = g.rhs_poisson(own["n"], own["p"], own["r"], own["o"])
    bn_own, bp_own = g.rhs_semiconductor(own["n"], own["p"], own["P"], own["r"], own["o"])
    br_own, bo_own = g.rhs_electrolyte(own["r"], own["o"], own["P"], own["n"], own["p"])
    x_own = g.solve(pecs_amd.POISSON, b_own)
    g.set_dof_layout(pecs_amd.LAYOUT_DEALII)
    dii = g.get_state()
    for c in "npro":
        assert np.array_equal(dii[c], own[c])
    assert np.array_equal(dii["P"], conv(own["P"]))
    b_d = g.rhs_poisson(own["n"], own["p"], own["r"], own["o"])
    assert np.array_equal(b_d, conv(b_own, rhs=True))
    bn, bp = g.rhs_semiconductor(dii["n"], dii["p"], dii["P"], dii["r"], dii["o"])
    br, bo = g.rhs_electrolyte(dii["r"], dii["o"], dii["P"], dii["n"], dii["p"])
    for a, b in ((bn, bn_own), (bp, bp_own), (br, br_own), (bo, bo_own)):
        assert _rel(a, b) < 1e-14
    x_d = g.solve(pecs_amd.POISSON, b_d)
    assert _rel(back(x_d), x_own) < 1e-13
    # a state set in deal.II's layout reads back the same in the library's own
    dii["P"] = conv(own["P"])
    g.set_state(dii)
    g.set_dof_layout(pecs_amd.LAYOUT_OWN)
    again = g.get_state()
    assert _rel(again["P"], own["P"]) < 1e-14
    with pytest.raises(RuntimeError, match="layout"):
        g.set_dof_layout(7)
    g.close()
