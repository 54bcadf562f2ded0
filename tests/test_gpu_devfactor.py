"""GPU tests of the device-side numeric factorisation (multifrontal Cholesky
on the GPU: front scatter, extend-add, batched rocSOLVER / rocBLAS dense
calls, sweep-layout kernels).  The device factors must give the same
trajectories as the host factorisation (PECS_HOST_FACTOR=1) to rounding, for
the scalar sweeps (single instance) and the MFMA sweeps over blocked copies
(batched instances), and the trajectories against the oracle are covered by
the other GPU tests, which run on the device factors by default.
"""
import numpy as np
import pytest

import pecs_amd

pytestmark = pytest.mark.gpu

STEPS = 5


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _run(prms, steps):
    g = pecs_amd.SolarCellGPU(prms)
    g.initial_state()
    g.step(steps)
    B = g.batch
    out = [g.get_state(i) for i in range(B)], g.residuals(), g.stats()
    g.close()
    return out


# (refinement, degree, conic interface, separate hole factor)
@pytest.mark.parametrize("r,k,conic,mu_p", [(4, 2, False, 1.0), (4, 3, True, 1.0), (5, 1, False, 0.5)])
def test_device_factor_matches_host_factor(r, k, conic, mu_p, monkeypatch):
    p = pecs_amd.workload_params(r, k, 5e-4 if k > 1 else 1e-3)
    if conic:
        p.update(r1=0.5, r2=0.5 + float(np.tan(np.radians(10.0))), L=1.5)
    # a built-in and applied bias keep the potential O(1) (at zero bias it is
    # a rounding-level residue of the charge balance and has no relative scale)
    p.update(mu_p=mu_p, phi_bi=0.4, phi_app=0.1)
    _say(f"[gpu] device vs host factor r={r} k={k} conic={conic} mu_p={mu_p}")
    monkeypatch.delenv("PECS_HOST_FACTOR", raising=False)
    dev, dres, dstats = _run(p, STEPS)
    monkeypatch.setenv("PECS_HOST_FACTOR", "1")
    host, hres, hstats = _run(p, STEPS)
    assert dstats["device_numeric_s"] > 0 and dstats["device_numeric_flops"] > 0
    assert hstats["device_numeric_s"] == 0
    assert dstats["factor_semi_doubles"] == hstats["factor_semi_doubles"]
    for c in "nproP":
        assert _rel(dev[0][c], host[0][c]) < 1e-11, (c, _rel(dev[0][c], host[0][c]))
    assert abs(dres[0] - hres[0]) <= 1e-8 * hres[0]


def test_device_factor_batched_mfma_matches_host_factor(monkeypatch):
    rng = np.random.default_rng(7)
    prms = []
    for _ in range(3):
        p = pecs_amd.workload_params(4, 2, 5e-4)
        p.update(phi_bi=0.4, phi_app=float(rng.uniform(-0.5, 0.5)))
        prms.append(p)
    _say("[gpu] device vs host factor, batch of 3 (MFMA sweeps over blocked copies)")
    monkeypatch.delenv("PECS_HOST_FACTOR", raising=False)
    dev, _, _ = _run(prms, STEPS)
    monkeypatch.setenv("PECS_HOST_FACTOR", "1")
    host, _, _ = _run(prms, STEPS)
    for i in range(3):
        for c in "nproP":
            assert _rel(dev[i][c], host[i][c]) < 1e-11, (i, c, _rel(dev[i][c], host[i][c]))
