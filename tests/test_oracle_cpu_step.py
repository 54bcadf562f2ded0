"""The compiled CPU restatement (oracle/cpu_step.cpp, the benchmark's CPU
baseline) against the numpy oracle: the same trajectories to 1e-10 relative.

It shares the project's host setup code (meshes, local operators, host
supernodal Cholesky) and writes every per-step operation itself, so this also
checks that setup code on the CPU against the oracle's directly assembled
systems.  The potential gets the trajectory tests' derived bar: 1e-10 of its
norm plus the potential that the density difference induces.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle.solarcell import OracleSolarCell, default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_build", "pecs_cpu_step")
GEN = {"ref": 0, "config": 1, "off": 2}

CASES = {
    "C1-config-physics": (dict(gen_mode="config", srh=True, ni=0.1), 30),
    "k2-r3-bias": (dict(k=2, r=3, phi_bi=1.0, phi_app=0.25, ne=2.0, rinf=30.0, oinf=29.0, dt=5e-4), 20),
    "k2-r3-conic-insulated-schottky": (dict(k=2, r=3, r2=0.5 + np.tan(np.radians(10.0)), L=1.5, insulated=True,
                                            schottky=True, v_rec_n=0.3, v_rec_p=0.2, mu_n=1.3, mu_p=0.7,
                                            mu_r=2.0, mu_o=0.5, eps_s=2.0, eps_e=3.0, lam2=0.5, phi_bi=0.4,
                                            phi_app=0.1, dt=5e-4), 15),
}


def _args(p, steps, out):
    keys = ("k", "r", "dt", "penalty", "mu_n", "mu_p", "mu_r", "mu_o", "eps_s", "eps_e", "lam2", "k_et", "k_ht",
            "v_rec_n", "v_rec_p", "ne", "pe", "rinf", "oinf", "phi_bi", "phi_app", "G0", "alpha", "ni", "tau_n",
            "tau_p", "r1", "r2", "H", "L")
    a = [f"{k}={float(p[k])!r}" for k in keys]
    a += [f"gen_mode={GEN[p['gen_mode']]}", f"srh={int(bool(p['srh']))}", f"insulated={int(bool(p['insulated']))}",
          f"schottky={int(bool(p['schottky']))}", f"steps={steps}", f"out={out}"]
    return a


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_cpu_step_matches_oracle(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_build/pecs_cpu_step not built (run __graft_entry__.build())")
    upd, steps = CASES[name]
    p = default_params()
    p.update(upd)
    out = str(tmp_path / "state.bin")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([EXE] + _args(p, steps, out), capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    o = OracleSolarCell(p)
    ref = o.run(steps)
    raw = np.fromfile(out)
    st = {}
    off = 0
    for key in "npro":
        n = len(ref[key])
        st[key] = raw[off:off + n]
        off += n
    st["P"] = raw[off:]
    assert len(st["P"]) == len(ref["P"])
    for key in "npro":
        assert _rel(st[key], ref[key]) < 1e-10, (key, _rel(st[key], ref[key]))
    P_own = o.solve("P", o.rhs_poisson(st))
    assert _rel(st["P"], P_own) < 1e-10
    dP = np.linalg.norm(st["P"] - ref["P"])
    assert dP <= 1e-10 * np.linalg.norm(ref["P"]) + np.linalg.norm(P_own - ref["P"])
