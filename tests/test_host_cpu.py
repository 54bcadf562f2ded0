"""CPU tests of the C++ host layer (pecs_amd/host): the input_file.prm reader and
its mapping onto the C-ABI structs (ParameterReader.cpp:29-208,
Parameters.hpp:117-287), the reference loop's step count (SolarCell.cpp:1978-2082)
and the restart .dofs format (CarrierPair.cpp:89-121, deal.II Vector::block_write).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle.solarcell import scale_parameters

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HOST_TEST = os.path.join(ROOT, "pecs_amd", "_build", "pecs_host_test")
RUNNER = os.path.join(ROOT, "pecs_amd", "_build", "pecs_run")

PRM = """# Listing of Parameters
subsection computational
  # time to steady state
  set end time           = 5e1
  set end time 2         = 1e5
  set global refinements = 3
  set local refinements  = 0
  set restart status     = false
  set time stamps        = 100
  set time step size     = 0.05
end

subsection electrons
  set mobility               = 1350.0   # [v/cm^2]
  set recombination time     = 5e-5
  set recombination velocity = 3e5
  set transfer rate          = 1e-19
end

subsection holes
  set mobility               = 480.0
  set recombination time     = 5e-5
  set recombination velocity = 2.9e5
  set transfer rate          = 1e-14
end

subsection mesh
  set boundary layer = 0.1
  set mesh height    = 1
  set mesh length    = 1.0
  set radius one     = 0.3
  set radius two     = 0.6
end

subsection oxidants
  set mobility = 1.0
end

subsection physical
  set absorption coefficient     = 1.74974e5
  set applied bias               = 0.0
  set built in bias              = 0.41
  set characteristic density     = 1.0e16
  set characteristic length      = 1.0e-4
  set characteristic time        = 1.0e-12
  set electrolyte permittivity   = 1000
  set illumination status        = true
  set insulated                  = true
  set intrinsic density          = 2.564e9
  set photon flux                = 1.2e17
  set schottky bias              = 0.0
  set schottky status            = true
  set semiconductor permittivity = 11.9
end

subsection reductants
  set mobility = 1.0
end
"""


def _host(*args, check=True):
    if not os.path.exists(HOST_TEST):
        pytest.fail("pecs_host_test is not built (run __graft_entry__.build())")
    r = subprocess.run([HOST_TEST] + [str(a) for a in args], capture_output=True, text=True)
    if check and r.returncode != 0:
        raise AssertionError(r.stderr)
    return r


def _kv(text):
    return {k: float(v) for k, v in (line.split("=") for line in text.strip().splitlines())}


def test_prm_mapping(tmp_path):
    f = tmp_path / "input_file.prm"
    f.write_text(PRM)
    kv = _kv(_host("prm", f, 2).stdout)
    phys = dict(char_length=1e-4, char_time=1e-12, char_density=1e16, electron_mobility=1350.0,
                hole_mobility=480.0, reductant_mobility=1.0, oxidant_mobility=1.0, electron_recombo_t=5e-5,
                hole_recombo_t=5e-5, electron_recombo_v=3e5, hole_recombo_v=2.9e5, k_et=1e-19, k_ht=1e-14,
                intrinsic_density=2.564e9, photon_flux=1.2e17, absorption_coeff=1.74974e5, applied_bias=0.0,
                built_in_bias=0.41, schottky_bias=0.0)
    s = scale_parameters(phys)
    assert (kv["r1"], kv["r2"], kv["H"], kv["L"]) == (0.3, 0.6, 1.0, 1.0)
    assert kv["n_refine"] == 3 and kv["insulated"] == 1 and kv["schottky"] == 1 and kv["degree"] == 2
    assert kv["dt"] == 0.05 and kv["t_end"] == 50.0 and kv["t_end_2"] == 1e5 and kv["time_stamps"] == 100
    for i, key in enumerate(("electron_mobility", "hole_mobility", "reductant_mobility", "oxidant_mobility")):
        assert kv["mu%d" % i] == pytest.approx(s[key], rel=1e-14)
    assert kv["lambda2"] == pytest.approx(s["debye_length"], rel=1e-14)
    assert kv["k_et"] == pytest.approx(s["k_et"], rel=1e-14)
    assert kv["k_ht"] == pytest.approx(s["k_ht"], rel=1e-14)
    assert kv["v_rec_n"] == pytest.approx(s["electron_recombo_v"], rel=1e-14)
    assert kv["phi_bi"] == pytest.approx(0.41 / 0.02585, rel=1e-3)  # V / thermal voltage at 300 K
    assert kv["G0"] == pytest.approx(1.2e17 * 1e-12 / 1e16, rel=1e-14)
    assert kv["alpha"] == pytest.approx(1.74974e5 * 1e-4, rel=1e-14)
    assert kv["eps_s"] == 11.9 and kv["eps_e"] == 1000.0
    assert kv["gen_mode"] == 0  # illuminated: reference generation alpha G0 exp(alpha (y - H))
    # equilibria hard-coded in InitialConditions.cpp
    assert (kv["n_e"], kv["p_e"], kv["r_inf"], kv["o_inf"]) == (2.0, 0.0, 30.0, 29.0)


def test_prm_defaults_and_errors(tmp_path):
    f = tmp_path / "empty.prm"
    f.write_text("# nothing set\n")
    kv = _kv(_host("prm", f, 1).stdout)
    # ParameterReader.cpp declared defaults
    assert kv["n_refine"] == 4 and kv["dt"] == 0.05 and kv["t_end"] == 10 and kv["time_stamps"] == 50
    assert kv["eps_e"] == 100 and kv["gen_mode"] == 2  # dark by default
    for bad in ("subsection nowhere\nend\n", "subsection mesh\n  set radius three = 1\nend\n",
                "subsection mesh\n  set radius one = abc\nend\n", "subsection mesh\n", "set mesh height = 1\n",
                "subsection computational\n set global refinements = 2.5\nend\n"):
        f.write_text(bad)
        r = _host("prm", f, 1, check=False)
        assert r.returncode == 1 and "error" in r.stderr, bad


def test_reference_step_count():
    # `while (time < stamp) time += dt` in double arithmetic: 0.1 summed ten
    # times is 0.9999999999999999 < 1, so an interval of length 1 takes 11 steps.
    assert _host("steps", 0, 1.0, 2.0, 0.1, 1).stdout.split() == ["11"]
    assert _host("steps", 0, 0.3, 2.0, 0.1, 1).stdout.split() == ["3"]
    assert _host("steps", 0, 0.7, 2.0, 0.1, 1).stdout.split() == ["7"]
    # 200 steps of 1e-3 to t = 0.2 in 4 stamps (the measured small workload)
    n = [int(x) for x in _host("steps", 0, 0.2, 1.0, 1e-3, 4).stdout.split()]
    assert sum(n) in (200, 201) and len(n) == 4
    # restart: stamps t_end + (i+1)(t_end_2 - t_end)/stamps
    n = [int(x) for x in _host("steps", 1, 1.0, 2.0, 0.25, 2).stdout.split()]
    assert n == [2, 2]


def test_dofs_format(tmp_path):
    f = tmp_path / "Electrons.dofs"
    assert _host("dofs", f, 7).stdout.strip() == "ok 7"
    raw = f.read_bytes()
    v = np.arange(7) + 0.25
    assert raw == b"7\n[" + v.astype("<f8").tobytes() + b"]"
    # a file written by hand in the same format reads back
    g = tmp_path / "Holes.dofs"
    w = np.array([1.5, -2.0, 3.25e-300])
    g.write_bytes(b"3\n[" + struct.pack("<3d", *w) + b"]")
    got = np.array([float(x) for x in _host("readdofs", g).stdout.split()])
    assert np.array_equal(got, w)
    g.write_bytes(b"3\n[" + struct.pack("<2d", *w[:2]))
    assert _host("readdofs", g, check=False).returncode == 1


def test_runner_rejects_bad_input(tmp_path):
    f = tmp_path / "in.prm"
    # one level of boundary-layer refinement runs; deeper local refinement is refused
    f.write_text("subsection computational\n  set local refinements = 2\nend\n")
    r = subprocess.run([RUNNER, str(f)], capture_output=True, text=True)
    assert r.returncode == 1 and "local refinements" in r.stderr and "Aborting!" in r.stderr
    r = subprocess.run([RUNNER, str(tmp_path / "missing.prm")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
