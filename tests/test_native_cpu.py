"""CPU checks of the native library (no GPU needed).

* the library loads and exports every function declared in include/pecs_abi.h;
* the constant matrices rebuilt from the library's local operators (the
  matrices its Schur / hybridised factorisations are built from) equal the
  oracle's directly assembled matrices, including C = -B^T, trapezoidal
  cells, Neumann / Schottky faces and the Neumann D.n constraints;
* the parameter scaling equals the oracle's;
* the native self-test (nested dissection + multifrontal Cholesky of the
  Schur and multiplier matrices) reaches residual < 1e-12;
* creating a context without a GPU fails loudly (no CPU fallback).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import pecs_amd
from oracle.solarcell import OracleSolarCell, default_params, scale_parameters

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = pecs_amd.lib()
    hdr = open(os.path.join(ROOT, "include", "pecs_abi.h")).read()
    names = sorted(set(re.findall(r"\b(pecs_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", pecs_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (pecs_\w+)", out))
    assert set(names) <= exported, set(names) - exported


CASES = [
    dict(k=1, r=4),
    dict(k=2, r=3, r2=0.5 + np.tan(np.radians(10.0)), L=1.5, insulated=True, schottky=True, mu_n=1.3, mu_p=0.7,
         mu_r=2.0, mu_o=0.5, eps_s=2.0, eps_e=3.0, lam2=0.5),
    dict(k=3, r=2, insulated=True, dt=1e-4, penalty=2.0),
    # boundary-layer local refinement: hanging-face fluxes and the face-block Poisson layout
    dict(k=1, r=3, local_refine=1, bl=0.05),
    dict(k=2, r=2, r1=0.3, r2=0.6, L=1.2, local_refine=1, bl=0.1, insulated=True, schottky=True, mu_n=1.3,
         mu_p=0.7, eps_s=2.0, eps_e=3.0, lam2=0.5, penalty=2.0),
    dict(k=3, r=2, local_refine=1, bl=0.1, dt=1e-4),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_constant_matrices_match_oracle(case):
    p = default_params()
    p.update(CASES[case])
    o = OracleSolarCell(p)
    for which, A in ((pecs_amd.ELECTRONS, o.A_n), (pecs_amd.HOLES, o.A_p), (pecs_amd.REDUCTANTS, o.A_r),
                     (pecs_amd.OXIDANTS, o.A_o), (pecs_amd.MASS_SEMI, o.M_s), (pecs_amd.MASS_ELEC, o.M_e)):
        B = pecs_amd.export_matrix(p, which)
        assert abs(B - A).max() <= 1e-14 * abs(A).max(), which
    B, cons = pecs_amd.export_matrix(p, pecs_amd.POISSON)
    assert abs(B - o.A_P).max() <= 1e-14 * abs(o.A_P).max()
    assert set(np.nonzero(cons)[0]) == set(o.cons)


def test_scaling_matches_oracle():
    phys = dict(char_length=1.0e-4, char_time=1.0e-12, char_density=1.0e16, electron_mobility=1350.0,
                hole_mobility=480.0, reductant_mobility=1.0, oxidant_mobility=1.0, electron_recombo_t=5e-5,
                hole_recombo_t=5e-5, electron_recombo_v=3e5, hole_recombo_v=2.9e5, k_et=1e-19, k_ht=1e-14,
                intrinsic_density=2.564e9, photon_flux=1.2e17, absorption_coeff=1.74974e5, applied_bias=0.0,
                built_in_bias=0.41, schottky_bias=0.0)
    a = pecs_amd.scale_parameters(phys)
    b = scale_parameters(phys)
    pairs = [("debye_length", "debye_length"), ("built_in_bias", "built_in_bias"), ("k_et", "k_et"),
             ("k_ht", "k_ht"), ("photon_flux", "photon_flux"), ("absorption_coeff", "absorption_coeff"),
             ("intrinsic_density", "intrinsic_density"), ("rescale_current", "rescale_current")]
    for ka, kb in pairs:
        assert a[ka] == pytest.approx(b[kb], rel=1e-15), ka
    assert a["mobility"][0] == pytest.approx(b["electron_mobility"], rel=1e-15)
    assert a["recombo_v"][1] == pytest.approx(b["hole_recombo_v"], rel=1e-15)


@pytest.mark.parametrize("args", [["4", "1"], ["4", "2", "16", "1"], ["5", "2", "4"], ["3", "2", "16", "1", "1"],
                                  ["4", "1", "16", "0", "1"]])
def test_native_selftest(args):
    exe = os.path.join(ROOT, "pecs_amd", "_build", "pecs_selftest")
    if not os.path.exists(exe):
        pytest.skip("native self-test not built (run __graft_entry__.build())")
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "SELFTEST OK" in r.stdout, r.stdout + r.stderr


def test_create_without_gpu_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(RuntimeError, match="GPU"):
        pecs_amd.SolarCellGPU(default_params())
