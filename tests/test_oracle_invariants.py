"""Oracle pins: structural invariants of the reference's LDG / mixed forms.

* the lower-index rule for interior faces (LDG.cpp:390-391) does not change
  the matrix (assembling from the higher-index side gives the same matrix);
* the LDG matrix has the block structure [[A, B], [-B^T, D]] with A SPD and
  D symmetric, so its symmetric part is positive definite;
* the electrolyte interface current enters b_r and b_o with opposite signs
  (SolarCell.cpp:1711-1712): b_r + b_o does not depend on k_et, k_ht;
* an electrically neutral equilibrium with no generation is a fixed point of
  the time step (consistency of the Dirichlet terms in matrix and RHS);
* the parameter scaling reproduces the closed-form Debye length of the
  shipped input_file.prm values.
"""
import numpy as np
import pytest

from oracle.ldg import assemble_ldg
from oracle.solarcell import OracleSolarCell, default_params, scale_parameters


def _split(A, nb, nc):
    n = 2 * nb * nc
    A = A.toarray()
    return A[:n, :n], A[:n, n:], A[n:, :n], A[n:, n:]


@pytest.mark.parametrize("k,conic", [(1, False), (2, True)])
def test_orientation_and_block_structure(k, conic):
    p = default_params()
    p.update(k=k, r=3, mu_n=1.7, mu_p=0.4)
    if conic:
        p.update(r2=0.5 + np.tan(np.radians(12.0)), L=1.5, insulated=True, schottky=True)
    o = OracleSolarCell(p)
    A1, A2, M = assemble_ldg(o.mesh.semi, o.cfe, 1.7, 0.4, p["dt"], 1.0, 1.0, flip_orientation=True)
    assert abs(A1 - o.A_n).max() < 1e-13 * abs(o.A_n).max()
    assert abs(A2 - o.A_p).max() < 1e-13 * abs(o.A_p).max()
    Aqq, B, C, D = _split(o.A_n, o.cfe.nb, o.mesh.semi.ncell)
    assert np.abs(C + B.T).max() < 1e-13 * np.abs(B).max()
    assert np.abs(D - D.T).max() < 1e-13 * np.abs(D).max()
    assert np.abs(Aqq - Aqq.T).max() < 1e-13 * np.abs(Aqq).max()
    sym = 0.5 * (o.A_n.toarray() + o.A_n.toarray().T)
    assert np.linalg.eigvalsh(sym).min() > 0


def test_interface_current_antisymmetry(seed):
    p = default_params()
    p.update(k=2, r=3, k_et=0.3, k_ht=0.7)
    o = OracleSolarCell(p)
    rng = np.random.default_rng(seed)
    st = o.initial_state()
    for key in "npro":
        st[key] = st[key] + 0.1 * rng.standard_normal(st[key].shape)
    br, bo = o.rhs_elec(st)
    p0 = dict(p, k_et=0.0, k_ht=0.0)
    o0 = OracleSolarCell(p0)
    br0, bo0 = o0.rhs_elec(st)
    assert np.abs((br + bo) - (br0 + bo0)).max() < 1e-13 * np.abs(br).max()
    assert np.abs(br - br0).max() > 1e-6   # the current is really there


@pytest.mark.parametrize("k", [1, 2])
def test_neutral_equilibrium_is_a_fixed_point(k):
    p = default_params()
    p.update(k=k, r=3, gen_mode="off", srh=False, ne=1.3, pe=0.0, rinf=2.0, oinf=2.0)
    o = OracleSolarCell(p)
    st = o.initial_state()
    assert np.abs(st["P"]).max() < 1e-13
    st2 = o.run(5, st)
    for key in "npro":
        assert np.abs(st2[key] - st[key]).max() < 1e-12, key


def test_scaling_closed_form():
    phys = dict(char_length=1.0e-4, char_time=1.0e-12, char_density=1.0e16, electron_mobility=1350.0,
                hole_mobility=480.0, reductant_mobility=1.0, oxidant_mobility=1.0, electron_recombo_t=5e-5,
                hole_recombo_t=5e-5, electron_recombo_v=3e5, hole_recombo_v=2.9e5, k_et=1e-19, k_ht=1e-14,
                intrinsic_density=2.564e9, photon_flux=1.2e17, absorption_coeff=1.74974e5, applied_bias=0.0,
                built_in_bias=0.41, schottky_bias=0.0)
    s = scale_parameters(phys)
    # lambda^2 = U_T eps_0 / (q C L^2) with the constants of Parameters.hpp:7-12
    assert abs(s["debye_length"] - 0.02585 * 8.85e-14 / (1.62e-19 * 1e16 * 1e-8)) < 1e-18
    assert abs(s["debye_length"] - 1.41218e-4) < 1e-8
    assert abs(s["built_in_bias"] - 0.41 / 0.02585) < 1e-12
    assert abs(s["electron_mobility"] - 1350.0 * 1e-12 * 0.02585 / 1e-8) < 1e-15
    assert abs(s["absorption_coeff"] - 17.4974) < 1e-12
