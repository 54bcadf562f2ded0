"""The oracle's per-step right-hand sides against high-precision golden vectors.

tests/golden/rhs_*.json hold the carrier right-hand sides of both regions and
the Poisson right-hand side on 2 x 2-cell regions (rectangles and trapezoids),
computed at 40 digits by tests/golden/make_rhs_fixtures.py -- a scalar
restatement of SolarCell.cpp:1146-1404 (semiconductor), 1519-1716
(electrolyte) and 487-815 (Poisson) that shares no code with the oracle.
Between them the cases switch on every term: generation in both modes, SRH,
drift under a non-zero field, Dirichlet, both interface sides, Schottky,
and the Phi_bi - Phi_app datum at x == 0.

Planted mistakes (each made in a copy of the oracle function, on the CPU)
must each be caught by at least one case: a flipped hole-drift sign,
generation along the wrong axis, a flipped Schottky sign, and others.
"""
import inspect
import json
import os
import textwrap

import numpy as np
import pytest

import oracle.ldg as oldg
import oracle.mixed as omixed
import oracle.solarcell as osc
from oracle import fe
from oracle.ldg import CarrierFE
from oracle.mesh import DIRICHLET, SolarCellMesh
from oracle.mixed import PoissonFE

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["rect_k1_config", "trap_k2_insulated_schottky", "trap_k2_dirichlet"]
TOL = 1e-13


def _load(case):
    with open(os.path.join(HERE, f"rhs_{case}.json")) as f:
        return json.load(f)


def _setup(fx):
    c = fx["case"]
    k = int(c["k"])
    geo = dict(r1=c["r1"], r2=c["r2"], H=c["H"], L=c["L"])
    mesh = SolarCellMesh(1, geo, bool(c["insulated"]), bool(c["schottky"]))
    assert np.abs(mesh.semi.V - np.array(fx["semi_vertices"])).max() <= 1e-15
    assert np.abs(mesh.elec.V - np.array(fx["elec_vertices"])).max() <= 1e-15
    assert mesh.semi.bid.tolist() == fx["semi_bid"]
    assert mesh.elec.bid.tolist() == fx["elec_bid"]
    cfe = CarrierFE(k)
    pfe = PoissonFE(k, mesh)
    prm = osc.default_params()
    prm.update({key: c[key] for key in ("dt", "penalty", "eps_s", "eps_e", "ne", "pe", "rinf", "oinf", "k_et",
                                        "k_ht", "v_rec_n", "v_rec_p", "G0", "alpha", "tau_n", "tau_p",
                                        "phi_bi", "phi_app", "H")})
    prm.update(k=k, r=1, gen_mode=c["gen_mode"], srh=bool(c["srh"]), ni=c["ni"], **geo)
    nb = cfe.nb
    u = {}
    for key in "npro":
        v = np.zeros(3 * nb * 4)
        v[2 * nb * 4:] = np.array(fx["density"][key]).ravel()
        u[key] = v
    # the reference fields Dhat (polynomials in RT_{k-1}) interpolated in the
    # oracle's nodal RT basis: x-components at (Xn[a], Gn[b]), y at (Gn[a], Xn[b])
    rt = pfe.rt
    deg = rt.deg
    pco = {}
    for reg in ("semi", "elec"):
        rows = []
        for ax, ay in fx["field"][reg]:
            ax, ay = np.array(ax), np.array(ay)
            co = np.zeros(pfe.nloc)
            for b in range(deg + 1):
                for a in range(deg + 2):
                    xi, eta = rt.Xn[a], rt.Gn[b]
                    co[a + (deg + 2) * b] = sum(ax[q, p] * xi ** p * eta ** q
                                                for q in range(deg + 1) for p in range(deg + 2))
            for b in range(deg + 2):
                for a in range(deg + 1):
                    xi, eta = rt.Gn[a], rt.Xn[b]
                    co[rt.nx + a + (deg + 1) * b] = sum(ay[q, p] * xi ** p * eta ** q
                                                        for q in range(deg + 2) for p in range(deg + 1))
            rows.append(co)
        pco[reg] = np.array(rows)
    return mesh, cfe, pfe, prm, u, pco


def _mass(mesh_region, cfe, dt):
    _, _, M = oldg.assemble_ldg(mesh_region, cfe, 1.0, 1.0, dt, 1.0, 1.0)
    return M


def _carrier(fx, carrier_rhs=oldg.carrier_rhs, make_physics=osc.make_physics):
    mesh, cfe, pfe, prm, u, pco = _setup(fx)
    p = make_physics(prm)
    Ms = _mass(mesh.semi, cfe, prm["dt"])
    Me = _mass(mesh.elec, cfe, prm["dt"])
    bn, bp = carrier_rhs(mesh.semi, cfe, pfe, mesh, Ms, u["n"], u["p"], pco["semi"], p, "semi",
                         u["r"], u["o"], mesh.elec)
    br, bo = carrier_rhs(mesh.elec, cfe, pfe, mesh, Me, u["r"], u["o"], pco["elec"], p, "elec",
                         u["n"], u["p"], mesh.semi)
    return {"n": bn, "p": bp, "r": br, "o": bo}


def _carrier_err(fx, out):
    worst = 0.0
    for key in "npro":
        ref = np.array(fx["rhs_" + key])
        worst = max(worst, np.abs(out[key] - ref).max() / np.abs(ref).max())
    return worst


def _poisson(fx, poisson_rhs=omixed.poisson_rhs, make_physics=osc.make_physics):
    mesh, cfe, pfe, prm, u, _ = _setup(fx)
    p = make_physics(prm)
    return poisson_rhs(mesh, pfe, cfe, u["n"], u["p"], u["r"], u["o"], p), mesh, pfe


def _poisson_expected(fx, mesh, pfe):
    """Expected [D | Phi] vector: Phi part per cell; D part on the semiconductor
    Dirichlet faces from the basis-free moments, through the reference normal
    flux of the oracle's face functions (+-G_b, the Lagrange polynomials on the
    deg+1 Gauss nodes; sign -1 on faces 0 and 2, whose outward normal points
    against the reference coordinate)."""
    rt = pfe.rt
    deg = rt.deg
    nx = rt.nx
    exp = np.zeros(pfe.n)
    for pc, loc in enumerate(fx["poisson_phi"]):
        for l, v in enumerate(loc):
            exp[pfe.idx[pc, 2 * nx + l]] = v
    # monomial coefficients of G_b: G_b(s) = sum_m C[b, m] s^m
    Vm = np.vander(rt.Gn, deg + 1, increasing=True)
    C = np.linalg.inv(Vm).T
    for cell, f, mom in fx["poisson_moments"]:
        assert mesh.pbid[cell, f] == DIRICHLET
        sign = -1.0 if f in (0, 2) else 1.0
        for b in range(deg + 1):
            if f == 0:
                li = (deg + 2) * b
            elif f == 1:
                li = deg + 1 + (deg + 2) * b
            elif f == 2:
                li = nx + b
            else:
                li = nx + b + (deg + 1) * (deg + 1)
            exp[pfe.idx[cell, li]] += sign * float(np.dot(C[b], mom))
    return exp


def _poisson_err(fx, b, mesh, pfe):
    exp = _poisson_expected(fx, mesh, pfe)
    return np.abs(b - exp).max() / np.abs(exp).max()


def test_face_functions_have_the_stated_normal_flux():
    """The trace property the Poisson check relies on: on its own face, an
    oracle face function has reference normal flux +-G_b; every other RT
    function has none there."""
    for deg in (0, 1):
        rt = fe.RT(deg)
        s = np.linspace(0.05, 0.95, 7)
        Vg, _ = fe.lagrange1d(rt.Gn, s)
        for f in range(4):
            val, _ = rt.tab(fe.face_points(f, s))
            flux = val[:, :, 0] if f < 2 else val[:, :, 1]
            for i in range(rt.n):
                if f == 0 and i < rt.nx and i % (deg + 2) == 0:
                    assert np.allclose(flux[i], Vg[i // (deg + 2)], atol=1e-14)
                elif f == 1 and i < rt.nx and i % (deg + 2) == deg + 1:
                    assert np.allclose(flux[i], Vg[i // (deg + 2)], atol=1e-14)
                elif f == 2 and rt.nx <= i < rt.nx + deg + 1:
                    assert np.allclose(flux[i], Vg[i - rt.nx], atol=1e-14)
                elif f == 3 and i >= rt.n - (deg + 1):
                    assert np.allclose(flux[i], Vg[i - (rt.n - deg - 1)], atol=1e-14)
                else:
                    assert np.abs(flux[i]).max() < 1e-14


@pytest.mark.parametrize("case", CASES)
def test_carrier_rhs_matches_golden(case):
    fx = _load(case)
    assert _carrier_err(fx, _carrier(fx)) <= TOL


@pytest.mark.parametrize("case", CASES)
def test_poisson_rhs_matches_golden(case):
    fx = _load(case)
    b, mesh, pfe = _poisson(fx)
    assert _poisson_err(fx, b, mesh, pfe) <= TOL


def _planted(module, name, old, new):
    src = textwrap.dedent(inspect.getsource(getattr(module, name)))
    assert old in src, f"planted-mistake pattern not found in {name}: {old!r}"
    ns = dict(module.__dict__)
    exec(compile(src.replace(old, new, 1), module.__file__, "exec"), ns)
    return ns[name]


CARRIER_MISTAKES = {
    "hole drift sign": ("s1, s2 = -1.0, 1.0\n        inv_perm = 1.0 / params[\"eps_s\"]",
                        "s1, s2 = -1.0, -1.0\n        inv_perm = 1.0 / params[\"eps_s\"]"),
    "oxidant drift sign": ("s1, s2 = -1.0, 1.0\n        inv_perm = 1.0 / params[\"eps_e\"]",
                           "s1, s2 = -1.0, -1.0\n        inv_perm = 1.0 / params[\"eps_e\"]"),
    "Schottky hole sign": ('params["v_rec_p"] * (own2 - params["pe"])', '-params["v_rec_p"] * (own2 - params["pe"])'),
    "Schottky electron sign": ('-params["v_rec_n"] * (own1', 'params["v_rec_n"] * (own1'),
    "interface hole sign": ('t2 = params["k_ht"]', 't2 = -params["k_ht"]'),
    "interface traces swapped": ("* r_nb * jw", "* o_nb * jw"),
    "electrolyte current sign": ("loc2[cells] -= I", "loc2[cells] += I"),
    "Dirichlet normal sign": ("t = (-pin + ", "t = (pin + "),
    "permittivity of the wrong region": ('inv_perm = 1.0 / params["eps_e"]', 'inv_perm = 1.0 / params["eps_s"]'),
}
PHYSICS_MISTAKES = {
    "generation along the wrong axis (config)": ("np.exp(-alpha * x[..., 0])", "np.exp(-alpha * x[..., 1])"),
    "generation along the wrong axis (ref)": ("np.exp(alpha * (x[..., 1] - H))", "np.exp(alpha * (x[..., 0] - H))"),
    "SRH with (n - n_i)": ("tn * (n + ni)", "tn * (n - ni)"),
}
POISSON_MISTAKES = {
    "bias sign": ("(bi - other)", "(bi + other)"),
    "semiconductor charge sign": ("(-1.0 * n_q + 1.0 * p_q)", "(1.0 * n_q + 1.0 * p_q)"),
    "electrolyte charge sign": ("(-1.0 * r_q + 1.0 * o_q)", "(-1.0 * r_q - 1.0 * o_q)"),
    "doping sign": ('(params["ne"] - params["pe"])', '(params["pe"] - params["ne"])'),
}


@pytest.mark.parametrize("mistake", sorted(CARRIER_MISTAKES))
def test_planted_carrier_mistake_is_caught(mistake):
    old, new = CARRIER_MISTAKES[mistake]
    bad = _planted(oldg, "carrier_rhs", old, new)
    worst = max(_carrier_err(_load(c), _carrier(_load(c), carrier_rhs=bad)) for c in CASES)
    assert worst > 1e-6, mistake


@pytest.mark.parametrize("mistake", sorted(PHYSICS_MISTAKES))
def test_planted_physics_mistake_is_caught(mistake):
    old, new = PHYSICS_MISTAKES[mistake]
    bad = _planted(osc, "make_physics", old, new)
    worst = max(_carrier_err(_load(c), _carrier(_load(c), make_physics=bad)) for c in CASES)
    assert worst > 1e-6, mistake


@pytest.mark.parametrize("mistake", sorted(POISSON_MISTAKES))
def test_planted_poisson_mistake_is_caught(mistake):
    old, new = POISSON_MISTAKES[mistake]
    bad = _planted(omixed, "poisson_rhs", old, new)
    worst = 0.0
    for c in CASES:
        fx = _load(c)
        b, mesh, pfe = _poisson(fx, poisson_rhs=bad)
        worst = max(worst, _poisson_err(fx, b, mesh, pfe))
    assert worst > 1e-6, mistake
