"""Oracle pins by physics and closed forms (no restated formula involved).

* Mass conservation with insulated boundaries: with every boundary face of a
  region Neumann and no sources, the LDG step conserves the integral of each
  density exactly (v = 1 is a test function and the interior fluxes
  telescope, LDG.cpp:283-678).
* Drift sign and scaling: in an insulated region under a constant field
  D, zero current mu^{-1} q = -grad u + s_c (D / eps) u = 0 holds for the
  Boltzmann profile u = C exp(s_c D.x / eps) (SolarCell.cpp:1160-1190: drift
  s_c p.D/eps u; LDG.cpp:147-171: mu^{-1} p.q - div(p) u), so the stepper
  must converge to it, with C fixed by the conserved mass, at rate k+1.
* Generation placement: the per-step source rhs - M u_old, paired with
  v = 1, x, y (so the integral of the density grows by dt int G per step),
  equals the first moments int v G, checked against
  closed-form integrals of the reference generation alpha G0 exp(alpha (y-H))
  (Generation.cpp:29-44) and the Beer-Lambert G0 exp(-alpha x).
* Poisson right-hand side (SolarCell.cpp:551-681, 754-760): paired with the
  Raviart-Thomas interpolant of the constant field w = (1, 0), the Dirichlet
  terms give H (Phi_bi - Phi_app) (the datum lives on x = 0, whose outward
  normal is -x); paired with the constant 1 in the potential space, the cell
  terms give -int rho over both regions, rho = (n^e - p^e) - n + p in the
  semiconductor and -r + o in the electrolyte.
"""
import copy

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.ldg import CarrierFE, assemble_ldg, carrier_rhs
from oracle.mesh import INTERIOR, NEUMANN, SolarCellMesh
from oracle.mixed import PoissonFE, poisson_rhs
from oracle.solarcell import default_params, make_physics

GEO = dict(r1=0.5, r2=0.5, H=1.0, L=1.0)


def _insulated(region):
    reg = copy.copy(region)
    reg.bid = np.where(region.bid == INTERIOR, INTERIOR, NEUMANN)
    return reg


def _const_field(pfe, V, D):
    """Oracle RT coefficients of a constant physical field D on axis-aligned
    rectangles: Dhat = det(J) J^{-1} D = (b Dx, a Dy) for a cell a x b."""
    a = V[:, 1, 0] - V[:, 0, 0]
    b = V[:, 2, 1] - V[:, 0, 1]
    co = np.zeros((len(V), pfe.nloc))
    co[:, :pfe.rt.nx] = (b * D[0])[:, None]
    co[:, pfe.rt.nx:pfe.nrt] = (a * D[1])[:, None]
    return co


def _density_interp(cfe, V, fun):
    """Nodal interpolant of fun(x, y) in FE_DGQ(k) (exact for bilinear fun)."""
    from oracle import fe
    nodes = fe.tensor_points(cfe.dgq.nodes)
    x, _ = fe.bilinear(V, nodes)
    nc = len(V)
    u = np.zeros(3 * cfe.nb * nc)
    u[2 * cfe.nb * nc:] = fun(x[..., 0], x[..., 1]).ravel()
    return u


def _mass_of(cfe, M, u, dt):
    """int u = dt * (1^T M u) over the density block (M carries 1/dt)."""
    nc = len(u) // (3 * cfe.nb)
    one = np.zeros_like(u)
    one[2 * cfe.nb * nc:] = 1.0
    return dt * one @ (M @ u)


def _region_stepper(k, r, which, D, eps, dt, gen_mode="off"):
    p = default_params()
    p.update(k=k, r=r, gen_mode=gen_mode, srh=False, eps_s=eps, eps_e=eps, G0=1.7, alpha=1.3, dt=dt)
    prm = make_physics(p)
    mesh = SolarCellMesh(r, GEO, False, False)
    reg = _insulated(mesh.semi if which == "semi" else mesh.elec)
    cfe = CarrierFE(k)
    pfe = PoissonFE(k, mesh)
    A1, A2, M = assemble_ldg(reg, cfe, 1.0, 1.0, dt, 1.0, 1.0)
    lu1, lu2 = spla.splu(A1.tocsc()), spla.splu(A2.tocsc())
    pco = _const_field(pfe, reg.V, D)

    def step(u1, u2):
        b1, b2 = carrier_rhs(reg, cfe, pfe, mesh, M, u1, u2, pco, prm, which, u1, u2, reg)
        return lu1.solve(b1), lu2.solve(b2)

    return reg, cfe, M, step, prm


@pytest.mark.parametrize("which", ["semi", "elec"])
def test_insulated_mass_conservation(which):
    dt = 0.01
    reg, cfe, M, step, _ = _region_stepper(2, 3, which, np.array([0.6, -0.4]), 1.5, dt)
    rng = np.random.default_rng(5)
    u1 = _density_interp(cfe, reg.V, lambda x, y: 1.0 + 0.3 * np.sin(3 * x) * np.cos(2 * y))
    u2 = _density_interp(cfe, reg.V, lambda x, y: 0.5 + 0.2 * x * y)
    u1[2 * cfe.nb * reg.ncell:] += 0.01 * rng.standard_normal(cfe.nb * reg.ncell)
    m1, m2 = _mass_of(cfe, M, u1, dt), _mass_of(cfe, M, u2, dt)
    for _ in range(20):
        u1, u2 = step(u1, u2)
    assert abs(_mass_of(cfe, M, u1, dt) - m1) < 1e-13 * m1
    assert abs(_mass_of(cfe, M, u2, dt) - m2) < 1e-13 * m2


@pytest.mark.parametrize("gen_mode", ["ref", "config"])
def test_generation_moments_closed_form(gen_mode):
    """(rhs - M u) paired with v = 1, x, y equals int v G over the
    semiconductor [0, 1/2] x [0, 1] (insulated, no field)."""
    dt = 0.01
    reg, cfe, M, _, prm = _region_stepper(2, 3, "semi", np.zeros(2), 1.0, dt, gen_mode)
    mesh = SolarCellMesh(3, GEO, False, False)
    pfe = PoissonFE(2, mesh)
    u = _density_interp(cfe, reg.V, lambda x, y: 1.0 + 0.0 * x)
    b, _ = carrier_rhs(reg, cfe, pfe, mesh, M, u, u, _const_field(pfe, reg.V, np.zeros(2)), prm, "semi",
                       u, u, reg)
    src = b - M @ u
    G0, al = 1.7, 1.3
    if gen_mode == "config":   # G0 exp(-al x), x in [0, 1/2], y in [0, 1]
        ix0 = (1 - np.exp(-al / 2)) / al
        ix1 = (1 - np.exp(-al / 2) * (1 + al / 2)) / al ** 2
        exact = {"1": G0 * ix0, "x": G0 * ix1, "y": G0 * ix0 * 0.5}
    else:                      # al G0 exp(al (y - 1))
        iy0 = (1 - np.exp(-al)) / al
        iy1 = (1 / al) - (1 - np.exp(-al)) / al ** 2
        exact = {"1": al * G0 * 0.5 * iy0, "x": al * G0 * 0.125 * iy0, "y": al * G0 * 0.5 * iy1}
    for name, fun in (("1", lambda x, y: 1.0 + 0 * x), ("x", lambda x, y: x), ("y", lambda x, y: y)):
        v = _density_interp(cfe, reg.V, fun)
        assert abs(v @ src - exact[name]) < 1e-11 * abs(exact[name]), name


def _boltzmann_error(k, r, which, s1, s2):
    dt = 0.02
    D = np.array([0.6, -0.4])
    eps = 1.5
    reg, cfe, M, step, _ = _region_stepper(k, r, which, D, eps, dt)
    u1 = _density_interp(cfe, reg.V, lambda x, y: 1.0 + 0 * x)
    u2 = u1.copy()
    for _ in range(400):
        u1, u2 = step(u1, u2)
    from oracle import fe
    x, _, det, _, JxW = cfe.cell_geometry(reg.V)
    nb, nc = cfe.nb, reg.ncell
    errs = []
    for u, s in ((u1, s1), (u2, s2)):
        uq = u[2 * nb * nc:].reshape(nc, nb) @ cfe.phi
        prof = np.exp(s * (x @ D) / eps)
        C = (uq * JxW).sum() / (prof * JxW).sum()
        errs.append(np.sqrt((((uq - C * prof) ** 2) * JxW).sum() / ((uq ** 2) * JxW).sum()))
    del fe, det
    return errs


@pytest.mark.parametrize("which", ["semi", "elec"])
def test_drift_reaches_boltzmann_profile(which):
    """Charge numbers s = -1 (electrons, reductants), +1 (holes, oxidants):
    the steady profile is exp(s D.x / eps); the opposite sign is far off."""
    e3 = _boltzmann_error(1, 3, which, -1.0, 1.0)
    e4 = _boltzmann_error(1, 4, which, -1.0, 1.0)
    for a, b in zip(e3, e4):
        assert b < 3e-5
        assert a / b > 3.5       # second order (k + 1 = 2) in h
    wrong = _boltzmann_error(1, 3, which, 1.0, -1.0)
    assert min(wrong) > 0.05


def test_poisson_rhs_closed_forms():
    k, r = 2, 3
    p = default_params()
    p.update(k=k, r=r, phi_bi=0.7, phi_app=0.25, ne=1.3, pe=0.2)
    prm = make_physics(p)
    mesh = SolarCellMesh(r, GEO, False, False)
    cfe = CarrierFE(k)
    pfe = PoissonFE(k, mesh)
    # densities: n = 1 + x, p = 0.5 y (semiconductor), r = 2 - y, o = 1 + x y (electrolyte)
    un = _density_interp(cfe, mesh.semi.V, lambda x, y: 1.0 + x)
    up = _density_interp(cfe, mesh.semi.V, lambda x, y: 0.5 * y)
    ur = _density_interp(cfe, mesh.elec.V, lambda x, y: 2.0 - y)
    uo = _density_interp(cfe, mesh.elec.V, lambda x, y: 1.0 + x * y)
    b = poisson_rhs(mesh, pfe, cfe, un, up, ur, uo, prm)
    # RT interpolant of w = (1, 0): Dhat = (cell height, 0) at every x node
    c = np.zeros(pfe.n)
    hgt = mesh.pV[:, 2, 1] - mesh.pV[:, 0, 1]
    for pc in range(len(mesh.pV)):
        c[pfe.idx[pc, :pfe.rt.nx]] = hgt[pc]
    assert abs(b @ c - 1.0 * (0.7 - 0.25)) < 1e-13
    # constant 1 in DGQ(k-1)
    one = np.zeros(pfe.n)
    one[pfe.nD:] = 1.0
    # semiconductor [0,1/2]x[0,1]: int (1.3-0.2) - (1+x) + 0.5y = 0.55 - 0.5 - 0.125 + 0.125
    semi = 0.55 - 0.5 - 0.125 + 0.125
    # electrolyte [1/2,1]x[0,1]: int -(2-y) + (1+xy) = -1 + 0.25 + 0.5 + 0.1875
    elec = -1.0 + 0.25 + 0.5 + 0.1875
    assert abs(b @ one - (-(semi + elec))) < 1e-13
