"""The Poisson vector in deal.II's layout (pecs_poisson_layout_convert, CPU).

The library converts [D | Phi] between its own nodal RT basis and the
reference's FE_RaviartThomas(k-1) basis with deal.II's DoF numbering
(SolarCell.cpp:23-24, Poisson.cpp:18-30).  Pinned here by:
  * deal.II's RT0 shape functions in closed form, (1-x, 0), (x, 0), (0, 1-y),
    (0, y), and the duality l_i(phi_j) = delta_ij of the restatement
    (oracle/dealii_rt.py, written independently of the C++ conversion);
  * D as a function: the converted coefficients, expanded in the restated
    deal.II basis with the restated numbering, give the same field at random
    points of every cell as the library's coefficients in its own basis;
  * the pairing b . x of a right-hand side with a solution is the same in both
    layouts (which ties the right-hand-side conversion to the solution's), and
    round trips are the identity;
  * Neumann D.n = 0 DoFs stay zero in the deal.II layout.
deal.II itself cannot be built here, so the restatement is of its published
definition (the orthonormal scaling of its Legendre polynomials included),
not a comparison with deal.II's own output.
"""
import numpy as np
import pytest

import pecs_amd
from oracle import fe
from oracle.dealii_rt import DealiiRT, dealii_numbering, field_at, legendre01
from oracle.mesh import SolarCellMesh
from oracle.mixed import PoissonFE, constrained_dofs
from oracle.solarcell import OracleSolarCell, default_params

CASES = {
    "k1-r2": dict(k=1, r=2),
    "k2-r2-conic": dict(k=2, r=2, r2=0.62, L=1.4, insulated=True, schottky=True),
    "k2-r3": dict(k=2, r=3),
    "k3-r2": dict(k=3, r=2),
}


def _prm(name):
    p = default_params()
    p.update(CASES[name])
    return p


def test_rt0_closed_form():
    rt = DealiiRT(0)
    pts = np.array([[0.1, 0.7], [0.55, 0.2], [0.9, 0.95]])
    v = rt.tab(pts)
    x, y = pts[:, 0], pts[:, 1]
    exact = [np.stack([1 - x, 0 * x], -1), np.stack([x, 0 * x], -1), np.stack([0 * x, 1 - y], -1),
             np.stack([0 * x, y], -1)]
    for j in range(4):
        assert np.abs(v[j] - exact[j]).max() < 1e-14


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_duality_and_orthonormal_legendre(deg):
    rt = DealiiRT(deg)
    # l_i(phi_j) = delta_ij
    D = np.zeros((rt.n, rt.n))
    for i, (pts, w, comp) in enumerate(rt.fun):
        D[i] = np.sum(w[None, :] * rt.tab(pts)[:, :, comp], axis=1)
    assert np.abs(D - np.eye(rt.n)).max() < 1e-12
    s, w = fe.gauss01(deg + 3)
    G = np.array([[np.sum(w * legendre01(i, s) * legendre01(j, s)) for j in range(deg + 1)] for i in range(deg + 1)])
    assert np.abs(G - np.eye(deg + 1)).max() < 1e-14


def _random_solution(o, rng):
    """A random Poisson coefficient vector (own layout) with D.n = 0 on Neumann faces."""
    x = rng.standard_normal(o.pfe.n)
    x[o.cons] = 0.0
    return x


@pytest.mark.parametrize("name", list(CASES))
def test_field_is_the_same_function(name, seed):
    p = _prm(name)
    o = OracleSolarCell(p)
    rng = np.random.default_rng(seed)
    x = _random_solution(o, rng)
    xd = pecs_amd.poisson_layout_convert(p, x, to_dealii=True)
    deg = p["k"] - 1
    rt = DealiiRT(deg)
    gd, nD = dealii_numbering(o.mesh, deg)
    assert nD == o.pfe.nD
    pts = rng.uniform(0.0, 1.0, (7, 2))
    cells = np.arange(len(o.mesh.pV))
    own = field_at(o.mesh, cells, pts, x[o.pfe.idx[:, :o.pfe.nrt]], o.pfe.rt.tab(pts)[0])
    dii = field_at(o.mesh, cells, pts, xd[gd], rt.tab(pts))
    assert np.abs(own - dii).max() <= 1e-13 * np.abs(own).max()
    # Phi is the same FE_DGQ(k-1) block in both
    assert np.array_equal(xd[o.pfe.nD:], x[o.pfe.nD:])


@pytest.mark.parametrize("name", list(CASES))
def test_round_trips_and_pairing(name, seed):
    p = _prm(name)
    o = OracleSolarCell(p)
    rng = np.random.default_rng(seed + 1)
    x = _random_solution(o, rng)
    b = o.rhs_poisson(o.initial_state()) + rng.standard_normal(o.pfe.n)
    xd = pecs_amd.poisson_layout_convert(p, x, to_dealii=True)
    bd = pecs_amd.poisson_layout_convert(p, b, to_dealii=True, rhs=True)
    assert abs(b @ x - bd @ xd) <= 1e-13 * np.abs(b).max() * np.abs(x).sum()
    x2 = pecs_amd.poisson_layout_convert(p, xd, to_dealii=False)
    b2 = pecs_amd.poisson_layout_convert(p, bd, to_dealii=False, rhs=True)
    assert np.abs(x2 - x).max() <= 1e-13 * np.abs(x).max()
    assert np.abs(b2 - b).max() <= 1e-13 * np.abs(b).max()
    if p["k"] == 1:  # RT0: one face DoF per face, the flux at the midpoint in both bases
        perm = dealii_numbering(o.mesh, 0)[0]
        assert np.array_equal(np.sort(np.abs(xd[:o.pfe.nD])), np.sort(np.abs(x[:o.pfe.nD])))
        del perm


def test_neumann_dofs_stay_zero(seed):
    p = _prm("k2-r2-conic")
    o = OracleSolarCell(p)
    x = _random_solution(o, np.random.default_rng(seed))
    xd = pecs_amd.poisson_layout_convert(p, x, to_dealii=True)
    gd, _ = dealii_numbering(o.mesh, 1)
    nf = 2
    zeros = 0
    for c in range(len(o.mesh.pV)):
        for f in range(4):
            if o.mesh.pbid[c, f] == 2:  # Neumann
                assert np.all(xd[gd[c, f * nf:(f + 1) * nf]] == 0.0)
                zeros += 1
    assert zeros > 0
