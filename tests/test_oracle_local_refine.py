"""Boundary-layer local refinement in the oracle (Grid.cpp:72-111,
LDG.cpp:338-383, Poisson.cpp:33-34), pinned by what the mathematics fixes:

* the mesh: cell areas add up to the domain, every hanging pair's fine face
  is exactly the coarse face's half, cells come level by level;
* the LDG discretisation is consistent, so an exact solution inside the
  discrete space (u of total degree <= k, current -grad u) is reproduced to
  rounding on a mesh with hanging faces -- any mistake in the subface
  quadrature points, normal, JxW or pairing breaks it (one such mistake is
  planted below and must fail);
* the mixed RT_{k-1} x DGQ_{k-1} method with hanging-node constraints
  reproduces Phi of total degree <= k-1 (D = -grad Phi in RT_{k-1}); planting
  a wrong constraint factor breaks it;
* convergence rates on locally refined meshes match the textbook ones.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import fe
from oracle import manufactured as mm
from oracle.ldg import CarrierFE, assemble_ldg
from oracle.mesh import DIRICHLET, INTERFACE, LocalRefinedMesh, SolarCellMesh
from oracle.mixed import PoissonFE, assemble_mixed, hanging_constraints, reduction

PLANAR = dict(r1=0.5, r2=0.5, H=1.0, L=1.0)
CONIC = dict(r1=0.3, r2=0.6, H=1.0, L=1.2)


def _area(V):
    x, J = fe.bilinear(V, np.array([[0.5, 0.5]]))
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    return det[:, 0].sum()  # exact for bilinear quads (det is linear, midpoint rule)


@pytest.mark.parametrize("geo", [PLANAR, CONIC])
def test_mesh_invariants(geo):
    r, bl = 3, 0.1
    m = LocalRefinedMesh(r, geo, bl)
    N = 1 << r
    H, L = geo["H"], geo["L"]
    semi_area = 0.5 * (geo["r1"] + geo["r2"]) * H
    assert abs(_area(m.semi.V) - semi_area) < 1e-13
    assert abs(_area(m.elec.V) - (L * H - semi_area)) < 1e-13
    assert abs(_area(m.pV) - L * H) < 1e-13
    for R in (m.semi, m.elec):
        assert R.ncell == N * N + 4 * N * N
        # level-r cells first (deal.II active-cell order), then the layer
        assert (R.level[:N * N] == 0).all() and (R.level[N * N:] == 1).all()
        assert len(R.hang) == 2 * N
        for c, f, s, e, fe_ in R.hang:
            a, b = {0: (0, 2), 1: (1, 3)}[f]
            lo = R.V[c, a] + s * (R.V[c, b] - R.V[c, a]) / 2
            hi = R.V[c, a] + (s + 1) * (R.V[c, b] - R.V[c, a]) / 2
            fa, fb = {0: (0, 2), 1: (1, 3)}[fe_]
            assert np.allclose(R.V[e, fa], lo, atol=1e-15) and np.allclose(R.V[e, fb], hi, atol=1e-15)
            assert R.nbr[c, f] < 0 and R.nbr[e, fe_] < 0 and R.bid[c, f] < 0 and R.bid[e, fe_] < 0
    assert len(m.phang) == 4 * N
    # interface pairs: 2N conforming layer faces with identical vertices
    assert len(m.sigma_semi) == 2 * N
    for s, e in zip(m.sigma_semi, m.sigma_elec):
        assert (m.semi.V[s, 1] == m.elec.V[e, 0]).all() and (m.semi.V[s, 3] == m.elec.V[e, 2]).all()


def _poly(k):
    # exact solutions of total degree k: u, -grad u, -lap u
    if k == 1:
        return (lambda x: 1.0 + 2.0 * x[..., 0] - 3.0 * x[..., 1],
                lambda x: np.stack([-2.0 + 0 * x[..., 0], 3.0 + 0 * x[..., 0]], -1),
                lambda x: 0.0 * x[..., 0])
    return (lambda x: 0.5 + x[..., 0] - x[..., 1] + x[..., 0] ** 2 - 0.75 * x[..., 0] * x[..., 1] + 2.0 * x[..., 1] ** 2,
            lambda x: -np.stack([1.0 + 2 * x[..., 0] - 0.75 * x[..., 1], -1.0 - 0.75 * x[..., 0] + 4.0 * x[..., 1]], -1),
            lambda x: -(2.0 + 4.0) + 0 * x[..., 0])


def _steady_ldg(reg, k, u, J, f, penalty=1.0):
    cfe = CarrierFE(k)
    A, _, _ = assemble_ldg(reg, cfe, 1.0, 1.0, 1.0, 0.0, penalty)
    loc, _ = mm._cell_rhs(reg, cfe, f_vals=f)
    mm._dirichlet_rhs(reg, cfe, loc, u, penalty)
    mm._flux_rhs(reg, cfe, loc, lambda x, n: np.sum(J(x) * n, -1))
    b = mm._assemble_vec(reg, cfe, loc)
    sol = spla.spsolve(A.tocsc(), b)
    return mm.ldg_errors(reg, cfe, sol, u, J)


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("geo", [PLANAR, CONIC])
def test_ldg_exact_for_polynomials_on_hanging_mesh(k, geo):
    m = LocalRefinedMesh(2, geo, 0.1)
    u, J, f = _poly(min(k, 2))
    for reg in (m.semi, m.elec):
        eu, eJ = _steady_ldg(reg, k, u, J, f)
        assert eu < 1e-11 and eJ < 1e-10, (eu, eJ)


def test_ldg_planted_subface_swap_fails():
    m = LocalRefinedMesh(2, CONIC, 0.1)
    reg = m.semi
    reg.hang = reg.hang.copy()
    reg.hang[:, 2] = 1 - reg.hang[:, 2]  # fine cell paired with the wrong half
    u, J, f = _poly(1)
    eu, eJ = _steady_ldg(reg, 1, u, J, f)
    assert eu > 1e-4 or eJ > 1e-4


def _mixed(m, k, phi, D, f, mutate=None):
    pfe = PoissonFE(k, m)
    A = assemble_mixed(m, pfe, 1.0, 1.0, 1.0)
    b = mm._mixed_rhs(m, pfe, f, phi)
    T = reduction(m, pfe)
    if mutate is not None:
        T = mutate(T, pfe)
    y = spla.spsolve((T.T @ A @ T).tocsc(), T.T @ b)
    return mm.mixed_errors(m, pfe, T @ y, phi, D)


@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("geo", [PLANAR, CONIC])
def test_mixed_exact_for_polynomials_with_hanging_constraints(k, geo):
    m = LocalRefinedMesh(2, geo, 0.1)
    phi, D, f = _poly(k - 1)
    ep, eD = _mixed(m, k, phi, D, f)
    assert ep < 1e-11 and eD < 1e-10, (ep, eD)


def test_mixed_planted_constraint_factor_fails():
    m = LocalRefinedMesh(2, CONIC, 0.1)
    phi, D, f = _poly(1)

    def one_instead_of_half(T, pfe):
        slaves = [h[0] for h in hanging_constraints(m, pfe)]
        T = T.tolil()
        for s in slaves:
            T[s] = T[s] * 2.0
        return T.tocsr()
    ep, eD = _mixed(m, 2, phi, D, f, one_instead_of_half)
    assert ep > 1e-6 or eD > 1e-6


def test_constraint_count():
    m = LocalRefinedMesh(2, PLANAR, 0.1)
    pfe = PoissonFE(2, m)
    assert len(hanging_constraints(m, pfe)) == len(m.phang) * pfe.rt.Gn.size


@pytest.mark.parametrize("k", [1, 2])
def test_convergence_rates_on_locally_refined_meshes(k):
    # smooth solution of the reference's steady study (test_functions.cpp:13-93)
    u, J, f = mm.poisson_g, mm.poisson_D, mm.poisson_f
    errs, ep, hs = [], [], []
    for r in (2, 3, 4):
        m = LocalRefinedMesh(r, CONIC, 0.1)
        eu, eJ = _steady_ldg(m.semi, k, u, J, f)
        errs.append(eu)
        if k >= 2:
            ep.append(_mixed(m, k, u, J, f)[0])
        hs.append(fe.diameter(m.semi.V[:(1 << r) ** 2]).max())
    rates = np.log(np.array(errs[:-1]) / errs[1:]) / np.log(np.array(hs[:-1]) / hs[1:])
    assert rates[-1] > k + 1 - 0.3, rates
    if k >= 2:
        pr = np.log(np.array(ep[:-1]) / ep[1:]) / np.log(np.array(hs[:-1]) / hs[1:])
        assert pr[-1] > k - 0.3, pr  # Phi: h^k for RT_{k-1} x DGQ_{k-1}


def test_solarcell_step_on_locally_refined_mesh():
    from oracle.solarcell import OracleSolarCell, default_params
    p = default_params()
    p.update(k=2, r=2, r1=0.3, r2=0.6, local_refine=1, bl=0.1, phi_app=0.1, k_et=0.1)
    o = OracleSolarCell(p)
    st = o.run(3)
    # hanging fine-face D DoFs follow the coarse faces exactly
    for sl, ms, cs in hanging_constraints(o.mesh, o.pfe):
        assert abs(st["P"][sl] - cs @ st["P"][ms]) < 1e-12 * max(1.0, np.abs(st["P"]).max())
    assert all(np.isfinite(v).all() for v in st.values())
