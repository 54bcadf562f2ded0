"""Oracle pins: quadrature, Lagrange / RT tabulations, geometry, mesh order.

Each check is fixed by mathematics independent of the oracle's formulas
(exactness of Gauss rules, nodal and flux properties, divergence theorem,
areas of known polygons, deal.II's documented child numbering).
"""
import numpy as np
import pytest

from oracle import fe
from oracle.mesh import SolarCellMesh, morton_ij, vertex_grid


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_gauss_exactness(n):
    x, w = fe.gauss01(n)
    for m in range(2 * n):
        assert abs(np.sum(w * x ** m) - 1.0 / (m + 1)) < 1e-14
    # and not exact one degree higher (the rule really has n points)
    m = 2 * n
    assert abs(np.sum(w * x ** m) - 1.0 / (m + 1)) > 1e-8


@pytest.mark.parametrize("k", [1, 2, 3])
def test_dgq_nodal_and_partition(k):
    d = fe.DGQ(k)
    pts = fe.tensor_points(d.nodes)
    V, G = d.tab(pts)
    assert np.allclose(V, np.eye(d.nb), atol=1e-13)       # nodal
    rnd = np.random.default_rng(1).random((7, 2))
    V, G = d.tab(rnd)
    assert np.allclose(V.sum(0), 1.0, atol=1e-13)          # partition of unity
    assert np.allclose(G.sum(0), 0.0, atol=1e-12)
    # exact gradient of a Q_k polynomial interpolated at the nodes
    f = lambda x, y: x ** k * y + 0.5 * y ** k - 3 * x * y
    fx = lambda x, y: k * x ** (k - 1) * y - 3 * y
    c = f(pts[:, 0], pts[:, 1])
    assert np.allclose(c @ V, f(rnd[:, 0], rnd[:, 1]), atol=1e-12)
    assert np.allclose(np.einsum("a,aq->q", c, G[..., 0]), fx(rnd[:, 0], rnd[:, 1]), atol=1e-11)


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_rt_face_fluxes_and_divergence_theorem(deg):
    rt = fe.RT(deg)
    g, _ = fe.gauss01(deg + 1)
    nrm = {0: (-1, 0), 1: (1, 0), 2: (0, -1), 3: (0, 1)}
    face_dofs = {0: [0 + (deg + 2) * m for m in range(deg + 1)],
                 1: [deg + 1 + (deg + 2) * m for m in range(deg + 1)],
                 2: [rt.nx + m for m in range(deg + 1)],
                 3: [rt.nx + m + (deg + 1) * (deg + 1) for m in range(deg + 1)]}
    for f in range(4):
        V, _ = rt.tab(fe.face_points(f, g))
        flux = V @ np.array(nrm[f], float)                # [n, nfq]
        expect = np.zeros_like(flux)
        sign = 1.0 if f in (1, 3) else -1.0
        for m, i in enumerate(face_dofs[f]):
            expect[i, m] = sign
        assert np.allclose(flux, expect, atol=1e-13)
    # divergence theorem on the reference square, for every basis function
    qp, qw = fe.cell_quadrature(deg + 3)
    _, div = rt.tab(qp)
    lhs = div @ qw
    s, w = fe.gauss01(deg + 3)
    rhs = np.zeros(rt.n)
    for f in range(4):
        V, _ = rt.tab(fe.face_points(f, s))
        rhs += (V @ np.array(nrm[f], float)) @ w
    assert np.allclose(lhs, rhs, atol=1e-13)
    assert rt.n == 2 * (deg + 1) * (deg + 2)


def test_bilinear_geometry_trapezoid():
    V = np.array([[[0.0, 0.0], [0.7, 0.0], [0.0, 1.0], [0.5, 1.0]]])
    pts, w = fe.cell_quadrature(3)
    x, J = fe.bilinear(V, pts)
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    assert abs(np.sum(det * w) - 0.6) < 1e-14              # trapezoid area (0.7+0.5)/2
    for f, n_expect in ((0, (-1, 0)), (2, (0, -1)), (3, (0, 1))):
        _, n, ln = fe.face_geometry(V, f, np.array([0.3]))
        assert np.allclose(n[0, 0], n_expect)
    _, n, ln = fe.face_geometry(V, 1, np.array([0.3]))
    t = np.array([-0.2, 1.0]) / np.hypot(0.2, 1.0)
    assert abs(n[0, 0] @ t) < 1e-15 and n[0, 0, 0] > 0   # outward, orthogonal to the slanted edge
    assert abs(ln[0, 0] - np.hypot(0.2, 1.0)) < 1e-15
    assert abs(fe.diameter(V)[0] - max(np.hypot(0.5, 1.0), np.hypot(0.7, 1.0))) < 1e-15


def test_morton_matches_deal_ii_child_numbering():
    # level index = 4*parent + child, child (0,0),(1,0),(0,1),(1,1)
    for r in (1, 2, 3):
        N = 1 << r
        for c in range(N * N):
            i, j = morton_ij(c, r)
            parent, child = c // 4, c % 4
            pi, pj = morton_ij(parent, r - 1) if r > 1 else (0, 0)
            assert (i, j) == (2 * pi + (child & 1), 2 * pj + (child >> 1))


def test_vertex_grid_boundaries_exact():
    G = vertex_grid([(0, 0), (0.61, 0), (0, 1.0), (0.5, 1.0)], 5)
    assert np.all(G[0, :, 0] == 0.0) and np.all(G[:, 0, 1] == 0.0) and np.all(G[:, -1, 1] == 1.0)
    m = SolarCellMesh(3, dict(r1=0.5, r2=0.6, H=1.0, L=1.5))
    # interface pairing matches face centres exactly (SolarCell.cpp:232-240)
    for s, e in zip(m.sigma_semi, m.sigma_elec):
        cs = (m.semi.V[s, 1] + m.semi.V[s, 3]) / 2
        ce = (m.elec.V[e, 0] + m.elec.V[e, 2]) / 2
        assert np.linalg.norm(cs - ce) < 1e-13
