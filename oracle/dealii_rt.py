"""deal.II's Raviart-Thomas element and Poisson DoF numbering (oracle, test
infrastructure).

Restated from deal.II's published definitions (deal.II 8.x, the reference's
FE_RaviartThomas(k-1) x FE_DGQ(k-1) Poisson element, SolarCell.cpp:23-24,
numbered by DoFHandler::distribute_dofs and DoFRenumbering::component_wise,
Poisson.cpp:18-30), written independently of the library's C++ conversion:
  * the space RT_deg on the reference square, spanned here by monomials:
    x-component xi^a eta^b (a <= deg+1, b <= deg), y-component xi^a eta^b
    (a <= deg, b <= deg+1);
  * the node functionals: on face f (0: xi=0, 1: xi=1, 2: eta=0, 3: eta=1)
    l_{f,i}(v) = sum_q w_q P_i(s_q) v_d, d = x on faces 0, 1 and y on 2, 3,
    QGauss(deg+1) on the face, P_i = sqrt(2i+1) L_i(2t-1) (Legendre on [0,1],
    orthonormal); interior l_{j,d}(v) = sum_q w_q Q_{d,j}(x_q) v_d(x_q) over
    QGauss(deg+1)^2, Q_{d,j} = P_jx(xi) P_jy(eta) with degree <= deg-1 in
    direction d and <= deg in the other, j = jx + nx jy, local index
    4(deg+1) + 2j + d;
  * shape functions: the basis dual to the functionals (node matrix inverse),
    mapped by the contravariant Piola transform;
  * numbering: active cells in order, each cell's unnumbered lines 0..3 first
    (deg+1 DoFs each), then its interior; Phi after all of D.
"""
import numpy as np

from . import fe


def legendre01(n, t):
    c = np.zeros(n + 1)
    c[n] = 1.0
    return np.sqrt(2 * n + 1) * np.polynomial.legendre.legval(2.0 * np.asarray(t) - 1.0, c)


def monomials(deg):
    """(component, a, b) of the monomial basis of RT_deg."""
    out = [(0, a, b) for b in range(deg + 1) for a in range(deg + 2)]
    out += [(1, a, b) for b in range(deg + 2) for a in range(deg + 1)]
    return out


def _mono_val(m, xi, eta):
    comp, a, b = m
    v = np.zeros(np.broadcast(xi, eta).shape + (2,))
    v[..., comp] = np.asarray(xi) ** a * np.asarray(eta) ** b
    return v


def functionals(deg):
    """Node functionals as (points[np, 2], weights[np], component) triples."""
    s, w = fe.gauss01(deg + 1)
    out = []
    for f in range(4):
        pts = fe.face_points(f, s)
        comp = 0 if f < 2 else 1
        for i in range(deg + 1):
            out.append((pts, w * legendre01(i, s), comp))
    cp, cw = fe.cell_quadrature(deg + 1)
    for j in range(deg * (deg + 1)):
        for d in range(2):
            nx = deg if d == 0 else deg + 1
            jx, jy = j % nx, j // nx
            out.append((cp, cw * legendre01(jx, cp[:, 0]) * legendre01(jy, cp[:, 1]), d))
    return out


class DealiiRT:
    """deal.II's FE_RaviartThomas(deg) shape functions on the reference square."""

    def __init__(self, deg):
        self.deg = deg
        self.mono = monomials(deg)
        self.fun = functionals(deg)
        n = len(self.mono)
        assert len(self.fun) == n
        C = np.zeros((n, n))
        for i, (pts, w, comp) in enumerate(self.fun):
            for m, mo in enumerate(self.mono):
                C[i, m] = np.sum(w * _mono_val(mo, pts[:, 0], pts[:, 1])[:, comp])
        self.C = C
        self.coef = np.linalg.inv(C)  # shape function j = sum_m coef[m, j] monomial m
        self.n = n

    def tab(self, pts):
        """Reference values [n, np, 2] of the shape functions at pts[np, 2]."""
        M = np.stack([_mono_val(m, pts[:, 0], pts[:, 1]) for m in self.mono])  # [nmono, np, 2]
        return np.einsum("mj,mpd->jpd", self.coef, M)


def dealii_numbering(mesh, deg):
    """Per Poisson cell the deal.II global index of each local RT DoF, and nD."""
    nf = deg + 1
    nint = 2 * deg * (deg + 1)
    NX = mesh.NX
    vb, hb = {}, {}
    nxt = 0
    idx = []
    for (I, J) in mesh.pIJ:
        keys = [(vb, (I, J)), (vb, (I + 1, J)), (hb, (I, J)), (hb, (I, J + 1))]
        loc = []
        for tbl, key in keys:
            if key not in tbl:
                tbl[key] = nxt
                nxt += nf
            loc += [tbl[key] + i for i in range(nf)]
        loc += [nxt + i for i in range(nint)]
        nxt += nint
        idx.append(loc)
    del NX
    return np.array(idx, dtype=np.int64), nxt


def field_at(mesh, cells, ref_pts, coef_rt, tab):
    """Physical D at reference points of the given cells: Piola of sum_j c_j psi_j."""
    V = mesh.pV[cells]
    x, J = fe.bilinear(V, ref_pts)
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    Dhat = np.einsum("cj,jqd->cqd", coef_rt, tab)
    return np.einsum("cqrd,cqd->cqr", J, Dhat) / det[..., None]
