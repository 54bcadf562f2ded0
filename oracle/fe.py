"""Finite-element tabulations and MappingQ1 geometry (oracle, test infrastructure).

Restates what deal.II's FEValues provide to the reference:
  * QGauss<dim>(k+2) / QGauss<dim-1>(k+2) for every cell and face integral
    (Assembly.cpp:9-68, e.g. SolarCell.cpp:402-403);
  * FE_DGQ(k) -- Lagrange basis on equispaced support points, lexicographic
    local numbering (x fastest); the carrier element is
    FESystem(FE_DGQ(k)^dim, FE_DGQ(k)) (SolarCell.cpp:29-30) with local order
    [Jx | Jy | u];
  * FE_RaviartThomas(k-1) x FE_DGQ(k-1) for the Poisson problem
    (SolarCell.cpp:23-24).  The RT basis here is a nodal basis of the same
    space: x-component nodes = equispaced points (including 0 and 1) in x
    times Gauss points in y, y-component the transpose.  Face nodes carry the
    reference-coordinate normal flux, so a face DoF is shared by the two cells
    of a face exactly as deal.II shares its RT face DoFs; the coefficient
    vector differs from deal.II's by a change of basis, the functions D and
    Phi do not;
  * MappingQ1: bilinear map of the 4 cell vertices (deal.II vertex order
    v0=(0,0), v1=(1,0), v2=(0,1), v3=(1,1)), covariant gradients, contravariant
    Piola transform for RT, face normals and JxW.
"""
import numpy as np


def gauss01(n):
    """n-point Gauss-Legendre rule on [0,1] (deal.II QGauss<1>(n))."""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def lagrange1d(nodes, x):
    """Values and derivatives of the Lagrange polynomials on `nodes` at `x`.

    Returns (V, D) with V[a, j] = l_a(x_j), D[a, j] = l_a'(x_j).
    """
    nodes = np.asarray(nodes, dtype=float)
    x = np.atleast_1d(np.asarray(x, dtype=float))
    n = len(nodes)
    V = np.ones((n, len(x)))
    D = np.zeros((n, len(x)))
    for a in range(n):
        for b in range(n):
            if b == a:
                continue
            V[a] *= (x - nodes[b]) / (nodes[a] - nodes[b])
        # derivative: sum over the dropped factor
        for c in range(n):
            if c == a:
                continue
            term = np.full(len(x), 1.0 / (nodes[a] - nodes[c]))
            for b in range(n):
                if b == a or b == c:
                    continue
                term = term * (x - nodes[b]) / (nodes[a] - nodes[b])
            D[a] += term
    return V, D


def equispaced(k):
    """Support points of FE_DGQ(k) in 1D (equidistant, k>=1); k=0 -> centre."""
    if k == 0:
        return np.array([0.5])
    return np.linspace(0.0, 1.0, k + 1)


def tensor_points(x1d):
    """2D tensor-product points, x fastest: q = i + n*j."""
    n = len(x1d)
    X = np.empty((n * n, 2))
    for j in range(n):
        for i in range(n):
            X[i + n * j] = (x1d[i], x1d[j])
    return X


def cell_quadrature(nq1):
    x, w = gauss01(nq1)
    pts = tensor_points(x)
    W = np.array([w[i] * w[j] for j in range(nq1) for i in range(nq1)])
    return pts, W


def face_points(face, s):
    """Reference points on face `face` (deal.II 2D numbering) at params s."""
    s = np.asarray(s, dtype=float)
    z = np.zeros_like(s)
    o = np.ones_like(s)
    if face == 0:
        return np.stack([z, s], -1)
    if face == 1:
        return np.stack([o, s], -1)
    if face == 2:
        return np.stack([s, z], -1)
    return np.stack([s, o], -1)


class DGQ:
    """FE_DGQ(k) scalar basis, nb=(k+1)^2 functions, lexicographic (x fastest)."""

    def __init__(self, k):
        self.k = k
        self.nodes = equispaced(k)
        self.n1 = len(self.nodes)
        self.nb = self.n1 ** 2

    def tab(self, pts):
        """values [nb, np], grads [nb, np, 2] at reference points pts[np,2]."""
        Vx, Dx = lagrange1d(self.nodes, pts[:, 0])
        Vy, Dy = lagrange1d(self.nodes, pts[:, 1])
        n1 = self.n1
        val = np.empty((self.nb, len(pts)))
        grd = np.empty((self.nb, len(pts), 2))
        for b in range(n1):
            for a in range(n1):
                i = a + n1 * b
                val[i] = Vx[a] * Vy[b]
                grd[i, :, 0] = Dx[a] * Vy[b]
                grd[i, :, 1] = Vx[a] * Dy[b]
        return val, grd


class RT:
    """Nodal Raviart-Thomas RT_deg basis on the reference square.

    x-component functions: L_a(xi) G_b(eta), a over deg+2 equispaced nodes
    in [0,1] (a=0: left face, a=deg+1: right face), b over deg+1 Gauss
    nodes; local index a + (deg+2)*b.  y-component functions:
    G_a(xi) L_b(eta), local index nx + a + (deg+1)*b.
    """

    def __init__(self, deg):
        self.deg = deg
        self.Xn = np.linspace(0.0, 1.0, deg + 2)
        self.Gn, _ = gauss01(deg + 1)
        self.nx = (deg + 2) * (deg + 1)
        self.n = 2 * self.nx

    def tab(self, pts):
        """vals [n, np, 2] (reference), div [n, np] (reference divergence)."""
        LX, DLX = lagrange1d(self.Xn, pts[:, 0])
        GY, DGY = lagrange1d(self.Gn, pts[:, 1])
        GX, DGX = lagrange1d(self.Gn, pts[:, 0])
        LY, DLY = lagrange1d(self.Xn, pts[:, 1])
        d = self.deg
        val = np.zeros((self.n, len(pts), 2))
        div = np.zeros((self.n, len(pts)))
        for b in range(d + 1):
            for a in range(d + 2):
                i = a + (d + 2) * b
                val[i, :, 0] = LX[a] * GY[b]
                div[i] = DLX[a] * GY[b]
        for b in range(d + 2):
            for a in range(d + 1):
                i = self.nx + a + (d + 1) * b
                val[i, :, 1] = GX[a] * LY[b]
                div[i] = GX[a] * DLY[b]
        return val, div


def bilinear(V, pts):
    """MappingQ1 of cells V[nc,4,2] at reference pts[np,2].

    Returns x[nc,np,2], J[nc,np,2,2] with J[...,r,c] = d x_r / d xi_c.
    """
    xi = pts[:, 0][None, :, None]
    et = pts[:, 1][None, :, None]
    v0, v1, v2, v3 = (V[:, i][:, None, :] for i in range(4))
    x = v0 * (1 - xi) * (1 - et) + v1 * xi * (1 - et) + v2 * (1 - xi) * et + v3 * xi * et
    dxi = (v1 - v0) * (1 - et) + (v3 - v2) * et
    det_ = (v2 - v0) * (1 - xi) + (v3 - v1) * xi
    J = np.stack([dxi, det_], -1)  # [nc,np,2(r),2(c)]
    return x, J


def face_geometry(V, face, s):
    """Physical points, outward unit normals and |dx/ds| on face `face`."""
    pts = face_points(face, s)
    x, J = bilinear(V, pts)
    if face in (0, 1):
        t = J[..., 1]
    else:
        t = J[..., 0]
    ln = np.linalg.norm(t, axis=-1)
    if face in (1, 2):
        n = np.stack([t[..., 1], -t[..., 0]], -1)
    else:
        n = np.stack([-t[..., 1], t[..., 0]], -1)
    n = n / ln[..., None]
    return x, n, ln


def diameter(V):
    """deal.II 2D cell->diameter(): the longer diagonal."""
    d1 = np.linalg.norm(V[:, 3] - V[:, 0], axis=-1)
    d2 = np.linalg.norm(V[:, 2] - V[:, 1], axis=-1)
    return np.maximum(d1, d2)
