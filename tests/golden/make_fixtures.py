"""Generate exact golden vectors for the oracle's constant matrices (sympy).

The reference holds no golden vectors, so these fixtures are derived here by
exact symbolic integration, restating the reference's bilinear forms directly
from its source:

* ldg_<case>.json -- the LDG system matrix A_1 and the mass matrix M of one
  carrier pair on a 2 x 2 rectangular region: cell and boundary terms
  (LDG.cpp:85-281), interior fluxes assembled once from the lower-index cell
  with its normal and beta = (1,1)/sqrt(2) (LDG.cpp:283-678), mass
  (LDG.cpp:40-83).  FE_DGQ(k) equispaced Lagrange basis, lexicographic,
  component-wise [Jx | Jy | u] numbering in deal.II's cell order.
* mixed_<case>.json -- the potential Schur complement S = C A^{-1} B of the
  mixed Poisson matrix [[A, B], [C, 0]] (MixedFEM.cpp:29-164) on the 4 x 2
  rectangular Poisson mesh, with D.n = 0 on Neumann faces (Poisson.cpp:32-49).
  S does not depend on the basis of the Raviart-Thomas space (A -> T^T A T,
  B -> T^T B, C -> C T leaves C A^{-1} B unchanged), so it pins the oracle's own
  RT basis against the monomial basis used here: the conforming space is the
  null space of the normal-continuity constraints on the broken space
  Q_{k,k-1} x Q_{k-1,k} per cell.

On rectangles every integrand is a polynomial, so Gauss quadrature with k+2
points is exact and the oracle must agree to rounding.

    python tests/golden/make_fixtures.py     (writes the JSON files next to it)
"""
import json
import os
import sys

import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
X, Y = sp.symbols("x y")

DIRICHLET, INTERFACE, NEUMANN = 1, 0, 2


def lagrange(k, t):
    nodes = [sp.Rational(i, k) for i in range(k + 1)] if k > 0 else [sp.Rational(1, 2)]
    out = []
    for a, na in enumerate(nodes):
        e = sp.Integer(1)
        for b, nb_ in enumerate(nodes):
            if b != a:
                e *= (t - nb_) / (na - nb_)
        out.append(sp.expand(e))
    return out


class Rect:
    def __init__(self, x0, y0, a, b):
        self.x0, self.y0, self.a, self.b = x0, y0, a, b

    def integrate(self, f):
        return sp.integrate(sp.integrate(f, (X, self.x0, self.x0 + self.a)), (Y, self.y0, self.y0 + self.b))

    def face(self, f):
        """(fixed variable, value, free variable, lo, hi, outward normal)"""
        if f == 0:
            return X, self.x0, Y, self.y0, self.y0 + self.b, (-1, 0)
        if f == 1:
            return X, self.x0 + self.a, Y, self.y0, self.y0 + self.b, (1, 0)
        if f == 2:
            return Y, self.y0, X, self.x0, self.x0 + self.a, (0, -1)
        return Y, self.y0 + self.b, X, self.x0, self.x0 + self.a, (0, 1)

    def face_integrate(self, f, g):
        var, val, free, lo, hi, _ = self.face(f)
        return sp.integrate(g.subs(var, val), (free, lo, hi))

    def center(self):
        return (self.x0 + self.a / 2, self.y0 + self.b / 2)


def dgq_basis(k, R):
    """FE_DGQ(k) on an axis-aligned rectangle: lexicographic, x fastest."""
    lx = lagrange(k, (X - R.x0) / R.a)
    ly = lagrange(k, (Y - R.y0) / R.b)
    return [sp.expand(lx[i] * ly[j]) for j in range(k + 1) for i in range(k + 1)]


def morton(c, r):
    i = j = 0
    for b in range(r):
        i |= ((c >> (2 * b)) & 1) << b
        j |= ((c >> (2 * b + 1)) & 1) << b
    return i, j


def region_cells(x0, y0, w, h, r):
    x0, y0, w, h = (sp.nsimplify(v) for v in (x0, y0, w, h))
    n = sp.Integer(1 << r)
    cells = []
    for c in range(int(n * n)):
        i, j = morton(c, r)
        cells.append(Rect(x0 + i * w / n, y0 + j * h / n, w / n, h / n))
    return cells


def ldg_fixture(k, r, mu, dt, theta, pen, insulated):
    """Semiconductor region [0, 1/2] x [0, 1]: x = 0 Dirichlet (Neumann when
    insulated), x = 1/2 the interface, y = 0 and y = 1 Dirichlet (Grid.cpp:340-461
    with r1 = r2 = 1/2: insulated top/bottom faces with x < r1 stay Dirichlet)."""
    half = sp.Rational(1, 2)
    cells = region_cells(0, 0, half, 1, r)
    nc = len(cells)
    nb = (k + 1) ** 2
    n = 3 * nb * nc
    beta = (1 / sp.sqrt(2), 1 / sp.sqrt(2))
    A = {}
    M = {}

    def add(D, i, j, v):
        v = sp.nsimplify(v) if v.is_Rational else v
        if v != 0:
            D[(i, j)] = D.get((i, j), 0) + v

    def gidx(c, comp, a):
        return comp * nb * nc + c * nb + a

    # test/trial functions as (vector part, scalar part) per local index
    def funcs(R):
        ph = dgq_basis(k, R)
        F = []
        for comp in range(3):
            for a in range(nb):
                if comp == 0:
                    F.append(((ph[a], 0), 0, comp, a))
                elif comp == 1:
                    F.append(((0, ph[a]), 0, comp, a))
                else:
                    F.append(((0, 0), ph[a], comp, a))
        return F

    def dot(u, v):
        return u[0] * v[0] + u[1] * v[1]

    def div(p):
        return sp.diff(p[0], X) + sp.diff(p[1], Y)

    def grad(v):
        return (sp.diff(v, X), sp.diff(v, Y))

    diam = [sp.sqrt(R.a ** 2 + R.b ** 2) for R in cells]
    N = 1 << r
    for c, R in enumerate(cells):
        F = funcs(R)
        for (pi, vi, ci, ai) in F:
            for (qj, uj, cj, aj) in F:
                e = (theta / dt) * vi * uj + dot(pi, qj) / mu - div(pi) * uj - dot(grad(vi), qj)
                add(A, gidx(c, ci, ai), gidx(c, cj, aj), R.integrate(sp.expand(e)))
                add(M, gidx(c, ci, ai), gidx(c, cj, aj), R.integrate(sp.expand(vi * uj / dt)))
        i, j = morton(c, r)
        for f in range(4):
            di, dj = {0: (-1, 0), 1: (1, 0), 2: (0, -1), 3: (0, 1)}[f]
            ni, nj = i + di, j + dj
            nrm = R.face(f)[5]
            if 0 <= ni < N and 0 <= nj < N:
                nbr = next(q for q in range(nc) if morton(q, r) == (ni, nj))
                if nbr < c:
                    continue
                h = sp.Min(diam[c], diam[nbr])
                sig = pen / h
                G = funcs(cells[nbr])
                for (pi, vi, ci, ai) in F:  # minus-side tests
                    for (qj, uj, cj, aj) in F:
                        e = (sp.Rational(1, 2) * (dot(pi, nrm) * uj + vi * dot(nrm, qj)) + dot(beta, pi) * uj
                             - vi * dot(beta, qj) + sig * vi * uj)
                        add(A, gidx(c, ci, ai), gidx(c, cj, aj), R.face_integrate(f, sp.expand(e)))
                    for (qj, uj, cj, aj) in G:
                        e = (sp.Rational(1, 2) * (dot(pi, nrm) * uj + vi * dot(nrm, qj)) - dot(beta, pi) * uj
                             + vi * dot(beta, qj) - sig * vi * uj)
                        add(A, gidx(c, ci, ai), gidx(nbr, cj, aj), R.face_integrate(f, sp.expand(e)))
                for (pi, vi, ci, ai) in G:  # plus-side tests
                    for (qj, uj, cj, aj) in F:
                        e = (-sp.Rational(1, 2) * (dot(pi, nrm) * uj + vi * dot(nrm, qj)) - dot(beta, pi) * uj
                             + vi * dot(beta, qj) - sig * vi * uj)
                        add(A, gidx(nbr, ci, ai), gidx(c, cj, aj), R.face_integrate(f, sp.expand(e)))
                    for (qj, uj, cj, aj) in G:
                        e = (-sp.Rational(1, 2) * (dot(pi, nrm) * uj + vi * dot(nrm, qj)) + dot(beta, pi) * uj
                             - vi * dot(beta, qj) + sig * vi * uj)
                        add(A, gidx(nbr, ci, ai), gidx(nbr, cj, aj), R.face_integrate(f, sp.expand(e)))
                continue
            # boundary face
            if f == 1:
                bid = INTERFACE
            elif f == 0:
                bid = NEUMANN if insulated else DIRICHLET
            else:
                bid = DIRICHLET
            for (pi, vi, ci, ai) in F:
                for (qj, uj, cj, aj) in F:
                    if bid == DIRICHLET:
                        e = vi * (dot(nrm, qj) + (pen / diam[c]) * uj)
                    else:
                        e = dot(pi, nrm) * uj
                    add(A, gidx(c, ci, ai), gidx(c, cj, aj), R.face_integrate(f, sp.expand(e)))
    return n, A, M


def rt_broken_basis(k, R):
    """Monomial basis of RT_{k-1} = Q_{k,k-1} x Q_{k-1,k} on a rectangle."""
    B = []
    for a in range(k + 1):
        for b in range(k):
            B.append((X ** a * Y ** b, sp.Integer(0)))
    for a in range(k):
        for b in range(k + 1):
            B.append((sp.Integer(0), X ** a * Y ** b))
    return B


def mixed_fixture(k, r, eps_s, eps_e, lam2, insulated):
    """Potential Schur complement on the Poisson mesh [0,1] x [0,1] at
    refinement r (both regions), semiconductor x < 1/2."""
    half = sp.Rational(1, 2)
    cells = region_cells(0, 0, half, 1, r) + region_cells(half, 0, half, 1, r)
    nc = len(cells)
    npc = k * k  # DGQ(k-1) per cell
    rt = [rt_broken_basis(k, R) for R in cells]
    nrt = len(rt[0])
    nbroken = nc * nrt
    # constraints: normal continuity on interior faces, D.n = 0 on Neumann faces
    rows = []
    tol_faces = []
    for c, R in enumerate(cells):
        for f in range(4):
            var, val, free, lo, hi, nrm = R.face(f)
            comp = 0 if f < 2 else 1
            # the other cell sharing this face
            other = None
            for d, S in enumerate(cells):
                if d == c:
                    continue
                for g in range(4):
                    v2, val2, fr2, lo2, hi2, n2 = S.face(g)
                    if v2 == var and val2 == val and lo2 == lo and hi2 == hi and n2 != nrm:
                        other = (d, g)
            if other is not None and other[0] < c:
                continue
            cx = val if var == X else (lo + hi) / 2
            cy = val if var == Y else (lo + hi) / 2
            neumann = insulated and (((cy == 0 or cy == 1) and cx > half) or cx == 0)
            if other is None and not neumann:
                continue
            for m in range(k):  # coefficients of free**m of the normal component
                row = [sp.Integer(0)] * nbroken
                for a, psi in enumerate(rt[c]):
                    p = sp.Poly(sp.expand(psi[comp].subs(var, val)), free)
                    row[c * nrt + a] = p.coeff_monomial(free ** m)
                if other is not None:
                    d, g = other
                    for a, psi in enumerate(rt[d]):
                        p = sp.Poly(sp.expand(psi[comp].subs(var, val)), free)
                        row[d * nrt + a] -= p.coeff_monomial(free ** m)
                rows.append(row)
    Cm = sp.Matrix(rows)
    T = sp.Matrix.hstack(*Cm.nullspace())
    Ab = sp.zeros(nbroken, nbroken)
    Bb = sp.zeros(nbroken, nc * npc)
    Cb = sp.zeros(nc * npc, nbroken)
    for c, R in enumerate(cells):
        eps = eps_s if R.x0 < half else eps_e
        ph = dgq_basis(k - 1, R)
        for a, pa in enumerate(rt[c]):
            for b, pb in enumerate(rt[c]):
                Ab[c * nrt + a, c * nrt + b] = R.integrate(sp.expand((pa[0] * pb[0] + pa[1] * pb[1]) / eps))
            dv = sp.diff(pa[0], X) + sp.diff(pa[1], Y)
            for j, phj in enumerate(ph):
                v = R.integrate(sp.expand(dv * phj))
                Bb[c * nrt + a, c * npc + j] = -v
                Cb[c * npc + j, c * nrt + a] = -lam2 * v
    A = T.T * Ab * T
    B = T.T * Bb
    C = Cb * T
    S = C * A.LUsolve(B)
    return S


def _dump_sparse(n, D):
    keys = sorted(D)
    return {"n": n, "i": [k[0] for k in keys], "j": [k[1] for k in keys],
            "v": [float(sp.N(D[k], 30)) for k in keys]}


CASES_LDG = {
    "k1_r1": dict(k=1, r=1, mu=sp.Rational(3, 2), dt=sp.Rational(1, 100), theta=1, pen=1, insulated=False),
    "k2_r1_insulated": dict(k=2, r=1, mu=sp.Rational(7, 10), dt=sp.Rational(1, 2000), theta=1, pen=2,
                            insulated=True),
}
CASES_MIXED = {
    "k1_r1": dict(k=1, r=1, eps_s=2, eps_e=3, lam2=sp.Rational(1, 2), insulated=False),
    "k2_r1_insulated": dict(k=2, r=1, eps_s=sp.Rational(119, 10), eps_e=100, lam2=sp.Rational(3, 4),
                            insulated=True),
}


def main():
    for name, c in CASES_LDG.items():
        n, A, M = ldg_fixture(**c)
        out = {"params": {k: (float(v) if not isinstance(v, bool) else v) for k, v in c.items()},
               "A": _dump_sparse(n, A), "M": _dump_sparse(n, M)}
        with open(os.path.join(HERE, f"ldg_{name}.json"), "w") as f:
            json.dump(out, f)
        print("ldg", name, n, len(A), flush=True)
    for name, c in CASES_MIXED.items():
        S = mixed_fixture(**c)
        out = {"params": {k: (float(v) if not isinstance(v, bool) else v) for k, v in c.items()},
               "S": [[float(sp.N(S[i, j], 30)) for j in range(S.shape[1])] for i in range(S.shape[0])]}
        with open(os.path.join(HERE, f"mixed_{name}.json"), "w") as f:
            json.dump(out, f)
        print("mixed", name, S.shape, flush=True)


if __name__ == "__main__":
    sys.exit(main())
