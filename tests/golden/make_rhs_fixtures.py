"""Generate golden vectors for the per-step right-hand sides (high precision).

The reference holds no golden vectors, so these are computed here, at 40
significant digits (mpmath, Gauss-Legendre nodes from sympy), by a scalar
restatement of the reference's per-step forms written directly from its
source, sharing no code with the oracle:

* carrier right-hand sides b = M u_old + sum_K b_K of both regions
  - semiconductor body (SolarCell.cpp:1146-1193):
      sum_q JxW [ v_i G(x_q) + v_i R(n_q, p_q) + s_c p_i . D_q / eps_S u_c,q ]
  - Dirichlet faces (SolarCell.cpp:1212-1261): sum_q ds (-p_i.n + pen/h v_i) u^e
  - interface faces, semiconductor side (SolarCell.cpp:1262-1343):
      b_n += -sum ds v_i k_et (n - n^e) o_nb,   b_p += +sum ds v_i k_ht (p - p^e) r_nb
  - Schottky faces (SolarCell.cpp:1344-1401):
      b_n += -sum ds v_i v_rec_n (n - n^e),     b_p += +sum ds v_i v_rec_p (p - p^e)
  - electrolyte body (SolarCell.cpp:1519-1550): sum_q JxW s_c p_i . D_q / eps_E u_c,q
  - electrolyte Dirichlet faces (SolarCell.cpp:1580-1624) with r^inf, o^inf
  - electrolyte interface faces (SolarCell.cpp:1640-1716):
      I_i = sum ds [ -v_i k_et (n_nb - n^e) o + v_i k_ht (p_nb - p^e) r ],
      b_r += I_i, b_o -= I_i
  - mass (LDG.cpp:40-83, applied as in SolarCell.cpp:1043-1047):
      (M u_old)_i = 1/dt sum_q JxW v_i u_old,q
  Neighbour traces are paired by quadrature-point index (SolarCell.cpp:1301-1311).
* Poisson right-hand side (SolarCell.cpp:430-815)
  - semiconductor cell: -sum_q JxW phi_j [ (n^e - p^e) + (-n + p) ]
  - electrolyte cell:   -sum_q JxW phi_j (-r + o)
  - semiconductor Dirichlet faces: -sum ds (psi.n)(Phi_bi(x) - Phi_app(x)),
    the biases being nonzero only where x == 0 (BiasValues.cpp:11-31, 75-95).
  The Raviart-Thomas part is stored basis-free: per boundary face and monomial
  s^m of the face parameter, -sum_q w_q s_q^m g(x_q) (the reference normal
  flux of any RT shape function on a face is a polynomial of degree k-1 in s,
  and the contravariant Piola map makes psi.n ds equal to that flux times ds_ref).

Generation: 'ref' alpha G0 exp(alpha (y - H)) (Generation.cpp:29-44),
'config' G0 exp(-alpha x).  Recombination: 0 (SolarCell.hpp:86-98), or
with srh the documented SRH formula (Documentation/README.md:266-271)
(n_i^2 - n p) / (tau_n (n + n_i) + tau_p (p + n_i)).

Meshes: one coarse quad per region (Grid.cpp:135-336) refined once (2 x 2
cells, deal.II order), boundary ids by the face-centre rules of
Grid.cpp:340-461.  Inputs are seeded doubles, used exactly; the outputs are
rounded to doubles at the end.

    python tests/golden/make_rhs_fixtures.py   (writes rhs_<case>.json next to it)
"""
import json
import os
import sys

import mpmath as mp
import numpy as np
from sympy.integrals.quadrature import gauss_legendre

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 40

INTERFACE, DIRICHLET, NEUMANN, SCHOTTKY = 0, 1, 2, 3


def gauss01(n):
    xs, ws = gauss_legendre(n, 45)
    return [(mp.mpf(str(x)) + 1) / 2 for x in xs], [mp.mpf(str(w)) / 2 for w in ws]


def lag(nodes, a, t):
    v = mp.mpf(1)
    for b, nb_ in enumerate(nodes):
        if b != a:
            v *= (t - nb_) / (nodes[a] - nb_)
    return v


def dlag(nodes, a, t):
    s = mp.mpf(0)
    for c in range(len(nodes)):
        if c == a:
            continue
        term = 1 / (nodes[a] - nodes[c])
        for b in range(len(nodes)):
            if b != a and b != c:
                term *= (t - nodes[b]) / (nodes[a] - nodes[b])
        s += term
    return s


def dgq_nodes(k):
    if k == 0:
        return [mp.mpf(1) / 2]
    return [mp.mpf(i) / k for i in range(k + 1)]


def dgq(k, xi, eta):
    """FE_DGQ(k) values, lexicographic (x fastest)."""
    nd = dgq_nodes(k)
    return [lag(nd, i, xi) * lag(nd, j, eta) for j in range(k + 1) for i in range(k + 1)]


def region_vertices(coarse):
    """Vertices of the 2 x 2 refinement: corners, edge midpoints, centre (doubles)."""
    v0, v1, v2, v3 = [np.array(v, dtype=float) for v in coarse]
    G = {}
    G[0, 0], G[2, 0], G[0, 2], G[2, 2] = v0, v1, v2, v3
    G[1, 0] = (v0 + v1) / 2.0
    G[1, 2] = (v2 + v3) / 2.0
    G[0, 1] = (v0 + v2) / 2.0
    G[2, 1] = (v1 + v3) / 2.0
    G[1, 1] = (v0 + v1 + v2 + v3) / 4.0
    cells = []
    for (i, j) in ((0, 0), (1, 0), (0, 1), (1, 1)):  # deal.II child order
        cells.append([G[i, j], G[i + 1, j], G[i, j + 1], G[i + 1, j + 1]])
    return np.array(cells)


def face_bid(cx, cy, geo, insulated, schottky):
    L, H, r1 = geo["L"], geo["H"], geo["r1"]
    b = INTERFACE
    if cx == 0.0 or cx == L or cy == 0.0 or cy == H:
        b = DIRICHLET
    if insulated:
        if cy == H or cy == 0.0:
            b = NEUMANN if cx > r1 else DIRICHLET
        if cx == 0.0:
            b = NEUMANN
    if schottky and cy == H:
        b = SCHOTTKY
    return b


FACE_NB = {0: (-1, 0), 1: (1, 0), 2: (0, -1), 3: (0, 1)}
FACE_V = {0: (0, 2), 1: (1, 3), 2: (0, 1), 3: (2, 3)}
IJ = [(0, 0), (1, 0), (0, 1), (1, 1)]


def bids(V, geo, insulated, schottky):
    out = []
    for c, (i, j) in enumerate(IJ):
        row = []
        for f in range(4):
            di, dj = FACE_NB[f]
            if 0 <= i + di < 2 and 0 <= j + dj < 2:
                row.append(-1)
            else:
                a, b = FACE_V[f]
                cen = (V[c, a] + V[c, b]) / 2.0
                row.append(face_bid(cen[0], cen[1], geo, insulated, schottky))
        out.append(row)
    return out


class Cell:
    def __init__(self, V):
        self.v = [[mp.mpf(float(V[a, d])) for d in range(2)] for a in range(4)]

    def x(self, xi, eta):
        v0, v1, v2, v3 = self.v
        return [v0[d] * (1 - xi) * (1 - eta) + v1[d] * xi * (1 - eta) + v2[d] * (1 - xi) * eta + v3[d] * xi * eta
                for d in range(2)]

    def jac(self, xi, eta):
        v0, v1, v2, v3 = self.v
        dxi = [(v1[d] - v0[d]) * (1 - eta) + (v3[d] - v2[d]) * eta for d in range(2)]
        deta = [(v2[d] - v0[d]) * (1 - xi) + (v3[d] - v1[d]) * xi for d in range(2)]
        return [[dxi[0], deta[0]], [dxi[1], deta[1]]]

    def diameter(self):
        v0, v1, v2, v3 = self.v
        d1 = mp.sqrt((v3[0] - v0[0]) ** 2 + (v3[1] - v0[1]) ** 2)
        d2 = mp.sqrt((v2[0] - v1[0]) ** 2 + (v2[1] - v1[1]) ** 2)
        return max(d1, d2)

    def face(self, f, s):
        """(reference point, physical point, outward unit normal, |dx/ds|)."""
        ref = {0: (mp.mpf(0), s), 1: (mp.mpf(1), s), 2: (s, mp.mpf(0)), 3: (s, mp.mpf(1))}[f]
        J = self.jac(*ref)
        col = 1 if f < 2 else 0
        t = [J[0][col], J[1][col]]
        ln = mp.sqrt(t[0] ** 2 + t[1] ** 2)
        if f in (1, 2):
            n = [t[1] / ln, -t[0] / ln]
        else:
            n = [-t[1] / ln, t[0] / ln]
        return ref, self.x(*ref), n, ln


def density(k, coef, xi, eta):
    ph = dgq(k, xi, eta)
    return sum(c * p for c, p in zip(coef, ph))


def field(deg, dcoef, cell, xi, eta):
    """Piola image D = J Dhat / det J of the reference field
    Dhat_x = sum a_{pq} xi^p eta^q (p <= deg+1, q <= deg),
    Dhat_y = sum b_{pq} xi^p eta^q (p <= deg, q <= deg+1)."""
    ax, ay = dcoef
    dx = sum(ax[q][p] * xi ** p * eta ** q for q in range(deg + 1) for p in range(deg + 2))
    dy = sum(ay[q][p] * xi ** p * eta ** q for q in range(deg + 2) for p in range(deg + 1))
    J = cell.jac(xi, eta)
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0]
    return [(J[0][0] * dx + J[0][1] * dy) / det, (J[1][0] * dx + J[1][1] * dy) / det]


def generation(P, x):
    if P["gen_mode"] == "ref":
        return P["alpha"] * P["G0"] * mp.exp(P["alpha"] * (x[1] - P["H"]))
    if P["gen_mode"] == "config":
        return P["G0"] * mp.exp(-P["alpha"] * x[0])
    return mp.mpf(0)


def recombination(P, n, p):
    if not P["srh"]:
        return mp.mpf(0)
    ni = P["ni"]
    return (ni * ni - n * p) / (P["tau_n"] * (n + ni) + P["tau_p"] * (p + ni))


def carrier_rhs(P, k, cells, bid, U1, U2, D, Onb1, Onb2, nbcell, which):
    """One region's two right-hand sides in [Jx | Jy | u] layout (4 cells)."""
    nb = (k + 1) ** 2
    nc = 4
    xq, wq = gauss01(k + 2)
    out1 = [mp.mpf(0)] * (3 * nb * nc)
    out2 = [mp.mpf(0)] * (3 * nb * nc)
    ndg = dgq_nodes(k)
    semi = which == "semi"
    eps = P["eps_s"] if semi else P["eps_e"]
    e1, e2 = (P["ne"], P["pe"]) if semi else (P["rinf"], P["oinf"])
    pen = P["penalty"]

    def g(c, comp, a):
        return comp * nb * nc + c * nb + a

    for c, cell in enumerate(cells):
        h = cell.diameter()
        for qy in range(k + 2):
            for qx in range(k + 2):
                xi, eta = xq[qx], xq[qy]
                J = cell.jac(xi, eta)
                jxw = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * wq[qx] * wq[qy]
                x = cell.x(xi, eta)
                u1 = density(k, U1[c], xi, eta)
                u2 = density(k, U2[c], xi, eta)
                E = field(k - 1, D[c], cell, xi, eta)
                ph = dgq(k, xi, eta)
                src = (generation(P, x) + recombination(P, u1, u2)) if semi else mp.mpf(0)
                for a in range(nb):
                    # density tests: source + mass of the old density
                    out1[g(c, 2, a)] += jxw * ph[a] * (src + u1 / P["dt"])
                    out2[g(c, 2, a)] += jxw * ph[a] * (src + u2 / P["dt"])
                    # current tests p = (phi_a, 0), (0, phi_a): drift s_c p.D/eps u
                    for comp in range(2):
                        out1[g(c, comp, a)] += jxw * (-1) * ph[a] * E[comp] / eps * u1
                        out2[g(c, comp, a)] += jxw * (+1) * ph[a] * E[comp] / eps * u2
        for f in range(4):
            b = bid[c][f]
            if b < 0 or b == NEUMANN or (b == SCHOTTKY and not semi):
                continue
            for q in range(k + 2):
                ref, x, n, ln = cell.face(f, xq[q])
                ds = ln * wq[q]
                ph = dgq(k, *ref)
                if b == DIRICHLET:
                    for a in range(nb):
                        for comp in range(2):
                            out1[g(c, comp, a)] += ds * (-ph[a] * n[comp]) * e1
                            out2[g(c, comp, a)] += ds * (-ph[a] * n[comp]) * e2
                        out1[g(c, 2, a)] += ds * (pen / h) * ph[a] * e1
                        out2[g(c, 2, a)] += ds * (pen / h) * ph[a] * e2
                    continue
                own1 = density(k, U1[c], *ref)
                own2 = density(k, U2[c], *ref)
                if b == SCHOTTKY:
                    t1 = -P["v_rec_n"] * (own1 - P["ne"])
                    t2 = P["v_rec_p"] * (own2 - P["pe"])
                else:
                    oc = nbcell[c]
                    of = 0 if f == 1 else 1
                    oref = {0: (mp.mpf(0), xq[q]), 1: (mp.mpf(1), xq[q])}[of]
                    nb1 = density(k, Onb1[oc], *oref)
                    nb2 = density(k, Onb2[oc], *oref)
                    if semi:  # nb1 = r, nb2 = o
                        t1 = -P["k_et"] * (own1 - P["ne"]) * nb2
                        t2 = P["k_ht"] * (own2 - P["pe"]) * nb1
                    else:     # nb1 = n, nb2 = p; own1 = r, own2 = o
                        cur = -P["k_et"] * (nb1 - P["ne"]) * own2 + P["k_ht"] * (nb2 - P["pe"]) * own1
                        t1, t2 = cur, -cur
                for a in range(nb):
                    out1[g(c, 2, a)] += ds * ph[a] * t1
                    out2[g(c, 2, a)] += ds * ph[a] * t2
    del ndg
    return out1, out2


def poisson_rhs(P, k, scells, ecells, sbid, Un, Up, Ur, Uo):
    """Phi part per Poisson cell (DGQ(k-1) lexicographic) and the basis-free
    Raviart-Thomas boundary moments of the semiconductor cells."""
    xq, wq = gauss01(k + 2)
    deg = k - 1
    phi_part = []
    for reg, cells in (("semi", scells), ("elec", ecells)):
        for c, cell in enumerate(cells):
            loc = [mp.mpf(0)] * ((deg + 1) ** 2)
            for qy in range(k + 2):
                for qx in range(k + 2):
                    xi, eta = xq[qx], xq[qy]
                    J = cell.jac(xi, eta)
                    jxw = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * wq[qx] * wq[qy]
                    if reg == "semi":
                        rho = (P["ne"] - P["pe"]) + (-density(k, Un[c], xi, eta) + density(k, Up[c], xi, eta))
                    else:
                        rho = -density(k, Ur[c], xi, eta) + density(k, Uo[c], xi, eta)
                    ph = dgq(deg, xi, eta)
                    for j in range(len(loc)):
                        loc[j] += -jxw * ph[j] * rho
            phi_part.append(loc)
    moments = []
    for c, cell in enumerate(scells):
        for f in range(4):
            if sbid[c][f] != DIRICHLET:
                continue
            mom = [mp.mpf(0)] * (deg + 1)
            for q in range(k + 2):
                ref, x, n, ln = cell.face(f, xq[q])
                gval = (P["phi_bi"] if x[0] == 0 else 0) - (P["phi_app"] if x[0] == 0 else 0)
                for m in range(deg + 1):
                    mom[m] += -wq[q] * xq[q] ** m * gval
            moments.append([c, f, [float(v) for v in mom]])
    return phi_part, moments


CASES = {
    # unit square split at x = 1/2, all outer faces Dirichlet, Beer-Lambert
    # generation along x, SRH on, biased: every term but Schottky
    "rect_k1_config": dict(k=1, r1=0.5, r2=0.5, H=1.0, L=1.0, insulated=False, schottky=False,
                           gen_mode="config", srh=True, seed=11),
    # trapezoids (a conic wire), insulated + Schottky top, reference generation
    # along y, SRH on
    "trap_k2_insulated_schottky": dict(k=2, r1=0.5, r2=0.7, H=1.0, L=1.5, insulated=True, schottky=True,
                                       gen_mode="ref", srh=True, seed=12),
    # trapezoids, all Dirichlet (bias on the left edge), reference generation, no SRH
    "trap_k2_dirichlet": dict(k=2, r1=0.5, r2=0.62, H=1.0, L=1.25, insulated=False, schottky=False,
                              gen_mode="ref", srh=False, seed=13),
}

PHYS = dict(dt=0.01, penalty=1.5, eps_s=1.7, eps_e=0.8, ne=1.3, pe=0.2, rinf=2.5, oinf=1.7,
            k_et=0.3, k_ht=0.45, v_rec_n=0.6, v_rec_p=0.9, G0=2.0, alpha=1.3, ni=0.1, tau_n=0.8,
            tau_p=1.1, phi_bi=0.7, phi_app=0.25)


def make_case(name, c):
    k = c["k"]
    geo = dict(r1=c["r1"], r2=c["r2"], H=c["H"], L=c["L"])
    P = {key: (mp.mpf(v) if isinstance(v, float) else v) for key, v in PHYS.items()}
    P.update(H=mp.mpf(c["H"]), gen_mode=c["gen_mode"], srh=c["srh"])
    Vs = region_vertices([(0.0, 0.0), (c["r2"], 0.0), (0.0, c["H"]), (c["r1"], c["H"])])
    Ve = region_vertices([(c["r2"], 0.0), (c["L"], 0.0), (c["r1"], c["H"]), (c["L"], c["H"])])
    sb = bids(Vs, geo, c["insulated"], c["schottky"])
    eb = bids(Ve, geo, c["insulated"], False)
    rng = np.random.default_rng(c["seed"])
    nb = (k + 1) ** 2
    # positive densities keep the SRH denominator away from zero
    st = {key: (base + 0.3 * rng.random((4, nb))) for key, base in (("n", 1.1), ("p", 0.4), ("r", 2.2), ("o", 1.5))}
    deg = k - 1
    Dc = {}
    for reg in ("semi", "elec"):
        Dc[reg] = [(rng.uniform(-1, 1, (deg + 1, deg + 2)), rng.uniform(-1, 1, (deg + 2, deg + 1))) for _ in range(4)]
    mpc = lambda a: [[mp.mpf(float(v)) for v in row] for row in a]
    U = {key: mpc(v) for key, v in st.items()}
    Dm = {reg: [(mpc(ax), mpc(ay)) for ax, ay in v] for reg, v in Dc.items()}
    sc = [Cell(V) for V in Vs]
    ec = [Cell(V) for V in Ve]
    # interface pairing: semiconductor (1, j) face 1 <-> electrolyte (0, j) face 0
    s2e = {1: 0, 3: 2}
    e2s = {0: 1, 2: 3}
    bn, bp = carrier_rhs(P, k, sc, sb, U["n"], U["p"], Dm["semi"], U["r"], U["o"], s2e, "semi")
    br, bo = carrier_rhs(P, k, ec, eb, U["r"], U["o"], Dm["elec"], U["n"], U["p"], e2s, "elec")
    phi_part, moments = poisson_rhs(P, k, sc, ec, sb, U["n"], U["p"], U["r"], U["o"])
    fl = lambda v: [float(x) for x in v]
    out = {
        "case": dict(c, **{key: float(v) if isinstance(v, float) else v for key, v in PHYS.items()}),
        "semi_vertices": Vs.tolist(), "elec_vertices": Ve.tolist(), "semi_bid": sb, "elec_bid": eb,
        "density": {key: v.tolist() for key, v in st.items()},
        "field": {reg: [[ax.tolist(), ay.tolist()] for ax, ay in v] for reg, v in Dc.items()},
        "rhs_n": fl(bn), "rhs_p": fl(bp), "rhs_r": fl(br), "rhs_o": fl(bo),
        "poisson_phi": [fl(v) for v in phi_part], "poisson_moments": moments,
    }
    with open(os.path.join(HERE, f"rhs_{name}.json"), "w") as f:
        json.dump(out, f)
    print("rhs", name, len(bn), flush=True)


def main():
    for name, c in CASES.items():
        make_case(name, c)


if __name__ == "__main__":
    sys.exit(main())
