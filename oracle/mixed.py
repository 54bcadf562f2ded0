"""Mixed RT Poisson matrix and right-hand side (oracle, test infrastructure).

Restates:
  * MixedFEM::assemble_local_Poisson_matrix (MixedFEM.cpp:29-164):
      sum_q JxW [ eps_K^{-1} psi_i.psi_j - div(psi_i) phi_j
                  - lambda^2 phi_i div(psi_j) ]
    with eps_K = eps_S on semiconductor cells, eps_E on electrolyte cells;
  * PoissonData::setup_dofs constraints (Poisson.cpp:32-49): the D.n DoFs of
    Neumann faces are fixed to zero (here: eliminated);
  * assemble_Poisson_rhs and the two local workers (SolarCell.cpp:430-815):
      semiconductor cell  -sum_q JxW phi_i [(N_D - N_A) + (s_n n + s_p p)]
      Dirichlet face      -sum_q JxW (psi_i.n) (Phi_bi(x) - Phi_app(x))
      Schottky face       -sum_q JxW (psi_i.n) (Phi_bi(x) - Phi_sch(x))
      electrolyte cell    -sum_q JxW phi_i (s_r r + s_o o)
      electrolyte Dirichlet face: datum Phi_inf = 0 (BiasValues.cpp:107-110).
"""
import numpy as np
import scipy.sparse as sp

from . import fe
from .mesh import DIRICHLET, NEUMANN, SCHOTTKY, SEMI_BL_ID, SEMI_ID, poisson_dofs, poisson_dofs_general


class PoissonFE:
    def __init__(self, k, mesh):
        self.k = k
        self.deg = k - 1
        self.rt = fe.RT(self.deg)
        self.dg = fe.DGQ(self.deg)
        self.nrt = self.rt.n
        self.nphi = self.dg.nb
        self.nloc = self.nrt + self.nphi
        self.nq1 = k + 2
        self.qpts, self.qw = fe.cell_quadrature(self.nq1)
        self.psi_q, self.div_q = self.rt.tab(self.qpts)       # [nrt,nq,2], [nrt,nq]
        self.phi_q, _ = self.dg.tab(self.qpts)                 # [nphi,nq]
        self.fs, self.fw = fe.gauss01(self.nq1)
        self.psi_f = {f: self.rt.tab(fe.face_points(f, self.fs))[0] for f in range(4)}
        if getattr(mesh, "local", 0):
            self.idx, self.nD, self.n, self.fb = poisson_dofs_general(mesh, self.deg)
        else:
            self.idx, self.nD, self.n = poisson_dofs(mesh, self.deg)

    def cell_parts(self, V):
        x, J = fe.bilinear(V, self.qpts)
        det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        JxW = det * self.qw[None, :]
        psi = np.einsum("cqrd,jqd->jcqr", J, self.psi_q) / det[None, :, :, None]
        div = self.div_q[:, None, :] / det[None]
        return x, JxW, psi, div


def constrained_dofs(mesh, pfe):
    """D.n DoFs on Neumann faces of the merged mesh (Poisson.cpp:41-47)."""
    cons = set()
    nx = pfe.rt.nx
    d = pfe.deg
    for p in range(len(mesh.pV)):
        for f in range(4):
            if mesh.pbid[p, f] != NEUMANN:
                continue
            if f == 0:
                li = [0 + (d + 2) * b for b in range(d + 1)]
            elif f == 1:
                li = [d + 1 + (d + 2) * b for b in range(d + 1)]
            elif f == 2:
                li = [nx + a for a in range(d + 1)]
            else:
                li = [nx + a + (d + 1) * (d + 1) for a in range(d + 1)]
            for l in li:
                cons.add(int(pfe.idx[p, l]))
    return np.array(sorted(cons), dtype=np.int64)


def hanging_constraints(mesh, pfe):
    """D DoFs of the fine faces of hanging faces as combinations of the coarse
    face's DoFs (DoFTools::make_hanging_node_constraints, Poisson.cpp:33-34).

    The constrained RT space has a single-valued normal flux: on subface s of
    a coarse face, the fine face's flux is the coarse face's flux restricted
    there.  Face DoFs here are reference normal components at the face's Gauss
    nodes; the Piola map gives flux density = reference component / face
    length, and a fine face is half as long, so fine DoF m = 1/2 * the coarse
    reference component at the subface point (s + G_m) / 2, i.e.
    sum_j G_j((s + G_m)/2) / 2 * coarse DoF j (G_j: Lagrange basis on the
    Gauss nodes).  Both sides' reference normal axes point the same physical
    way (faces 1 | 0 across vertical faces), so no sign.  Returns a list of
    (slave, masters[K], coefs[K]).
    """
    out = []
    if not getattr(mesh, "local", 0):
        return out
    G = pfe.rt.Gn
    K = len(G)
    for c, f, s, e, fe_ in mesh.phang:
        for m in range(K):
            Lv, _ = fe.lagrange1d(G, [(s + G[m]) / 2.0])
            out.append((int(pfe.fb[e, fe_] + m), pfe.fb[c, f] + np.arange(K), 0.5 * Lv[:, 0]))
    return out


def reduction(mesh, pfe):
    """Sparse T (n x nfree) with x = T y: Neumann D.n DoFs are zero and
    hanging fine-face DoFs follow their coarse face; y runs over the others."""
    cons = set(constrained_dofs(mesh, pfe).tolist())
    hang = hanging_constraints(mesh, pfe)
    slaves = {h[0] for h in hang}
    free = np.array([i for i in range(pfe.n) if i not in cons and i not in slaves], dtype=np.int64)
    col = -np.ones(pfe.n, dtype=np.int64)
    col[free] = np.arange(len(free))
    R, C, V = list(free), list(range(len(free))), [1.0] * len(free)
    for sl, ms, cs in hang:
        for mm, cc in zip(ms, cs):
            assert col[mm] >= 0
            R.append(sl)
            C.append(col[mm])
            V.append(cc)
    return sp.csr_matrix((V, (R, C)), shape=(pfe.n, len(free)))


def assemble_mixed(mesh, pfe, eps_s, eps_e, lam2):
    V = mesh.pV
    nc = len(V)
    x, JxW, psi, div = pfe.cell_parts(V)
    nrt, nphi, nl = pfe.nrt, pfe.nphi, pfe.nloc
    eps = np.where((mesh.pmat == SEMI_ID) | (mesh.pmat == SEMI_BL_ID), eps_s, eps_e)
    K = np.zeros((nc, nl, nl))
    K[:, :nrt, :nrt] = np.einsum("icqd,jcqd,cq->cij", psi, psi, JxW) / eps[:, None, None]
    B = np.einsum("icq,jq,cq->cji", div, pfe.phi_q, JxW)  # B[c, phi_j, psi_i] = (phi_j, div psi_i)
    K[:, :nrt, nrt:] = -np.transpose(B, (0, 2, 1))
    K[:, nrt:, :nrt] = -lam2 * B
    rows = np.repeat(pfe.idx[:, :, None], nl, 2)
    cols = np.repeat(pfe.idx[:, None, :], nl, 1)
    A = sp.coo_matrix((K.ravel(), (rows.ravel(), cols.ravel())), shape=(pfe.n, pfe.n)).tocsr()
    return A


def poisson_rhs(mesh, pfe, cfe, u_n, u_p, u_r, u_o, params, with_boundary=True):
    """Assembled Poisson RHS (SolarCell.cpp:430-815)."""
    from .ldg import density_at
    b = np.zeros(pfe.n)
    nps = mesh.semi.ncell
    V = mesh.pV
    x, JxW, psi, div = pfe.cell_parts(V)
    nb = cfe.nb
    ncs = mesh.semi.ncell
    nce = mesh.elec.ncell
    # carrier densities at the (shared) cell quadrature points
    n_q = density_at(cfe, u_n, ncs, cfe.phi)
    p_q = density_at(cfe, u_p, ncs, cfe.phi)
    r_q = density_at(cfe, u_r, nce, cfe.phi)
    o_q = density_at(cfe, u_o, nce, cfe.phi)
    loc = np.zeros((len(V), pfe.nloc))
    rho_s = (params["ne"] - params["pe"]) + (-1.0 * n_q + 1.0 * p_q)
    rho_e = (-1.0 * r_q + 1.0 * o_q)
    loc[:nps, pfe.nrt:] = -np.einsum("jq,cq->cj", pfe.phi_q, rho_s * JxW[:nps])
    loc[nps:, pfe.nrt:] = -np.einsum("jq,cq->cj", pfe.phi_q, rho_e * JxW[nps:])
    if with_boundary:
        for f in range(4):
            for kind in (DIRICHLET, SCHOTTKY):
                cells = np.nonzero((mesh.pbid[:nps, f] == kind))[0]
                if len(cells) == 0:
                    continue
                pts = fe.face_points(f, pfe.fs)
                xf, Jf = fe.bilinear(V[cells], pts)
                detf = Jf[..., 0, 0] * Jf[..., 1, 1] - Jf[..., 0, 1] * Jf[..., 1, 0]
                _, nrm, ln = fe.face_geometry(V[cells], f, pfe.fs)
                jw = pfe.fw[None, :] * ln
                psif = np.einsum("cqrd,jqd->jcqr", Jf, pfe.psi_f[f]) / detf[None, :, :, None]
                psin = np.einsum("jcqr,cqr->cjq", psif, nrm)
                bi = np.where(xf[..., 0] == 0.0, params["phi_bi"], 0.0)
                if kind == DIRICHLET:
                    other = np.where(xf[..., 0] == 0.0, params["phi_app"], 0.0)
                else:
                    other = np.where(xf[..., 1] == params.get("schottky_location", np.nan),
                                     params["phi_sch"], 0.0)
                loc[cells, :pfe.nrt] += -np.einsum("cjq,cq->cj", psin, (bi - other) * jw)
        # electrolyte Dirichlet datum is identically zero (Bulk_Bias)
    np.add.at(b, pfe.idx, loc)
    return b
