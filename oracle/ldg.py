"""LDG matrices and carrier right-hand sides (oracle, test infrastructure).

Restates, cell by cell and face by face, the reference's LDG forms:
  * assemble_local_LDG_mass_matrix      LDG.cpp:40-83
  * assemble_local_LDG_cell_and_bc_terms LDG.cpp:85-281
  * assemble_flux_terms / assemble_local_flux_terms /
    distribute_local_fluxes_to_global   LDG.cpp:283-678 (interior faces are
    assembled once, by the lower-index cell, with its outward normal and
    beta = (1,1)/sqrt(2) NOT multiplied by the normal)
  * assemble_semiconductor_rhs + assemble_local_semiconductor_rhs
                                        SolarCell.cpp:1037-1414
  * assemble_electrolyte_rhs + assemble_local_electrolyte_rhs
                                        SolarCell.cpp:1417-1726
  * mass_matrix.vmult of the old state  SolarCell.cpp:1043-1047, 1423-1427
"""
import numpy as np
import scipy.sparse as sp

from . import fe
from .mesh import DIRICHLET, INTERFACE, NEUMANN, SCHOTTKY, carrier_dofs

OPP = {0: 1, 1: 0, 2: 3, 3: 2}
BETA = np.array([1.0, 1.0]) / np.sqrt(2.0)


class CarrierFE:
    """Tabulations of FESystem(FE_DGQ(k)^2, FE_DGQ(k)) on QGauss(k+2)."""

    def __init__(self, k):
        self.k = k
        self.dgq = fe.DGQ(k)
        self.nb = self.dgq.nb
        self.nloc = 3 * self.nb
        self.nq1 = k + 2
        self.qpts, self.qw = fe.cell_quadrature(self.nq1)
        self.phi, self.gphi = self.dgq.tab(self.qpts)  # [nb,nq], [nb,nq,2]
        self.fs, self.fw = fe.gauss01(self.nq1)
        self.fphi = {f: self.dgq.tab(fe.face_points(f, self.fs))[0] for f in range(4)}

    def cell_geometry(self, V):
        x, J = fe.bilinear(V, self.qpts)
        det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        Jinv = np.linalg.inv(J)
        JxW = det * self.qw[None, :]
        return x, J, det, Jinv, JxW

    def local_parts(self, Jinv):
        """P[i,c,q,2], DIVP[i,c,q], VV[i,q], GV[i,c,q,2] for the 3nb local DoFs."""
        nb = self.nb
        nc, nq = Jinv.shape[:2]
        # physical gradient: grad_x phi = J^{-T} grad_xi phi
        g = np.einsum("aqr,cqrd->acqd", self.gphi, Jinv)  # [nb,nc,nq,2]
        P = np.zeros((3 * nb, nc, nq, 2))
        DIVP = np.zeros((3 * nb, nc, nq))
        VV = np.zeros((3 * nb, nq))
        GV = np.zeros((3 * nb, nc, nq, 2))
        P[:nb, :, :, 0] = self.phi[:, None, :]
        P[nb:2 * nb, :, :, 1] = self.phi[:, None, :]
        DIVP[:nb] = g[..., 0]
        DIVP[nb:2 * nb] = g[..., 1]
        VV[2 * nb:] = self.phi
        GV[2 * nb:] = g
        return P, DIVP, VV, GV

    def face_parts(self, f, ph=None):
        """Face tabulation: PF[i,q,2] (current), VF[i,q] (density); ph
        replaces the basis values at the face points (subface points)."""
        nb = self.nb
        ph = self.fphi[f] if ph is None else ph
        nq = ph.shape[1]
        PF = np.zeros((3 * nb, nq, 2))
        VF = np.zeros((3 * nb, nq))
        PF[:nb, :, 0] = ph
        PF[nb:2 * nb, :, 1] = ph
        VF[2 * nb:] = ph
        return PF, VF


def _scatter(rows, cols, vals, n):
    return sp.coo_matrix((vals.ravel(), (rows.ravel(), cols.ravel())), shape=(n, n)).tocsr()


def _flux_blocks(PFm, VFm, Pm_n, PFp, VFp, Pp_n, sig, jw):
    """Local flux matrices (mm, mp, pm, pp) of one interior face
    (assemble_local_flux_terms, LDG.cpp:437-626): minus side m assembles with
    its outward normal n (Pm_n, Pp_n = p.n of each side's current basis)."""
    bPm = PFm @ BETA
    bPp = PFp @ BETA

    def blk(Pa_n, Va, bPa, Pb_n, Vb, bPb, s_half, s_beta, s_pen):
        # s_half * 1/2 (p_i.n u_j + v_i n.q_j) + s_beta*((b.p_i) u_j - v_i (b.q_j))
        # + s_pen * sigma v_i u_j
        t = 0.5 * s_half * (np.einsum("ciq,jq->cijq", Pa_n, Vb) + np.einsum("iq,cjq->cijq", Va, Pb_n))
        t = t + s_beta * (np.einsum("iq,jq->ijq", bPa, Vb)[None] - np.einsum("iq,jq->ijq", Va, bPb)[None])
        t = t + s_pen * np.einsum("iq,jq->ijq", Va, Vb)[None] * sig[:, None, None, :]
        return np.einsum("cijq,cq->cij", t, jw)

    return (blk(Pm_n, VFm, bPm, Pm_n, VFm, bPm, 1.0, 1.0, 1.0),
            blk(Pm_n, VFm, bPm, Pp_n, VFp, bPp, 1.0, -1.0, -1.0),
            blk(Pp_n, VFp, bPp, Pm_n, VFm, bPm, -1.0, -1.0, -1.0),
            blk(Pp_n, VFp, bPp, Pp_n, VFp, bPp, -1.0, 1.0, 1.0))


def assemble_ldg(region, cfe, mu1, mu2, dt, theta, penalty, flip_orientation=False):
    """Constant LDG system matrices A1, A2 and mass matrix M of one carrier pair.

    flip_orientation=True assembles each interior face from the HIGHER-index
    cell instead (used only to check that the reference's lower-index rule,
    LDG.cpp:390-391, does not change the matrix).
    """
    nb, nl = cfe.nb, cfe.nloc
    V = region.V
    nc = region.ncell
    n = 3 * nb * nc
    dofs = carrier_dofs(nc, nb)
    x, J, det, Jinv, JxW = cfe.cell_geometry(V)
    P, DIVP, VV, GV = cfe.local_parts(Jinv)
    h = fe.diameter(V)
    common = (theta / dt) * np.einsum("iq,jq,cq->cij", VV, VV, JxW) \
        - np.einsum("icq,jq,cq->cij", DIVP, VV, JxW) \
        - np.einsum("icqd,jcqd,cq->cij", GV, P, JxW)
    PP = np.einsum("icqd,jcqd,cq->cij", P, P, JxW)
    K1 = common + PP / mu1
    K2 = common + PP / mu2
    Mloc = (1.0 / dt) * np.einsum("iq,jq,cq->cij", VV, VV, JxW)
    # boundary faces
    for f in range(4):
        PF, VF = cfe.face_parts(f)
        bid = region.bid[:, f]
        for kind in (DIRICHLET, "flux"):
            if kind == DIRICHLET:
                cells = np.nonzero(bid == DIRICHLET)[0]
            else:
                cells = np.nonzero((bid == INTERFACE) | (bid == NEUMANN) | (bid == SCHOTTKY))[0]
            if len(cells) == 0:
                continue
            _, nrm, ln = fe.face_geometry(V[cells], f, cfe.fs)
            jw = cfe.fw[None, :] * ln
            if kind == DIRICHLET:
                npj = np.einsum("cqd,jqd->cjq", nrm, PF)
                loc = np.einsum("iq,cjq,cq->cij", VF, npj, jw) \
                    + np.einsum("iq,jq,cq->cij", VF, VF, jw * (penalty / h[cells])[:, None])
            else:
                pin = np.einsum("iqd,cqd->ciq", PF, nrm)
                loc = np.einsum("ciq,jq,cq->cij", pin, VF, jw)
            K1[cells] += loc
            K2[cells] += loc
    rows = [np.repeat(dofs[:, :, None], nl, 2)]
    cols = [np.repeat(dofs[:, None, :], nl, 1)]
    vals1 = [K1]
    vals2 = [K2]
    # interior faces: lower index cell assembles (LDG.cpp:385-425)
    for f in range(4):
        nbr = region.nbr[:, f]
        if flip_orientation:
            cells = np.nonzero((nbr >= 0) & (nbr < np.arange(nc)))[0]
        else:
            cells = np.nonzero(nbr > np.arange(nc))[0]
        if len(cells) == 0:
            continue
        e = nbr[cells]
        fo = OPP[f]
        PFm, VFm = cfe.face_parts(f)
        PFp, VFp = cfe.face_parts(fo)
        _, nrm, ln = fe.face_geometry(V[cells], f, cfe.fs)
        jw = cfe.fw[None, :] * ln
        hmin = np.minimum(h[cells], h[e])
        sig = (penalty / hmin)[:, None]
        Pm_n = np.einsum("iqd,cqd->ciq", PFm, nrm)
        Pp_n = np.einsum("iqd,cqd->ciq", PFp, nrm)
        dm, dp = dofs[cells], dofs[e]
        for B, r_, c_ in zip(_flux_blocks(PFm, VFm, Pm_n, PFp, VFp, Pp_n, sig, jw), (dm, dm, dp, dp), (dm, dp, dm, dp)):
            rows.append(np.repeat(r_[:, :, None], nl, 2))
            cols.append(np.repeat(c_[:, None, :], nl, 1))
            vals1.append(B)
            vals2.append(B)
    # hanging faces (LDG.cpp:338-383): the coarse cell assembles each subface
    # with its own outward normal, its basis at the subface's quadrature
    # points and the fine cell's basis at its face's points (the same points),
    # h = min(coarse diameter, diameter of the fine cell's parent)
    hang = getattr(region, "hang", None)
    for f in range(4):
        for s in range(2):
            if hang is None or len(hang) == 0:
                break
            sel = hang[(hang[:, 1] == f) & (hang[:, 2] == s)]
            if len(sel) == 0:
                continue
            cells, e, fo = sel[:, 0], sel[:, 3], int(sel[0, 4])
            sub = (s + cfe.fs) / 2.0
            PFm, VFm = cfe.face_parts(f, cfe.dgq.tab(fe.face_points(f, sub))[0])
            PFp, VFp = cfe.face_parts(fo)
            _, nrm, ln = fe.face_geometry(V[cells], f, sub)
            jw = cfe.fw[None, :] * ln / 2.0
            sig = (penalty / np.minimum(h[cells], region.pdiam[e]))[:, None]
            Pm_n = np.einsum("iqd,cqd->ciq", PFm, nrm)
            Pp_n = np.einsum("iqd,cqd->ciq", PFp, nrm)
            dm, dp = dofs[cells], dofs[e]
            for B, r_, c_ in zip(_flux_blocks(PFm, VFm, Pm_n, PFp, VFp, Pp_n, sig, jw), (dm, dm, dp, dp),
                                 (dm, dp, dm, dp)):
                rows.append(np.repeat(r_[:, :, None], nl, 2))
                cols.append(np.repeat(c_[:, None, :], nl, 1))
                vals1.append(B)
                vals2.append(B)
    R = np.concatenate([r.ravel() for r in rows])
    C = np.concatenate([c.ravel() for c in cols])
    A1 = sp.coo_matrix((np.concatenate([v.ravel() for v in vals1]), (R, C)), shape=(n, n)).tocsr()
    A2 = sp.coo_matrix((np.concatenate([v.ravel() for v in vals2]), (R, C)), shape=(n, n)).tocsr()
    Mr = np.repeat(dofs[:, :, None], nl, 2)
    Mc = np.repeat(dofs[:, None, :], nl, 1)
    M = _scatter(Mr, Mc, Mloc, n)
    return A1, A2, M


def density_at(cfe, u, nc, pts_tab):
    """Density of carrier vector u at tabulated points: pts_tab[nb, np] -> [nc, np]."""
    nb = cfe.nb
    ud = u[2 * nc * nb:].reshape(nc, nb)
    return ud @ pts_tab


def field_at_carrier_q(cfe, pfe, Vcells, pcoef):
    """Poisson vector field D at the carrier cell quadrature points.

    pcoef[nc, nloc_P] are the Poisson cell coefficients of the matching
    Poisson cells (same geometry, SolarCell.cpp:1102-1140).  Piola:
    D = (1/detJ) J sum_j d_j psi_hat_j.
    """
    x, J = fe.bilinear(Vcells, cfe.qpts)
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    nrt = pfe.rt.n
    Dhat = np.einsum("cj,jqd->cqd", pcoef[:, :nrt], pfe.psi_q)
    return np.einsum("cqrd,cqd->cqr", J, Dhat) / det[..., None]


def carrier_rhs(region, cfe, pfe, mesh, M, u1, u2, pcoef, params, which,
                other_u1=None, other_u2=None, other_region=None):
    """Right-hand sides of one carrier pair (SolarCell.cpp:1037-1726).

    which = 'semi' (u1=n, u2=p; other = (r, o)) or 'elec' (u1=r, u2=o;
    other = (n, p)).  Returns (rhs1, rhs2) including M u_old.
    """
    nb, nl = cfe.nb, cfe.nloc
    nc = region.ncell
    dofs = carrier_dofs(nc, nb)
    rhs1 = M @ u1
    rhs2 = M @ u2
    V = region.V
    x, J, det, Jinv, JxW = cfe.cell_geometry(V)
    P, DIVP, VV, GV = cfe.local_parts(Jinv)
    h = fe.diameter(V)
    d1 = density_at(cfe, u1, nc, cfe.phi)
    d2 = density_at(cfe, u2, nc, cfe.phi)
    E = field_at_carrier_q(cfe, pfe, V, pcoef)
    pen = params["penalty"]
    if which == "semi":
        s1, s2 = -1.0, 1.0
        inv_perm = 1.0 / params["eps_s"]
        G = params["generation"](x)
        R = params["recombination"](d1, d2)
        src = G + R
        loc1 = np.einsum("iq,cq->ci", VV, src * JxW) + s1 * inv_perm * np.einsum("icqd,cqd,cq->ci", P, E, d1 * JxW)
        loc2 = np.einsum("iq,cq->ci", VV, src * JxW) + s2 * inv_perm * np.einsum("icqd,cqd,cq->ci", P, E, d2 * JxW)
        e1, e2 = params["ne"], params["pe"]
    else:
        s1, s2 = -1.0, 1.0
        inv_perm = 1.0 / params["eps_e"]
        loc1 = s1 * inv_perm * np.einsum("icqd,cqd,cq->ci", P, E, d1 * JxW)
        loc2 = s2 * inv_perm * np.einsum("icqd,cqd,cq->ci", P, E, d2 * JxW)
        e1, e2 = params["rinf"], params["oinf"]
    for f in range(4):
        PF, VF = cfe.face_parts(f)
        bid = region.bid[:, f]
        cells = np.nonzero(bid == DIRICHLET)[0]
        if len(cells):
            _, nrm, ln = fe.face_geometry(V[cells], f, cfe.fs)
            jw = cfe.fw[None, :] * ln
            pin = np.einsum("iqd,cqd->ciq", PF, nrm)
            t = (-pin + (pen / h[cells])[:, None, None] * VF[None]) * jw[:, None, :]
            loc1[cells] += e1 * t.sum(-1)
            loc2[cells] += e2 * t.sum(-1)
        cells = np.nonzero(bid == INTERFACE)[0]
        if len(cells):
            _, nrm, ln = fe.face_geometry(V[cells], f, cfe.fs)
            jw = cfe.fw[None, :] * ln
            fo = OPP[f]
            own1 = density_at(cfe, u1, nc, cfe.fphi[f])[cells]
            own2 = density_at(cfe, u2, nc, cfe.fphi[f])[cells]
            if which == "semi":
                nbc = np.array([mesh.semi_sigma_nb[c] for c in cells])
                onc = other_region.ncell
                r_nb = density_at(cfe, other_u1, onc, cfe.fphi[fo])[nbc]
                o_nb = density_at(cfe, other_u2, onc, cfe.fphi[fo])[nbc]
                t1 = -params["k_et"] * (own1 - params["ne"]) * o_nb * jw
                t2 = params["k_ht"] * (own2 - params["pe"]) * r_nb * jw
                loc1[cells] += np.einsum("iq,cq->ci", VF, t1)
                loc2[cells] += np.einsum("iq,cq->ci", VF, t2)
            else:
                nbc = np.array([mesh.elec_sigma_nb[c] for c in cells])
                onc = other_region.ncell
                n_nb = density_at(cfe, other_u1, onc, cfe.fphi[fo])[nbc]
                p_nb = density_at(cfe, other_u2, onc, cfe.fphi[fo])[nbc]
                cur = -params["k_et"] * (n_nb - params["ne"]) * own2 \
                    + params["k_ht"] * (p_nb - params["pe"]) * own1
                I = np.einsum("iq,cq->ci", VF, cur * jw)
                loc1[cells] += I
                loc2[cells] -= I
        cells = np.nonzero(bid == SCHOTTKY)[0]
        if len(cells) and which == "semi":
            _, nrm, ln = fe.face_geometry(V[cells], f, cfe.fs)
            jw = cfe.fw[None, :] * ln
            own1 = density_at(cfe, u1, nc, cfe.fphi[f])[cells]
            own2 = density_at(cfe, u2, nc, cfe.fphi[f])[cells]
            loc1[cells] += np.einsum("iq,cq->ci", VF, -params["v_rec_n"] * (own1 - params["ne"]) * jw)
            loc2[cells] += np.einsum("iq,cq->ci", VF, params["v_rec_p"] * (own2 - params["pe"]) * jw)
    np.add.at(rhs1, dofs, loc1)
    np.add.at(rhs2, dofs, loc2)
    return rhs1, rhs2
