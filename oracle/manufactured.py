"""The reference's four manufactured-solution studies (oracle, test infrastructure).

Restates the test drivers and closed forms:
  * test_steady_state   SolarCell.cpp:2749-2860, test_functions.cpp:13-93,
                        LDG.cpp:681-812 (assemble_local_test_rhs),
                        MixedFEM.cpp:166-254, grid Grid.cpp:464-503
  * test_transient      SolarCell.cpp:2865-2965, test_functions.cpp:99-220,
                        LDG.cpp:814-982, grid Grid.cpp:531-568
  * test_DD_Poisson     SolarCell.cpp:2972-3117, test_functions.cpp:227-361,
                        SolarCell.cpp:2105-2260, grid Grid.cpp:570-598
  * test_interface_coupling SolarCell.cpp:3123-3278,
                        test_functions.cpp:369-527, SolarCell.cpp:2262-2745
The reference prints L2 errors and rates and asserts nothing; these return
the errors so tests can assert the textbook rates (LDG u: k+1, mixed Phi and
D: k).  The transient studies project 1 + cos(2 pi x) cos(2 pi y) as initial
data (test_functions.cpp:429-448), which is not the exact u(0); `ic='exact'`
uses the exact initial value instead, `ic='reference'` reproduces the quirk.

geometry='unit' is the reference's grids (unit squares with the boundary
ids of Grid.cpp:464-598).  geometry='cell' poses the same exact solutions on
the solar-cell grid the library's GPU path is built for (r1 = r2 = 0.5,
H = L = 1, outer faces Dirichlet): the carriers live on the semiconductor
half [0, 0.5] x [0, 1], whose face x = 0.5 (the interface) takes the exact
normal flux J.n as a natural boundary condition, and the mixed Poisson
problem covers the merged unit square with Dirichlet data on all of its
boundary.  The interface study is on that grid in the reference already.
"""
import numpy as np
import scipy.sparse.linalg as spla

from . import fe
from .ldg import CarrierFE, assemble_ldg, density_at
from .mesh import DIRICHLET, INTERFACE, NEUMANN, SEMI_ID, ELEC_ID, Region, carrier_dofs
from .mixed import PoissonFE, assemble_mixed, constrained_dofs

TP = 2.0 * np.pi


class _PMesh:
    """A single region seen as a Poisson mesh (NX = NY = N)."""

    def __init__(self, reg):
        self.pV = reg.V
        self.pmat = np.full(reg.ncell, SEMI_ID)
        self.pIJ = reg.ij
        self.NX = self.NY = reg.N
        self.pbid = reg.bid


def unit_square(r, classify):
    return Region([(0, 0), (1, 0), (0, 1), (1, 1)], r, SEMI_ID, None, False, False, classify)


def test_grid_classify(cx, cy):  # Grid.cpp:464-503
    return NEUMANN if (cy == 0.0 or cy == 1.0) else DIRICHLET


def tran_grid_classify(cx, cy):  # Grid.cpp:531-568
    bid = INTERFACE
    if cx != 1.0:
        bid = DIRICHLET
    if cy == 0.0 or cy == 1.0:
        bid = NEUMANN
    return bid


def dd_grid_classify(cx, cy):  # Grid.cpp:570-598
    return DIRICHLET


CELL_GEO = dict(r1=0.5, r2=0.5, H=1.0, L=1.0)


def study_mesh(r, geometry, classify):
    """(carrier region, Poisson mesh) of a study on `geometry`."""
    if geometry == "unit":
        reg = unit_square(r, classify)
        return reg, _PMesh(reg)
    if geometry == "cell":
        from .mesh import SolarCellMesh
        m = SolarCellMesh(r, CELL_GEO)
        return m.semi, m
    raise ValueError(geometry)


def _flux_rhs(reg, cfe, loc, h):
    """-<v, h(x, n)> on interface faces: a prescribed normal flux J.n."""
    for f in range(4):
        cells = np.nonzero(reg.bid[:, f] == INTERFACE)[0]
        if len(cells) == 0:
            continue
        _, VF = cfe.face_parts(f)
        xf, nrm, ln = fe.face_geometry(reg.V[cells], f, cfe.fs)
        loc[cells] += np.einsum("iq,cq->ci", VF, -h(xf, nrm) * cfe.fw[None] * ln)


def _cell_rhs(reg, cfe, f_vals=None, p_field=None, u_old=None):
    """sum_q JxW [v_i f - p_i.F u_old] per cell (LDG.cpp:876-896 pattern)."""
    x, J, det, Jinv, JxW = cfe.cell_geometry(reg.V)
    P, DIVP, VV, GV = cfe.local_parts(Jinv)
    loc = np.zeros((reg.ncell, cfe.nloc))
    if f_vals is not None:
        loc += np.einsum("iq,cq->ci", VV, f_vals(x) * JxW)
    if p_field is not None:
        ud = density_at(cfe, u_old, reg.ncell, cfe.phi)
        loc -= np.einsum("icqd,cqd,cq->ci", P, p_field, ud * JxW)
    return loc, x


def _dirichlet_rhs(reg, cfe, loc, g, penalty=1.0):
    h = fe.diameter(reg.V)
    for f in range(4):
        cells = np.nonzero(reg.bid[:, f] == DIRICHLET)[0]
        if len(cells) == 0:
            continue
        PF, VF = cfe.face_parts(f)
        xf, nrm, ln = fe.face_geometry(reg.V[cells], f, cfe.fs)
        jw = cfe.fw[None, :] * ln
        pin = np.einsum("iqd,cqd->ciq", PF, nrm)
        t = (-pin + (penalty / h[cells])[:, None, None] * VF[None]) * (g(xf) * jw)[:, None, :]
        loc[cells] += t.sum(-1)


def _assemble_vec(reg, cfe, loc):
    b = np.zeros(3 * cfe.nb * reg.ncell)
    np.add.at(b, carrier_dofs(reg.ncell, cfe.nb), loc)
    return b


def ldg_errors(reg, cfe, u, exact_u, exact_J, nq=None):
    """L2 errors of density and current of carrier vector u."""
    nq = nq or (cfe.k + 4)
    pts, w = fe.cell_quadrature(nq)
    x, J = fe.bilinear(reg.V, pts)
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    jw = det * w[None]
    ph, _ = cfe.dgq.tab(pts)
    nb, nc = cfe.nb, reg.ncell
    jx = u[:nc * nb].reshape(nc, nb) @ ph
    jy = u[nc * nb:2 * nc * nb].reshape(nc, nb) @ ph
    ud = u[2 * nc * nb:].reshape(nc, nb) @ ph
    eu = np.sqrt(np.sum((ud - exact_u(x)) ** 2 * jw))
    EJ = exact_J(x)
    eJ = np.sqrt(np.sum(((jx - EJ[..., 0]) ** 2 + (jy - EJ[..., 1]) ** 2) * jw))
    return eu, eJ


def mixed_errors(pm, pfe, sol, exact_phi, exact_D, nq=None):
    nq = nq or (pfe.k + 4)
    pts, w = fe.cell_quadrature(nq)
    x, J = fe.bilinear(pm.pV, pts)
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    jw = det * w[None]
    psi, _ = pfe.rt.tab(pts)
    ph, _ = pfe.dg.tab(pts)
    coef = sol[pfe.idx]
    Dhat = np.einsum("cj,jqd->cqd", coef[:, :pfe.nrt], psi)
    D = np.einsum("cqrd,cqd->cqr", J, Dhat) / det[..., None]
    Phi = coef[:, pfe.nrt:] @ ph
    ep = np.sqrt(np.sum((Phi - exact_phi(x)) ** 2 * jw))
    ED = exact_D(x)
    eD = np.sqrt(np.sum(np.sum((D - ED) ** 2, -1) * jw))
    return ep, eD


def _mixed_solve(pm, pfe, b, eps=1.0, lam2=1.0):
    A = assemble_mixed(pm, pfe, eps, eps, lam2)
    cons = constrained_dofs(pm, pfe)
    free = np.setdiff1d(np.arange(pfe.n), cons)
    x = np.zeros(pfe.n)
    x[free] = spla.spsolve(A[free][:, free].tocsc(), b[free])
    return x


def _mixed_rhs(pm, pfe, f_vals, g):
    """-phi_i f on cells, -(psi_i.n) g on Dirichlet faces (MixedFEM.cpp:166-254)."""
    x, JxW, psi, div = pfe.cell_parts(pm.pV)
    loc = np.zeros((len(pm.pV), pfe.nloc))
    loc[:, pfe.nrt:] = -np.einsum("jq,cq->cj", pfe.phi_q, f_vals(x) * JxW)
    for f in range(4):
        cells = np.nonzero(pm.pbid[:, f] == DIRICHLET)[0]
        if len(cells) == 0:
            continue
        pts = fe.face_points(f, pfe.fs)
        xf, Jf = fe.bilinear(pm.pV[cells], pts)
        detf = Jf[..., 0, 0] * Jf[..., 1, 1] - Jf[..., 0, 1] * Jf[..., 1, 0]
        _, nrm, ln = fe.face_geometry(pm.pV[cells], f, pfe.fs)
        jw = pfe.fw[None, :] * ln
        psif = np.einsum("cqrd,jqd->jcqr", Jf, pfe.psi_f[f]) / detf[None, :, :, None]
        psin = np.einsum("jcqr,cqr->cjq", psif, nrm)
        loc[cells, :pfe.nrt] += -np.einsum("cjq,cq->cj", psin, g(xf) * jw)
    b = np.zeros(pfe.n)
    np.add.at(b, pfe.idx, loc)
    return b


# ---- closed forms (test_functions.cpp) ------------------------------------
def poisson_f(x):
    return 4 * np.pi ** 2 * (np.cos(TP * x[..., 1]) - np.sin(TP * x[..., 0]))


def poisson_g(x):
    return np.cos(TP * x[..., 1]) - np.sin(TP * x[..., 0]) - x[..., 0]


def poisson_D(x):
    return np.stack([1 + TP * np.cos(TP * x[..., 0]), TP * np.sin(TP * x[..., 1])], -1)


def _dot_n(vec):
    return lambda x, n: np.sum(vec(x) * n, -1)


def steady_state(k, r, geometry="unit", out=None):
    """test_steady_state: mixed Poisson and steady LDG (theta = 0) errors.
    `out` (a dict) receives the right-hand sides and solutions."""
    reg, pm = study_mesh(r, geometry, test_grid_classify)
    pfe = PoissonFE(k, pm)
    bP = _mixed_rhs(pm, pfe, poisson_f, poisson_g)
    sol = _mixed_solve(pm, pfe, bP)
    ephi, eD = mixed_errors(pm, pfe, sol, poisson_g, poisson_D)
    cfe = CarrierFE(k)
    A, _, _ = assemble_ldg(reg, cfe, 1.0, 1.0, 1.0, 0.0, 1.0)
    loc, _ = _cell_rhs(reg, cfe, f_vals=poisson_f)
    _dirichlet_rhs(reg, cfe, loc, poisson_g)
    _flux_rhs(reg, cfe, loc, _dot_n(poisson_D))  # J = -grad u = D (interface faces of 'cell' only)
    b = _assemble_vec(reg, cfe, loc)
    u = spla.spsolve(A.tocsc(), b)
    eu, eJ = ldg_errors(reg, cfe, u, poisson_g, lambda x: -(-poisson_D(x)))
    h = fe.diameter(reg.V).max()
    if out is not None:
        out.update(bP=bP, P=sol, b=b, u=u)
    return dict(h=h, Phi=ephi, D=eD, u=eu, J=eJ)


def _u_exact(x, t):
    return np.exp(-t) + np.cos(TP * x[..., 0]) + np.cos(TP * x[..., 1])


def _ic(ic):
    if ic == "reference":
        return lambda x: 1 + np.cos(TP * x[..., 0]) * np.cos(TP * x[..., 1])
    return lambda x: _u_exact(x, 0.0)


def project_density(reg, cfe, func):
    """L2 projection of a scalar onto the density block (VectorTools::project)."""
    x, J, det, Jinv, JxW = cfe.cell_geometry(reg.V)
    Mloc = np.einsum("aq,bq,cq->cab", cfe.phi, cfe.phi, JxW)
    rhs = np.einsum("aq,cq->ca", cfe.phi, func(x) * JxW)
    ud = np.linalg.solve(Mloc, rhs[..., None])[..., 0]
    u = np.zeros(3 * cfe.nb * reg.ncell)
    u[2 * cfe.nb * reg.ncell:] = ud.ravel()
    return u


def transient_J(x, t):
    """Exact current -grad u - E u of the transient study, E = (1, 0)."""
    return np.stack([TP * np.sin(TP * x[..., 0]) - _u_exact(x, t), TP * np.sin(TP * x[..., 1])], -1)


def transient_f(x, t):
    """Source of the transient study (test_functions.cpp:148-164)."""
    return -np.exp(-t) + TP * TP * np.cos(TP * x[..., 0]) + TP * TP * np.cos(TP * x[..., 1]) \
        + TP * np.sin(TP * x[..., 0])


def transient(k, r, ic="exact", geometry="unit"):
    """test_transient: u_t + div(-grad u - E u) = f, E = (1,0), dt = h^(k+1), T = 1."""
    reg, _ = study_mesh(r, geometry, tran_grid_classify)
    cfe = CarrierFE(k)
    h = fe.diameter(reg.V).max()
    dt = h ** (k + 1)
    A, _, M = assemble_ldg(reg, cfe, 1.0, 1.0, dt, 1.0, 1.0)
    lu = spla.splu(A.tocsc())
    u = project_density(reg, cfe, _ic(ic))
    Efield = np.zeros((reg.ncell, cfe.qpts.shape[0], 2))
    Efield[..., 0] = 1.0
    t = 0.0
    f = transient_f

    def gI(x, n, t):
        if geometry == "unit":  # x = 1 (test_functions.cpp:184-196)
            return -np.exp(-t) - np.cos(TP * x[..., 1]) - 1.0
        return np.sum(transient_J(x, t) * n, -1)

    while t < 1.0:
        loc, _ = _cell_rhs(reg, cfe, f_vals=lambda x: f(x, t), p_field=Efield, u_old=u)
        _dirichlet_rhs(reg, cfe, loc, lambda x: _u_exact(x, t))
        _flux_rhs(reg, cfe, loc, lambda x, n: gI(x, n, t))
        u = lu.solve(M @ u + _assemble_vec(reg, cfe, loc))
        t += dt
    eu, eJ = ldg_errors(reg, cfe, u, lambda x: _u_exact(x, t), lambda x: transient_J(x, t))
    return dict(h=h, u=eu, J=eJ, t=t, dt=dt, u_h=u)


def dd_J(x, t):
    """Exact current -grad u - D u of the DD-Poisson study."""
    ut = _u_exact(x, t)
    return np.stack([TP * np.sin(TP * x[..., 0]) - (TP * np.cos(TP * x[..., 0]) + 1) * ut,
                     TP * np.sin(TP * x[..., 1]) - TP * np.sin(TP * x[..., 1]) * ut], -1)


def dd_poisson_f(x, t):
    """Poisson source of the DD-Poisson study (test_functions.cpp:300-314)."""
    return poisson_f(x) + _u_exact(x, t)


def dd_f(x, t):
    """Drift-diffusion source of the DD-Poisson study (test_functions.cpp:276-298)."""
    X, Y = x[..., 0], x[..., 1]
    ut = _u_exact(x, t)
    div_E_u = TP * TP * (np.cos(TP * Y) - np.sin(TP * X)) * ut \
        - TP * (TP * np.cos(TP * X) + 1) * np.sin(TP * X) - TP * TP * np.sin(TP * Y) ** 2
    return -np.exp(-t) + TP * TP * np.cos(TP * X) + TP * TP * np.cos(TP * Y) - div_E_u


def dd_poisson(k, r, ic="exact", geometry="unit"):
    """test_DD_Poisson: mixed Poisson coupled to drift-diffusion via D, all Dirichlet."""
    reg, pm = study_mesh(r, geometry, dd_grid_classify)
    ncs = reg.ncell  # the carrier cells come first in the Poisson mesh
    pfe = PoissonFE(k, pm)
    cfe = CarrierFE(k)
    h = fe.diameter(reg.V).max()
    dt = h ** (k + 1)
    A, _, M = assemble_ldg(reg, cfe, 1.0, 1.0, dt, 1.0, 1.0)
    lu = spla.splu(A.tocsc())
    AP = assemble_mixed(pm, pfe, 1.0, 1.0, 1.0)
    luP = spla.splu(AP.tocsc())
    u = project_density(reg, cfe, _ic(ic))
    from .ldg import field_at_carrier_q
    t = 0.0

    fP, fDD = dd_poisson_f, dd_f
    sol = None
    while t < 1.0:
        # Poisson with the old density (SolarCell.cpp:2105-2185)
        x, JxW, psi, div = pfe.cell_parts(pm.pV)
        ud = density_at(cfe, u, reg.ncell, cfe.phi)
        # the data carries u only where the density lives (the whole square
        # on 'unit', the semiconductor half on 'cell')
        bP = _mixed_rhs(pm, pfe, lambda xx: fP(xx, t) if geometry == "unit" else poisson_f(xx), poisson_g)
        loc = np.zeros((ncs, pfe.nloc))
        if geometry != "unit":
            loc[:, pfe.nrt:] = -np.einsum("jq,cq->cj", pfe.phi_q, _u_exact(x[:ncs], t) * JxW[:ncs])
        loc[:, pfe.nrt:] += np.einsum("jq,cq->cj", pfe.phi_q, ud * JxW[:ncs])
        np.add.at(bP, pfe.idx[:ncs], loc)
        sol = luP.solve(bP)
        Ef = field_at_carrier_q(cfe, pfe, reg.V, sol[pfe.idx[:ncs]])
        loc, _ = _cell_rhs(reg, cfe, f_vals=lambda xx: fDD(xx, t), p_field=Ef, u_old=u)
        _dirichlet_rhs(reg, cfe, loc, lambda xx: _u_exact(xx, t))
        _flux_rhs(reg, cfe, loc, lambda xx, n: np.sum(dd_J(xx, t) * n, -1))
        u = lu.solve(M @ u + _assemble_vec(reg, cfe, loc))
        t += dt

    eu, eJ = ldg_errors(reg, cfe, u, lambda x: _u_exact(x, t), lambda x: dd_J(x, t))
    ephi, eD = mixed_errors(pm, pfe, sol, poisson_g, poisson_D)
    return dict(h=h, u=eu, J=eJ, Phi=ephi, D=eD, t=t, dt=dt, u_h=u, P=sol)


def interface_f(x, t):
    """Source of the interface study (test_functions.cpp:450-464)."""
    return -np.exp(-t) + TP * TP * np.cos(TP * x[..., 0]) + TP * TP * np.cos(TP * x[..., 1])


def interface_gI(x, t):
    """Interface datum I = u^2 (test_functions.cpp:484-505)."""
    return _u_exact(x, t) ** 2


def u_exact(x, t):
    """exp(-t) + cos(2 pi x) + cos(2 pi y), the transient studies' solution."""
    return _u_exact(x, t)


def interface(k, r, ic="exact"):
    """test_interface_coupling: two heat equations coupled on x = 1/2 by
    -grad u . eta = u_s u_e - I (SolarCell.cpp:2262-2745, 3123-3278)."""
    from .mesh import SolarCellMesh
    geo = dict(r1=0.5, r2=0.5, H=1.0, L=1.0)
    m = SolarCellMesh(r, geo)
    cfe = CarrierFE(k)
    h = fe.diameter(m.semi.V).max()
    dt = h ** (k + 1)
    As, _, Ms = assemble_ldg(m.semi, cfe, 1.0, 1.0, dt, 1.0, 1.0)
    Ae, _, Me = assemble_ldg(m.elec, cfe, 1.0, 1.0, dt, 1.0, 1.0)
    lus, lue = spla.splu(As.tocsc()), spla.splu(Ae.tocsc())
    us = project_density(m.semi, cfe, _ic(ic))
    ue = project_density(m.elec, cfe, _ic(ic))

    f, gI = interface_f, interface_gI
    t = 0.0
    while t < 1.0:
        new = []
        for reg, u, other, oreg, nbmap, M, lu in (
                (m.semi, us, ue, m.elec, m.semi_sigma_nb, Ms, lus),
                (m.elec, ue, us, m.semi, m.elec_sigma_nb, Me, lue)):
            loc, _ = _cell_rhs(reg, cfe, f_vals=lambda xx: f(xx, t))
            _dirichlet_rhs(reg, cfe, loc, lambda xx: _u_exact(xx, t))
            for fc in range(4):
                cells = np.nonzero(reg.bid[:, fc] == INTERFACE)[0]
                if len(cells) == 0:
                    continue
                PF, VF = cfe.face_parts(fc)
                xf, nrm, ln = fe.face_geometry(reg.V[cells], fc, cfe.fs)
                fo = {0: 1, 1: 0, 2: 3, 3: 2}[fc]
                own = density_at(cfe, u, reg.ncell, cfe.fphi[fc])[cells]
                nbc = np.array([nbmap[c] for c in cells])
                oth = density_at(cfe, other, oreg.ncell, cfe.fphi[fo])[nbc]
                loc[cells] += np.einsum("iq,cq->ci", VF, -(own * oth - gI(xf, t)) * cfe.fw[None] * ln)
            new.append(lu.solve(M @ u + _assemble_vec(reg, cfe, loc)))
        us, ue = new
        t += dt
    J_ex = lambda x: np.stack([TP * np.sin(TP * x[..., 0]), TP * np.sin(TP * x[..., 1])], -1)
    e1 = ldg_errors(m.semi, cfe, us, lambda x: _u_exact(x, t), J_ex)
    e2 = ldg_errors(m.elec, cfe, ue, lambda x: _u_exact(x, t), J_ex)
    return dict(h=h, u1=e1[0], J1=e1[1], u2=e2[0], J2=e2[1], t=t, dt=dt, us=us, ue=ue)
