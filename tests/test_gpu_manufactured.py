"""The reference's four manufactured-solution studies through the HIP path.

tests/Poisson_test.cpp, IMEX_LDG_test.cpp, DD_Poisson_test.cpp and
interface_test.cpp (driven by SolarCell.cpp:2749-3278 with the data of
test_functions.cpp:13-527) run here with every right-hand side assembled by
the library's kernels (the manufactured source, Dirichlet datum and flux
given at the quadrature points through pecs_set_manufactured) and every
solve done by its device factors.  They are posed on the solar-cell grid the
GPU path is built for (r1 = r2 = 0.5, H = L = 1, outer faces Dirichlet, the
exact normal flux on the interface x = 0.5; the interface study is on that
grid in the reference already), and the oracle runs the same studies on the
same grid (oracle/manufactured.py, geometry='cell').  Asserted: the right-
hand sides and solutions agree with the oracle's, the L2 errors agree to
1e-10 relative, and the rates are the textbook ones (LDG density k+1, current
k; mixed Phi k, D k+1 with this grid's Dirichlet data).
"""
import numpy as np
import pytest

from oracle import fe
from oracle import manufactured as mm
from oracle.ldg import CarrierFE, field_at_carrier_q
from oracle.mesh import INTERFACE, SolarCellMesh
from oracle.mixed import PoissonFE

pytestmark = pytest.mark.gpu

REL = 1e-10


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _close(name, a, b, tol):
    d = _rel(a, b)
    _say(f"  {name}: {d:.1e} relative to the oracle's")
    assert d <= tol, (name, d)


def _params(k, r, dt, theta, **kw):
    import pecs_amd
    p = pecs_amd.default_params()
    p.update(mm.CELL_GEO)
    p.update(k=k, r=r, dt=dt, theta=theta, penalty=1.0, insulated=False, schottky=False,
             mu_n=1.0, mu_p=1.0, mu_r=1.0, mu_o=1.0, eps_s=1.0, eps_e=1.0, lam2=1.0,
             k_et=0.0, k_ht=0.0, v_rec_n=0.0, v_rec_p=0.0, ne=0.0, pe=0.0, rinf=0.0, oinf=0.0,
             phi_bi=0.0, phi_app=0.0, phi_sch=0.0, gen_mode="off", srh=False)
    p.update(kw)
    return p


class _Points:
    """The library's quadrature points of one target, checked against the oracle's."""

    def __init__(self, g, target, V, bid, cfe):
        self.cx, self.fx = g.quadrature_points(target)
        self.nrm = np.zeros_like(self.fx)
        self.iface = np.zeros(self.fx.shape[:3], bool)
        x, _, _, _, _ = cfe.cell_geometry(V)
        assert np.abs(self.cx - x).max() <= 1e-14
        for f in range(4):
            xf, nrm, _ = fe.face_geometry(V, f, cfe.fs)
            assert np.abs(self.fx[:, f] - xf).max() <= 1e-14
            self.nrm[:, f] = nrm
            self.iface[:, f] = (bid[:, f] == INTERFACE)[:, None]

    def flux(self, J):
        """J.n at the interface face points, 0 elsewhere."""
        return np.where(self.iface, np.sum(J(self.fx) * self.nrm, -1), 0.0)


def _errors_match(gpu, ora, keys):
    _say("  L2 errors " + ", ".join(f"{key} {gpu[key]:.6e} (oracle {abs(gpu[key] - ora[key]) / ora[key]:.1e} off)"
                                    for key in keys))
    for key in keys:
        assert abs(gpu[key] - ora[key]) <= REL * ora[key], (key, gpu[key], ora[key])


def _rates(rows, key):
    r = [float(np.log2(rows[i - 1][key] / rows[i][key])) for i in range(1, len(rows))]
    _say(f"  rates of {key}: " + ", ".join(f"{x:.3f}" for x in r))
    return r


# --- Poisson_test.cpp: mixed Poisson and steady LDG (theta = 0) -----------------------------
def _steady_gpu(k, r):
    import pecs_amd
    m = SolarCellMesh(r, mm.CELL_GEO)
    cfe, pfe = CarrierFE(k), PoissonFE(k, m)
    ora_out = {}
    ora = mm.steady_state(k, r, geometry="cell", out=ora_out)
    g = pecs_amd.SolarCellGPU(_params(k, r, 1.0, 0.0))
    try:
        pp = _Points(g, 2, m.pV, m.pbid, cfe)
        g.set_manufactured(2, src=mm.poisson_f(pp.cx), dir=mm.poisson_g(pp.fx))
        zs, ze = np.zeros(g.n_semi), np.zeros(g.n_elec)
        bP = g.rhs_poisson(zs, zs, ze, ze)
        _close("bP", bP, ora_out["bP"], 1e-13)
        P = g.solve(pecs_amd.POISSON, bP)
        _close("P", P, ora_out["P"], REL)
        sp = _Points(g, 0, m.semi.V, m.semi.bid, cfe)
        g.set_manufactured(0, src=mm.poisson_f(sp.cx), dir=mm.poisson_g(sp.fx), flux=sp.flux(mm.poisson_D))
        bn, bp = g.rhs_semiconductor(zs, zs, np.zeros(g.n_poisson), ze, ze)
        _close("bn", bn, ora_out["b"], 1e-13)
        assert np.array_equal(bn, bp)  # both carriers take the region's data; no drift, no state
        u = g.solve(pecs_amd.ELECTRONS, bn)
        _close("u", u, ora_out["u"], REL)
    finally:
        g.close()
    ephi, eD = mm.mixed_errors(m, pfe, P, mm.poisson_g, mm.poisson_D)
    eu, eJ = mm.ldg_errors(m.semi, cfe, u, mm.poisson_g, mm.poisson_D)
    return dict(Phi=ephi, D=eD, u=eu, J=eJ), ora


@pytest.mark.parametrize("k", [1, 2])
def test_steady_state_study(k):
    rows, orows = [], []
    for r in (2, 3, 4):
        _say(f"[gpu] manufactured steady study: k={k} refinement {r}")
        gpu, ora = _steady_gpu(k, r)
        _errors_match(gpu, ora, ("Phi", "D", "u", "J"))
        rows.append(gpu)
        orows.append(ora)
    for key, rate in (("Phi", k), ("D", k + 1), ("u", k + 1), ("J", k)):
        got = _rates(rows, key)
        assert np.allclose(got, _rates(orows, key), rtol=0, atol=1e-8)
        assert got[-1] > rate - 0.15, (key, got)


# --- IMEX_LDG_test.cpp: transient drift-diffusion, E = (1, 0) ---------------------------------
def _constant_field(pfe, V, D):
    """Poisson coefficient vector (library layout) whose RT part is the constant
    field D on every cell: nodal values of the reference field det(J) J^{-1} D."""
    rt = pfe.rt
    deg = rt.deg
    nodes = [(rt.Xn[a], rt.Gn[b], 0) for b in range(deg + 1) for a in range(deg + 2)]
    nodes += [(rt.Gn[a], rt.Xn[b], 1) for b in range(deg + 2) for a in range(deg + 1)]
    assert len(nodes) == rt.n
    x = np.zeros(pfe.n)
    pts = np.array([[xi, eta] for xi, eta, _ in nodes])
    _, J = fe.bilinear(V, pts)
    adj = np.stack([np.stack([J[..., 1, 1], -J[..., 0, 1]], -1), np.stack([-J[..., 1, 0], J[..., 0, 0]], -1)], -2)
    Dhat = np.einsum("cqij,j->cqi", adj, np.asarray(D, float))  # det(J) J^{-1} D
    for i, (_, _, comp) in enumerate(nodes):
        x[pfe.idx[:len(V), i]] = Dhat[:, i, comp]
    return x


def _transient_gpu(k, r):
    import pecs_amd
    m = SolarCellMesh(r, mm.CELL_GEO)
    reg = m.semi
    cfe, pfe = CarrierFE(k), PoissonFE(k, m)
    ora = mm.transient(k, r, geometry="cell")
    dt = ora["dt"]
    P = _constant_field(pfe, reg.V, (1.0, 0.0))
    E = field_at_carrier_q(cfe, pfe, reg.V, P[pfe.idx[:reg.ncell]])
    assert np.abs(E - np.array([1.0, 0.0])).max() <= 1e-13  # the input field is E = (1, 0)
    g = pecs_amd.SolarCellGPU(_params(k, r, dt, 1.0))
    try:
        sp = _Points(g, 0, reg.V, reg.bid, cfe)
        u = mm.project_density(reg, cfe, lambda x: mm.u_exact(x, 0.0))
        zs, ze = np.zeros(g.n_semi), np.zeros(g.n_elec)
        t = 0.0
        while t < 1.0:
            g.set_manufactured(0, src=mm.transient_f(sp.cx, t), dir=mm.u_exact(sp.fx, t),
                               flux=sp.flux(lambda x: mm.transient_J(x, t)))
            bn, _ = g.rhs_semiconductor(u, zs, P, ze, ze)
            u = g.solve(pecs_amd.ELECTRONS, bn)
            t += dt
    finally:
        g.close()
    assert t == ora["t"]
    _close("u", u, ora["u_h"], REL)
    eu, eJ = mm.ldg_errors(reg, cfe, u, lambda x: mm.u_exact(x, t), lambda x: mm.transient_J(x, t))
    return dict(u=eu, J=eJ), ora


def test_transient_study():
    rows, orows = [], []
    for r in (2, 3, 4):
        _say(f"[gpu] manufactured transient study: k=1 refinement {r}")
        gpu, ora = _transient_gpu(1, r)
        _errors_match(gpu, ora, ("u", "J"))
        rows.append(gpu)
        orows.append(ora)
    assert np.allclose(_rates(rows, "u"), _rates(orows, "u"), rtol=0, atol=1e-8)
    assert _rates(rows, "u")[-1] > 1.85
    assert _rates(rows, "J")[-1] > 0.8


# --- DD_Poisson_test.cpp: drift-diffusion coupled to the mixed Poisson problem -----------------
def _dd_gpu(k, r):
    import pecs_amd
    m = SolarCellMesh(r, mm.CELL_GEO)
    reg = m.semi
    ncs = reg.ncell
    cfe, pfe = CarrierFE(k), PoissonFE(k, m)
    ora = mm.dd_poisson(k, r, geometry="cell")
    dt = ora["dt"]
    g = pecs_amd.SolarCellGPU(_params(k, r, dt, 1.0))
    try:
        sp = _Points(g, 0, reg.V, reg.bid, cfe)
        pp = _Points(g, 2, m.pV, m.pbid, cfe)
        u = mm.project_density(reg, cfe, lambda x: mm.u_exact(x, 0.0))
        zs, ze = np.zeros(g.n_semi), np.zeros(g.n_elec)
        dirP = mm.poisson_g(pp.fx)
        t = 0.0
        while t < 1.0:
            # the charge is u_exact - u_h + (Poisson source) on the semiconductor, the source elsewhere
            srcP = mm.poisson_f(pp.cx)
            srcP[:ncs] += mm.u_exact(pp.cx[:ncs], t)
            g.set_manufactured(2, src=srcP, dir=dirP)
            P = g.solve(pecs_amd.POISSON, g.rhs_poisson(u, zs, ze, ze))
            g.set_manufactured(0, src=mm.dd_f(sp.cx, t), dir=mm.u_exact(sp.fx, t),
                               flux=sp.flux(lambda x: mm.dd_J(x, t)))
            bn, _ = g.rhs_semiconductor(u, zs, P, ze, ze)
            u = g.solve(pecs_amd.ELECTRONS, bn)
            t += dt
    finally:
        g.close()
    _close("u", u, ora["u_h"], REL)
    _close("P", P, ora["P"], REL)
    eu, eJ = mm.ldg_errors(reg, cfe, u, lambda x: mm.u_exact(x, t), lambda x: mm.dd_J(x, t))
    ephi, eD = mm.mixed_errors(m, pfe, P, mm.poisson_g, mm.poisson_D)
    return dict(u=eu, J=eJ, Phi=ephi, D=eD), ora


def test_dd_poisson_study():
    rows, orows = [], []
    for r in (2, 3, 4):
        _say(f"[gpu] manufactured DD-Poisson study: k=1 refinement {r}")
        gpu, ora = _dd_gpu(1, r)
        _errors_match(gpu, ora, ("u", "J", "Phi", "D"))
        rows.append(gpu)
        orows.append(ora)
    assert _rates(rows, "u")[-1] > 1.8
    assert _rates(rows, "Phi")[-1] > 0.9
    assert _rates(rows, "D")[-1] > 1.8


# --- interface_test.cpp: two heat equations coupled through -grad u . n = u_s u_e - I ----------
def _interface_gpu(k, r):
    import pecs_amd
    m = SolarCellMesh(r, mm.CELL_GEO)
    cfe = CarrierFE(k)
    ora = mm.interface(k, r)
    dt = ora["dt"]
    # the hole / reductant interface term k_ht (p - p_e) r becomes -u_s u_e with
    # k_ht = -1 and p_e = 0 (k_et = 0 switches the electron / oxidant one off)
    g = pecs_amd.SolarCellGPU(_params(k, r, dt, 1.0, k_ht=-1.0))
    try:
        sp = _Points(g, 0, m.semi.V, m.semi.bid, cfe)
        ep = _Points(g, 1, m.elec.V, m.elec.bid, cfe)
        us = mm.project_density(m.semi, cfe, lambda x: mm.u_exact(x, 0.0))
        ue = mm.project_density(m.elec, cfe, lambda x: mm.u_exact(x, 0.0))
        zs, ze, zP = np.zeros(g.n_semi), np.zeros(g.n_elec), np.zeros(g.n_poisson)
        t = 0.0
        while t < 1.0:
            for tgt, pts in ((0, sp), (1, ep)):
                # -<v, u_s u_e - I>: the kernel's interface term plus the flux h = -I
                g.set_manufactured(tgt, src=mm.interface_f(pts.cx, t), dir=mm.u_exact(pts.fx, t),
                                   flux=np.where(pts.iface, -mm.interface_gI(pts.fx, t), 0.0))
            _, bp = g.rhs_semiconductor(zs, us, zP, ue, ze)
            br, _ = g.rhs_electrolyte(ue, ze, zP, zs, us)
            us, ue = g.solve(pecs_amd.HOLES, bp), g.solve(pecs_amd.REDUCTANTS, br)
            t += dt
    finally:
        g.close()
    _close("us", us, ora["us"], REL)
    _close("ue", ue, ora["ue"], REL)
    ex = lambda x: mm.u_exact(x, t)
    J_ex = lambda x: np.stack([2 * np.pi * np.sin(2 * np.pi * x[..., 0]), 2 * np.pi * np.sin(2 * np.pi * x[..., 1])], -1)
    e1 = mm.ldg_errors(m.semi, cfe, us, ex, J_ex)
    e2 = mm.ldg_errors(m.elec, cfe, ue, ex, J_ex)
    return dict(u1=e1[0], J1=e1[1], u2=e2[0], J2=e2[1]), ora


def test_interface_study():
    rows, orows = [], []
    for r in (1, 2, 3):
        _say(f"[gpu] manufactured interface study: k=2 refinement {r}")
        gpu, ora = _interface_gpu(2, r)
        _errors_match(gpu, ora, ("u1", "J1", "u2", "J2"))
        rows.append(gpu)
        orows.append(ora)
    assert _rates(rows, "u1")[-1] > 2.8
    assert _rates(rows, "u2")[-1] > 2.8


def test_manufactured_argument_checks():
    import pecs_amd
    g = pecs_amd.SolarCellGPU(_params(1, 2, 1e-2, 1.0))
    try:
        nc, nq, nfq = g._qsizes(2)
        with pytest.raises(RuntimeError, match="no manufactured flux"):
            g.set_manufactured(2, flux=np.zeros(nc * 4 * nfq))
        with pytest.raises(RuntimeError, match="target"):
            g.set_manufactured(3)
        # switched off again: the solar-cell right-hand side returns
        zs, ze = np.zeros(g.n_semi), np.zeros(g.n_elec)
        b0 = g.rhs_semiconductor(zs, zs, np.zeros(g.n_poisson), ze, ze)[0]
        cx, fx = g.quadrature_points(0)
        g.set_manufactured(0, src=np.ones(cx.shape[:2]))
        b1 = g.rhs_semiconductor(zs, zs, np.zeros(g.n_poisson), ze, ze)[0]
        assert np.abs(b1 - b0).max() > 0
        g.set_manufactured(0)
        b2 = g.rhs_semiconductor(zs, zs, np.zeros(g.n_poisson), ze, ze)[0]
        assert np.array_equal(b2, b0)
    finally:
        g.close()
