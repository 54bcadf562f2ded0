"""Parameters, physics point functions and the IMEX time loop (oracle).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

Restates:
  * Parameters::parse_and_scale_parameters (Parameters.hpp:117-287);
  * Generation::value (Generation.cpp:29-44) -- reference form
    alpha*G0*exp(alpha*(y-H)); the configuration form G0*exp(-alpha*x)
    (Beer-Lambert from the left contact) is selected by gen_mode='config';
  * SRH_Recombination (SolarCell.hpp:86-98): identically 0 in the reference;
    srh=True selects the documented formula (Documentation/README.md:266-271);
  * equilibria n^e, p^e, r^inf, o^inf (InitialConditions.cpp:5-81) as
    parameters (reference values 2, 0, 30, 29);
  * SolarCellProblem::run_full_system setup and time loop
    (SolarCell.cpp:1898-2095): Poisson matrix, LDG matrices, factorisation,
    constant initial densities (the L2 projection of a constant is the
    constant), the first Poisson solve, then per step
      RHS_S(old), RHS_E(old) -> solve n, p, r, o -> RHS_P(new) -> solve P.
    The step count is explicit (the reference's float loop
    `while (time < timeStamps[k])`, SolarCell.cpp:2055, is not reproduced).
  * Carrier::solve / PoissonData::solve (Carrier.cpp:34-40,
    Poisson.cpp:98-105) with scipy's SuperLU in the role of UMFPACK.
"""
import numpy as np
import scipy.sparse.linalg as spla

from .ldg import CarrierFE, assemble_ldg, carrier_rhs
from .mesh import LocalRefinedMesh, SolarCellMesh
from .mixed import PoissonFE, assemble_mixed, constrained_dofs, poisson_rhs, reduction

THERMAL_VOLTAGE = 0.02585
ELECTRON_CHARGE = 1.62e-19
VACUUM_PERMITTIVITY = 8.85e-14


def default_params():
    """Scaled parameters of the measured configurations (unit square, 0.5 split)."""
    return dict(
        k=1, r=4, r1=0.5, r2=0.5, H=1.0, L=1.0, insulated=False, schottky=False,
        dt=1e-3, penalty=1.0, theta=1.0,
        mu_n=1.0, mu_p=1.0, mu_r=1.0, mu_o=1.0,
        eps_s=1.0, eps_e=1.0, lam2=1.0,
        k_et=1e-2, k_ht=1e-2, v_rec_n=0.0, v_rec_p=0.0,
        ne=1.0, pe=0.0, rinf=1.0, oinf=1.0,
        phi_bi=0.0, phi_app=0.0, phi_sch=0.0,
        gen_mode="ref", G0=1.0, alpha=1.0,
        srh=False, ni=0.1, tau_n=1.0, tau_p=1.0,
    )


def scale_parameters(phys):
    """Parameters.hpp:164-251 -- singular-perturbation scaling of physical inputs."""
    s = dict(phys)
    s["intrinsic_density"] = phys["intrinsic_density"] / phys["char_density"]
    s["electron_recombo_t"] = phys["electron_recombo_t"] / phys["char_time"]
    s["hole_recombo_t"] = phys["hole_recombo_t"] / phys["char_time"]
    s["electron_recombo_v"] = phys["electron_recombo_v"] * (phys["char_time"] / phys["char_length"])
    s["hole_recombo_v"] = phys["hole_recombo_v"] * (phys["char_time"] / phys["char_length"])
    s["photon_flux"] = phys["photon_flux"] * (phys["char_time"] / phys["char_density"])
    s["absorption_coeff"] = phys["absorption_coeff"] * phys["char_length"]
    s["debye_length"] = (THERMAL_VOLTAGE * VACUUM_PERMITTIVITY) / (
        ELECTRON_CHARGE * phys["char_density"] * phys["char_length"] * phys["char_length"])
    mob = (phys["char_time"] * THERMAL_VOLTAGE) / (phys["char_length"] * phys["char_length"])
    for kname in ("electron_mobility", "hole_mobility", "reductant_mobility", "oxidant_mobility"):
        s[kname] = phys[kname] * mob
    s["rescaled_k_et"] = ELECTRON_CHARGE * phys["k_et"] * phys["char_density"] ** 2
    s["rescaled_k_ht"] = ELECTRON_CHARGE * phys["k_ht"] * phys["char_density"] ** 2
    s["k_et"] = phys["k_et"] * (phys["char_time"] * phys["char_density"] / phys["char_length"])
    s["k_ht"] = phys["k_ht"] * (phys["char_time"] * phys["char_density"] / phys["char_length"])
    s["applied_bias"] = phys["applied_bias"] / THERMAL_VOLTAGE
    s["built_in_bias"] = phys["built_in_bias"] / THERMAL_VOLTAGE
    s["schottky_bias"] = phys["schottky_bias"] / THERMAL_VOLTAGE
    s["rescale_current"] = ELECTRON_CHARGE * phys["char_density"] * phys["char_length"] / phys["char_time"]
    return s


def make_physics(prm):
    """Attach generation / recombination callables to a parameter dict."""
    p = dict(prm)
    G0, alpha, H = p["G0"], p["alpha"], p["H"]
    mode = p["gen_mode"]
    if mode == "ref":
        p["generation"] = lambda x: alpha * G0 * np.exp(alpha * (x[..., 1] - H))
    elif mode == "config":
        p["generation"] = lambda x: G0 * np.exp(-alpha * x[..., 0])
    else:
        p["generation"] = lambda x: np.zeros(x.shape[:-1])
    if p["srh"]:
        ni, tn, tp = p["ni"], p["tau_n"], p["tau_p"]
        p["recombination"] = lambda n, pp: (ni * ni - n * pp) / (tn * (n + ni) + tp * (pp + ni))
    else:
        p["recombination"] = lambda n, pp: np.zeros_like(n)
    return p


class OracleSolarCell:
    """CPU FP64 restatement of SolarCellProblem::run_full_system."""

    def __init__(self, prm):
        self.prm = make_physics(prm)
        p = self.prm
        k, r = p["k"], p["r"]
        geo = dict(r1=p["r1"], r2=p["r2"], H=p["H"], L=p["L"])
        if p.get("local_refine", 0):
            if p["local_refine"] != 1:
                raise ValueError("only one level of boundary-layer refinement is restated")
            self.mesh = LocalRefinedMesh(r, geo, p["bl"], p["insulated"], p["schottky"])
        else:
            self.mesh = SolarCellMesh(r, geo, p["insulated"], p["schottky"])
        self.cfe = CarrierFE(k)
        self.pfe = PoissonFE(k, self.mesh)
        m = self.mesh
        self.A_n, self.A_p, self.M_s = assemble_ldg(m.semi, self.cfe, p["mu_n"], p["mu_p"], p["dt"], p["theta"], p["penalty"])
        self.A_r, self.A_o, self.M_e = assemble_ldg(m.elec, self.cfe, p["mu_r"], p["mu_o"], p["dt"], p["theta"], p["penalty"])
        self.A_P = assemble_mixed(m, self.pfe, p["eps_s"], p["eps_e"], p["lam2"])
        self.cons = constrained_dofs(m, self.pfe)
        self.free = np.setdiff1d(np.arange(self.pfe.n), self.cons)
        # with hanging faces the constraints are x = T y (Neumann D.n = 0,
        # hanging fine faces following their coarse face)
        self.mesh_local = bool(getattr(m, "local", 0))
        if self.mesh_local:
            self.T = reduction(m, self.pfe)
        self.lu = {}
        self.factor()
        nb = self.cfe.nb
        self.Nc = 3 * nb * m.semi.ncell
        self.Ne = 3 * nb * m.elec.ncell

    def factor(self):
        # The LDG matrices have a positive definite symmetric part (the
        # off-diagonal blocks satisfy C = -B^T), so LU without pivoting on a
        # minimum-degree ordering of A + A^T is stable; the mixed matrix is a
        # saddle point and keeps SuperLU's threshold pivoting.
        for name in ("n", "p", "r", "o"):
            self.lu[name] = spla.splu(getattr(self, "A_" + name).tocsc(), permc_spec="MMD_AT_PLUS_A",
                                      diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))
        if self.mesh_local:
            Af = (self.T.T @ self.A_P @ self.T).tocsr()
        else:
            Af = self.A_P[self.free][:, self.free]
        self.lu["P"] = spla.splu(Af.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.1,
                                 options=dict(SymmetricMode=True))

    def solve(self, which, b):
        if which == "P":
            if self.mesh_local:
                return self.T @ self.lu["P"].solve(self.T.T @ b)
            x = np.zeros(self.pfe.n)
            x[self.free] = self.lu["P"].solve(b[self.free])
            return x
        return self.lu[which].solve(b)

    def poisson_coef(self, phi, region):
        pc = phi[self.pfe.idx]
        if region == "semi":
            return pc[self.mesh.s2p]
        return pc[self.mesh.e2p]

    def initial_state(self):
        p = self.prm
        nb = self.cfe.nb
        ns, ne_ = self.mesh.semi.ncell, self.mesh.elec.ncell
        st = {}
        for name, val, nc in (("n", p["ne"], ns), ("p", p["pe"], ns), ("r", p["rinf"], ne_), ("o", p["oinf"], ne_)):
            u = np.zeros(3 * nb * nc)
            u[2 * nb * nc:] = val
            st[name] = u
        st["P"] = self.solve("P", self.rhs_poisson(st))
        return st

    def rhs_semi(self, st):
        m = self.mesh
        return carrier_rhs(m.semi, self.cfe, self.pfe, m, self.M_s, st["n"], st["p"],
                           self.poisson_coef(st["P"], "semi"), self.prm, "semi",
                           st["r"], st["o"], m.elec)

    def rhs_elec(self, st):
        m = self.mesh
        return carrier_rhs(m.elec, self.cfe, self.pfe, m, self.M_e, st["r"], st["o"],
                           self.poisson_coef(st["P"], "elec"), self.prm, "elec",
                           st["n"], st["p"], m.semi)

    def rhs_poisson(self, st):
        return poisson_rhs(self.mesh, self.pfe, self.cfe, st["n"], st["p"], st["r"], st["o"], self.prm)

    def step(self, st):
        bn, bp = self.rhs_semi(st)
        br, bo = self.rhs_elec(st)
        new = {"n": self.solve("n", bn), "p": self.solve("p", bp),
               "r": self.solve("r", br), "o": self.solve("o", bo)}
        new["P"] = self.solve("P", self.rhs_poisson(new))
        return new

    def run(self, nsteps, st=None):
        if st is None:
            st = self.initial_state()
        for _ in range(nsteps):
            st = self.step(st)
        return st
