"""Python binding of the MI355X-native PECS IMEX path (ctypes over include/pecs_abi.h).

The product is the C-ABI library libpecs_amd.so (HIP kernels for gfx950 plus
the C++ setup code).  This module only marshals numpy arrays to it; every
right-hand side, solve and time step runs in the library on the GPU.  If the
library is missing, or no GPU is present, the calls raise -- there is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpecs_amd.so")
_lib = None

ELECTRONS, HOLES, REDUCTANTS, OXIDANTS, POISSON, MASS_SEMI, MASS_ELEC = range(7)

ABI_FUNCTIONS = [
    "pecs_scale_parameters", "pecs_create", "pecs_destroy", "pecs_last_error", "pecs_global_error",
    "pecs_set_stream", "pecs_sizes", "pecs_stats", "pecs_export_matrix", "pecs_rhs_semiconductor",
    "pecs_rhs_electrolyte", "pecs_rhs_poisson", "pecs_solve", "pecs_set_state", "pecs_get_state",
    "pecs_initial_state", "pecs_step", "pecs_last_residual", "pecs_rhs_semiconductor_dev",
    "pecs_rhs_electrolyte_dev", "pecs_rhs_poisson_dev", "pecs_solve_dev", "pecs_state_dev",
    "pecs_enable_timing", "pecs_phase_times", "pecs_factor_bytes", "pecs_sweep_config", "pecs_uniform_lists",
    "pecs_set_mixed_precision", "pecs_set_manufactured", "pecs_quadrature_points", "pecs_sample_grid",
    "pecs_set_state_from_grid", "pecs_last_density_change",
    "pecs_create_batch", "pecs_batch_size", "pecs_set_state_batch", "pecs_get_state_batch",
    "pecs_last_residual_batch", "pecs_write_vtu", "pecs_use_graph", "pecs_sweep_create", "pecs_sweep_set_active",
    "pecs_sweep_step", "pecs_sweep_destroy", "pecs_mlp_create", "pecs_mlp_destroy", "pecs_mlp_num_params",
    "pecs_mlp_get_params", "pecs_mlp_set_params", "pecs_mlp_gradients", "pecs_mlp_train_step", "pecs_mlp_predict",
    "pecs_mlp_last_error", "pecs_mlp_create_decoder", "pecs_mlp_num_outputs",
]


class CMesh(ctypes.Structure):
    _fields_ = [("r1", ctypes.c_double), ("r2", ctypes.c_double), ("H", ctypes.c_double), ("L", ctypes.c_double),
                ("n_refine", ctypes.c_int), ("insulated", ctypes.c_int), ("schottky", ctypes.c_int),
                ("boundary_layer", ctypes.c_double), ("local_refine", ctypes.c_int)]


class CParams(ctypes.Structure):
    _fields_ = [("degree", ctypes.c_int), ("dt", ctypes.c_double), ("theta", ctypes.c_double),
                ("penalty", ctypes.c_double), ("mu", ctypes.c_double * 4), ("eps_s", ctypes.c_double),
                ("eps_e", ctypes.c_double), ("lambda2", ctypes.c_double), ("k_et", ctypes.c_double),
                ("k_ht", ctypes.c_double), ("v_rec_n", ctypes.c_double), ("v_rec_p", ctypes.c_double),
                ("n_e", ctypes.c_double), ("p_e", ctypes.c_double), ("r_inf", ctypes.c_double),
                ("o_inf", ctypes.c_double), ("phi_bi", ctypes.c_double), ("phi_app", ctypes.c_double),
                ("phi_sch", ctypes.c_double), ("schottky_location", ctypes.c_double),
                ("schottky_location_set", ctypes.c_int), ("gen_mode", ctypes.c_int), ("G0", ctypes.c_double),
                ("alpha", ctypes.c_double), ("srh", ctypes.c_int), ("n_i", ctypes.c_double),
                ("tau_n", ctypes.c_double), ("tau_p", ctypes.c_double), ("leaf_cells", ctypes.c_int)]


class CPhysical(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "char_length", "char_time", "char_density", "electron_mobility", "hole_mobility", "reductant_mobility",
        "oxidant_mobility", "electron_recombo_t", "hole_recombo_t", "electron_recombo_v", "hole_recombo_v",
        "k_et", "k_ht", "intrinsic_density", "photon_flux", "absorption_coeff", "applied_bias",
        "built_in_bias", "schottky_bias")]


class CScaled(ctypes.Structure):
    _fields_ = [("mobility", ctypes.c_double * 4), ("recombo_t", ctypes.c_double * 2),
                ("recombo_v", ctypes.c_double * 2)] + [(n, ctypes.c_double) for n in (
                    "k_et", "k_ht", "intrinsic_density", "photon_flux", "absorption_coeff", "debye_length",
                    "applied_bias", "built_in_bias", "schottky_bias", "rescaled_k_et", "rescaled_k_ht",
                    "rescale_current")]


def lib():
    """Load libpecs_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.pecs_last_error.restype = ctypes.c_char_p
        L.pecs_global_error.restype = ctypes.c_char_p
        L.pecs_create.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(CParams), ctypes.POINTER(ctypes.c_void_p)]
        for name in ("pecs_destroy", "pecs_initial_state"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.pecs_last_error.argtypes = [ctypes.c_void_p]
        L.pecs_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.pecs_sizes.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int64)] * 3
        L.pecs_stats.argtypes = [ctypes.c_void_p, dp, ctypes.c_int]
        L.pecs_last_residual.argtypes = [ctypes.c_void_p, dp]
        L.pecs_rhs_semiconductor.argtypes = [ctypes.c_void_p] + [dp] * 7
        L.pecs_rhs_electrolyte.argtypes = [ctypes.c_void_p] + [dp] * 7
        L.pecs_rhs_poisson.argtypes = [ctypes.c_void_p] + [dp] * 5
        L.pecs_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.pecs_set_state.argtypes = [ctypes.c_void_p] + [dp] * 5
        L.pecs_get_state.argtypes = [ctypes.c_void_p] + [dp] * 5
        L.pecs_create_batch.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(CParams), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.pecs_batch_size.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.pecs_set_state_batch.argtypes = [ctypes.c_void_p, ctypes.c_int] + [dp] * 5
        L.pecs_get_state_batch.argtypes = [ctypes.c_void_p, ctypes.c_int] + [dp] * 5
        L.pecs_last_residual_batch.argtypes = [ctypes.c_void_p, dp]
        L.pecs_export_matrix.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(CParams), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32), dp,
                                         ctypes.POINTER(ctypes.c_int8)]
        L.pecs_scale_parameters.argtypes = [ctypes.POINTER(CPhysical), ctypes.POINTER(CScaled)]
        _lib = L
    return _lib


GEN_MODES = {"ref": 0, "config": 1, "off": 2}


def default_params():
    """Scaled parameters of the measured configurations (unit square split at x = 0.5).

    Keys: k (LDG degree), r (global refinement), r1, r2, H, L (Grid.cpp
    geometry), insulated, schottky (boundary passes), dt, theta, penalty,
    mu_{n,p,r,o}, eps_s, eps_e, lam2 (Debye length^2), k_et, k_ht, v_rec_{n,p},
    ne, pe, rinf, oinf (equilibria / doping), phi_bi, phi_app, phi_sch,
    gen_mode ('ref' | 'config' | 'off'), G0, alpha, srh, ni, tau_n, tau_p;
    optional: local_refine (0 | 1) with bl (boundary-layer width).
    """
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


def workload_params(refine, degree, dt):
    """The measured workload: constant donor doping 1, Beer-Lambert generation
    G = exp(-alpha x) with alpha = 1, SRH recombination (tau = 1, n_i = 0.1),
    interface transfer rates 1e-2, no bias, all other coefficients 1."""
    p = default_params()
    p.update(k=degree, r=refine, dt=dt, gen_mode="config", G0=1.0, alpha=1.0, srh=True, ni=0.1,
             tau_n=1.0, tau_p=1.0, ne=1.0, pe=0.0, rinf=1.0, oinf=1.0, k_et=1e-2, k_ht=1e-2)
    return p


def to_c(prm):
    """Build (CMesh, CParams) from a parameter dict (keys as oracle.solarcell.default_params)."""
    m = CMesh(prm["r1"], prm["r2"], prm["H"], prm["L"], prm["r"], int(prm["insulated"]), int(prm["schottky"]),
              float(prm.get("bl", 0.0)), int(prm.get("local_refine", 0)))
    p = CParams()
    p.degree = prm["k"]
    p.dt, p.theta, p.penalty = prm["dt"], prm.get("theta", 1.0), prm["penalty"]
    for i, key in enumerate(("mu_n", "mu_p", "mu_r", "mu_o")):
        p.mu[i] = prm[key]
    p.eps_s, p.eps_e, p.lambda2 = prm["eps_s"], prm["eps_e"], prm["lam2"]
    p.k_et, p.k_ht, p.v_rec_n, p.v_rec_p = prm["k_et"], prm["k_ht"], prm["v_rec_n"], prm["v_rec_p"]
    p.n_e, p.p_e, p.r_inf, p.o_inf = prm["ne"], prm["pe"], prm["rinf"], prm["oinf"]
    p.phi_bi, p.phi_app, p.phi_sch = prm["phi_bi"], prm["phi_app"], prm["phi_sch"]
    if "schottky_location" in prm:
        p.schottky_location = prm["schottky_location"]
        p.schottky_location_set = 1
    p.gen_mode = GEN_MODES[prm["gen_mode"]]
    p.G0, p.alpha = prm["G0"], prm["alpha"]
    p.srh = int(prm["srh"])
    p.n_i, p.tau_n, p.tau_p = prm["ni"], prm["tau_n"], prm["tau_p"]
    p.leaf_cells = prm.get("leaf_cells", 0)
    return m, p


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc, ctx=None):
    if rc != 0:
        L = lib()
        msg = L.pecs_last_error(ctx) if ctx else L.pecs_global_error()
        raise RuntimeError("pecs: " + (msg.decode() if msg else f"error {rc}"))


def export_matrix(prm, which):
    """Constant matrix `which` as scipy CSR (+ constrained flags for POISSON); host only."""
    import scipy.sparse as sp
    L = lib()
    m, p = to_c(prm)
    n = ctypes.c_int64()
    nnz = ctypes.c_int64()
    _check(L.pecs_export_matrix(ctypes.byref(m), ctypes.byref(p), which, ctypes.byref(n), ctypes.byref(nnz),
                                None, None, None, None))
    rp = np.zeros(n.value + 1, np.int64)
    col = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value, np.float64)
    cons = np.zeros(n.value, np.int8)
    _check(L.pecs_export_matrix(ctypes.byref(m), ctypes.byref(p), which, None, None,
                                rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _ptr(val),
                                cons.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))))
    A = sp.csr_matrix((val, col, rp), shape=(n.value, n.value))
    return (A, cons.astype(bool)) if which == POISSON else A


LAYOUT_OWN, LAYOUT_DEALII = 0, 1


def poisson_layout_convert(prm, v, to_dealii=True, rhs=False):
    """One Poisson vector between this library's layout and deal.II's
    (FE_RaviartThomas node functionals, distribute_dofs + component_wise
    numbering); rhs=True for a right-hand side (dual vector)."""
    L = lib()
    L.pecs_poisson_layout_convert.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(CParams), ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double)]
    m, p = to_c(prm)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(v)
    _check(L.pecs_poisson_layout_convert(ctypes.byref(m), ctypes.byref(p), int(bool(to_dealii)), int(bool(rhs)),
                                         _ptr(v), _ptr(out)))
    return out


def write_vtu(prm, state, step, out_dir=".", char_length=1.0, char_density=1.0, char_time=1.0):
    """print_results: Semiconductor-NNN.vtu, Electrolyte-NNN.vtu, Poisson-NNN.vtu
    in out_dir from a state dict (keys n, p, r, o, P); host only."""
    L = lib()
    L.pecs_write_vtu.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(CParams)] + \
        [ctypes.POINTER(ctypes.c_double)] * 5 + [ctypes.c_int, ctypes.c_char_p] + [ctypes.c_double] * 3
    m, p = to_c(prm)
    v = [np.ascontiguousarray(state[c], dtype=np.float64) for c in "nproP"]
    _check(L.pecs_write_vtu(ctypes.byref(m), ctypes.byref(p), *[_ptr(a) for a in v], int(step),
                            os.fsencode(out_dir), float(char_length), float(char_density), float(char_time)))


def scale_parameters(phys):
    """Parameters::parse_and_scale_parameters through the library (dict in, dict out)."""
    c = CPhysical(**phys)
    out = CScaled()
    _check(lib().pecs_scale_parameters(ctypes.byref(c), ctypes.byref(out)))
    d = {name: getattr(out, name) for name, _ in CScaled._fields_}
    d["mobility"] = list(out.mobility)
    d["recombo_t"] = list(out.recombo_t)
    d["recombo_v"] = list(out.recombo_v)
    return d


class SolarCellGPU:
    """One problem instance on the GPU (SolarCellProblem roles, SolarCell.hpp:105-757).

    Given a list of parameter dicts, a batch of instances on one mesh that
    share the constant matrices and factors (pecs_create_batch): they may
    differ only in right-hand-side data (biases, generation, doping, rates).
    get_state / set_state / residuals then take the instance index."""

    def __init__(self, prm):
        L = lib()
        prms = [dict(q) for q in prm] if isinstance(prm, (list, tuple)) else [dict(prm)]
        self.prm = prms[0]
        self.prms = prms
        self.batch = len(prms)
        cm = [to_c(q) for q in prms]
        m = cm[0][0]
        params = (CParams * len(prms))(*[q[1] for q in cm])
        ctx = ctypes.c_void_p()
        _check(L.pecs_create_batch(ctypes.byref(m), params, len(prms), ctypes.byref(ctx)))
        self.ctx = ctx
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.pecs_sizes(ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        self.n_semi, self.n_elec, self.n_poisson = a.value, b.value, c.value

    def close(self):
        if self.ctx:
            lib().pecs_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.ctx)

    def stats(self):
        out = np.zeros(12)
        self._c(lib().pecs_stats(self.ctx, _ptr(out), 12))
        keys = ["factor_semi_doubles", "factor_elec_doubles", "factor_poisson_doubles", "setup_s", "factor_s",
                "levels_semi", "levels_elec", "levels_poisson", "n_schur_semi", "n_multipliers",
                "device_numeric_s", "device_numeric_flops"]
        return dict(zip(keys, out.tolist()))

    def sweep_config(self, which):
        """Tile shapes of the triangular sweeps per level: (fwd lanes, fwd passes,
        fwd rows/tile, bwd lanes, bwd passes, bwd rows/tile)."""
        out = (ctypes.c_int * (6 * 64))()
        nl = ctypes.c_int()
        L = lib()
        L.pecs_sweep_config.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int)]
        self._c(L.pecs_sweep_config(self.ctx, which, out, 6 * 64, ctypes.byref(nl)))
        return [tuple(out[6 * l:6 * l + 6]) for l in range(nl.value)]

    def set_mixed_precision(self, enable=True):
        """FP32 cell-local contractions (FP64 state, assembly and solves);
        needs the wave-uniform cell lists (planar meshes)."""
        L = lib()
        L.pecs_set_mixed_precision.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self._c(L.pecs_set_mixed_precision(self.ctx, 1 if enable else 0))

    def set_dof_layout(self, layout):
        """Poisson layout of the host-pointer calls: LAYOUT_OWN or LAYOUT_DEALII."""
        L = lib()
        L.pecs_set_dof_layout.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self._c(L.pecs_set_dof_layout(self.ctx, int(layout)))

    def set_manufactured(self, target, src=None, dir=None, flux=None):
        """Manufactured-solution data of the right-hand sides (test hooks):
        target 0 / 1 the semiconductor / electrolyte carriers (src at the
        cell points, dir the Dirichlet datum, flux a normal flux on
        interface and Neumann faces), 2 the Poisson right-hand side (src
        added to the charge, dir the Dirichlet datum).  None switches a term
        off; arrays shaped as quadrature_points returns the points."""
        L = lib()
        dp = ctypes.POINTER(ctypes.c_double)
        L.pecs_set_manufactured.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
        nc, nq, nfq = self._qsizes(target)
        arrs = []
        for a, n in ((src, nc * nq), (dir, nc * 4 * nfq), (flux, nc * 4 * nfq)):
            if a is None:
                arrs.append(None)
                continue
            a = self._f(a).ravel()
            assert a.size == n, (a.size, n)
            arrs.append(a)
        self._c(L.pecs_set_manufactured(self.ctx, int(target), *[None if a is None else _ptr(a) for a in arrs]))

    def sample_grid(self, G, instance=0):
        """Densities of the four carriers at the G x G logical cell centres of
        their regions: array [4, G, G] (n, p, r, o; row = y index)."""
        L = lib()
        L.pecs_sample_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        out = np.zeros((4, G, G))
        self._c(L.pecs_sample_grid(self.ctx, int(instance), int(G), _ptr(out)))
        return out

    def density_change(self):
        """Per instance, the relative change of the four densities over the
        last step (pecs_last_density_change)."""
        L = lib()
        L.pecs_last_density_change.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        out = np.zeros(self.batch)
        self._c(L.pecs_last_density_change(self.ctx, _ptr(out)))
        return out

    def set_state_from_grid(self, grid, instance=0):
        """Densities from a [4, G, G] grid (nodal values of its bilinear
        interpolant), currents zero, the potential solved from them."""
        L = lib()
        L.pecs_set_state_from_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double)]
        g = self._f(grid)
        assert g.ndim == 3 and g.shape[0] == 4 and g.shape[1] == g.shape[2], g.shape
        self._c(L.pecs_set_state_from_grid(self.ctx, int(instance), int(g.shape[1]), _ptr(g)))

    def _qsizes(self, target):
        L = lib()
        L.pecs_quadrature_points.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        nc, nq, nfq = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        self._c(L.pecs_quadrature_points(self.ctx, int(target), ctypes.byref(nc), ctypes.byref(nq),
                                         ctypes.byref(nfq), None, None))
        return nc.value, nq.value, nfq.value

    def quadrature_points(self, target):
        """(cell_xy [ncell, nq, 2], face_xy [ncell, 4, nfq, 2]) of target 0 / 1 / 2."""
        nc, nq, nfq = self._qsizes(target)
        cx = np.zeros((nc, nq, 2))
        fx = np.zeros((nc, 4, nfq, 2))
        self._c(lib().pecs_quadrature_points(self.ctx, int(target), None, None, None, _ptr(cx), _ptr(fx)))
        return cx, fx

    def uniform_lists(self):
        """Bit flags of the cell-local kernels that take wave-uniform cell
        lists: 1/2 semiconductor Schur recovery/right-hand side, 4/8 the same
        for the electrolyte, 16 Poisson condensation and recovery."""
        f = ctypes.c_int()
        L = lib()
        L.pecs_uniform_lists.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        self._c(L.pecs_uniform_lists(self.ctx, ctypes.byref(f)))
        return f.value

    @staticmethod
    def _f(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def rhs_semiconductor(self, un, up, P, ur, uo):
        bn = np.zeros(self.n_semi)
        bp = np.zeros(self.n_semi)
        a = [self._f(x) for x in (un, up, P, ur, uo)]
        self._c(lib().pecs_rhs_semiconductor(self.ctx, *[_ptr(x) for x in a], _ptr(bn), _ptr(bp)))
        return bn, bp

    def rhs_electrolyte(self, ur, uo, P, un, up):
        br = np.zeros(self.n_elec)
        bo = np.zeros(self.n_elec)
        a = [self._f(x) for x in (ur, uo, P, un, up)]
        self._c(lib().pecs_rhs_electrolyte(self.ctx, *[_ptr(x) for x in a], _ptr(br), _ptr(bo)))
        return br, bo

    def rhs_poisson(self, un, up, ur, uo):
        b = np.zeros(self.n_poisson)
        a = [self._f(x) for x in (un, up, ur, uo)]
        self._c(lib().pecs_rhs_poisson(self.ctx, *[_ptr(x) for x in a], _ptr(b)))
        return b

    def solve(self, which, b):
        n = {ELECTRONS: self.n_semi, HOLES: self.n_semi, REDUCTANTS: self.n_elec, OXIDANTS: self.n_elec,
             POISSON: self.n_poisson}[which]
        b = self._f(b)
        assert b.shape == (n,)
        x = np.zeros(n)
        self._c(lib().pecs_solve(self.ctx, which, _ptr(b), _ptr(x)))
        return x

    def set_state(self, st, instance=0):
        a = [self._f(st[k]) for k in "npro"]
        P = _ptr(self._f(st["P"])) if st.get("P") is not None else None
        self._c(lib().pecs_set_state_batch(self.ctx, int(instance), *[_ptr(x) for x in a], P))

    def get_state(self, instance=0):
        st = {k: np.zeros(self.n_semi if k in "np" else self.n_elec) for k in "npro"}
        st["P"] = np.zeros(self.n_poisson)
        self._c(lib().pecs_get_state_batch(self.ctx, int(instance), *[_ptr(st[k]) for k in "npro"],
                                           _ptr(st["P"])))
        return st

    def residuals(self):
        """Relative steady residual of each instance after the last step call."""
        r = np.zeros(self.batch)
        self._c(lib().pecs_last_residual_batch(self.ctx, _ptr(r)))
        return r

    def initial_state(self):
        self._c(lib().pecs_initial_state(self.ctx))

    def step(self, nsteps=1):
        self._c(lib().pecs_step(self.ctx, int(nsteps)))

    def enable_timing(self, enable):
        """Per-phase hipEvent timing of pecs_step (steps then launch kernel by kernel)."""
        L = lib()
        L.pecs_enable_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self._c(L.pecs_enable_timing(self.ctx, int(bool(enable))))

    def use_graph(self, enable):
        """Replay the step as a captured HIP graph (default) or launch kernel by kernel."""
        L = lib()
        L.pecs_use_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self._c(L.pecs_use_graph(self.ctx, int(bool(enable))))

    def residual(self):
        r = ctypes.c_double()
        self._c(lib().pecs_last_residual(self.ctx, ctypes.byref(r)))
        return r.value


class SolarCellSweep:
    """Contexts with different meshes stepped together (pecs_sweep_*): every
    active member's captured step graph per time step, spread over nstreams
    streams so that the members' steps overlap."""

    def __init__(self, members, nstreams=1):
        L = lib()
        L.pecs_sweep_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.pecs_sweep_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.pecs_sweep_set_active.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.pecs_sweep_destroy.argtypes = [ctypes.c_void_p]
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.ctx for m in self.members])
        self.h = ctypes.c_void_p()
        _check(L.pecs_sweep_create(arr, len(self.members), int(nstreams), ctypes.byref(self.h)))

    def set_active(self, mask):
        flags = (ctypes.c_int * len(self.members))(*[int(bool(a)) for a in mask])
        _check(lib().pecs_sweep_set_active(self.h, flags))

    def step(self, nsteps=1):
        _check(lib().pecs_sweep_step(self.h, int(nsteps)))

    def residuals(self):
        """Steady residual of every member (max over its batch instances)."""
        return np.array([m.residuals().max() for m in self.members])

    def close(self):
        if self.h:
            lib().pecs_sweep_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WarmStartNet:
    """The warm-start network (pecs_mlp_*): nin -> nlayers x nhidden (ReLU) ->
    nout, trained in FP32 on the GPU, evaluated in BF16 (or FP32)."""

    def __init__(self, nin, nhidden, nlayers, nout, seed=0, decoder=None):
        """decoder=(c0, g0, [(cout, up), ...]): the encoder's last layer is a
        c0 x g0 x g0 feature map and each stage upsamples by 2 (up) and
        applies a 3x3 convolution to cout channels; nout is then ignored (the
        output is cout_last x G x G per sample)."""
        L = lib()
        fp = ctypes.POINTER(ctypes.c_float)
        L.pecs_mlp_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_uint, ctypes.POINTER(ctypes.c_void_p)]
        L.pecs_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.pecs_mlp_num_params.argtypes = [ctypes.c_void_p]
        L.pecs_mlp_num_params.restype = ctypes.c_int64
        L.pecs_mlp_get_params.argtypes = [ctypes.c_void_p, fp]
        L.pecs_mlp_set_params.argtypes = [ctypes.c_void_p, fp]
        L.pecs_mlp_gradients.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_int, fp, fp]
        L.pecs_mlp_train_step.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_int, ctypes.c_float, fp]
        L.pecs_mlp_predict.argtypes = [ctypes.c_void_p, fp, ctypes.c_int, fp, ctypes.c_int]
        L.pecs_mlp_last_error.restype = ctypes.c_char_p
        L.pecs_mlp_create_decoder.argtypes = [ctypes.c_int] * 6 + [ctypes.POINTER(ctypes.c_int)] * 2 + \
            [ctypes.c_uint, ctypes.POINTER(ctypes.c_void_p)]
        L.pecs_mlp_num_outputs.argtypes = [ctypes.c_void_p]
        L.pecs_mlp_num_outputs.restype = ctypes.c_int64
        self.h = ctypes.c_void_p()
        if decoder is None:
            self._c(L.pecs_mlp_create(nin, nhidden, nlayers, nout, seed, ctypes.byref(self.h)))
        else:
            c0, g0, stages = decoder
            co = (ctypes.c_int * len(stages))(*[int(c) for c, _ in stages])
            up = (ctypes.c_int * len(stages))(*[int(u) for _, u in stages])
            self._c(L.pecs_mlp_create_decoder(nin, nhidden, nlayers, c0, g0, len(stages), co, up, seed,
                                              ctypes.byref(self.h)))
        self.nin = nin
        self.nout = int(L.pecs_mlp_num_outputs(self.h))
        self.nparams = int(L.pecs_mlp_num_params(self.h))

    @staticmethod
    def _c(rc):
        if rc != 0:
            msg = lib().pecs_mlp_last_error()
            raise RuntimeError("pecs_mlp: " + (msg.decode() if msg else f"error {rc}"))

    @staticmethod
    def _f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def params(self):
        out = np.zeros(self.nparams, np.float32)
        self._c(lib().pecs_mlp_get_params(self.h, self._f(out)[1]))
        return out

    def set_params(self, p):
        a, ptr = self._f(p)
        assert a.size == self.nparams
        self._c(lib().pecs_mlp_set_params(self.h, ptr))

    def gradients(self, x, y):
        xa, xp = self._f(x)
        ya, yp = self._f(y)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_float()
        self._c(lib().pecs_mlp_gradients(self.h, xp, yp, xa.shape[0], self._f(g)[1], ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, x, y, lr):
        xa, xp = self._f(x)
        ya, yp = self._f(y)
        loss = ctypes.c_float()
        self._c(lib().pecs_mlp_train_step(self.h, xp, yp, xa.shape[0], float(lr), ctypes.byref(loss)))
        return loss.value

    def predict(self, x, bf16=True):
        xa, xp = self._f(x)
        y = np.zeros((xa.shape[0], self.nout), np.float32)
        self._c(lib().pecs_mlp_predict(self.h, xp, xa.shape[0], self._f(y)[1], int(bool(bf16))))
        return y

    def close(self):
        if self.h:
            lib().pecs_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
