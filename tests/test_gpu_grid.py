"""Densities on the warm-start network's regular grid (pecs_sample_grid,
pecs_set_state_from_grid) against a NumPy evaluation of the same fields.

The grid is each region's logical G x G lattice of cell centres; the library
evaluates the DG densities there (sampling) and sets densities from grid
values as the nodal values of the grid's bilinear interpolant (projection).
Checked against the oracle's DGQ(k) tabulation and the Morton numbering of
its regions, on a planar and a conic mesh; the projected state's potential
equals the one pecs_set_state solves for the same densities.
"""
import numpy as np
import pytest

from oracle.fe import DGQ
from oracle.mesh import SolarCellMesh

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _prm(conic, k, r):
    import pecs_amd
    p = pecs_amd.workload_params(r, k, 5e-4)
    if conic:
        p.update(r2=0.62, L=1.4)
    return p


def _sample_numpy(u, reg, k, G):
    """Density block of carrier vector u at the logical cell centres of region reg."""
    dg = DGQ(k)
    N, nb = reg.N, dg.nb
    d = u[2 * reg.ncell * nb:].reshape(reg.ncell, nb)
    out = np.zeros((G, G))
    for b in range(G):
        for a in range(G):
            s, t = (a + 0.5) / G * N, (b + 0.5) / G * N
            i, j = min(int(s), N - 1), min(int(t), N - 1)
            val, _ = dg.tab(np.array([[s - i, t - j]]))
            out[b, a] = d[reg.cell_of[i, j]] @ val[:, 0]
    return out


def _project_numpy(grid, reg, k):
    """Nodal values of the bilinear interpolant of a cell-centred grid."""
    G = grid.shape[0]
    N = reg.N
    nodes = np.linspace(0.0, 1.0, k + 1)
    d = np.zeros((reg.ncell, (k + 1) ** 2))
    for c, (i, j) in enumerate(reg.ij):
        for q in range(k + 1):
            for p in range(k + 1):
                gx = np.clip((i + nodes[p]) / N * G - 0.5, 0, G - 1)
                gy = np.clip((j + nodes[q]) / N * G - 0.5, 0, G - 1)
                x0, y0 = min(int(gx), G - 2), min(int(gy), G - 2)
                fx, fy = gx - x0, gy - y0
                d[c, p + (k + 1) * q] = ((1 - fy) * ((1 - fx) * grid[y0, x0] + fx * grid[y0, x0 + 1])
                                         + fy * ((1 - fx) * grid[y0 + 1, x0] + fx * grid[y0 + 1, x0 + 1]))
    return d.ravel()


@pytest.mark.parametrize("conic,k,r,G", [(False, 2, 3, 16), (True, 1, 4, 12), (True, 3, 2, 8)])
def test_grid_sample_and_project(conic, k, r, G):
    import pecs_amd
    _say(f"[gpu] grid sample / project: conic={conic} k={k} r={r} G={G}")
    p = _prm(conic, k, r)
    m = SolarCellMesh(r, dict(r1=p["r1"], r2=p["r2"], H=p["H"], L=p["L"]))
    g = pecs_amd.SolarCellGPU(p)
    try:
        g.initial_state()
        g.step(20)
        st = g.get_state()
        grid = g.sample_grid(G)
        for ch, (c, reg) in enumerate(zip("npro", (m.semi, m.semi, m.elec, m.elec))):
            ref = _sample_numpy(st[c], reg, k, G)
            assert np.abs(grid[ch] - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0), c
        rng = np.random.default_rng(r)
        target = grid * (1.0 + 0.1 * rng.standard_normal(grid.shape))
        g.set_state_from_grid(target)
        got = g.get_state()
        for ch, (c, reg) in enumerate(zip("npro", (m.semi, m.semi, m.elec, m.elec))):
            nb = (k + 1) ** 2
            ref = _project_numpy(target[ch], reg, k)
            assert np.abs(got[c][2 * reg.ncell * nb:] - ref).max() <= 1e-14 * np.abs(ref).max(), c
            assert not got[c][:2 * reg.ncell * nb].any()  # currents zero
        # the potential is the one pecs_set_state solves for these densities
        again = dict(got)
        again["P"] = None
        g.set_state(again)
        P2 = g.get_state()["P"]
        assert np.abs(got["P"] - P2).max() <= 1e-13 * np.abs(P2).max()
    finally:
        g.close()


def test_grid_projection_reproduces_bilinear_fields():
    # a field linear in the logical coordinates is reproduced exactly at every
    # node away from the half-cell band along the edges, and sampling the
    # projection at a finer refinement returns it
    import pecs_amd
    k, r, G = 2, 4, 16
    p = _prm(False, k, r)
    m = SolarCellMesh(r, dict(r1=p["r1"], r2=p["r2"], H=p["H"], L=p["L"]))
    yy, xx = np.meshgrid((np.arange(G) + 0.5) / G, (np.arange(G) + 0.5) / G, indexing="ij")
    lin = 1.0 + 2.0 * xx - 0.5 * yy + 0.75 * xx * yy
    grid = np.stack([lin, 2 * lin, 3 * lin, 4 * lin])
    g = pecs_amd.SolarCellGPU(p)
    try:
        g.set_state_from_grid(grid)
        st = g.get_state()
        nb = (k + 1) ** 2
        nodes = np.linspace(0.0, 1.0, k + 1)
        for ch, (c, reg) in enumerate(zip("npro", (m.semi, m.semi, m.elec, m.elec))):
            d = st[c][2 * reg.ncell * nb:].reshape(reg.ncell, nb)
            for cell, (i, j) in enumerate(reg.ij):
                for q in range(k + 1):
                    for pp in range(k + 1):
                        s, t = (i + nodes[pp]) / reg.N, (j + nodes[q]) / reg.N
                        if 0.5 / G <= s <= 1 - 0.5 / G and 0.5 / G <= t <= 1 - 0.5 / G:
                            exact = (ch + 1) * (1.0 + 2.0 * s - 0.5 * t + 0.75 * s * t)
                            assert abs(d[cell, pp + (k + 1) * q] - exact) <= 1e-13
        back = g.sample_grid(G)
        assert np.abs(back - grid).max() <= 1e-13 * np.abs(grid).max()
    finally:
        g.close()


def test_last_density_change_matches_host():
    import pecs_amd
    p = [_prm(True, 2, 3), dict(_prm(True, 2, 3), phi_app=0.3)]
    g = pecs_amd.SolarCellGPU(p)
    try:
        g.initial_state()
        g.step(3)
        before = [g.get_state(i) for i in range(2)]
        g.step(1)
        after = [g.get_state(i) for i in range(2)]
        got = g.density_change()
        for i in range(2):
            num = den = 0.0
            for c in "npro":
                m = len(before[i][c]) // 3
                num += np.sum((after[i][c][2 * m:] - before[i][c][2 * m:]) ** 2)
                den += np.sum(before[i][c][2 * m:] ** 2)
            ref = np.sqrt(num / den)
            assert abs(got[i] - ref) <= 1e-12 * ref, (i, got[i], ref)
    finally:
        g.close()
