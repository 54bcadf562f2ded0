"""Structured region meshes, boundary ids and maps (oracle, test infrastructure).

Restates:
  * Grid_Maker::Grid::make_semiconductor_grid / make_electrolyte_grid /
    make_merged_grid + refine_global (Grid.cpp:44-336): one coarse quad per
    region, refined r times.  New vertices are edge midpoints and the cell
    centre (average of the 4 vertices), as global refinement of a straight
    quad produces them.
  * make_Dirichlet/Neumann/Schottky_boundaries (Grid.cpp:340-461): boundary
    ids set by exact comparisons of the face centre with 0, L, H and r1.
  * deal.II cell order after global refinement of one coarse cell: level
    index 4*parent + child, child 0..3 = (0,0),(1,0),(0,1),(1,1), i.e. the
    Morton (Z) order of (i, j) with the x bit lowest.  The merged Poisson mesh
    lists the semiconductor cells first, then the electrolyte cells.
  * setup_mappings (SolarCell.cpp:156-372): the interface pairing
    (semiconductor cell (N-1, j) face 1 <-> electrolyte cell (0, j) face 0)
    and the region -> Poisson cell maps, which for these meshes are index
    maps (the reference finds them by an O(N^2) search on cell centres).
  * boundary-layer local refinement (n_local_refine = 1, Grid.cpp:11-44,
    72-111): LocalRefinedMesh below (deeper local refinement makes deal.II
    refine bulk cells too to keep one level between face neighbours, and is
    not restated).
"""
import numpy as np

INTERFACE, DIRICHLET, NEUMANN, SCHOTTKY = 0, 1, 2, 3
INTERIOR = -1
SEMI_ID, SEMI_BL_ID, ELEC_ID, ELEC_BL_ID = 0, 1, 2, 3


def morton_ij(c, nbits):
    i = 0
    j = 0
    for b in range(nbits):
        i |= ((c >> (2 * b)) & 1) << b
        j |= ((c >> (2 * b + 1)) & 1) << b
    return i, j


def vertex_grid(coarse, r):
    """Vertex array G[i, j] (shape (N+1, N+1, 2)) of a coarse quad refined r times.

    coarse = [v0, v1, v2, v3] in deal.II order.
    """
    G = np.empty((2, 2, 2))
    G[0, 0], G[1, 0], G[0, 1], G[1, 1] = coarse
    for _ in range(r):
        n = G.shape[0] - 1
        H = np.empty((2 * n + 1, 2 * n + 1, 2))
        H[0::2, 0::2] = G
        H[1::2, 0::2] = (G[:-1, :] + G[1:, :]) / 2.0
        H[0::2, 1::2] = (G[:, :-1] + G[:, 1:]) / 2.0
        H[1::2, 1::2] = (G[:-1, :-1] + G[1:, :-1] + G[:-1, 1:] + G[1:, 1:]) / 4.0
        G = H
    return G


def classify_face(cx, cy, geo, insulated, schottky):
    """Boundary id of a boundary face with centre (cx, cy) (Grid.cpp:340-461)."""
    L, H, r1 = geo["L"], geo["H"], geo["r1"]
    bid = INTERFACE
    if cx == 0.0 or cx == L or cy == 0.0 or cy == H:
        bid = DIRICHLET
    if insulated:
        if cy == H:
            bid = NEUMANN if cx > r1 else DIRICHLET
        if cy == 0.0:
            bid = NEUMANN if cx > r1 else DIRICHLET
        if cx == 0.0:
            bid = NEUMANN
    if schottky and cy == H:
        bid = SCHOTTKY
    return bid


# face f of cell (i,j): vertex pair and neighbour offset
FACE_VERTS = {0: (0, 2), 1: (1, 3), 2: (0, 1), 3: (2, 3)}
FACE_NBR = {0: (-1, 0), 1: (1, 0), 2: (0, -1), 3: (0, 1)}


class Region:
    """A refined region (one coarse quad), cells in Morton order."""

    def __init__(self, coarse, r, material, geo, insulated, schottky, classify=None):
        if classify is None:
            classify = lambda cx, cy: classify_face(cx, cy, geo, insulated, schottky)
        self.r = r
        self.N = 1 << r
        N = self.N
        self.ncell = N * N
        self.material = material
        G = vertex_grid(np.asarray(coarse, dtype=float), r)
        self.ij = np.array([morton_ij(c, r) for c in range(self.ncell)], dtype=np.int64)
        self.cell_of = np.empty((N, N), dtype=np.int64)
        for c, (i, j) in enumerate(self.ij):
            self.cell_of[i, j] = c
        V = np.empty((self.ncell, 4, 2))
        for c, (i, j) in enumerate(self.ij):
            V[c, 0] = G[i, j]
            V[c, 1] = G[i + 1, j]
            V[c, 2] = G[i, j + 1]
            V[c, 3] = G[i + 1, j + 1]
        self.V = V
        self.nbr = np.full((self.ncell, 4), -1, dtype=np.int64)
        self.bid = np.full((self.ncell, 4), INTERIOR, dtype=np.int64)
        for c, (i, j) in enumerate(self.ij):
            for f in range(4):
                di, dj = FACE_NBR[f]
                ii, jj = i + di, j + dj
                if 0 <= ii < N and 0 <= jj < N:
                    self.nbr[c, f] = self.cell_of[ii, jj]
                else:
                    a, b = FACE_VERTS[f]
                    cen = (V[c, a] + V[c, b]) / 2.0
                    self.bid[c, f] = classify(cen[0], cen[1])


class SolarCellMesh:
    """Semiconductor, electrolyte and merged Poisson meshes + maps."""

    def __init__(self, r, geo, insulated=False, schottky=False):
        r1, r2, H, L = geo["r1"], geo["r2"], geo["H"], geo["L"]
        self.geo = geo
        self.r = r
        self.N = 1 << r
        semi_coarse = [(0.0, 0.0), (r2, 0.0), (0.0, H), (r1, H)]
        elec_coarse = [(r2, 0.0), (L, 0.0), (r1, H), (L, H)]
        self.semi = Region(semi_coarse, r, SEMI_ID, geo, insulated, schottky)
        self.elec = Region(elec_coarse, r, ELEC_ID, geo, insulated, False)
        # merged Poisson mesh: Grid.cpp:322-336 -- ids re-derived on the merged
        # triangulation (Dirichlet, Neumann and Schottky all applied to it).
        semiP = Region(semi_coarse, r, SEMI_ID, geo, insulated, schottky)
        elecP = Region(elec_coarse, r, ELEC_ID, geo, insulated, schottky)
        N = self.N
        self.pV = np.concatenate([semiP.V, elecP.V])
        self.pmat = np.array([SEMI_ID] * semiP.ncell + [ELEC_ID] * elecP.ncell)
        self.pIJ = np.concatenate([semiP.ij, elecP.ij + np.array([N, 0])])
        self.NX, self.NY = 2 * N, N
        self.pbid = np.concatenate([semiP.bid, elecP.bid])
        # the region interface is interior in the merged mesh
        for c, (i, j) in enumerate(semiP.ij):
            if i == N - 1:
                self.pbid[c, 1] = INTERIOR
        for c, (i, j) in enumerate(elecP.ij):
            if i == 0:
                self.pbid[semiP.ncell + c, 0] = INTERIOR
        # interface pairing (SolarCell.cpp:163-255): ordered by semiconductor cells
        self.sigma_semi = []
        self.sigma_elec = []
        for c, (i, j) in enumerate(self.semi.ij):
            if self.semi.bid[c, 1] == INTERFACE and i == N - 1:
                self.sigma_semi.append(c)
                self.sigma_elec.append(self.elec.cell_of[0, j])
        self.semi_sigma_nb = {s: e for s, e in zip(self.sigma_semi, self.sigma_elec)}
        self.elec_sigma_nb = {e: s for s, e in zip(self.sigma_semi, self.sigma_elec)}
        # region -> Poisson cell maps (s_2_p_map / e_2_p_map)
        self.s2p = np.arange(self.semi.ncell)
        self.e2p = self.semi.ncell + np.arange(self.elec.ncell)


def carrier_dofs(ncell, nb):
    """Component-wise carrier layout [Jx | Jy | u] (CarrierPair.cpp:23-64).

    Returns idx[ncell, 3*nb] of global DoF indices in local order [Jx,Jy,u].
    """
    c = np.arange(ncell)[:, None]
    a = np.arange(nb)[None, :]
    jx = c * nb + a
    jy = ncell * nb + c * nb + a
    u = 2 * ncell * nb + c * nb + a
    return np.concatenate([jx, jy, u], axis=1)


def poisson_dofs(mesh, deg):
    """Global [D | Phi] DoF indices of every Poisson cell (own layout).

    D block: vertical faces (j*(NX+1)+i, deg+1 DoFs each), horizontal faces
    (j*NX+i), then per cell its interior x and y DoFs; Phi block: per cell
    (deg+1)^2 DoFs.  Returns (idx[ncellP, nloc], nD, ntotal).
    """
    NX, NY = mesh.NX, mesh.NY
    nf = deg + 1
    nV = (NX + 1) * NY * nf
    nH = NX * (NY + 1) * nf
    nint = 2 * deg * (deg + 1)
    ncp = len(mesh.pV)
    nD = nV + nH + ncp * nint
    nphi = (deg + 1) ** 2
    nx = (deg + 2) * (deg + 1)
    nloc = 2 * nx + nphi
    idx = np.empty((ncp, nloc), dtype=np.int64)
    for p in range(ncp):
        I, J = mesh.pIJ[p]
        ib = nV + nH + p * nint
        for b in range(deg + 1):
            for a in range(deg + 2):
                li = a + (deg + 2) * b
                if a == 0:
                    g = (J * (NX + 1) + I) * nf + b
                elif a == deg + 1:
                    g = (J * (NX + 1) + I + 1) * nf + b
                else:
                    g = ib + (a - 1) + deg * b
                idx[p, li] = g
        for b in range(deg + 2):
            for a in range(deg + 1):
                li = nx + a + (deg + 1) * b
                if b == 0:
                    g = nV + (J * NX + I) * nf + a
                elif b == deg + 1:
                    g = nV + ((J + 1) * NX + I) * nf + a
                else:
                    g = ib + deg * (deg + 1) + a + (deg + 1) * (b - 1)
                idx[p, li] = g
        for l in range(nphi):
            idx[p, 2 * nx + l] = nD + p * nphi + l
    return idx, nD, nD + ncp * nphi


# ---------------------------------------------------------------------------
# Boundary-layer local refinement (Grid.cpp:11-44, 72-111, 135-336)
# ---------------------------------------------------------------------------
# With n_local_refine = 1 and a boundary layer of width bl > 0 each region is
# two coarse quads side by side -- semiconductor: bulk [0, r-bl] then layer
# [r-bl, r]; electrolyte: layer [r, r+bl] then bulk [r+bl, L] (r running
# from r2 at y = 0 to r1 at y = H) -- both refined r times, after which the
# layer cells (material ids 1 and 3) are refined once more.  The bulk/layer
# face is then hanging: one bulk cell face meets two layer cell faces
# (subface 0 = the half at the face's first vertex, i.e. the lower half).
# deal.II lists active cells level by level, so a region's cells are its
# level-r block (Morton) followed by its level-(r+1) block (Morton over the
# 2N x 2N children, index 4*parent + child).  The merged Poisson mesh here
# lists the semiconductor cells then the electrolyte cells, each in its
# region's order, so the region -> Poisson maps stay index maps.
#
# Grid positions: every block is laid on one fine lattice (2 positions per
# level-r cell width, 1 per level-(r+1) cell); a cell's position is its lower
# left lattice point.  They are used by the nested-dissection ordering only.

class BlockRegion:
    """Structured blocks side by side (left to right), each one coarse quad
    refined r + lev times; cells listed in the block order `order`."""

    def __init__(self, blocks, r, order, classify):
        # blocks: list of (coarse quad, lev in {0, 1}, material)
        self.r = r
        self.N = 1 << r
        gscale = 2 if any(b[1] for b in blocks) else 1
        self.blocks = []
        Vs, mats, ijs, gij, levs, pd = [], [], [], [], [], []
        base = {}
        c0 = 0
        for bi in order:
            coarse, lev, mat = blocks[bi]
            rr = r + lev
            n = 1 << rr
            G = vertex_grid(np.asarray(coarse, dtype=float), rr)
            ij = np.array([morton_ij(c, rr) for c in range(n * n)], dtype=np.int64)
            V = np.stack([G[ij[:, 0], ij[:, 1]], G[ij[:, 0] + 1, ij[:, 1]],
                          G[ij[:, 0], ij[:, 1] + 1], G[ij[:, 0] + 1, ij[:, 1] + 1]], 1)
            if lev:
                I, J = 2 * (ij[:, 0] // 2), 2 * (ij[:, 1] // 2)
                Vp = np.stack([G[I, J], G[I + 2, J], G[I, J + 2], G[I + 2, J + 2]], 1)
                pd.append(_diam(Vp))
            else:
                pd.append(_diam(V))
            sc = gscale >> lev
            gx0 = bi * (gscale << r)
            gij.append(np.stack([gx0 + sc * ij[:, 0], sc * ij[:, 1]], 1))
            cell_of = np.empty((n, n), dtype=np.int64)
            cell_of[ij[:, 0], ij[:, 1]] = c0 + np.arange(n * n)
            base[bi] = dict(n=n, lev=lev, cell_of=cell_of, c0=c0)
            Vs.append(V)
            mats.append(np.full(n * n, mat))
            ijs.append(ij)
            levs.append(np.full(n * n, lev))
            c0 += n * n
        self.V = np.concatenate(Vs)
        self.mat = np.concatenate(mats)
        self.material = int(blocks[0][2])
        self.lij = np.concatenate(ijs)       # (i, j) inside the cell's block
        self.gij = np.concatenate(gij)       # lattice position
        self.level = np.concatenate(levs)
        self.pdiam = np.concatenate(pd)      # diameter of the level-r cell containing the cell
        self.NXg = len(blocks) * (gscale << r)
        self.NYg = gscale << r
        self.ncell = len(self.V)
        self.block_info = base
        nc = self.ncell
        self.nbr = np.full((nc, 4), -1, dtype=np.int64)
        self.bid = np.full((nc, 4), INTERIOR, dtype=np.int64)
        hang = []
        for bi in range(len(blocks)):
            B = base[bi]
            n, co = B["n"], B["cell_of"]
            for i in range(n):
                for j in range(n):
                    c = co[i, j]
                    for f in range(4):
                        di, dj = FACE_NBR[f]
                        ii, jj = i + di, j + dj
                        if 0 <= ii < n and 0 <= jj < n:
                            self.nbr[c, f] = co[ii, jj]
            # faces on the block's right edge meet block bi + 1
            if bi + 1 < len(blocks):
                R = base[bi + 1]
                if R["lev"] == B["lev"]:
                    for j in range(n):
                        a, b = co[n - 1, j], R["cell_of"][0, j]
                        self.nbr[a, 1], self.nbr[b, 0] = b, a
                elif B["lev"] < R["lev"]:      # coarse left, fine right
                    for j in range(n):
                        for s in range(2):
                            hang.append((co[n - 1, j], 1, s, R["cell_of"][0, 2 * j + s], 0))
                else:                          # fine left, coarse right
                    m = R["n"]
                    for j in range(m):
                        for s in range(2):
                            hang.append((R["cell_of"][0, j], 0, s, co[n - 1, 2 * j + s], 1))
        self.hang = np.array(hang, dtype=np.int64).reshape(-1, 5)
        hanging = np.zeros((nc, 4), dtype=bool)
        for c, f, s, e, fe_ in self.hang:
            hanging[c, f] = True
            hanging[e, fe_] = True
        for c in range(nc):
            for f in range(4):
                if self.nbr[c, f] < 0 and not hanging[c, f]:
                    a, b = FACE_VERTS[f]
                    cen = (self.V[c, a] + self.V[c, b]) / 2.0
                    self.bid[c, f] = classify(cen[0], cen[1])


def _diam(V):
    d1 = np.linalg.norm(V[:, 3] - V[:, 0], axis=-1)
    d2 = np.linalg.norm(V[:, 2] - V[:, 1], axis=-1)
    return np.maximum(d1, d2)


class LocalRefinedMesh:
    """SolarCellMesh with boundary-layer local refinement (n_local_refine = 1)."""

    def __init__(self, r, geo, bl, insulated=False, schottky=False):
        if not bl > 0.0:
            raise ValueError("local refinement needs a boundary layer of positive width")
        r1, r2, H, L = geo["r1"], geo["r2"], geo["H"], geo["L"]
        self.geo = geo
        self.r = r
        self.N = 1 << r
        self.local = 1
        sb = [(0.0, 0.0), (r2 - bl, 0.0), (0.0, H), (r1 - bl, H)]
        sl = [(r2 - bl, 0.0), (r2, 0.0), (r1 - bl, H), (r1, H)]
        el = [(r2, 0.0), (r2 + bl, 0.0), (r1, H), (r1 + bl, H)]
        eb = [(r2 + bl, 0.0), (L, 0.0), (r1 + bl, H), (L, H)]
        cs = lambda cx, cy: classify_face(cx, cy, geo, insulated, schottky)
        ce = lambda cx, cy: classify_face(cx, cy, geo, insulated, False)
        cp = lambda cx, cy: classify_face(cx, cy, geo, insulated, schottky)
        self.semi = BlockRegion([(sb, 0, SEMI_ID), (sl, 1, SEMI_BL_ID)], r, [0, 1], cs)
        self.elec = BlockRegion([(el, 1, ELEC_BL_ID), (eb, 0, ELEC_ID)], r, [1, 0], ce)
        P = BlockRegion([(sb, 0, SEMI_ID), (sl, 1, SEMI_BL_ID), (el, 1, ELEC_BL_ID), (eb, 0, ELEC_ID)],
                        r, [0, 1, 3, 2], cp)
        self.P = P
        self.pV = P.V
        self.pmat = P.mat
        self.pbid = P.bid
        self.pnbr = P.nbr
        self.phang = P.hang
        self.pIJ = P.gij
        self.NX, self.NY = P.NXg, P.NYg
        # interface pairs: semiconductor layer cells of the last column with
        # the electrolyte layer cells of the first (both level r+1)
        bs = self.semi.block_info[1]
        be = self.elec.block_info[0]
        n = bs["n"]
        self.sigma_semi, self.sigma_elec = [], []
        for j in range(n):
            s, e = bs["cell_of"][n - 1, j], be["cell_of"][0, j]
            if self.semi.bid[s, 1] == INTERFACE:
                self.sigma_semi.append(int(s))
                self.sigma_elec.append(int(e))
        self.semi_sigma_nb = {s: e for s, e in zip(self.sigma_semi, self.sigma_elec)}
        self.elec_sigma_nb = {e: s for s, e in zip(self.sigma_semi, self.sigma_elec)}
        self.s2p = np.arange(self.semi.ncell)
        self.e2p = self.semi.ncell + np.arange(self.elec.ncell)


def poisson_dofs_general(mesh, deg):
    """[D | Phi] DoF indices of every Poisson cell on a mesh with hanging faces.

    Face blocks (deg+1 DoFs) are numbered on first encounter, cells in order
    and faces 0..3, a conforming interior face sharing its neighbour's block;
    the two sides of a hanging face keep their own blocks (the coarse face,
    and each fine face).  Interior RT DoFs and Phi DoFs follow per cell as in
    poisson_dofs.  Returns (idx, nD, ntotal, face_base[nc, 4]).
    """
    nc = len(mesh.pV)
    nf = deg + 1
    fb = np.full((nc, 4), -1, dtype=np.int64)
    cnt = 0
    for c in range(nc):
        for f in range(4):
            e = mesh.pnbr[c, f]
            if e >= 0 and fb[e, OPP_FACE[f]] >= 0:
                fb[c, f] = fb[e, OPP_FACE[f]]
            else:
                fb[c, f] = cnt
                cnt += nf
    nint = 2 * deg * (deg + 1)
    nD = cnt + nc * nint
    nphi = (deg + 1) ** 2
    nx = (deg + 2) * (deg + 1)
    nloc = 2 * nx + nphi
    idx = np.empty((nc, nloc), dtype=np.int64)
    for p in range(nc):
        ib = cnt + p * nint
        for b in range(deg + 1):
            for a in range(deg + 2):
                li = a + (deg + 2) * b
                if a == 0:
                    g = fb[p, 0] + b
                elif a == deg + 1:
                    g = fb[p, 1] + b
                else:
                    g = ib + (a - 1) + deg * b
                idx[p, li] = g
        for b in range(deg + 2):
            for a in range(deg + 1):
                li = nx + a + (deg + 1) * b
                if b == 0:
                    g = fb[p, 2] + a
                elif b == deg + 1:
                    g = fb[p, 3] + a
                else:
                    g = ib + deg * (deg + 1) + a + (deg + 1) * (b - 1)
                idx[p, li] = g
        for l in range(nphi):
            idx[p, 2 * nx + l] = nD + p * nphi + l
    return idx, nD, nD + nc * nphi, fb


OPP_FACE = {0: 1, 1: 0, 2: 3, 3: 2}
