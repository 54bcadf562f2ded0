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
Boundary-layer local refinement (n_local_refine > 0, Grid.cpp:72-111) is not
restated: none of the measured configurations uses it.
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
