// Structured meshes and maps (host side); see mesh.hpp for the reference map.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace pecs {

int morton_i(int c, int r) {
  int i = 0;
  for (int b = 0; b < r; ++b) i |= ((c >> (2 * b)) & 1) << b;
  return i;
}
int morton_index(int i, int j) {
  int c = 0;
  for (int b = 0; b < 15; ++b) c |= (((i >> b) & 1) << (2 * b)) | (((j >> b) & 1) << (2 * b + 1));
  return c;
}
int morton_j(int c, int r) {
  int j = 0;
  for (int b = 0; b < r; ++b) j |= ((c >> (2 * b + 1)) & 1) << b;
  return j;
}

// Recursive midpoint subdivision (what refine_global does to a straight quad):
// edge vertices are midpoints, the new centre is the mean of the 4 vertices.
void vertex_grid(const double coarse[8], int r, std::vector<double>& G) {
  int n = 1;
  G.assign(2 * 2 * 2, 0.0);
  auto at = [](std::vector<double>& A, int nn, int i, int j) -> double* { return &A[2 * (i + (nn + 1) * j)]; };
  for (int c = 0; c < 2; ++c) {
    at(G, 1, 0, 0)[c] = coarse[0 + c];
    at(G, 1, 1, 0)[c] = coarse[2 + c];
    at(G, 1, 0, 1)[c] = coarse[4 + c];
    at(G, 1, 1, 1)[c] = coarse[6 + c];
  }
  for (int l = 0; l < r; ++l) {
    int m = 2 * n;
    std::vector<double> H(2 * (m + 1) * (m + 1));
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i)
        for (int c = 0; c < 2; ++c) at(H, m, 2 * i, 2 * j)[c] = at(G, n, i, j)[c];
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < 2; ++c)
          at(H, m, 2 * i + 1, 2 * j)[c] = (at(G, n, i, j)[c] + at(G, n, i + 1, j)[c]) / 2.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i <= n; ++i)
        for (int c = 0; c < 2; ++c)
          at(H, m, 2 * i, 2 * j + 1)[c] = (at(G, n, i, j)[c] + at(G, n, i, j + 1)[c]) / 2.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < 2; ++c)
          at(H, m, 2 * i + 1, 2 * j + 1)[c] =
              (at(G, n, i, j)[c] + at(G, n, i + 1, j)[c] + at(G, n, i, j + 1)[c] + at(G, n, i + 1, j + 1)[c]) / 4.0;
    G.swap(H);
    n = m;
  }
}

int classify_face(double cx, double cy, const GeometrySpec& g, bool insulated, bool schottky) {
  int bid = INTERFACE;
  if (cx == 0.0 || cx == g.L || cy == 0.0 || cy == g.H) bid = DIRICHLET;  // Grid.cpp:340-374
  if (insulated) {                                                           // Grid.cpp:376-430
    if (cy == g.H) bid = (cx > g.r1) ? NEUMANN : DIRICHLET;
    if (cy == 0.0) bid = (cx > g.r1) ? NEUMANN : DIRICHLET;
    if (cx == 0.0) bid = NEUMANN;
  }
  if (schottky && cy == g.H) bid = SCHOTTKY;                                 // Grid.cpp:432-461
  return bid;
}

static const int FV[4][2] = {{0, 2}, {1, 3}, {0, 1}, {2, 3}};

// deal.II 2D cell->diameter(): the longer diagonal
static double cell_diam(const double* v) {
  return std::max(std::hypot(v[6] - v[0], v[7] - v[1]), std::hypot(v[4] - v[2], v[5] - v[3]));
}
static const int FN[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};

Region make_region(const double coarse[8], int r, int material, const GeometrySpec& g,
                   bool insulated, bool schottky) {
  Region R;
  R.r = r;
  R.N = 1 << r;
  R.ncell = R.N * R.N;
  R.material = material;
  const int N = R.N;
  std::vector<double> G;
  vertex_grid(coarse, r, G);
  R.NXg = R.NYg = N;
  R.mat.assign(R.ncell, material);
  R.ci.resize(R.ncell);
  R.cj.resize(R.ncell);
  R.cell_of.resize(R.ncell);
  R.V.resize(8 * (size_t)R.ncell);
  R.nbr.assign(4 * (size_t)R.ncell, -1);
  R.bid.assign(4 * (size_t)R.ncell, INTERIOR);
  for (int c = 0; c < R.ncell; ++c) {
    int i = morton_i(c, r), j = morton_j(c, r);
    R.ci[c] = i;
    R.cj[c] = j;
    R.cell_of[i + N * j] = c;
  }
  for (int c = 0; c < R.ncell; ++c) {
    int i = R.ci[c], j = R.cj[c];
    const int vi[4] = {i, i + 1, i, i + 1}, vj[4] = {j, j, j + 1, j + 1};
    for (int v = 0; v < 4; ++v) {
      R.V[8 * c + 2 * v] = G[2 * (vi[v] + (N + 1) * vj[v])];
      R.V[8 * c + 2 * v + 1] = G[2 * (vi[v] + (N + 1) * vj[v]) + 1];
    }
  }
  R.pdiam.resize(R.ncell);
  for (int c = 0; c < R.ncell; ++c) R.pdiam[c] = cell_diam(&R.V[8 * c]);
  for (int c = 0; c < R.ncell; ++c) {
    int i = R.ci[c], j = R.cj[c];
    for (int f = 0; f < 4; ++f) {
      int ii = i + FN[f][0], jj = j + FN[f][1];
      if (ii >= 0 && ii < N && jj >= 0 && jj < N) {
        R.nbr[4 * c + f] = R.cell_of[ii + N * jj];
      } else {
        int a = FV[f][0], b = FV[f][1];
        double cx = (R.V[8 * c + 2 * a] + R.V[8 * c + 2 * b]) / 2.0;
        double cy = (R.V[8 * c + 2 * a + 1] + R.V[8 * c + 2 * b + 1]) / 2.0;
        R.bid[4 * c + f] = classify_face(cx, cy, g, insulated, schottky);
      }
    }
  }
  return R;
}

static SolarCellMesh make_local_mesh(int r, const GeometrySpec& g);

SolarCellMesh make_solarcell_mesh(int r, const GeometrySpec& g) {
  if (r < 1 || r > 11) throw std::runtime_error("global refinement must be in [1, 11]");
  if (g.local) return make_local_mesh(r, g);
  SolarCellMesh m;
  m.geo = g;
  m.r = r;
  m.N = 1 << r;
  const double sc[8] = {0.0, 0.0, g.r2, 0.0, 0.0, g.H, g.r1, g.H};
  const double ec[8] = {g.r2, 0.0, g.L, 0.0, g.r1, g.H, g.L, g.H};
  m.semi = make_region(sc, r, SEMI_ID, g, g.insulated, g.schottky);
  m.elec = make_region(ec, r, ELEC_ID, g, g.insulated, false);
  // merged Poisson triangulation: all three boundary passes apply to it
  Region sP = make_region(sc, r, SEMI_ID, g, g.insulated, g.schottky);
  Region eP = make_region(ec, r, ELEC_ID, g, g.insulated, g.schottky);
  const int N = m.N;
  m.NX = 2 * N;
  m.NY = N;
  m.ncellP = sP.ncell + eP.ncell;
  m.pV.resize(8 * (size_t)m.ncellP);
  m.pI.resize(m.ncellP);
  m.pJ.resize(m.ncellP);
  m.pmat.resize(m.ncellP);
  m.pbid.resize(4 * (size_t)m.ncellP);
  m.pcell_of.resize((size_t)m.NX * m.NY);
  for (int c = 0; c < sP.ncell; ++c) {
    for (int t = 0; t < 8; ++t) m.pV[8 * c + t] = sP.V[8 * c + t];
    m.pI[c] = sP.ci[c];
    m.pJ[c] = sP.cj[c];
    m.pmat[c] = SEMI_ID;
    for (int f = 0; f < 4; ++f) m.pbid[4 * c + f] = sP.bid[4 * c + f];
    if (sP.ci[c] == N - 1) m.pbid[4 * c + 1] = INTERIOR;
  }
  for (int c = 0; c < eP.ncell; ++c) {
    int p = sP.ncell + c;
    for (int t = 0; t < 8; ++t) m.pV[8 * p + t] = eP.V[8 * c + t];
    m.pI[p] = N + eP.ci[c];
    m.pJ[p] = eP.cj[c];
    m.pmat[p] = ELEC_ID;
    for (int f = 0; f < 4; ++f) m.pbid[4 * p + f] = eP.bid[4 * c + f];
    if (eP.ci[c] == 0) m.pbid[4 * p + 0] = INTERIOR;
  }
  for (int p = 0; p < m.ncellP; ++p) m.pcell_of[m.pI[p] + m.NX * m.pJ[p]] = p;
  m.pnbr.assign(4 * (size_t)m.ncellP, -1);
  for (int p = 0; p < m.ncellP; ++p)
    for (int f = 0; f < 4; ++f) {
      const int I = m.pI[p] + FN[f][0], J = m.pJ[p] + FN[f][1];
      if (I >= 0 && I < m.NX && J >= 0 && J < m.NY) m.pnbr[4 * p + f] = m.pcell_of[I + m.NX * J];
    }
  m.semi_sigma_nb.assign(m.semi.ncell, -1);
  m.elec_sigma_nb.assign(m.elec.ncell, -1);
  for (int c = 0; c < m.semi.ncell; ++c) {
    if (m.semi.ci[c] == N - 1 && m.semi.bid[4 * c + 1] == INTERFACE) {
      int e = m.elec.cell_of[0 + N * m.semi.cj[c]];
      m.sigma_semi.push_back(c);
      m.sigma_elec.push_back(e);
      m.semi_sigma_nb[c] = e;
      m.elec_sigma_nb[e] = c;
    }
  }
  return m;
}

// ---------------------------------------------------------------------------
// Boundary-layer local refinement (Grid.cpp:11-44, 72-111, 135-336)
// ---------------------------------------------------------------------------
// A region of structured blocks side by side, left to right, block b being
// the coarse quad blocks[b] refined r + lev[b] times; cells are listed block
// by block in the order `order` (each block in Morton order), and faces
// between neighbouring blocks of different levels are hanging.
struct BlockSpec {
  double coarse[8];
  int lev, material;
};

static Region make_block_region(const std::vector<BlockSpec>& blocks, int r, const std::vector<int>& order,
                                const GeometrySpec& g, bool insulated, bool schottky) {
  Region R;
  R.r = r;
  R.N = 1 << r;
  R.material = blocks[0].material;
  int gscale = 1;
  for (const BlockSpec& b : blocks)
    if (b.lev) gscale = 2;
  R.NXg = (int)blocks.size() * (gscale << r);
  R.NYg = gscale << r;
  const int nb = (int)blocks.size();
  std::vector<int> c0(nb), nside(nb);
  std::vector<std::vector<int>> cell_of(nb);
  int cnt = 0;
  for (int bi : order) {
    const BlockSpec& B = blocks[bi];
    const int rr = r + B.lev, n = 1 << rr;
    std::vector<double> G;
    vertex_grid(B.coarse, rr, G);
    auto gv = [&](int i, int j, int c) { return G[2 * (i + (n + 1) * j) + c]; };
    c0[bi] = cnt;
    nside[bi] = n;
    cell_of[bi].assign((size_t)n * n, -1);
    const int sc = gscale >> B.lev, gx0 = bi * (gscale << r);
    for (int c = 0; c < n * n; ++c) {
      const int i = morton_i(c, rr), j = morton_j(c, rr);
      cell_of[bi][i + n * j] = cnt + c;
      R.ci.push_back(gx0 + sc * i);
      R.cj.push_back(sc * j);
      R.mat.push_back(B.material);
      const int vi[4] = {i, i + 1, i, i + 1}, vj[4] = {j, j, j + 1, j + 1};
      for (int v = 0; v < 4; ++v) {
        R.V.push_back(gv(vi[v], vj[v], 0));
        R.V.push_back(gv(vi[v], vj[v], 1));
      }
      if (B.lev) {  // the level-r parent: vertices of the 2x2 patch
        const int I = 2 * (i / 2), J = 2 * (j / 2);
        const double pv[8] = {gv(I, J, 0),     gv(I, J, 1),     gv(I + 2, J, 0),     gv(I + 2, J, 1),
                              gv(I, J + 2, 0), gv(I, J + 2, 1), gv(I + 2, J + 2, 0), gv(I + 2, J + 2, 1)};
        R.pdiam.push_back(cell_diam(pv));
      } else {
        R.pdiam.push_back(cell_diam(&R.V[R.V.size() - 8]));
      }
    }
    cnt += n * n;
  }
  R.ncell = cnt;
  R.nbr.assign(4 * (size_t)cnt, -1);
  R.bid.assign(4 * (size_t)cnt, INTERIOR);
  std::vector<char> hanging(4 * (size_t)cnt, 0);
  for (int bi = 0; bi < nb; ++bi) {
    const int n = nside[bi];
    const std::vector<int>& co = cell_of[bi];
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (int f = 0; f < 4; ++f) {
          const int ii = i + FN[f][0], jj = j + FN[f][1];
          if (ii >= 0 && ii < n && jj >= 0 && jj < n) R.nbr[4 * co[i + n * j] + f] = co[ii + n * jj];
        }
    if (bi + 1 == nb) continue;
    const int m = nside[bi + 1];
    const std::vector<int>& cr = cell_of[bi + 1];
    if (blocks[bi].lev == blocks[bi + 1].lev) {
      for (int j = 0; j < n; ++j) {
        const int a = co[(n - 1) + n * j], b = cr[0 + m * j];
        R.nbr[4 * a + 1] = b;
        R.nbr[4 * b + 0] = a;
      }
    } else if (blocks[bi].lev < blocks[bi + 1].lev) {  // coarse left, fine right
      for (int j = 0; j < n; ++j)
        for (int s = 0; s < 2; ++s) R.hang.push_back({co[(n - 1) + n * j], 1, s, cr[0 + m * (2 * j + s)], 0});
    } else {  // fine left, coarse right
      for (int j = 0; j < m; ++j)
        for (int s = 0; s < 2; ++s) R.hang.push_back({cr[0 + m * j], 0, s, co[(n - 1) + n * (2 * j + s)], 1});
    }
  }
  for (const HangFace& h : R.hang) {
    hanging[4 * h.coarse + h.face] = 1;
    hanging[4 * h.fine + h.fine_face] = 1;
  }
  for (int c = 0; c < cnt; ++c)
    for (int f = 0; f < 4; ++f) {
      if (R.nbr[4 * c + f] >= 0 || hanging[4 * c + f]) continue;
      const int a = FV[f][0], b = FV[f][1];
      const double cx = (R.V[8 * c + 2 * a] + R.V[8 * c + 2 * b]) / 2.0;
      const double cy = (R.V[8 * c + 2 * a + 1] + R.V[8 * c + 2 * b + 1]) / 2.0;
      R.bid[4 * c + f] = classify_face(cx, cy, g, insulated, schottky);
    }
  return R;
}

// Semiconductor: bulk (0,0),(r2-bl,0),(0,H),(r1-bl,H) then layer up to the
// r line; electrolyte: layer from the r line to r+bl, then bulk up to L
// (make_semiconductor_grid / make_electrolyte_grid with use_boundary_layer).
// Region cells: level-r block then level-(r+1) block (deal.II's active-cell
// order); the Poisson mesh lists the semiconductor cells then the
// electrolyte cells in their regions' orders, so both region -> Poisson maps
// stay index maps.
static SolarCellMesh make_local_mesh(int r, const GeometrySpec& g) {
  if (g.local != 1) throw std::runtime_error("only one level of boundary-layer refinement is supported");
  if (!(g.bl > 0.0)) throw std::runtime_error("local refinement needs a boundary layer of positive width");
  if (r > 9) throw std::runtime_error("local refinement: global refinement must be at most 9");
  SolarCellMesh m;
  m.geo = g;
  m.r = r;
  m.N = 1 << r;
  m.local = 1;
  const BlockSpec sb{{0.0, 0.0, g.r2 - g.bl, 0.0, 0.0, g.H, g.r1 - g.bl, g.H}, 0, SEMI_ID};
  const BlockSpec sl{{g.r2 - g.bl, 0.0, g.r2, 0.0, g.r1 - g.bl, g.H, g.r1, g.H}, 1, SEMI_BL_ID};
  const BlockSpec el{{g.r2, 0.0, g.r2 + g.bl, 0.0, g.r1, g.H, g.r1 + g.bl, g.H}, 1, ELEC_BL_ID};
  const BlockSpec eb{{g.r2 + g.bl, 0.0, g.L, 0.0, g.r1 + g.bl, g.H, g.L, g.H}, 0, ELEC_ID};
  m.semi = make_block_region({sb, sl}, r, {0, 1}, g, g.insulated, g.schottky);
  m.elec = make_block_region({el, eb}, r, {1, 0}, g, g.insulated, false);
  Region P = make_block_region({sb, sl, el, eb}, r, {0, 1, 3, 2}, g, g.insulated, g.schottky);
  m.NX = P.NXg;
  m.NY = P.NYg;
  m.ncellP = P.ncell;
  m.pV = P.V;
  m.pI = P.ci;
  m.pJ = P.cj;
  m.pmat = P.mat;
  m.pbid = P.bid;
  m.pnbr = P.nbr;
  m.phang = P.hang;
  // interface pairs: the layer cells of the semiconductor's last column and
  // the electrolyte's first (both level r+1, conforming)
  const int n = 2 * m.N, ns0 = m.N * m.N, ne0 = m.N * m.N;
  m.semi_sigma_nb.assign(m.semi.ncell, -1);
  m.elec_sigma_nb.assign(m.elec.ncell, -1);
  for (int j = 0; j < n; ++j) {
    const int s = ns0 + morton_index(n - 1, j), e = ne0 + morton_index(0, j);
    if (m.semi.bid[4 * s + 1] != INTERFACE) continue;
    m.sigma_semi.push_back(s);
    m.sigma_elec.push_back(e);
    m.semi_sigma_nb[s] = e;
    m.elec_sigma_nb[e] = s;
  }
  return m;
}

PoissonLayout poisson_layout(const SolarCellMesh& m, int deg) {
  if (m.local) {
    PoissonLayout L;
    L.deg = deg;
    L.nf = deg + 1;
    const int nc = m.ncellP;
    L.face_base.assign(4 * (size_t)nc, -1);
    static const int OPPF[4] = {1, 0, 3, 2};
    int cnt = 0;
    for (int c = 0; c < nc; ++c)
      for (int f = 0; f < 4; ++f) {
        const int e = m.pnbr[4 * c + f];
        if (e >= 0 && L.face_base[4 * e + OPPF[f]] >= 0) {
          L.face_base[4 * c + f] = L.face_base[4 * e + OPPF[f]];
        } else {
          L.face_base[4 * c + f] = cnt;
          cnt += L.nf;
        }
      }
    L.nV = cnt;
    L.nH = 0;
    L.nint = 2 * deg * (deg + 1);
    L.nD = cnt + nc * L.nint;
    L.nphi = (deg + 1) * (deg + 1);
    L.n = L.nD + nc * L.nphi;
    const int nx = (deg + 2) * (deg + 1);
    L.nloc = 2 * nx + L.nphi;
    L.idx.resize((size_t)nc * L.nloc);
    for (int p = 0; p < nc; ++p) {
      const int ib = cnt + p * L.nint;
      const int* fb = &L.face_base[4 * (size_t)p];
      int* o = &L.idx[(size_t)p * L.nloc];
      for (int b = 0; b <= deg; ++b)
        for (int a = 0; a <= deg + 1; ++a)
          o[a + (deg + 2) * b] = a == 0 ? fb[0] + b : (a == deg + 1 ? fb[1] + b : ib + (a - 1) + deg * b);
      for (int b = 0; b <= deg + 1; ++b)
        for (int a = 0; a <= deg; ++a)
          o[nx + a + (deg + 1) * b] =
              b == 0 ? fb[2] + a : (b == deg + 1 ? fb[3] + a : ib + deg * (deg + 1) + a + (deg + 1) * (b - 1));
      for (int l = 0; l < L.nphi; ++l) o[2 * nx + l] = L.nD + p * L.nphi + l;
    }
    return L;
  }
  PoissonLayout L;
  L.deg = deg;
  L.nf = deg + 1;
  L.nV = (m.NX + 1) * m.NY * L.nf;
  L.nH = m.NX * (m.NY + 1) * L.nf;
  L.nint = 2 * deg * (deg + 1);
  L.nD = L.nV + L.nH + m.ncellP * L.nint;
  L.nphi = (deg + 1) * (deg + 1);
  L.n = L.nD + m.ncellP * L.nphi;
  const int nx = (deg + 2) * (deg + 1);
  L.nloc = 2 * nx + L.nphi;
  L.idx.resize((size_t)m.ncellP * L.nloc);
  for (int p = 0; p < m.ncellP; ++p) {
    const int I = m.pI[p], J = m.pJ[p];
    const int ib = L.nV + L.nH + p * L.nint;
    int* o = &L.idx[(size_t)p * L.nloc];
    for (int b = 0; b <= deg; ++b)
      for (int a = 0; a <= deg + 1; ++a) {
        int li = a + (deg + 2) * b, g;
        if (a == 0) g = (J * (m.NX + 1) + I) * L.nf + b;
        else if (a == deg + 1) g = (J * (m.NX + 1) + I + 1) * L.nf + b;
        else g = ib + (a - 1) + deg * b;
        o[li] = g;
      }
    for (int b = 0; b <= deg + 1; ++b)
      for (int a = 0; a <= deg; ++a) {
        int li = nx + a + (deg + 1) * b, g;
        if (b == 0) g = L.nV + (J * m.NX + I) * L.nf + a;
        else if (b == deg + 1) g = L.nV + ((J + 1) * m.NX + I) * L.nf + a;
        else g = ib + deg * (deg + 1) + a + (deg + 1) * (b - 1);
        o[li] = g;
      }
    for (int l = 0; l < L.nphi; ++l) o[2 * nx + l] = L.nD + p * L.nphi + l;
  }
  return L;
}

}  // namespace pecs
