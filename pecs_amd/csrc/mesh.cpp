// Structured meshes and maps (host side); see mesh.hpp for the reference map.
#include "mesh.hpp"

#include <stdexcept>

namespace pecs {

int morton_i(int c, int r) {
  int i = 0;
  for (int b = 0; b < r; ++b) i |= ((c >> (2 * b)) & 1) << b;
  return i;
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

SolarCellMesh make_solarcell_mesh(int r, const GeometrySpec& g) {
  if (r < 1 || r > 11) throw std::runtime_error("global refinement must be in [1, 11]");
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

PoissonLayout poisson_layout(const SolarCellMesh& m, int deg) {
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
