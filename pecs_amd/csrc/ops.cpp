// Constant local operators (host side, setup); see ops.hpp.
#include "ops.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <unordered_map>

namespace pecs {

static const int OPP[4] = {1, 0, 3, 2};
static const double BETA = 0.70710678118654752440;  // (1,1)/sqrt(2), LDG.cpp:455-458

bool dense_inverse(int n, const double* A, double* Ainv) {
  std::vector<double> a(A, A + (size_t)n * n);
  std::vector<int> piv(n);
  for (int i = 0; i < n * n; ++i) Ainv[i] = 0.0;
  for (int i = 0; i < n; ++i) Ainv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) return false;
    if (p != c)
      for (int t = 0; t < n; ++t) {
        std::swap(a[p * n + t], a[c * n + t]);
        std::swap(Ainv[p * n + t], Ainv[c * n + t]);
      }
    const double inv = 1.0 / a[c * n + c];
    for (int t = 0; t < n; ++t) {
      a[c * n + t] *= inv;
      Ainv[c * n + t] *= inv;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int t = 0; t < n; ++t) {
        a[r * n + t] -= f * a[c * n + t];
        Ainv[r * n + t] -= f * Ainv[c * n + t];
      }
    }
  }
  return true;
}

// FNV-1a over 64-bit words (equal records give equal hashes; the table
// lookup still compares whole records).
static uint64_t fnv1a(const double* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) {
    uint64_t w;
    std::memcpy(&w, p + i, sizeof w);
    h ^= w;
    h *= 1099511628211ull;
  }
  return h;
}

// De-duplicate per-cell records of `stride` doubles (bitwise identical cells
// share one table entry).
static void dedup(const std::vector<double>& rec, int ncell, int stride, std::vector<int>& type,
                  std::vector<double>& table, int& ntype) {
  std::unordered_map<uint64_t, std::vector<int>> seen;
  type.assign(ncell, -1);
  table.clear();
  ntype = 0;
  std::vector<uint64_t> hash(ncell);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < ncell; ++c) hash[c] = fnv1a(&rec[(size_t)c * stride], stride);
  for (int c = 0; c < ncell; ++c) {
    const double* r = &rec[(size_t)c * stride];
    const uint64_t h = hash[c];
    auto& cand = seen[h];
    int found = -1;
    for (int t : cand)
      if (std::memcmp(&table[(size_t)t * stride], r, stride * sizeof(double)) == 0) {
        found = t;
        break;
      }
    if (found < 0) {
      found = ntype++;
      table.insert(table.end(), r, r + stride);
      cand.push_back(found);
    }
    type[c] = found;
  }
}

LdgRegionOps build_ldg_ops(const Region& R, const RefTables& T, const Params& P) {
  const int nb = T.nb, nq = T.nq, nfq = T.nfq;
  LdgRegionOps O;
  O.nb = nb;
  O.ncell = R.ncell;
  O.cell.resize(R.ncell);
  std::vector<double> hdiam(R.ncell);
  for (int c = 0; c < R.ncell; ++c) hdiam[c] = cell_diameter(&R.V[8 * c]);
  // hanging faces: the penalty uses h = min(coarse diameter, diameter of the
  // fine cell's parent) on both sides (LDG.cpp:366-367: cell->neighbor() is
  // the fine cells' level-r parent)
  std::vector<double> hsig(4 * (size_t)R.ncell, 0.0);
  for (const HangFace& hf : R.hang) {
    const double sg = P.penalty / std::min(hdiam[hf.coarse], R.pdiam[hf.fine]);
    hsig[4 * hf.coarse + hf.face] = sg;
    hsig[4 * hf.fine + hf.fine_face] = sg;
  }
  int singular = 0;  // the cells are independent: one OpenMP loop, errors raised after it
#pragma omp parallel for schedule(static) reduction(| : singular)
  for (int c = 0; c < R.ncell; ++c) {
    const double* v = &R.V[8 * c];
    const double h = hdiam[c];
    LdgCellOps& L = O.cell[c];
    L.Mk.assign(nb * nb, 0);
    L.Minv.assign(nb * nb, 0);
    L.Bself.assign(2 * nb * nb, 0);
    L.Bf.assign(4 * 2 * nb * nb, 0);
    L.Dself.assign(nb * nb, 0);
    L.Df.assign(4 * nb * nb, 0);
    for (int q = 0; q < nq; ++q) {
      Geo g = bilinear(v, T.qx[q], T.qy[q]);
      const double jxw = T.qw[q] * g.det;
      for (int a = 0; a < nb; ++a) {
        const double pa = T.phi[a * nq + q];
        const double dx = (g.J11 * T.dphx[a * nq + q] - g.J10 * T.dphy[a * nq + q]) / g.det;
        const double dy = (-g.J01 * T.dphx[a * nq + q] + g.J00 * T.dphy[a * nq + q]) / g.det;
        for (int b = 0; b < nb; ++b) {
          const double pb = T.phi[b * nq + q];
          L.Mk[a * nb + b] += jxw * pa * pb;
          L.Bself[a * nb + b] -= jxw * dx * pb;         // -(div p_i, u_j), p_i = (phi_a, 0)
          L.Bself[(nb + a) * nb + b] -= jxw * dy * pb;  // p_i = (0, phi_a)
        }
      }
    }
    for (int f = 0; f < 4; ++f) {
      double nx, ny, len;
      face_normal(v, f, nx, ny, len);
      const int e = R.nbr[4 * c + f];
      const int bid = R.bid[4 * c + f];
      const double* ps = &T.fphi[(size_t)f * nb * nfq];
      const double* po = &T.fphi[(size_t)OPP[f] * nb * nfq];
      const bool hanging = e < 0 && bid == INTERIOR;
      const double sig =
          hanging ? hsig[4 * c + f] : ((e >= 0) ? P.penalty / std::min(h, hdiam[e]) : P.penalty / h);
      for (int q = 0; q < nfq; ++q) {
        const double jw = T.fw[q] * len;
        for (int a = 0; a < nb; ++a) {
          const double pa = ps[a * nfq + q];
          for (int b = 0; b < nb; ++b) {
            const double pbs = ps[b * nfq + q];
            if (hanging) {  // own-side terms of an interior face, over the whole face
              L.Bself[a * nb + b] += jw * (0.5 * pa * nx + BETA * pa) * pbs;
              L.Bself[(nb + a) * nb + b] += jw * (0.5 * pa * ny + BETA * pa) * pbs;
              L.Dself[a * nb + b] += sig * jw * pa * pbs;
            } else if (e >= 0) {
              const double pbo = po[b * nfq + q];
              L.Bself[a * nb + b] += jw * (0.5 * pa * nx + BETA * pa) * pbs;
              L.Bself[(nb + a) * nb + b] += jw * (0.5 * pa * ny + BETA * pa) * pbs;
              L.Bf[(f * 2 * nb + a) * nb + b] += jw * (0.5 * pa * nx - BETA * pa) * pbo;
              L.Bf[(f * 2 * nb + nb + a) * nb + b] += jw * (0.5 * pa * ny - BETA * pa) * pbo;
              L.Dself[a * nb + b] += sig * jw * pa * pbs;
              L.Df[(f * nb + a) * nb + b] -= sig * jw * pa * pbo;
            } else if (bid == DIRICHLET) {
              L.Dself[a * nb + b] += sig * jw * pa * pbs;
            } else {  // Interface, Neumann, Schottky: p.n u (LDG.cpp:236-272)
              L.Bself[a * nb + b] += jw * pa * nx * pbs;
              L.Bself[(nb + a) * nb + b] += jw * pa * ny * pbs;
            }
          }
        }
      }
    }
    for (int t = 0; t < nb * nb; ++t) L.Dself[t] += (P.theta / P.dt) * L.Mk[t];
    if (!dense_inverse(nb, L.Mk.data(), L.Minv.data())) singular = 1;
  }
  if (singular) throw std::runtime_error("singular cell mass matrix");
  // hanging couplings: quadrature on each fine face, the coarse basis at the
  // matching subface points (s + t) / 2 of its face
  auto dgq_at = [&](double xi, double eta, std::vector<double>& out) {
    const auto nodes = equispaced(T.k);
    out.resize(nb);
    for (int bb = 0; bb < T.n1; ++bb)
      for (int aa = 0; aa < T.n1; ++aa) {
        double vx, dx, vy, dy;
        lagrange(nodes, aa, xi, vx, dx);
        lagrange(nodes, bb, eta, vy, dy);
        out[aa + T.n1 * bb] = vx * vy;
      }
  };
  for (const HangFace& hf : R.hang) {
    double nx, ny, len;
    face_normal(&R.V[8 * hf.coarse], hf.face, nx, ny, len);
    const double sg = hsig[4 * hf.coarse + hf.face];
    LdgHangBlock cf, fc;  // coarse rows / fine columns, fine rows / coarse columns
    cf.M = hf.coarse;
    cf.X = hf.fine;
    fc.M = hf.fine;
    fc.X = hf.coarse;
    for (LdgHangBlock* hb : {&cf, &fc}) {
      hb->B.assign(2 * nb * nb, 0.0);
      hb->D.assign(nb * nb, 0.0);
    }
    std::vector<double> pc, pf;
    for (int q = 0; q < nfq; ++q) {
      double xi, eta;
      face_point(hf.face, 0.5 * (hf.sub + T.fs[q]), xi, eta);
      dgq_at(xi, eta, pc);
      face_point(hf.fine_face, T.fs[q], xi, eta);
      dgq_at(xi, eta, pf);
      const double jw = T.fw[q] * 0.5 * len;
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b) {
          // coarse rows, outward normal n; fine rows, outward normal -n
          cf.B[a * nb + b] += jw * (0.5 * pc[a] * nx - BETA * pc[a]) * pf[b];
          cf.B[(nb + a) * nb + b] += jw * (0.5 * pc[a] * ny - BETA * pc[a]) * pf[b];
          cf.D[a * nb + b] -= sg * jw * pc[a] * pf[b];
          fc.B[a * nb + b] += jw * (-0.5 * pf[a] * nx - BETA * pf[a]) * pc[b];
          fc.B[(nb + a) * nb + b] += jw * (-0.5 * pf[a] * ny - BETA * pf[a]) * pc[b];
          fc.D[a * nb + b] -= sg * jw * pf[a] * pc[b];
        }
    }
    O.hang.push_back(std::move(cf));
    O.hang.push_back(std::move(fc));
  }
  O.type_stride = nb * nb + 5 * 2 * nb * nb;
  std::vector<double> rec((size_t)R.ncell * O.type_stride);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < R.ncell; ++c) {
    double* r = &rec[(size_t)c * O.type_stride];
    const LdgCellOps& L = O.cell[c];
    std::copy(L.Minv.begin(), L.Minv.end(), r);
    std::copy(L.Bself.begin(), L.Bself.end(), r + nb * nb);
    std::copy(L.Bf.begin(), L.Bf.end(), r + nb * nb + 2 * nb * nb);
  }
  dedup(rec, R.ncell, O.type_stride, O.type, O.table, O.ntype);
  return O;
}

static int offset_index(const BlockSPD& S, int di, int dj) {
  for (int o = 0; o < S.noff(); ++o)
    if (S.offi[o] == di && S.offj[o] == dj) return o;
  return -1;
}

// The per-cell coupling lists of the LDG operators: for cell M, (Y, B_{M,Y},
// D_{M,Y}) over Y = M, its conforming face neighbours and its hanging
// partners.
struct Coupling {
  int Y;
  const double* B;
  const double* D;
};
static std::vector<std::vector<Coupling>> ldg_couplings(const Region& R, const LdgRegionOps& ops) {
  const int nb = ops.nb;
  std::vector<std::vector<Coupling>> adj(R.ncell);
  for (int M = 0; M < R.ncell; ++M) {
    const LdgCellOps& L = ops.cell[M];
    adj[M].push_back({M, L.Bself.data(), L.Dself.data()});
    for (int f = 0; f < 4; ++f) {
      const int e = R.nbr[4 * M + f];
      if (e >= 0) adj[M].push_back({e, &L.Bf[(size_t)f * 2 * nb * nb], &L.Df[(size_t)f * nb * nb]});
    }
  }
  for (const LdgHangBlock& hb : ops.hang) adj[hb.M].push_back({hb.X, hb.B.data(), hb.D.data()});
  return adj;
}

// S = D + mu B^T blockdiag(Minv, Minv) B on the region's lattice (positions
// without a cell hold no DoFs), for meshes with hanging faces.
static BlockSPD schur_matrix_lattice(const Region& R, const LdgRegionOps& ops, double mu) {
  const int nb = ops.nb;
  const auto adj = ldg_couplings(R, ops);
  BlockSPD S;
  S.NXc = R.NXg;
  S.NYc = R.NYg;
  S.sep_width = 0;
  const size_t NG = (size_t)S.NXc * S.NYc;
  S.ndof.assign(NG, 0);
  auto gid = [&](int c) { return (size_t)R.ci[c] + (size_t)S.NXc * R.cj[c]; };
  for (int c = 0; c < R.ncell; ++c) S.ndof[gid(c)] = nb;
  // couplings X ~ Y through a shared M, and their lattice offsets
  std::map<std::pair<int, int>, int> offs;
  std::vector<std::vector<int>> nbh(R.ncell);
  for (int M = 0; M < R.ncell; ++M)
    for (const Coupling& a : adj[M])
      for (const Coupling& b : adj[M]) nbh[a.Y].push_back(b.Y);
  for (int X = 0; X < R.ncell; ++X) {
    std::sort(nbh[X].begin(), nbh[X].end());
    nbh[X].erase(std::unique(nbh[X].begin(), nbh[X].end()), nbh[X].end());
    for (int Y : nbh[X]) offs.emplace(std::make_pair(R.ci[Y] - R.ci[X], R.cj[Y] - R.cj[X]), 0);
  }
  for (auto& kv : offs) {
    kv.second = S.noff();
    S.offi.push_back(kv.first.first);
    S.offj.push_back(kv.first.second);
  }
  const int no = S.noff();
  S.blk.assign(NG * no, {});
  auto block = [&](int X, int Y) -> double* {
    const int o = offs.at(std::make_pair(R.ci[Y] - R.ci[X], R.cj[Y] - R.cj[X]));
    return S.blk[gid(X) * no + o].data();
  };
  for (int X = 0; X < R.ncell; ++X)
    for (int Y : nbh[X]) {
      const int o = offs.at(std::make_pair(R.ci[Y] - R.ci[X], R.cj[Y] - R.cj[X]));
      S.blk[gid(X) * no + o].assign((size_t)nb * nb, 0.0);
    }
  // row X: sum over the cells M coupled to X of B_{M,X}^T Minv_M B_{M,Y}
  std::vector<std::vector<std::pair<int, const double*>>> into(R.ncell);  // (M, B_{M,X})
  for (int M = 0; M < R.ncell; ++M)
    for (const Coupling& a : adj[M]) into[a.Y].push_back({M, a.B});
#pragma omp parallel for schedule(dynamic, 64)
  for (int X = 0; X < R.ncell; ++X) {
    std::vector<double> Pm(2 * nb * nb);
    for (const auto& mb : into[X]) {
      const int M = mb.first;
      const double* BX = mb.second;
      const LdgCellOps& L = ops.cell[M];
      for (const Coupling& y : adj[M]) {
        for (int half = 0; half < 2; ++half)
          for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) {
              double s = 0;
              for (int t = 0; t < nb; ++t) s += L.Minv[a * nb + t] * y.B[(half * nb + t) * nb + b];
              Pm[(half * nb + a) * nb + b] = s;
            }
        double* blk = block(X, y.Y);
        for (int a = 0; a < nb; ++a)
          for (int b = 0; b < nb; ++b) {
            double s = 0;
            for (int i = 0; i < 2 * nb; ++i) s += BX[i * nb + a] * Pm[i * nb + b];
            blk[a * nb + b] += mu * s;
          }
      }
    }
    for (const Coupling& y : adj[X]) {
      double* blk = block(X, y.Y);
      for (int t = 0; t < nb * nb; ++t) blk[t] += y.D[t];
    }
  }
  return S;
}

BlockSPD schur_matrix(const Region& R, const LdgRegionOps& ops, double mu) {
  if (!R.hang.empty()) return schur_matrix_lattice(R, ops, mu);
  const int nb = ops.nb, N = R.N;
  BlockSPD S;
  S.NXc = N;
  S.NYc = N;
  S.ndof.assign(N * N, nb);
  S.sep_width = 2;
  for (int dj = -2; dj <= 2; ++dj)
    for (int di = -2; di <= 2; ++di)
      if (std::abs(di) + std::abs(dj) <= 2) {
        S.offi.push_back(di);
        S.offj.push_back(dj);
      }
  const int no = S.noff();
  S.blk.assign((size_t)N * N * no, {});
  auto block = [&](int X, int Y) -> double* {
    int gx = R.ci[X] + N * R.cj[X];
    int o = offset_index(S, R.ci[Y] - R.ci[X], R.cj[Y] - R.cj[X]);
    auto& b = S.blk[(size_t)gx * no + o];
    if (b.empty()) b.assign(nb * nb, 0.0);
    return b.data();
  };
  // Cells of one colour (i mod 3, j mod 3) have disjoint closed neighbourhoods,
  // so the blocks they add to are disjoint: each colour runs in parallel.
  std::vector<std::vector<int>> colour(9);
  for (int M = 0; M < R.ncell; ++M) colour[(R.ci[M] % 3) + 3 * (R.cj[M] % 3)].push_back(M);
  for (int X = 0; X < R.ncell; ++X)  // allocate every block first (no allocation inside the parallel loops)
    for (int f = -1; f < 4; ++f) {
      const int Y = f < 0 ? X : R.nbr[4 * X + f];
      if (Y < 0) continue;
      for (int g = -1; g < 4; ++g) {
        const int Z = g < 0 ? Y : R.nbr[4 * Y + g];
        if (Z >= 0) block(X, Z);
      }
    }
  for (const std::vector<int>& cells_of_colour : colour)
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t mi = 0; mi < cells_of_colour.size(); ++mi) {
    const int M = cells_of_colour[mi];
    std::vector<double> Pm(2 * nb * nb);
    const LdgCellOps& L = ops.cell[M];
    int cells[5];
    const double* Bs[5];
    int nn = 0;
    cells[nn] = M;
    Bs[nn++] = L.Bself.data();
    for (int f = 0; f < 4; ++f)
      if (R.nbr[4 * M + f] >= 0) {
        cells[nn] = R.nbr[4 * M + f];
        Bs[nn++] = &L.Bf[(size_t)f * 2 * nb * nb];
      }
    for (int y = 0; y < nn; ++y) {
      // Pm = blockdiag(Minv, Minv) * B_{M,Y}
      for (int half = 0; half < 2; ++half)
        for (int a = 0; a < nb; ++a)
          for (int b = 0; b < nb; ++b) {
            double s = 0;
            for (int t = 0; t < nb; ++t) s += L.Minv[a * nb + t] * Bs[y][(half * nb + t) * nb + b];
            Pm[(half * nb + a) * nb + b] = s;
          }
      for (int x = 0; x < nn; ++x) {
        double* blk = block(cells[x], cells[y]);
        for (int a = 0; a < nb; ++a)
          for (int b = 0; b < nb; ++b) {
            double s = 0;
            for (int i = 0; i < 2 * nb; ++i) s += Bs[x][i * nb + a] * Pm[i * nb + b];
            blk[a * nb + b] += mu * s;
          }
      }
    }
    double* d = block(M, M);
    for (int t = 0; t < nb * nb; ++t) d[t] += L.Dself[t];
    for (int f = 0; f < 4; ++f) {
      int e = R.nbr[4 * M + f];
      if (e < 0) continue;
      double* b = block(M, e);
      for (int t = 0; t < nb * nb; ++t) b[t] += L.Df[(size_t)f * nb * nb + t];
    }
  }
  return S;
}

struct CsrBuilder {
  std::vector<std::map<int, double>> rows;
  explicit CsrBuilder(int n) : rows(n) {}
  void add(int r, int c, double v) { rows[r][c] += v; }
  void finish(std::vector<int64_t>& rowptr, std::vector<int>& col, std::vector<double>& val) {
    rowptr.assign(rows.size() + 1, 0);
    col.clear();
    val.clear();
    for (size_t r = 0; r < rows.size(); ++r) {
      for (auto& kv : rows[r]) {
        col.push_back(kv.first);
        val.push_back(kv.second);
      }
      rowptr[r + 1] = (int64_t)col.size();
    }
  }
};

void ldg_csr(const Region& R, const LdgRegionOps& ops, double mu, double dt, std::vector<int64_t>& rowptr,
             std::vector<int>& col, std::vector<double>& val, bool mass_only) {
  const int nb = ops.nb, nc = R.ncell;
  CsrBuilder C(3 * nb * nc);
  auto jx = [&](int c, int a) { return c * nb + a; };
  auto jy = [&](int c, int a) { return nc * nb + c * nb + a; };
  auto uu = [&](int c, int a) { return 2 * nc * nb + c * nb + a; };
  for (int M = 0; M < nc; ++M) {
    const LdgCellOps& L = ops.cell[M];
    if (mass_only) {
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b) C.add(uu(M, a), uu(M, b), L.Mk[a * nb + b] / dt);
      continue;
    }
    for (int a = 0; a < nb; ++a)
      for (int b = 0; b < nb; ++b) {
        C.add(jx(M, a), jx(M, b), L.Mk[a * nb + b] / mu);
        C.add(jy(M, a), jy(M, b), L.Mk[a * nb + b] / mu);
        C.add(uu(M, a), uu(M, b), L.Dself[a * nb + b]);
      }
    auto addB = [&](const double* B, int X) {
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b) {
          C.add(jx(M, a), uu(X, b), B[a * nb + b]);
          C.add(jy(M, a), uu(X, b), B[(nb + a) * nb + b]);
          C.add(uu(X, b), jx(M, a), -B[a * nb + b]);  // C = -B^T
          C.add(uu(X, b), jy(M, a), -B[(nb + a) * nb + b]);
        }
    };
    addB(L.Bself.data(), M);
    for (int f = 0; f < 4; ++f) {
      int e = R.nbr[4 * M + f];
      if (e < 0) continue;
      addB(&L.Bf[(size_t)f * 2 * nb * nb], e);
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b) C.add(uu(M, a), uu(e, b), L.Df[(size_t)(f * nb + a) * nb + b]);
    }
  }
  if (!mass_only)
    for (const LdgHangBlock& hb : ops.hang) {
      const int M = hb.M, X = hb.X;
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b) {
          C.add(jx(M, a), uu(X, b), hb.B[a * nb + b]);
          C.add(jy(M, a), uu(X, b), hb.B[(nb + a) * nb + b]);
          C.add(uu(X, b), jx(M, a), -hb.B[a * nb + b]);
          C.add(uu(X, b), jy(M, a), -hb.B[(nb + a) * nb + b]);
          C.add(uu(M, a), uu(X, b), hb.D[a * nb + b]);
        }
    }
  C.finish(rowptr, col, val);
}

// Multipliers of a mesh with hanging faces: one block of K = deg + 1 on every
// conforming interior face, Neumann face and fine side of a hanging face
// (none on the coarse side), owned by the cell that writes the face's D
// DoFs (O.own) and numbered by the owners' lattice positions.
static void poisson_multipliers_local(const SolarCellMesh& m, const RefTables& T, PoissonOps& O) {
  static const int OPPF[4] = {1, 0, 3, 2};
  const int nc = m.ncellP, kf = T.deg + 1;
  std::vector<char> coarse_hang(4 * (size_t)nc, 0), fine_hang(4 * (size_t)nc, 0);
  for (const HangFace& h : m.phang) {
    coarse_hang[4 * h.coarse + h.face] = 1;
    fine_hang[4 * h.fine + h.fine_face] = 1;
  }
  auto is_mult = [&](int K, int f) {
    const int b = m.pbid[4 * K + f];
    if (coarse_hang[4 * K + f]) return false;
    return fine_hang[4 * K + f] || (b == INTERIOR && m.pnbr[4 * K + f] >= 0) || b == NEUMANN;
  };
  const size_t NG = (size_t)m.NX * m.NY;
  std::vector<int> cell_at(NG, -1);
  for (int K = 0; K < nc; ++K) cell_at[m.pI[K] + (size_t)m.NX * m.pJ[K]] = K;
  O.ndof_owner.assign(NG, 0);
  O.owner_base.assign(NG, 0);
  O.mult_face.assign(4 * (size_t)nc, -1);
  int cnt = 0;
  for (size_t G = 0; G < NG; ++G) {
    O.owner_base[G] = cnt;
    const int K = cell_at[G];
    if (K < 0) continue;
    const int order[4] = {0, 2, 1, 3};
    for (int t = 0; t < 4; ++t) {
      const int f = order[t];
      if ((O.own[K] >> f & 1) && is_mult(K, f)) {
        O.mult_face[4 * K + f] = cnt;
        cnt += kf;
      }
    }
    O.ndof_owner[G] = cnt - O.owner_base[G];
  }
  O.nmult = cnt;
  for (int K = 0; K < nc; ++K)
    for (int f = 0; f < 4; ++f) {
      if (!is_mult(K, f) || O.mult_face[4 * K + f] >= 0) continue;
      const int e = m.pnbr[4 * K + f];
      if (e < 0 || O.mult_face[4 * e + OPPF[f]] < 0) throw std::runtime_error("multiplier ownership");
      O.mult_face[4 * K + f] = O.mult_face[4 * e + OPPF[f]];
    }
  for (const HangFace& h : m.phang) {
    const int mb = O.mult_face[4 * h.fine + h.fine_face];
    if (mb < 0) throw std::runtime_error("hanging face without multipliers");
    O.hang.push_back({h.coarse, h.face, h.sub, mb});
  }
  std::vector<double> gx, gw;
  gauss01(kf, gx, gw);
  O.hE.assign(2 * (size_t)kf * kf, 0.0);
  for (int s = 0; s < 2; ++s)
    for (int mm = 0; mm < kf; ++mm)
      for (int j = 0; j < kf; ++j) {
        double v, d;
        lagrange(gx, j, 0.5 * (s + gx[mm]), v, d);
        O.hE[((size_t)s * kf + mm) * kf + j] = 0.5 * gw[mm] * v;
      }
}

PoissonOps build_poisson_ops(const SolarCellMesh& m, const PoissonLayout& Lay, const RefTables& T, const Params& P) {
  PoissonOps O;
  const int nrt = T.nrt, nphi = T.nphi, nl = T.nlocP, nq = T.nq, nfq = T.nfq;
  O.nloc = nl;
  O.nrt = nrt;
  O.nphi = nphi;
  O.k = T.k;
  const int nc = m.ncellP;
  O.Ainv.assign((size_t)nc * nl * nl, 0);
  O.A.assign((size_t)nc * nl * nl, 0);
  O.ubias.assign((size_t)nc * 3 * nrt, 0);
  int singular = 0;  // the cells are independent: one OpenMP loop, errors raised after it
#pragma omp parallel for schedule(static) reduction(| : singular)
  for (int K = 0; K < nc; ++K) {
    std::vector<double> psx(nrt), psy(nrt);
    const double* v = &m.pV[8 * K];
    const double eps = (m.pmat[K] == SEMI_ID || m.pmat[K] == SEMI_BL_ID) ? P.eps_s : P.eps_e;
    double* A = &O.A[(size_t)K * nl * nl];
    for (int q = 0; q < nq; ++q) {
      Geo g = bilinear(v, T.qx[q], T.qy[q]);
      const double jxw = T.qw[q] * g.det;
      for (int i = 0; i < nrt; ++i) {
        const double hx = T.psix[i * nq + q], hy = T.psiy[i * nq + q];
        psx[i] = (g.J00 * hx + g.J01 * hy) / g.det;
        psy[i] = (g.J10 * hx + g.J11 * hy) / g.det;
      }
      for (int i = 0; i < nrt; ++i) {
        const double di = T.divpsi[i * nq + q] / g.det;
        for (int j = 0; j < nrt; ++j) A[i * nl + j] += jxw * (psx[i] * psx[j] + psy[i] * psy[j]) / eps;
        for (int l = 0; l < nphi; ++l) {
          const double pl = T.phiP[l * nq + q];
          A[i * nl + nrt + l] -= jxw * di * pl;           // -div(psi_i) phi_j
          A[(nrt + l) * nl + i] -= P.lam2 * jxw * pl * di;  // -lambda^2 phi_i div(psi_j)
        }
      }
    }
    if (!dense_inverse(nl, A, &O.Ainv[(size_t)K * nl * nl])) singular = 1;
    // constant Dirichlet / Schottky data of semiconductor cells (SolarCell.cpp:593-681)
    if (m.pmat[K] == SEMI_ID || m.pmat[K] == SEMI_BL_ID) {
      for (int f = 0; f < 4; ++f) {
        const int bid = m.pbid[4 * K + f];
        if (bid != DIRICHLET && bid != SCHOTTKY) continue;
        double nx, ny, len;
        face_normal(v, f, nx, ny, len);
        for (int q = 0; q < nfq; ++q) {
          double xi, eta;
          face_point(f, T.fs[q], xi, eta);
          Geo g = bilinear(v, xi, eta);
          const double jw = T.fw[q] * len;
          // datum = bi - other (BiasValues.cpp:11-27, 74-91): bi = phi_bi on
          // x = 0; other = phi_app on Dirichlet points with x = 0, phi_sch on
          // Schottky points at the Schottky location
          const double wbi = (g.x == 0.0) ? 1.0 : 0.0;
          double wapp = 0.0, wsch = 0.0;
          if (bid == DIRICHLET) wapp = (g.x == 0.0) ? 1.0 : 0.0;
          else wsch = (P.schottky_location_set && g.y == P.schottky_location) ? 1.0 : 0.0;
          if (wbi == 0.0 && wapp == 0.0 && wsch == 0.0) continue;
          for (int i = 0; i < nrt; ++i) {
            const double hx = T.fpsix[(f * nrt + i) * nfq + q], hy = T.fpsiy[(f * nrt + i) * nfq + q];
            const double px = (g.J00 * hx + g.J01 * hy) / g.det, py = (g.J10 * hx + g.J11 * hy) / g.det;
            const double unit = -(px * nx + py * ny) * jw;
            double* U = &O.ubias[(size_t)K * 3 * nrt];
            U[i] += unit * wbi;
            U[nrt + i] += unit * wapp;
            U[2 * nrt + i] += unit * wsch;
          }
        }
      }
    }
  }
  if (singular) throw std::runtime_error("singular local mixed matrix");
  const int kf = T.deg + 1;
  O.own.assign(nc, 0);
  for (int K = 0; K < nc; ++K)
    for (int f = 0; f < 4; ++f)
      if (f == 0 || f == 2 || m.pnbr[4 * K + f] < 0) O.own[K] |= (unsigned char)(1u << f);
  if (m.local) {
    poisson_multipliers_local(m, T, O);
  } else {
  // multipliers: interior and Neumann faces
  auto is_mult = [&](int K, int f) {
    int b = m.pbid[4 * K + f];
    return b == INTERIOR || b == NEUMANN;
  };
  const int NG = m.NX * m.NY;
  O.ndof_owner.assign(NG, 0);
  O.owner_base.assign(NG, 0);
  std::vector<int> own_face(4 * (size_t)nc, -1);  // owned face multiplier base
  int cnt = 0;
  for (int G = 0; G < NG; ++G) {
    const int K = m.pcell_of[G];
    const int I = m.pI[K], J = m.pJ[K];
    O.owner_base[G] = cnt;
    const int order[4] = {0, 2, 1, 3};
    for (int t = 0; t < 4; ++t) {
      const int f = order[t];
      bool owned = false;
      if (f == 0 || f == 2) owned = is_mult(K, f);
      if (f == 1) owned = (I == m.NX - 1) && is_mult(K, f);
      if (f == 3) owned = (J == m.NY - 1) && is_mult(K, f);
      if (owned) {
        own_face[4 * K + f] = cnt;
        cnt += kf;
      }
    }
    O.ndof_owner[G] = cnt - O.owner_base[G];
  }
  O.nmult = cnt;
  O.mult_face.assign(4 * (size_t)nc, -1);
  for (int K = 0; K < nc; ++K) {
    const int I = m.pI[K], J = m.pJ[K];
    for (int f = 0; f < 4; ++f) {
      if (!is_mult(K, f)) continue;
      int base = own_face[4 * K + f];
      if (base < 0) {
        if (f == 1) base = own_face[4 * m.pcell_of[(I + 1) + m.NX * J] + 0];
        if (f == 3) base = own_face[4 * m.pcell_of[I + m.NX * (J + 1)] + 2];
      }
      if (base < 0) throw std::runtime_error("multiplier ownership");
      O.mult_face[4 * K + f] = base;
    }
  }
  }
  O.constrained.assign(Lay.n, 0);
  for (int K = 0; K < nc; ++K)
    for (int f = 0; f < 4; ++f)
      if (m.pbid[4 * K + f] == NEUMANN)
        for (int mm = 0; mm < kf; ++mm) O.constrained[Lay.idx[(size_t)K * nl + T.rt_face_dof(f, mm)]] = 1;
  O.type_stride = nl * nl + 3 * nrt;
  std::vector<double> rec((size_t)nc * O.type_stride);
  for (int K = 0; K < nc; ++K) {
    std::copy(&O.Ainv[(size_t)K * nl * nl], &O.Ainv[(size_t)(K + 1) * nl * nl], &rec[(size_t)K * O.type_stride]);
    std::copy(&O.ubias[(size_t)K * 3 * nrt], &O.ubias[(size_t)(K + 1) * 3 * nrt],
              &rec[(size_t)K * O.type_stride + nl * nl]);
  }
  dedup(rec, nc, O.type_stride, O.type, O.table, O.ntype);
  return O;
}

// H = sum_K E_K A_K^{-1} E_K^T on the lattice of multiplier owners, for
// meshes with hanging faces (the coarse side of a hanging face couples to the
// fine faces' multipliers through the dense hE rows).
static BlockSPD multiplier_matrix_local(const SolarCellMesh& m, const PoissonOps& O, const RefTables& T) {
  const int nl = O.nloc, kf = T.deg + 1, nc = m.ncellP;
  std::vector<double> gx, gw;
  gauss01(kf, gx, gw);
  BlockSPD S;
  S.NXc = m.NX;
  S.NYc = m.NY;
  S.ndof = O.ndof_owner;
  S.sep_width = 0;
  const size_t NG = (size_t)m.NX * m.NY;
  std::vector<int> owner_of(O.nmult);
  for (size_t G = 0; G < NG; ++G)
    for (int t = 0; t < O.ndof_owner[G]; ++t) owner_of[O.owner_base[G] + t] = (int)G;
  // per cell: its multipliers and their E rows (local DoF, coefficient)
  struct Row {
    int mult;
    std::vector<std::pair<int, double>> e;
  };
  std::vector<std::vector<Row>> rows(nc);
  for (int K = 0; K < nc; ++K)
    for (int f = 0; f < 4; ++f) {
      const int bf = O.mult_face[4 * K + f];
      if (bf < 0) continue;
      const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
      for (int a = 0; a < kf; ++a) rows[K].push_back({bf + a, {{T.rt_face_dof(f, a), sf * gw[a]}}});
    }
  for (const PoissonOps::HangMult& h : O.hang) {
    const double sf = (h.face == 1 || h.face == 3) ? 1.0 : -1.0;
    for (int a = 0; a < kf; ++a) {
      Row r{h.mbase + a, {}};
      for (int j = 0; j < kf; ++j)
        r.e.push_back({T.rt_face_dof(h.face, j), sf * O.hE[((size_t)h.sub * kf + a) * kf + j]});
      rows[h.coarse].push_back(r);
    }
  }
  std::map<std::pair<int, int>, int> offs;
  for (int K = 0; K < nc; ++K)
    for (const Row& a : rows[K])
      for (const Row& b : rows[K]) {
        const int Ga = owner_of[a.mult], Gb = owner_of[b.mult];
        offs.emplace(std::make_pair(Gb % m.NX - Ga % m.NX, Gb / m.NX - Ga / m.NX), 0);
      }
  for (auto& kv : offs) {
    kv.second = S.noff();
    S.offi.push_back(kv.first.first);
    S.offj.push_back(kv.first.second);
  }
  const int no = S.noff();
  S.blk.assign(NG * no, {});
  for (int K = 0; K < nc; ++K) {
    const double* Ai = &O.Ainv[(size_t)K * nl * nl];
    for (const Row& a : rows[K])
      for (const Row& b : rows[K]) {
        const int Ga = owner_of[a.mult], Gb = owner_of[b.mult];
        const int o = offs.at(std::make_pair(Gb % m.NX - Ga % m.NX, Gb / m.NX - Ga / m.NX));
        auto& blk = S.blk[(size_t)Ga * no + o];
        const int ra = a.mult - O.owner_base[Ga], cb = b.mult - O.owner_base[Gb], rg = S.ndof[Gb];
        if (blk.empty()) blk.assign((size_t)S.ndof[Ga] * rg, 0.0);
        double v = 0.0;
        for (const auto& x : a.e)
          for (const auto& y : b.e) v += x.second * y.second * Ai[(size_t)x.first * nl + y.first];
        blk[(size_t)ra * rg + cb] += v;
      }
  }
  return S;
}

BlockSPD multiplier_matrix(const SolarCellMesh& m, const PoissonOps& O, const RefTables& T) {
  if (m.local) return multiplier_matrix_local(m, O, T);
  BlockSPD S;
  S.NXc = m.NX;
  S.NYc = m.NY;
  S.ndof = O.ndof_owner;
  S.sep_width = 1;
  const int oi[7] = {0, 1, -1, 0, 0, 1, -1}, oj[7] = {0, 0, 0, 1, -1, -1, 1};
  for (int o = 0; o < 7; ++o) {
    S.offi.push_back(oi[o]);
    S.offj.push_back(oj[o]);
  }
  const int no = 7, nl = O.nloc, kf = T.deg + 1;
  S.blk.assign((size_t)m.NX * m.NY * no, {});
  std::vector<double> gw, gx;
  gauss01(kf, gx, gw);
  // owner grid cell of a multiplier index
  std::vector<int> owner_of(O.nmult);
  for (int G = 0; G < m.NX * m.NY; ++G)
    for (int t = 0; t < O.ndof_owner[G]; ++t) owner_of[O.owner_base[G] + t] = G;
  for (int K = 0; K < m.ncellP; ++K) {
    const double* Ai = &O.Ainv[(size_t)K * nl * nl];
    for (int f = 0; f < 4; ++f) {
      const int bf = O.mult_face[4 * K + f];
      if (bf < 0) continue;
      const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
      for (int g = 0; g < 4; ++g) {
        const int bg = O.mult_face[4 * K + g];
        if (bg < 0) continue;
        const double sg = (g == 1 || g == 3) ? 1.0 : -1.0;
        const int Gf = owner_of[bf], Gg = owner_of[bg];
        const int di = (Gg % m.NX) - (Gf % m.NX), dj = (Gg / m.NX) - (Gf / m.NX);
        int o = -1;
        for (int t = 0; t < no; ++t)
          if (S.offi[t] == di && S.offj[t] == dj) o = t;
        if (o < 0) throw std::runtime_error("multiplier coupling outside stencil");
        auto& blk = S.blk[(size_t)Gf * no + o];
        const int rf = S.ndof[Gf], rg = S.ndof[Gg];
        if (blk.empty()) blk.assign((size_t)rf * rg, 0.0);
        for (int a = 0; a < kf; ++a)
          for (int b = 0; b < kf; ++b) {
            const int ra = bf - O.owner_base[Gf] + a, cb = bg - O.owner_base[Gg] + b;
            blk[(size_t)ra * rg + cb] +=
                sf * sg * gw[a] * gw[b] * Ai[T.rt_face_dof(f, a) * nl + T.rt_face_dof(g, b)];
          }
      }
    }
  }
  return S;
}

void mixed_csr(const SolarCellMesh& m, const PoissonLayout& L, const PoissonOps& O, std::vector<int64_t>& rowptr,
               std::vector<int>& col, std::vector<double>& val) {
  CsrBuilder C(L.n);
  const int nl = O.nloc;
  for (int K = 0; K < m.ncellP; ++K) {
    const int* id = &L.idx[(size_t)K * nl];
    const double* A = &O.A[(size_t)K * nl * nl];
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nl; ++j)
        if (A[i * nl + j] != 0.0) C.add(id[i], id[j], A[i * nl + j]);
  }
  C.finish(rowptr, col, val);
}

}  // namespace pecs
