// CPU FP64 restatement of the PECS per-time-step IMEX path -- TEST
// INFRASTRUCTURE ONLY: the timed CPU baseline of the benchmark (own C++
// restatement, not deal.II) and a second CPU check of the host setup code.
// Nothing in the product links or calls it.
//
// One step (SolarCell.cpp:2057-2075): right-hand sides of both carrier pairs
// from the old state, the four carrier solves, the Poisson right-hand side
// from the new densities, the Poisson solve.
//   * carrier right-hand sides per cell (SolarCell.cpp:1073-1414 semiconductor,
//     1451-1726 electrolyte; mass term 1037-1071), threaded over cells;
//   * Poisson right-hand side per cell (SolarCell.cpp:487-815);
//   * solves: the reference factors the full LDG and mixed matrices with
//     UMFPACK (Carrier.cpp:26-40, Poisson.cpp:90-105), which is not available
//     here.  This program solves the same systems through the SPD density
//     Schur complement S = D + B^T A^-1 B and the hybridised Poisson
//     multiplier system H, factored once on the host by the project's
//     supernodal Cholesky (cholesky.cpp), and applies the factors each step
//     with supernodal sweeps that run the nodes of one elimination-tree level
//     in parallel (OpenMP).
// The operators, the meshes and the factors come from the project's host
// setup code (mesh.cpp, ops.cpp, cholesky.cpp); every per-step operation is
// written here.
//
// Usage:  pecs_cpu_step key=value ...   (r k dt steps gen_mode srh ni phi_bi
//         phi_app r1 r2 H L insulated schottky mu_n mu_p mu_r mu_o eps_s eps_e
//         lam2 k_et k_ht v_rec_n v_rec_p ne pe rinf oinf G0 alpha tau_n tau_p
//         penalty leaf out=<file>)
// Prints one JSON line (setup and per-step seconds, threads); with out= it
// writes the final n, p, r, o, Poisson vectors as raw doubles.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../pecs_amd/csrc/cholesky.hpp"
#include "../pecs_amd/csrc/fe.hpp"
#include "../pecs_amd/csrc/mesh.hpp"
#include "../pecs_amd/csrc/ops.hpp"

using namespace pecs;

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Geo {
  double x, y, J00, J01, J10, J11, det;
};

Geo bilinear(const double* v, double xi, double eta) {
  Geo g;
  const double a = (1 - xi) * (1 - eta), b = xi * (1 - eta), c = (1 - xi) * eta, d = xi * eta;
  g.x = v[0] * a + v[2] * b + v[4] * c + v[6] * d;
  g.y = v[1] * a + v[3] * b + v[5] * c + v[7] * d;
  g.J00 = (v[2] - v[0]) * (1 - eta) + (v[6] - v[4]) * eta;
  g.J10 = (v[3] - v[1]) * (1 - eta) + (v[7] - v[5]) * eta;
  g.J01 = (v[4] - v[0]) * (1 - xi) + (v[6] - v[2]) * xi;
  g.J11 = (v[5] - v[1]) * (1 - xi) + (v[7] - v[3]) * xi;
  g.det = g.J00 * g.J11 - g.J01 * g.J10;
  return g;
}

// outward unit normal and length of face f (straight edges)
void face_normal(const double* v, int f, double& nx, double& ny, double& len) {
  static const int va[4] = {0, 1, 0, 2}, vb[4] = {2, 3, 1, 3};
  const double tx = v[2 * vb[f]] - v[2 * va[f]], ty = v[2 * vb[f] + 1] - v[2 * va[f] + 1];
  len = std::sqrt(tx * tx + ty * ty);
  if (f == 1 || f == 2) {
    nx = ty / len;
    ny = -tx / len;
  } else {
    nx = -ty / len;
    ny = tx / len;
  }
}

double diameter(const double* v) {  // deal.II cell->diameter(): the longer diagonal
  const double d1 = std::hypot(v[6] - v[0], v[7] - v[1]), d2 = std::hypot(v[4] - v[2], v[5] - v[3]);
  return std::max(d1, d2);
}

// ---------------------------------------------------------------------------
// supernodal factor with level-parallel sweeps
// ---------------------------------------------------------------------------
struct LevelFactor {
  SupernodalFactor F;
  std::vector<int64_t> z_off;  // per node: offset of its panel updates in z
  int64_t zlen = 0;
  std::vector<std::vector<int>> level_rows;  // own rows of each level's nodes
  std::vector<int64_t> cz;                   // per contribution: z index

  explicit LevelFactor(SupernodalFactor&& f) : F(std::move(f)) {
    const int nn = (int)F.nodes.size();
    z_off.resize(nn);
    for (int s = 0; s < nn; ++s) {
      z_off[s] = zlen;
      zlen += F.nodes[s].ms;
    }
    level_rows.resize(F.nlevels);
    for (int l = 0; l < F.nlevels; ++l)
      for (int s : F.level_nodes[l])
        for (int i = 0; i < F.nodes[s].ns; ++i) level_rows[l].push_back(F.nodes[s].p0 + i);
    cz.resize(F.contrib_node.size());
    for (size_t c = 0; c < cz.size(); ++c) cz[c] = z_off[F.contrib_node[c]] + F.contrib_row[c];
  }

  // forward: t_s = Linv_s (y_s - gathered updates), z_s = L21_s t_s
  void node_forward(int s, std::vector<double>& y, std::vector<double>& z, bool inner) const {
    const SNode& N = F.nodes[s];
    std::vector<double> t(N.ns);
    const double* Li = F.Linv.data() + N.linv_off;
#pragma omp parallel for schedule(static) if (inner)
    for (int i = 0; i < N.ns; ++i) {
      double v = 0;
      for (int j = 0; j <= i; ++j) v += Li[(int64_t)i * N.ns + j] * y[N.p0 + j];
      t[i] = v;
    }
    for (int i = 0; i < N.ns; ++i) y[N.p0 + i] = t[i];
    const double* L2 = F.L21.data() + N.l21_off;
#pragma omp parallel for schedule(static) if (inner)
    for (int r = 0; r < N.ms; ++r) {
      double v = 0;
      for (int j = 0; j < N.ns; ++j) v += L2[(int64_t)r * N.ns + j] * t[j];
      z[z_off[s] + r] = v;
    }
  }

  // backward: x_s = Linv_s^T (y_s - L21_s^T x_rows)
  void node_backward(int s, std::vector<double>& x, bool inner) const {
    const SNode& N = F.nodes[s];
    const double* L2 = F.L21.data() + N.l21_off;
    const double* Li = F.Linv.data() + N.linv_off;
    std::vector<double> t(N.ns);
    for (int j = 0; j < N.ns; ++j) t[j] = x[N.p0 + j];
    if (inner) {
      const int nt = omp_get_max_threads();
      std::vector<double> part((size_t)nt * N.ns, 0.0);
#pragma omp parallel
      {
        double* u = part.data() + (size_t)omp_get_thread_num() * N.ns;
#pragma omp for schedule(static)
        for (int r = 0; r < N.ms; ++r) {
          const double xr = x[F.rows[N.rows_off + r]];
          for (int j = 0; j < N.ns; ++j) u[j] += L2[(int64_t)r * N.ns + j] * xr;
        }
      }
      for (int p = 0; p < nt; ++p)
        for (int j = 0; j < N.ns; ++j) t[j] -= part[(size_t)p * N.ns + j];
    } else {
      for (int r = 0; r < N.ms; ++r) {
        const double xr = x[F.rows[N.rows_off + r]];
        for (int j = 0; j < N.ns; ++j) t[j] -= L2[(int64_t)r * N.ns + j] * xr;
      }
    }
    std::vector<double> o(N.ns, 0.0);
    for (int i = 0; i < N.ns; ++i)
      for (int j = 0; j <= i; ++j) o[j] += Li[(int64_t)i * N.ns + j] * t[i];
    for (int j = 0; j < N.ns; ++j) x[N.p0 + j] = o[j];
  }

  // factor entries a node's sweep reads (its work, up to a constant)
  int64_t node_work(int s) const {
    const SNode& N = F.nodes[s];
    return (int64_t)N.ns * (N.ns + 1) / 2 + (int64_t)N.ms * N.ns;
  }

  // x = S^{-1} b, b and x in the factor's permuted order.  A level's nodes run
  // in parallel when there are enough of them; a large node on a level with
  // few nodes splits its rows over the threads; small work stays serial (a
  // parallel region costs microseconds).
  void solve(const std::vector<double>& b, std::vector<double>& x, std::vector<double>& z) const {
    constexpr int64_t kParallel = 1 << 16;  // entries below which a loop stays serial
    std::vector<double> y(b);
    z.assign(zlen, 0.0);
    const int nt = omp_get_max_threads();
    for (int l = 0; l < F.nlevels; ++l) {
      const std::vector<int>& rows = level_rows[l];
      const int64_t ncontrib = rows.empty() ? 0 : (int64_t)rows.size() * 8;
#pragma omp parallel for schedule(static) if (ncontrib > kParallel)
      for (size_t i = 0; i < rows.size(); ++i) {
        const int p = rows[i];
        double v = 0;
        for (int64_t c = F.contrib_ptr[p]; c < F.contrib_ptr[p + 1]; ++c) v += z[cz[c]];
        y[p] -= v;
      }
      const std::vector<int>& nodes = F.level_nodes[l];
      int64_t work = 0;
      for (int s : nodes) work += node_work(s);
      if ((int)nodes.size() >= nt) {
#pragma omp parallel for schedule(dynamic) if (work > kParallel)
        for (size_t q = 0; q < nodes.size(); ++q) node_forward(nodes[q], y, z, false);
      } else {
        for (int s : nodes) node_forward(s, y, z, node_work(s) > kParallel);
      }
    }
    x = y;
    for (int l = F.nlevels - 1; l >= 0; --l) {
      const std::vector<int>& nodes = F.level_nodes[l];
      int64_t work = 0;
      for (int s : nodes) work += node_work(s);
      if ((int)nodes.size() >= nt) {
#pragma omp parallel for schedule(dynamic) if (work > kParallel)
        for (size_t q = 0; q < nodes.size(); ++q) node_backward(nodes[q], x, false);
      } else {
        for (int s : nodes) node_backward(s, x, node_work(s) > kParallel);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// the problem
// ---------------------------------------------------------------------------
struct Problem {
  Params P;
  int k;
  SolarCellMesh m;
  RefTables T;
  PoissonLayout L;
  LdgRegionOps os, oe;
  PoissonOps op;
  std::vector<double> hs, he;  // cell diameters
  std::unique_ptr<LevelFactor> Fs[2], Fe[2], FH;
  bool share_s, share_e;
  int64_t Ns, Ne, NP;
  // multiplier maps: per multiplier its (<= 2) cells, local RT DoF and weight
  std::vector<int> mcell, mloc, mperm;
  std::vector<double> mcoef, gw;

  Problem(const Params& p, const GeometrySpec& g, int r, int leaf)
      : P(p), k(p.k), m(make_solarcell_mesh(r, g)), T(p.k) {
    L = poisson_layout(m, k - 1);
    os = build_ldg_ops(m.semi, T, P);
    oe = build_ldg_ops(m.elec, T, P);
    op = build_poisson_ops(m, L, T, P);
    Ns = 3LL * T.nb * m.semi.ncell;
    Ne = 3LL * T.nb * m.elec.ncell;
    NP = L.n;
    for (int c = 0; c < m.semi.ncell; ++c) hs.push_back(diameter(&m.semi.V[8 * (size_t)c]));
    for (int c = 0; c < m.elec.ncell; ++c) he.push_back(diameter(&m.elec.V[8 * (size_t)c]));
    share_s = P.mu[0] == P.mu[1];
    share_e = P.mu[2] == P.mu[3];
    auto fac = [&](const BlockSPD& S) { return std::make_unique<LevelFactor>(cholesky_factor(S, leaf)); };
    Fs[0] = fac(schur_matrix(m.semi, os, P.mu[0]));
    if (!share_s) Fs[1] = fac(schur_matrix(m.semi, os, P.mu[1]));
    Fe[0] = fac(schur_matrix(m.elec, oe, P.mu[2]));
    if (!share_e) Fe[1] = fac(schur_matrix(m.elec, oe, P.mu[3]));
    FH = fac(multiplier_matrix(m, op, T));
    mcell.assign(2 * (size_t)op.nmult, -1);
    mloc.assign(2 * (size_t)op.nmult, 0);
    mcoef.assign(2 * (size_t)op.nmult, 0.0);
    mperm.assign(op.nmult, 0);
    for (int G = 0; G < m.NX * m.NY; ++G)
      for (int t = 0; t < op.ndof_owner[G]; ++t) mperm[op.owner_base[G] + t] = FH->F.cell_base[G] + t;
    std::vector<double> gx;
    gauss01(k, gx, gw);
    for (int K = 0; K < m.ncellP; ++K)
      for (int f = 0; f < 4; ++f) {
        const int base = op.mult_face[4 * K + f];
        if (base < 0) continue;
        const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
        for (int q = 0; q < k; ++q) {
          const int id = base + q;
          const int e = mcell[2 * id] < 0 ? 0 : 1;
          mcell[2 * id + e] = K;
          mloc[2 * id + e] = T.rt_face_dof(f, q);
          mcoef[2 * id + e] = sf * gw[q];
        }
      }
  }

  double generation(double x, double y) const {
    if (P.gen_mode == 0) return P.alpha * P.G0 * std::exp(P.alpha * (y - m.geo.H));
    if (P.gen_mode == 1) return P.G0 * std::exp(-P.alpha * x);
    return 0.0;
  }

  // b1, b2 = M u_old + cell and face terms of one region (SolarCell.cpp:1037-1726)
  void rhs_carrier(bool semi, const double* u1, const double* u2, const double* ou1, const double* ou2,
                   const double* pvec, double* b1, double* b2) const {
    const Region& R = semi ? m.semi : m.elec;
    const Region& O = semi ? m.elec : m.semi;
    const std::vector<int>& snb = semi ? m.semi_sigma_nb : m.elec_sigma_nb;
    const std::vector<double>& hc = semi ? hs : he;
    const int nb = T.nb, nq = T.nq, nf = T.nfq, nrt = T.nrt, nc = R.ncell;
    const double inv_perm = 1.0 / (semi ? P.eps_s : P.eps_e), idt = 1.0 / P.dt;
    const double e1 = semi ? P.ne : P.rinf, e2 = semi ? P.pe : P.oinf;
#pragma omp parallel for schedule(static) if (nc >= 1024)
    for (int c = 0; c < nc; ++c) {
      const double* v = &R.V[8 * (size_t)c];
      std::vector<double> d1(nb), d2(nb), dp(nrt), a1p(2 * nb, 0.0), a2p(2 * nb, 0.0), a1v(nb, 0.0),
          a2v(nb, 0.0);
      const size_t ub = 2 * (size_t)nc * nb + (size_t)c * nb;
      for (int a = 0; a < nb; ++a) {
        d1[a] = u1[ub + a];
        d2[a] = u2[ub + a];
      }
      const int pc = semi ? c : m.semi.ncell + c;
      for (int i = 0; i < nrt; ++i) dp[i] = pvec[L.idx[(size_t)pc * L.nloc + i]];
      for (int q = 0; q < nq; ++q) {
        const Geo g = bilinear(v, T.qx[q], T.qy[q]);
        const double jxw = T.qw[q] * g.det;
        double n1 = 0, n2 = 0, hx = 0, hy = 0;
        for (int a = 0; a < nb; ++a) {
          n1 += d1[a] * T.phi[a * nq + q];
          n2 += d2[a] * T.phi[a * nq + q];
        }
        for (int i = 0; i < nrt; ++i) {
          hx += dp[i] * T.psix[i * nq + q];
          hy += dp[i] * T.psiy[i * nq + q];
        }
        // Piola: D = J Dhat / det J
        const double Ex = (g.J00 * hx + g.J01 * hy) / g.det, Ey = (g.J10 * hx + g.J11 * hy) / g.det;
        double src = 0;
        if (semi) {
          src = generation(g.x, g.y);
          if (P.srh) src += (P.ni * P.ni - n1 * n2) / (P.tau_n * (n1 + P.ni) + P.tau_p * (n2 + P.ni));
        }
        for (int a = 0; a < nb; ++a) {
          const double ph = T.phi[a * nq + q];
          a1v[a] += jxw * ph * (n1 * idt + src);
          a2v[a] += jxw * ph * (n2 * idt + src);
          // drift: s_c p.D/eps u with s = -1 (electrons, reductants), +1 (holes, oxidants)
          a1p[a] += -inv_perm * n1 * jxw * Ex * ph;
          a1p[nb + a] += -inv_perm * n1 * jxw * Ey * ph;
          a2p[a] += inv_perm * n2 * jxw * Ex * ph;
          a2p[nb + a] += inv_perm * n2 * jxw * Ey * ph;
        }
      }
      for (int f = 0; f < 4; ++f) {
        const int bid = R.bid[4 * (size_t)c + f];
        if (bid == INTERIOR || bid == NEUMANN || (!semi && bid == SCHOTTKY)) continue;
        double nx, ny, len;
        face_normal(v, f, nx, ny, len);
        const double* fph = &T.fphi[(size_t)f * nb * nf];
        if (bid == DIRICHLET) {
          const double pen = P.penalty / hc[c];
          for (int q = 0; q < nf; ++q) {
            const double jw = T.fw[q] * len;
            for (int a = 0; a < nb; ++a) {
              const double fa = fph[a * nf + q];
              a1p[a] -= fa * nx * jw * e1;
              a1p[nb + a] -= fa * ny * jw * e1;
              a1v[a] += pen * fa * jw * e1;
              a2p[a] -= fa * nx * jw * e2;
              a2p[nb + a] -= fa * ny * jw * e2;
              a2v[a] += pen * fa * jw * e2;
            }
          }
          continue;
        }
        const int e = bid == INTERFACE ? snb[c] : -1;
        if (bid == INTERFACE && e < 0) continue;
        const double* foph = &T.fphi[(size_t)(f ^ 1) * nb * nf];
        const size_t ob = 2 * (size_t)O.ncell * nb + (size_t)(e < 0 ? 0 : e) * nb;
        for (int q = 0; q < nf; ++q) {
          const double jw = T.fw[q] * len;
          double t1 = 0, t2 = 0;
          for (int a = 0; a < nb; ++a) {
            t1 += d1[a] * fph[a * nf + q];
            t2 += d2[a] * fph[a * nf + q];
          }
          double w1, w2;
          if (bid == INTERFACE) {  // lagged traces of the other region, paired by point index
            double r1 = 0, r2 = 0;
            for (int a = 0; a < nb; ++a) {
              r1 += ou1[ob + a] * foph[a * nf + q];
              r2 += ou2[ob + a] * foph[a * nf + q];
            }
            if (semi) {
              w1 = -P.k_et * (t1 - P.ne) * r2 * jw;
              w2 = P.k_ht * (t2 - P.pe) * r1 * jw;
            } else {
              const double cur = (-P.k_et * (r1 - P.ne) * t2 + P.k_ht * (r2 - P.pe) * t1) * jw;
              w1 = cur;
              w2 = -cur;
            }
          } else {  // Schottky
            w1 = -P.v_rec_n * (t1 - P.ne) * jw;
            w2 = P.v_rec_p * (t2 - P.pe) * jw;
          }
          for (int a = 0; a < nb; ++a) {
            a1v[a] += fph[a * nf + q] * w1;
            a2v[a] += fph[a * nf + q] * w2;
          }
        }
      }
      for (int a = 0; a < nb; ++a) {
        b1[(size_t)c * nb + a] = a1p[a];
        b1[(size_t)nc * nb + (size_t)c * nb + a] = a1p[nb + a];
        b1[ub + a] = a1v[a];
        b2[(size_t)c * nb + a] = a2p[a];
        b2[(size_t)nc * nb + (size_t)c * nb + a] = a2p[nb + a];
        b2[ub + a] = a2v[a];
      }
    }
  }

  // one carrier solve: g = f_v + B^T mu A^-1 f_p, S u = g, q = mu A^-1 (f_p - B u)
  void solve_carrier(const Region& R, const LdgRegionOps& O, const LevelFactor& F, double mu, const double* b,
                     double* x) const {
    const int nb = T.nb, nc = R.ncell;
    std::vector<double> W((size_t)nc * 2 * nb), g(F.F.n), u, z;
#pragma omp parallel for schedule(static) if (nc >= 1024)
    for (int c = 0; c < nc; ++c) {
      const LdgCellOps& C = O.cell[c];
      for (int h = 0; h < 2; ++h)
        for (int a = 0; a < nb; ++a) {
          double s = 0;
          for (int t = 0; t < nb; ++t) s += C.Minv[a * nb + t] * b[(size_t)h * nc * nb + (size_t)c * nb + t];
          W[(size_t)c * 2 * nb + h * nb + a] = mu * s;
        }
    }
#pragma omp parallel for schedule(static) if (nc >= 1024)
    for (int X = 0; X < nc; ++X) {
      std::vector<double> acc(nb);
      for (int a = 0; a < nb; ++a) acc[a] = b[2 * (size_t)nc * nb + (size_t)X * nb + a];
      for (int f = -1; f < 4; ++f) {
        const int M = f < 0 ? X : R.nbr[4 * (size_t)X + f];
        if (M < 0) continue;
        // the block of M's current rows that couples to X's density
        const double* B = f < 0 ? O.cell[M].Bself.data() : O.cell[M].Bf.data() + (size_t)(f ^ 1) * 2 * nb * nb;
        for (int i = 0; i < 2 * nb; ++i)
          for (int a = 0; a < nb; ++a) acc[a] += B[i * nb + a] * W[(size_t)M * 2 * nb + i];
      }
      const int pb = F.F.cell_base[R.ci[X] + R.N * R.cj[X]];
      for (int a = 0; a < nb; ++a) g[pb + a] = acc[a];
    }
    F.solve(g, u, z);
#pragma omp parallel for schedule(static) if (nc >= 1024)
    for (int M = 0; M < nc; ++M) {
      const LdgCellOps& C = O.cell[M];
      std::vector<double> s(2 * nb, 0.0);
      const int pm = F.F.cell_base[R.ci[M] + R.N * R.cj[M]];
      for (int i = 0; i < 2 * nb; ++i)
        for (int a = 0; a < nb; ++a) s[i] += C.Bself[i * nb + a] * u[pm + a];
      for (int f = 0; f < 4; ++f) {
        const int X = R.nbr[4 * (size_t)M + f];
        if (X < 0) continue;
        const int px = F.F.cell_base[R.ci[X] + R.N * R.cj[X]];
        const double* B = C.Bf.data() + (size_t)f * 2 * nb * nb;
        for (int i = 0; i < 2 * nb; ++i)
          for (int a = 0; a < nb; ++a) s[i] += B[i * nb + a] * u[px + a];
      }
      for (int h = 0; h < 2; ++h)
        for (int a = 0; a < nb; ++a) {
          double t = 0;
          for (int r = 0; r < nb; ++r) t += C.Minv[a * nb + r] * s[h * nb + r];
          x[(size_t)h * nc * nb + (size_t)M * nb + a] = W[(size_t)M * 2 * nb + h * nb + a] - mu * t;
        }
      for (int a = 0; a < nb; ++a) x[2 * (size_t)nc * nb + (size_t)M * nb + a] = u[pm + a];
    }
  }

  // Poisson: right-hand side per cell, condensation, multiplier solve, recovery
  void poisson(const double* un, const double* up, const double* ur, const double* uo, double* out) const {
    const int nl = op.nloc, nrt = op.nrt, nphi = op.nphi, nb = T.nb, nq = T.nq, ncs = m.semi.ncell;
    std::vector<double> y((size_t)m.ncellP * nl), r(FH->F.n), lam, z;
#pragma omp parallel for schedule(static) if (m.ncellP >= 1024)
    for (int K = 0; K < m.ncellP; ++K) {
      std::vector<double> f(nl, 0.0);
      const double* U = &op.ubias[(size_t)K * 3 * nrt];
      for (int i = 0; i < nrt; ++i) f[i] = P.phi_bi * U[i] - P.phi_app * U[nrt + i] - P.phi_sch * U[2 * nrt + i];
      const bool semi = K < ncs;
      const int c = semi ? K : K - ncs, nc = semi ? ncs : m.elec.ncell;
      const double* u1 = semi ? un : ur;
      const double* u2 = semi ? up : uo;
      const double rho0 = semi ? (P.ne - P.pe) : 0.0;
      const double* v = &m.pV[8 * (size_t)K];
      for (int q = 0; q < nq; ++q) {
        const Geo g = bilinear(v, T.qx[q], T.qy[q]);
        double n1 = 0, n2 = 0;
        for (int a = 0; a < nb; ++a) {
          n1 += u1[2 * (size_t)nc * nb + (size_t)c * nb + a] * T.phi[a * nq + q];
          n2 += u2[2 * (size_t)nc * nb + (size_t)c * nb + a] * T.phi[a * nq + q];
        }
        const double rho = rho0 + (-n1 + n2);
        for (int l = 0; l < nphi; ++l) f[nrt + l] -= T.phiP[l * nq + q] * rho * T.qw[q] * g.det;
      }
      const double* A = &op.Ainv[(size_t)K * nl * nl];
      for (int i = 0; i < nl; ++i) {
        double s = 0;
        for (int j = 0; j < nl; ++j) s += A[i * nl + j] * f[j];
        y[(size_t)K * nl + i] = s;
      }
    }
#pragma omp parallel for schedule(static) if (op.nmult >= 4096)
    for (int mm = 0; mm < op.nmult; ++mm) {
      double s = 0;
      for (int e = 0; e < 2; ++e) {
        const int c = mcell[2 * mm + e];
        if (c >= 0) s += mcoef[2 * mm + e] * y[(size_t)c * nl + mloc[2 * mm + e]];
      }
      r[mperm[mm]] = s;
    }
    FH->solve(r, lam, z);
#pragma omp parallel for schedule(static) if (m.ncellP >= 1024)
    for (int K = 0; K < m.ncellP; ++K) {
      std::vector<double> e(4 * k, 0.0);
      for (int f = 0; f < 4; ++f) {
        const int base = op.mult_face[4 * K + f];
        const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
        for (int q = 0; q < k; ++q) e[f * k + q] = base < 0 ? 0.0 : sf * gw[q] * lam[mperm[base + q]];
      }
      const double* A = &op.Ainv[(size_t)K * nl * nl];
      const int I = m.pI[K], J = m.pJ[K];
      for (int i = 0; i < nl; ++i) {
        double s = y[(size_t)K * nl + i];
        for (int f = 0; f < 4; ++f)
          for (int q = 0; q < k; ++q) s -= A[i * nl + T.rt_face_dof(f, q)] * e[f * k + q];
        if (i < nrt) {
          // the face a local RT DoF sits on: left/bottom faces belong to the
          // cell, right/top faces only on the boundary; Neumann D.n = 0
          int fc = -1;
          if (i < T.nrtx) {
            const int a = i % (k + 1);
            fc = a == 0 ? 0 : (a == k ? 1 : -1);
          } else {
            const int b = (i - T.nrtx) / k;
            fc = b == 0 ? 2 : (b == k ? 3 : -1);
          }
          if (fc == 1 && I != m.NX - 1) continue;
          if (fc == 3 && J != m.NY - 1) continue;
          if (fc >= 0 && m.pbid[4 * (size_t)K + fc] == NEUMANN) s = 0.0;
        }
        out[L.idx[(size_t)K * nl + i]] = s;
      }
    }
  }

  struct State {
    std::vector<double> u[4], pv;
  };

  void initial(State& s) const {
    const double vals[4] = {P.ne, P.pe, P.rinf, P.oinf};
    const int64_t n[4] = {Ns, Ns, Ne, Ne};
    for (int c = 0; c < 4; ++c) {
      s.u[c].assign(n[c], 0.0);
      for (int64_t j = 2 * n[c] / 3; j < n[c]; ++j) s.u[c][j] = vals[c];
    }
    s.pv.assign(NP, 0.0);
    poisson(s.u[0].data(), s.u[1].data(), s.u[2].data(), s.u[3].data(), s.pv.data());
  }

  void step(State& s) const {
    std::vector<double> b[4] = {std::vector<double>(Ns), std::vector<double>(Ns), std::vector<double>(Ne),
                                std::vector<double>(Ne)};
    rhs_carrier(true, s.u[0].data(), s.u[1].data(), s.u[2].data(), s.u[3].data(), s.pv.data(), b[0].data(),
                b[1].data());
    rhs_carrier(false, s.u[2].data(), s.u[3].data(), s.u[0].data(), s.u[1].data(), s.pv.data(), b[2].data(),
                b[3].data());
    solve_carrier(m.semi, os, *Fs[0], P.mu[0], b[0].data(), s.u[0].data());
    solve_carrier(m.semi, os, share_s ? *Fs[0] : *Fs[1], P.mu[1], b[1].data(), s.u[1].data());
    solve_carrier(m.elec, oe, *Fe[0], P.mu[2], b[2].data(), s.u[2].data());
    solve_carrier(m.elec, oe, share_e ? *Fe[0] : *Fe[1], P.mu[3], b[3].data(), s.u[3].data());
    poisson(s.u[0].data(), s.u[1].data(), s.u[2].data(), s.u[3].data(), s.pv.data());
  }
};

}  // namespace

int main(int argc, char** argv) {
  std::map<std::string, std::string> kv;
  for (int i = 1; i < argc; ++i) {
    const char* e = std::strchr(argv[i], '=');
    if (!e) {
      std::fprintf(stderr, "argument without '=': %s\n", argv[i]);
      return 2;
    }
    kv[std::string(argv[i], e - argv[i])] = e + 1;
  }
  auto num = [&](const char* key, double def) { return kv.count(key) ? std::atof(kv[key].c_str()) : def; };
  Params P;
  P.k = (int)num("k", 1);
  P.dt = num("dt", 1e-3);
  P.penalty = num("penalty", 1.0);
  P.mu[0] = num("mu_n", 1.0);
  P.mu[1] = num("mu_p", 1.0);
  P.mu[2] = num("mu_r", 1.0);
  P.mu[3] = num("mu_o", 1.0);
  P.eps_s = num("eps_s", 1.0);
  P.eps_e = num("eps_e", 1.0);
  P.lam2 = num("lam2", 1.0);
  P.k_et = num("k_et", 1e-2);
  P.k_ht = num("k_ht", 1e-2);
  P.v_rec_n = num("v_rec_n", 0.0);
  P.v_rec_p = num("v_rec_p", 0.0);
  P.ne = num("ne", 1.0);
  P.pe = num("pe", 0.0);
  P.rinf = num("rinf", 1.0);
  P.oinf = num("oinf", 1.0);
  P.phi_bi = num("phi_bi", 0.0);
  P.phi_app = num("phi_app", 0.0);
  P.gen_mode = (int)num("gen_mode", 0);
  P.G0 = num("G0", 1.0);
  P.alpha = num("alpha", 1.0);
  P.srh = (int)num("srh", 0);
  P.ni = num("ni", 0.1);
  P.tau_n = num("tau_n", 1.0);
  P.tau_p = num("tau_p", 1.0);
  GeometrySpec g;
  g.r1 = num("r1", 0.5);
  g.r2 = num("r2", 0.5);
  g.H = num("H", 1.0);
  g.L = num("L", 1.0);
  g.insulated = num("insulated", 0) != 0;
  g.schottky = num("schottky", 0) != 0;
  const int r = (int)num("r", 4), max_steps = (int)num("steps", 5), leaf = (int)num("leaf", 16);
  // tol > 0: stop once the relative steady residual ||x^{n+1} - x^n|| /
  // (||x^n|| dt) over the five state vectors is at most tol
  const double tol = num("tol", 0.0);
  const double t0 = now();
  Problem pb(P, g, r, leaf);
  Problem::State s;
  pb.initial(s);
  const double t1 = now();
  int steps = 0;
  double res = -1.0;
  while (steps < max_steps) {
    Problem::State old;
    if (tol > 0) old = s;
    pb.step(s);
    ++steps;
    if (tol > 0) {
      double dd = 0, nn = 0;
      for (int c = 0; c < 5; ++c) {
        const std::vector<double>& a = c < 4 ? s.u[c] : s.pv;
        const std::vector<double>& b = c < 4 ? old.u[c] : old.pv;
        for (size_t i = 0; i < a.size(); ++i) {
          dd += (a[i] - b[i]) * (a[i] - b[i]);
          nn += b[i] * b[i];
        }
      }
      res = std::sqrt(dd) / (std::sqrt(nn) * P.dt);
      if (res <= tol) break;
    }
  }
  const double t2 = now();
  const double dofs = 2.0 * (double)pb.Ns + 2.0 * (double)pb.Ne + (double)pb.NP;  // DOF-updates per step
  std::printf("{\"setup_s\": %.6f, \"steps\": %d, \"run_s\": %.6f, \"step_s\": %.6f, \"threads\": %d, "
              "\"dof_updates_per_step\": %.0f, \"residual\": %.6e}\n",
              t1 - t0, steps, t2 - t1, steps ? (t2 - t1) / steps : 0.0, omp_get_max_threads(), dofs, res);
  if (kv.count("out")) {
    FILE* fo = std::fopen(kv["out"].c_str(), "wb");
    if (!fo) return 3;
    for (int c = 0; c < 4; ++c) std::fwrite(s.u[c].data(), sizeof(double), s.u[c].size(), fo);
    std::fwrite(s.pv.data(), sizeof(double), s.pv.size(), fo);
    std::fclose(fo);
  }
  return 0;
}
