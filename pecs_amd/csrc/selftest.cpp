// Native self-test of the setup path (CPU only, not part of the library):
// builds the meshes, local operators, the SPD Schur / multiplier matrices,
// factors them and checks ||S x - b|| / ||b|| of the stored factor through a
// plain host triangular solve written only for this test.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "cholesky.hpp"
#include "mesh.hpp"
#include "ops.hpp"

using namespace pecs;

static void block_matvec(const BlockSPD& S, const std::vector<int>& base, const double* x, double* y) {
  const int no = S.noff();
  for (int X = 0; X < S.NXc * S.NYc; ++X) {
    const int nx = S.ndof[X];
    for (int a = 0; a < nx; ++a) y[base[X] + a] = 0.0;
    const int xi = X % S.NXc, xj = X / S.NXc;
    for (int o = 0; o < no; ++o) {
      const int yi = xi + S.offi[o], yj = xj + S.offj[o];
      if (yi < 0 || yi >= S.NXc || yj < 0 || yj >= S.NYc) continue;
      const int Y = yi + S.NXc * yj;
      const auto& B = S.blk[(size_t)X * no + o];
      if (B.empty()) continue;
      const int ny = S.ndof[Y];
      for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) y[base[X] + a] += B[(size_t)a * ny + b] * x[base[Y] + b];
    }
  }
}

static void factor_solve_host(const SupernodalFactor& F, const double* b, double* x) {
  std::vector<double> y(b, b + F.n);
  for (size_t s = 0; s < F.nodes.size(); ++s) {  // forward, postorder
    const SNode& N = F.nodes[s];
    std::vector<double> t(N.ns);
    for (int i = 0; i < N.ns; ++i) {
      double v = 0;
      for (int j = 0; j <= i; ++j) v += F.Linv[N.linv_off + (int64_t)i * N.ns + j] * y[N.p0 + j];
      t[i] = v;
    }
    for (int i = 0; i < N.ns; ++i) y[N.p0 + i] = t[i];
    for (int r = 0; r < N.ms; ++r) {
      double v = 0;
      for (int j = 0; j < N.ns; ++j) v += F.L21[N.l21_off + (int64_t)r * N.ns + j] * t[j];
      y[F.rows[N.rows_off + r]] -= v;
    }
  }
  for (int i = 0; i < F.n; ++i) x[i] = y[i];
  for (int s = (int)F.nodes.size() - 1; s >= 0; --s) {  // backward
    const SNode& N = F.nodes[s];
    std::vector<double> t(N.ns);
    for (int j = 0; j < N.ns; ++j) {
      double v = x[N.p0 + j];
      for (int r = 0; r < N.ms; ++r) v -= F.L21[N.l21_off + (int64_t)r * N.ns + j] * x[F.rows[N.rows_off + r]];
      t[j] = v;
    }
    for (int j = 0; j < N.ns; ++j) {
      double v = 0;
      for (int i = j; i < N.ns; ++i) v += F.Linv[N.linv_off + (int64_t)i * N.ns + j] * t[i];
      x[N.p0 + j] = v;
    }
  }
}

static double check(const char* name, const BlockSPD& S, int leaf) {
  auto t0 = std::chrono::steady_clock::now();
  SupernodalFactor F = cholesky_factor(S, leaf);
  double ts = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::mt19937_64 rng(20261202);
  std::uniform_real_distribution<double> U(-1, 1);
  std::vector<double> x(F.n), b(F.n), xs(F.n), r(F.n);
  for (auto& v : x) v = U(rng);
  block_matvec(S, F.cell_base, x.data(), b.data());
  factor_solve_host(F, b.data(), xs.data());
  block_matvec(S, F.cell_base, xs.data(), r.data());
  double nr = 0, nb = 0, ne = 0, nx = 0;
  for (int i = 0; i < F.n; ++i) {
    nr += (r[i] - b[i]) * (r[i] - b[i]);
    nb += b[i] * b[i];
    ne += (xs[i] - x[i]) * (xs[i] - x[i]);
    nx += x[i] * x[i];
  }
  double rel = std::sqrt(nr / nb);
  std::printf("%s: n=%d nodes=%zu levels=%d nnzL=%lld (%.1f MB) flops=%.3g factor %.2fs  resid=%.3e err=%.3e\n", name,
              F.n, F.nodes.size(), F.nlevels, (long long)F.nnz_L(), F.nnz_L() * 8e-6, F.flops, ts, rel,
              std::sqrt(ne / nx));
  std::printf("   time: assemble %.2fs factor %.2fs inverse %.2fs\n", F.t_assemble, F.t_factor, F.t_inverse);
  return rel;
}

int main(int argc, char** argv) {
  int r = argc > 1 ? std::atoi(argv[1]) : 4;
  int k = argc > 2 ? std::atoi(argv[2]) : 1;
  int leaf = argc > 3 ? std::atoi(argv[3]) : 16;
  int conic = argc > 4 ? std::atoi(argv[4]) : 0;
  GeometrySpec g;
  if (conic) {
    g.r1 = 0.5;
    g.r2 = 0.5 + std::tan(10.0 * M_PI / 180.0);
    g.L = 1.5;
  }
  Params P;
  P.k = k;
  P.dt = (k == 1) ? 1e-3 : 5e-4;
  auto t0 = std::chrono::steady_clock::now();
  SolarCellMesh m = make_solarcell_mesh(r, g);
  RefTables T(k);
  LdgRegionOps os = build_ldg_ops(m.semi, T, P);
  LdgRegionOps oe = build_ldg_ops(m.elec, T, P);
  PoissonLayout L = poisson_layout(m, k - 1);
  PoissonOps op = build_poisson_ops(m, L, T, P);
  double ts = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("setup ops %.2fs: ldg types semi=%d elec=%d poisson types=%d nmult=%d\n", ts, os.ntype, oe.ntype,
              op.ntype, op.nmult);
  t0 = std::chrono::steady_clock::now();
  BlockSPD Ss = schur_matrix(m.semi, os, 1.0);
  BlockSPD Se = schur_matrix(m.elec, oe, 1.0);
  BlockSPD H = multiplier_matrix(m, op, T);
  std::printf("assemble S/H %.2fs\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  double worst = 0;
  worst = std::max(worst, check("S_semi", Ss, leaf));
  worst = std::max(worst, check("S_elec", Se, leaf));
  worst = std::max(worst, check("H_poisson", H, leaf));
  std::printf("%s\n", worst < 1e-12 ? "SELFTEST OK" : "SELFTEST FAIL");
  return worst < 1e-12 ? 0 : 1;
}
