// The Poisson vector in deal.II's basis and numbering (see dealii_layout.hpp).
#include "dealii_layout.hpp"

#include <cmath>
#include <stdexcept>

#include "ops.hpp"

namespace pecs {

namespace {

// Legendre polynomial of degree n on [0, 1], scaled orthonormal
double legendre01(int n, double t) {
  const double x = 2.0 * t - 1.0;
  double p0 = 1.0, p1 = x;
  if (n == 0) return 1.0;
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  return std::sqrt(2.0 * n + 1.0) * p1;
}

// which face a deal.II local DoF sits on (-1: interior)
int dii_face(int i, int nf) { return i < 4 * nf ? i / nf : -1; }

}  // namespace

DealiiPoisson dealii_poisson(const SolarCellMesh& m, const PoissonLayout& L, const RefTables& T) {
  DealiiPoisson X;
  const int deg = T.deg;
  X.deg = deg;
  X.nf = deg + 1;
  X.nint = 2 * deg * (deg + 1);
  X.nrt = T.nrt;
  X.nphi = T.nphi;
  X.ncell = m.ncellP;
  X.n = L.n;
  X.nD = L.nD;
  X.phi_off = L.nD;
  const int nrt = X.nrt, nf = X.nf;
  if (4 * nf + X.nint != nrt) throw std::runtime_error("dealii_poisson: RT DoF count mismatch");
  // which face each of this library's local DoFs sits on
  X.own_face.assign(nrt, -1);
  for (int f = 0; f < 4; ++f)
    for (int q = 0; q < nf; ++q) X.own_face[T.rt_face_dof(f, q)] = f;
  const std::vector<int>& own_face = X.own_face;
  std::vector<double> gs, gw;
  gauss01(deg + 1, gs, gw);
  X.N.assign((size_t)nrt * nrt, 0.0);
  for (int j = 0; j < nrt; ++j) {
    // face functionals
    for (int f = 0; f < 4; ++f) {
      if (own_face[j] >= 0 && own_face[j] != f) continue;  // no normal trace on face f
      if (own_face[j] < 0) continue;                       // interior functions: no normal trace
      for (int i = 0; i < nf; ++i) {
        double s = 0.0;
        for (int q = 0; q < deg + 1; ++q) {
          double xi, eta, vx, vy, dv;
          face_point(f, gs[q], xi, eta);
          T.rt_eval(j, xi, eta, vx, vy, dv);
          s += gw[q] * legendre01(i, gs[q]) * (f < 2 ? vx : vy);
        }
        X.N[(size_t)(f * nf + i) * nrt + j] = s;
      }
    }
    // interior functionals: degree <= deg-1 in direction d, <= deg in the other
    for (int d = 0; d < 2; ++d) {
      const int nx = d == 0 ? deg : deg + 1;
      for (int jj = 0; jj < deg * (deg + 1); ++jj) {
        const int jx = jj % nx, jy = jj / nx;
        double s = 0.0;
        for (int qy = 0; qy < deg + 1; ++qy)
          for (int qx = 0; qx < deg + 1; ++qx) {
            double vx, vy, dv;
            T.rt_eval(j, gs[qx], gs[qy], vx, vy, dv);
            s += gw[qx] * gw[qy] * legendre01(jx, gs[qx]) * legendre01(jy, gs[qy]) * (d == 0 ? vx : vy);
          }
        X.N[(size_t)(4 * nf + 2 * jj + d) * nrt + j] = s;
      }
    }
  }
  X.Ninv.assign((size_t)nrt * nrt, 0.0);
  if (!dense_inverse(nrt, X.N.data(), X.Ninv.data()))
    throw std::runtime_error("dealii_poisson: singular change of basis");
  // own face coefficients depend on deal.II's coefficients of the same face only
  for (int j = 0; j < nrt; ++j)
    for (int i = 0; i < nrt; ++i)
      if (own_face[j] >= 0 && dii_face(i, nf) != own_face[j]) X.Ninv[(size_t)j * nrt + i] = 0.0;
  // deal.II's numbering: lines in cell order, each cell's interior after its new lines
  const int nV = (m.NX + 1) * m.NY, nH = m.NX * (m.NY + 1);
  std::vector<int64_t> vbase(nV, -1), hbase(nH, -1);
  X.gd.assign((size_t)m.ncellP * nrt, -1);
  X.go.assign((size_t)m.ncellP * nrt, -1);
  X.owns.assign((size_t)m.ncellP * 4, 0);
  int64_t next = 0;
  for (int p = 0; p < m.ncellP; ++p) {
    const int I = m.pI[p], J = m.pJ[p];
    int64_t* base[4] = {&vbase[J * (m.NX + 1) + I], &vbase[J * (m.NX + 1) + I + 1], &hbase[J * m.NX + I],
                        &hbase[(J + 1) * m.NX + I]};
    for (int f = 0; f < 4; ++f) {
      if (*base[f] < 0) {
        *base[f] = next;
        next += nf;
      }
      for (int i = 0; i < nf; ++i) X.gd[(size_t)p * nrt + f * nf + i] = *base[f] + i;
    }
    for (int i = 0; i < X.nint; ++i) X.gd[(size_t)p * nrt + 4 * nf + i] = next + i;
    next += X.nint;
    for (int j = 0; j < nrt; ++j) X.go[(size_t)p * nrt + j] = L.idx[(size_t)p * L.nloc + j];
    X.owns[4 * p + 0] = X.owns[4 * p + 2] = 1;
    X.owns[4 * p + 1] = I == m.NX - 1;
    X.owns[4 * p + 3] = J == m.NY - 1;
  }
  if (next != X.nD) throw std::runtime_error("dealii_poisson: D DoF count mismatch");
  return X;
}

void poisson_to_dealii(const DealiiPoisson& X, const double* own, double* dii) {
  const int nrt = X.nrt, nf = X.nf;
  std::vector<double> c(nrt);
  for (int p = 0; p < X.ncell; ++p) {
    for (int j = 0; j < nrt; ++j) c[j] = own[X.go[(size_t)p * nrt + j]];
    for (int i = 0; i < nrt; ++i) {
      const int f = dii_face(i, nf);
      if (f >= 0 && !X.owns[4 * p + f]) continue;  // written by the face's owner
      double s = 0.0;
      for (int j = 0; j < nrt; ++j) s += X.N[(size_t)i * nrt + j] * c[j];
      dii[X.gd[(size_t)p * nrt + i]] = s;
    }
  }
  for (int64_t g = X.nD; g < X.n; ++g) dii[g] = own[g];
}

void poisson_from_dealii(const DealiiPoisson& X, const double* dii, double* own) {
  const int nrt = X.nrt;
  std::vector<double> c(nrt);
  for (int p = 0; p < X.ncell; ++p) {
    for (int i = 0; i < nrt; ++i) c[i] = dii[X.gd[(size_t)p * nrt + i]];
    for (int j = 0; j < nrt; ++j) {
      const int f = X.own_face[j];
      if (f >= 0 && !X.owns[4 * p + f]) continue;  // written by the face's owner
      double s = 0.0;
      for (int i = 0; i < nrt; ++i) s += X.Ninv[(size_t)j * nrt + i] * c[i];
      own[X.go[(size_t)p * nrt + j]] = s;
    }
  }
  for (int64_t g = X.nD; g < X.n; ++g) own[g] = dii[g];
}

void poisson_rhs_from_dealii(const DealiiPoisson& X, const double* dii, double* own) {
  // b_own = G^T b_dii, G the global change of coefficients: per cell N^T u,
  // u = the cell's deal.II entries with the faces it does not own left out
  // (their N_FF^T part comes from the owner), added into b_own
  const int nrt = X.nrt, nf = X.nf;
  for (int64_t g = 0; g < X.nD; ++g) own[g] = 0.0;
  std::vector<double> u(nrt);
  for (int p = 0; p < X.ncell; ++p) {
    for (int i = 0; i < nrt; ++i) {
      const int f = dii_face(i, nf);
      u[i] = (f >= 0 && !X.owns[4 * p + f]) ? 0.0 : dii[X.gd[(size_t)p * nrt + i]];
    }
    for (int j = 0; j < nrt; ++j) {
      double s = 0.0;
      for (int i = 0; i < nrt; ++i) s += X.N[(size_t)i * nrt + j] * u[i];
      own[X.go[(size_t)p * nrt + j]] += s;
    }
  }
  for (int64_t g = X.nD; g < X.n; ++g) own[g] = dii[g];
}

void poisson_rhs_to_dealii(const DealiiPoisson& X, const double* own, double* dii) {
  // solve G^T b_dii = b_own: interior entries per cell (N_II^{-T}), then the
  // face entries once the interior ones of both neighbours are known
  const int nrt = X.nrt, nf = X.nf;
  std::vector<double> r((size_t)X.nD, 0.0), bi((size_t)X.ncell * nrt, 0.0);
  for (int p = 0; p < X.ncell; ++p) {
    double* b = &bi[(size_t)p * nrt];
    for (int i = 4 * nf; i < nrt; ++i) {  // (N_II^{-1})^T: the interior block of Ninv, transposed
      double s = 0.0;
      for (int j = 0; j < nrt; ++j)
        if (X.own_face[j] < 0) s += X.Ninv[(size_t)j * nrt + i] * own[X.go[(size_t)p * nrt + j]];
      b[i] = s;
      dii[X.gd[(size_t)p * nrt + i]] = s;
    }
    for (int j = 0; j < nrt; ++j) {  // the interior entries' share of this library's face entries
      if (X.own_face[j] < 0) continue;
      double s = 0.0;
      for (int i = 4 * nf; i < nrt; ++i) s += X.N[(size_t)i * nrt + j] * b[i];
      r[X.go[(size_t)p * nrt + j]] += s;
    }
  }
  for (int p = 0; p < X.ncell; ++p)
    for (int f = 0; f < 4; ++f) {
      if (!X.owns[4 * p + f]) continue;
      for (int i = f * nf; i < (f + 1) * nf; ++i) {  // N_FF^{-T} restricted to face f
        double s = 0.0;
        for (int j = 0; j < nrt; ++j) {
          if (X.own_face[j] != f) continue;  // the face-face block of Ninv is N_FF^{-1}
          const int64_t g = X.go[(size_t)p * nrt + j];
          s += X.Ninv[(size_t)j * nrt + i] * (own[g] - r[g]);
        }
        dii[X.gd[(size_t)p * nrt + i]] = s;
      }
    }
  for (int64_t g = X.nD; g < X.n; ++g) dii[g] = own[g];
}

}  // namespace pecs
