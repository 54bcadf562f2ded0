// Reference-cell tabulations for the PECS IMEX path (host side).
//
// What deal.II's FEValues give the reference (Assembly.cpp:9-68):
//   * QGauss<2>(k+2) cell rule and QGauss<1>(k+2) face rule, used for every
//     integral of the per-step path (e.g. SolarCell.cpp:402-403);
//   * FE_DGQ(k): Lagrange basis on equispaced support points, lexicographic
//     (x fastest); the carrier element FESystem(FE_DGQ(k)^2, FE_DGQ(k))
//     (SolarCell.cpp:29-30) has local order [Jx | Jy | u];
//   * FE_RaviartThomas(k-1) x FE_DGQ(k-1) for the mixed Poisson problem
//     (SolarCell.cpp:23-24).  The RT basis is a nodal basis of the same space:
//     x-component = Lagrange on deg+2 equispaced points in xi times Lagrange on
//     the deg+1 Gauss points in eta (y-component transposed).  Face nodes are
//     reference normal fluxes at the face Gauss points, so a face DoF is shared
//     by the two cells of a face as in deal.II.
// Face numbering follows deal.II in 2D: 0 xi=0, 1 xi=1, 2 eta=0, 3 eta=1.
#pragma once
#include <cmath>
#include <vector>

namespace pecs {

constexpr int KMAX = 3;

inline void gauss01(int n, std::vector<double>& x, std::vector<double>& w) {
  // Gauss-Legendre nodes by Newton on P_n, mapped to [0,1].
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
      }
      double dp = n * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = 0.0;
    for (int j = 1; j <= n; ++j) {
      double p2 = p1;
      p1 = p0;
      p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
    }
    double dp = n * (z * p0 - p1) / (z * z - 1.0);
    // node order ascending on [0,1]
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);  // 2/((1-z^2)P'^2) * 1/2
  }
}

// Lagrange basis on `nodes`: value and derivative of l_a at t.
inline void lagrange(const std::vector<double>& nodes, int a, double t, double& v, double& d) {
  const int n = (int)nodes.size();
  v = 1.0;
  for (int b = 0; b < n; ++b)
    if (b != a) v *= (t - nodes[b]) / (nodes[a] - nodes[b]);
  d = 0.0;
  for (int c = 0; c < n; ++c) {
    if (c == a) continue;
    double term = 1.0 / (nodes[a] - nodes[c]);
    for (int b = 0; b < n; ++b)
      if (b != a && b != c) term *= (t - nodes[b]) / (nodes[a] - nodes[b]);
    d += term;
  }
}

inline std::vector<double> equispaced(int k) {
  if (k == 0) return {0.5};
  std::vector<double> x(k + 1);
  for (int i = 0; i <= k; ++i) x[i] = (double)i / k;
  return x;
}

inline void face_point(int f, double s, double& xi, double& eta) {
  switch (f) {
    case 0: xi = 0.0; eta = s; break;
    case 1: xi = 1.0; eta = s; break;
    case 2: xi = s; eta = 0.0; break;
    default: xi = s; eta = 1.0; break;
  }
}

// All reference tabulations of one polynomial degree k (carrier degree).
struct RefTables {
  int k = 0, n1 = 0, nb = 0, nq1 = 0, nq = 0, nfq = 0;
  int deg = 0, nrt = 0, nrtx = 0, nphi = 0, nlocP = 0;
  std::vector<double> qx, qy, qw;          // [nq]
  std::vector<double> fs, fw;              // [nfq]
  std::vector<double> phi, dphx, dphy;     // [nb*nq]  (a*nq + q)
  std::vector<double> fphi;                // [4*nb*nfq] (f*nb + a)*nfq + q
  // Poisson: RT_{deg} x DGQ_{deg}
  std::vector<double> psix, psiy, divpsi;  // [nrt*nq]
  std::vector<double> phiP;                // [nphi*nq]
  std::vector<double> fpsix, fpsiy;        // [4*nrt*nfq]
  std::vector<double> rtX, rtG;            // RT nodes: deg+2 equispaced (normal), deg+1 Gauss (tangential)

  explicit RefTables(int kk) { build(kk); }

  void build(int kk) {
    k = kk;
    n1 = k + 1;
    nb = n1 * n1;
    nq1 = k + 2;
    nq = nq1 * nq1;
    nfq = nq1;
    std::vector<double> g, gw;
    gauss01(nq1, g, gw);
    qx.resize(nq); qy.resize(nq); qw.resize(nq);
    for (int j = 0; j < nq1; ++j)
      for (int i = 0; i < nq1; ++i) {
        qx[i + nq1 * j] = g[i];
        qy[i + nq1 * j] = g[j];
        qw[i + nq1 * j] = gw[i] * gw[j];
      }
    fs = g;
    fw = gw;
    auto nodes = equispaced(k);
    phi.assign(nb * nq, 0); dphx.assign(nb * nq, 0); dphy.assign(nb * nq, 0);
    for (int b = 0; b < n1; ++b)
      for (int a = 0; a < n1; ++a) {
        int i = a + n1 * b;
        for (int q = 0; q < nq; ++q) {
          double vx, dx, vy, dy;
          lagrange(nodes, a, qx[q], vx, dx);
          lagrange(nodes, b, qy[q], vy, dy);
          phi[i * nq + q] = vx * vy;
          dphx[i * nq + q] = dx * vy;
          dphy[i * nq + q] = vx * dy;
        }
      }
    fphi.assign(4 * nb * nfq, 0);
    for (int f = 0; f < 4; ++f)
      for (int b = 0; b < n1; ++b)
        for (int a = 0; a < n1; ++a)
          for (int q = 0; q < nfq; ++q) {
            double xi, eta, vx, dx, vy, dy;
            face_point(f, fs[q], xi, eta);
            lagrange(nodes, a, xi, vx, dx);
            lagrange(nodes, b, eta, vy, dy);
            fphi[(f * nb + a + n1 * b) * nfq + q] = vx * vy;
          }
    // Poisson element
    deg = k - 1;
    nrtx = (deg + 2) * (deg + 1);
    nrt = 2 * nrtx;
    nphi = (deg + 1) * (deg + 1);
    nlocP = nrt + nphi;
    rtX.assign(deg + 2, 0.0);
    for (int i = 0; i < deg + 2; ++i) rtX[i] = (double)i / (deg + 1);
    std::vector<double> GW;
    gauss01(deg + 1, rtG, GW);
    auto pnodes = equispaced(deg);
    psix.assign(nrt * nq, 0); psiy.assign(nrt * nq, 0); divpsi.assign(nrt * nq, 0);
    fpsix.assign(4 * nrt * nfq, 0); fpsiy.assign(4 * nrt * nfq, 0);
    for (int i = 0; i < nrt; ++i) {
      for (int q = 0; q < nq; ++q) {
        double vx, vy, dv;
        rt_eval(i, qx[q], qy[q], vx, vy, dv);
        psix[i * nq + q] = vx;
        psiy[i * nq + q] = vy;
        divpsi[i * nq + q] = dv;
      }
      for (int f = 0; f < 4; ++f)
        for (int q = 0; q < nfq; ++q) {
          double xi, eta, vx, vy, dv;
          face_point(f, fs[q], xi, eta);
          rt_eval(i, xi, eta, vx, vy, dv);
          fpsix[(f * nrt + i) * nfq + q] = vx;
          fpsiy[(f * nrt + i) * nfq + q] = vy;
        }
    }
    phiP.assign(nphi * nq, 0);
    for (int b = 0; b <= deg; ++b)
      for (int a = 0; a <= deg; ++a)
        for (int q = 0; q < nq; ++q) {
          double vx, dx, vy, dy;
          lagrange(pnodes, a, qx[q], vx, dx);
          lagrange(pnodes, b, qy[q], vy, dy);
          phiP[(a + (deg + 1) * b) * nq + q] = vx * vy;
        }
  }

  // Reference RT basis function i at (xi, eta): value (vx, vy) and divergence.
  void rt_eval(int i, double xi, double eta, double& vx, double& vy, double& dv) const {
    vx = vy = dv = 0.0;
    if (i < nrtx) {
      int a = i % (deg + 2), b = i / (deg + 2);
      double l, dl, gg, dg;
      lagrange(rtX, a, xi, l, dl);
      lagrange(rtG, b, eta, gg, dg);
      vx = l * gg;
      dv = dl * gg;
    } else {
      int ii = i - nrtx;
      int a = ii % (deg + 1), b = ii / (deg + 1);
      double gg, dg, l, dl;
      lagrange(rtG, a, xi, gg, dg);
      lagrange(rtX, b, eta, l, dl);
      vy = gg * l;
      dv = gg * dl;
    }
  }

  // Local RT index of face-node m of face f (the DoF carrying the flux there)
  int rt_face_dof(int f, int m) const {
    switch (f) {
      case 0: return 0 + (deg + 2) * m;
      case 1: return (deg + 1) + (deg + 2) * m;
      case 2: return nrtx + m;
      default: return nrtx + m + (deg + 1) * (deg + 1);
    }
  }
};

// MappingQ1 of a quad with vertices v[4][2] (deal.II order) at (xi, eta).
struct Geo {
  double x, y, J00, J01, J10, J11, det;
};
inline Geo bilinear(const double* v, double xi, double eta) {
  Geo g;
  const double a = (1 - xi) * (1 - eta), b = xi * (1 - eta), c = (1 - xi) * eta, d = xi * eta;
  g.x = v[0] * a + v[2] * b + v[4] * c + v[6] * d;
  g.y = v[1] * a + v[3] * b + v[5] * c + v[7] * d;
  g.J00 = (v[2] - v[0]) * (1 - eta) + (v[6] - v[4]) * eta;  // dx/dxi
  g.J10 = (v[3] - v[1]) * (1 - eta) + (v[7] - v[5]) * eta;  // dy/dxi
  g.J01 = (v[4] - v[0]) * (1 - xi) + (v[6] - v[2]) * xi;    // dx/deta
  g.J11 = (v[5] - v[1]) * (1 - xi) + (v[7] - v[3]) * xi;    // dy/deta
  g.det = g.J00 * g.J11 - g.J01 * g.J10;
  return g;
}

// Outward unit normal and |dx/ds| of face f of a straight-edged quad.
inline void face_normal(const double* v, int f, double& nx, double& ny, double& len) {
  int a, b;
  switch (f) {
    case 0: a = 0; b = 2; break;
    case 1: a = 1; b = 3; break;
    case 2: a = 0; b = 1; break;
    default: a = 2; b = 3; break;
  }
  double tx = v[2 * b] - v[2 * a], ty = v[2 * b + 1] - v[2 * a + 1];
  len = std::sqrt(tx * tx + ty * ty);
  if (f == 1 || f == 2) { nx = ty / len; ny = -tx / len; }
  else { nx = -ty / len; ny = tx / len; }
}

inline double cell_diameter(const double* v) {
  double d1 = std::hypot(v[6] - v[0], v[7] - v[1]);
  double d2 = std::hypot(v[4] - v[2], v[5] - v[3]);
  return d1 > d2 ? d1 : d2;
}

}  // namespace pecs
