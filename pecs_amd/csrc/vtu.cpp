// VTU output (ASCII XML UnstructuredGrid, one VTK_QUAD patch per cell).
#include "vtu.hpp"

#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <vector>

namespace pecs {

VtuScales vtu_scales(double char_length, double char_density, double char_time) {
  VtuScales s;
  s.potential = 0.02585;
  s.field = 0.2585 / char_length;
  s.current = 1.6e-19 * char_density * char_length / char_time;
  s.density = 1.0;
  return s;
}

namespace {

const double kCorner[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};  // patch points, x fastest

struct Field {
  std::string name;
  int ncomp;                  // 1 or 3 (vectors padded with z = 0)
  std::vector<double> vals;   // [npoints * ncomp]
};

std::string file_name(const std::string& dir, const std::string& stem, int step) {
  char num[32];
  std::snprintf(num, sizeof num, "%03d", step);
  std::string d = dir.empty() ? std::string(".") : dir;
  return d + "/" + stem + num + ".vtu";
}

void write_file(const std::string& path, const double* V, int ncell, const std::vector<Field>& fields) {
  std::ofstream o(path);
  if (!o) throw std::runtime_error("cannot write " + path);
  o.precision(17);
  const int np = 4 * ncell;
  o << "<?xml version=\"1.0\" ?>\n";
  o << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
  o << "<UnstructuredGrid>\n";
  o << "<Piece NumberOfPoints=\"" << np << "\" NumberOfCells=\"" << ncell << "\">\n";
  o << "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
  for (int c = 0; c < ncell; ++c)
    for (int v = 0; v < 4; ++v) o << V[8 * c + 2 * v] << ' ' << V[8 * c + 2 * v + 1] << " 0\n";
  o << "</DataArray>\n</Points>\n<Cells>\n";
  o << "<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
  for (int c = 0; c < ncell; ++c) o << 4 * c << ' ' << 4 * c + 1 << ' ' << 4 * c + 3 << ' ' << 4 * c + 2 << '\n';
  o << "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
  for (int c = 0; c < ncell; ++c) o << 4 * (c + 1) << '\n';
  o << "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
  for (int c = 0; c < ncell; ++c) o << "9\n";
  o << "</DataArray>\n</Cells>\n<PointData>\n";
  for (const Field& f : fields) {
    o << "<DataArray type=\"Float64\" Name=\"" << f.name << "\" NumberOfComponents=\"" << f.ncomp
      << "\" format=\"ascii\">\n";
    for (int p = 0; p < np; ++p) {
      for (int d = 0; d < f.ncomp; ++d) o << (d ? " " : "") << f.vals[(size_t)p * f.ncomp + d];
      o << '\n';
    }
    o << "</DataArray>\n";
  }
  o << "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
  if (!o) throw std::runtime_error("error writing " + path);
}

// The two carriers of one region: "<Name> Current" (vector) and
// "<Name> Density" per carrier (LDG.cpp:1197-1232 with PostProcessor names).
void carrier_pair_file(const Region& R, const RefTables& T, const double* u1, const double* u2, const char* n1,
                       const char* n2, const std::string& path, const VtuScales& s) {
  const int nc = R.ncell, nb = T.nb;
  const std::vector<double> nodes = equispaced(T.k);
  // nodal basis at the four corners
  std::vector<double> bc(4 * nb);
  for (int v = 0; v < 4; ++v)
    for (int b = 0; b <= T.k; ++b)
      for (int a = 0; a <= T.k; ++a) {
        double lx, dx, ly, dy;
        lagrange(nodes, a, kCorner[v][0], lx, dx);
        lagrange(nodes, b, kCorner[v][1], ly, dy);
        bc[v * nb + a + (T.k + 1) * b] = lx * ly;
      }
  std::vector<Field> fields;
  const double* us[2] = {u1, u2};
  const char* names[2] = {n1, n2};
  for (int w = 0; w < 2; ++w) {
    Field J{std::string(names[w]) + " Current", 3, std::vector<double>((size_t)12 * nc, 0.0)};
    Field D{std::string(names[w]) + " Density", 1, std::vector<double>((size_t)4 * nc, 0.0)};
    const double* jx = us[w];
    const double* jy = us[w] + (size_t)nc * nb;
    const double* du = us[w] + (size_t)2 * nc * nb;
    for (int c = 0; c < nc; ++c)
      for (int v = 0; v < 4; ++v) {
        double a = 0, b = 0, d = 0;
        for (int i = 0; i < nb; ++i) {
          const double phi = bc[v * nb + i];
          a += jx[(size_t)c * nb + i] * phi;
          b += jy[(size_t)c * nb + i] * phi;
          d += du[(size_t)c * nb + i] * phi;
        }
        const size_t p = 4 * (size_t)c + v;
        J.vals[3 * p] = s.current * a;
        J.vals[3 * p + 1] = s.current * b;
        D.vals[p] = s.density * d;
      }
    fields.push_back(std::move(J));
    fields.push_back(std::move(D));
  }
  write_file(path, R.V.data(), nc, fields);
}

// "Field" (D with the Piola map, D = J D_hat / det J) and "Potential"
// (MixedFEM.cpp:297-319 with PostProcessor.cpp:126-138).
void poisson_file(const SolarCellMesh& m, const PoissonLayout& L, const RefTables& T, const double* pv,
                  const std::string& path, const VtuScales& s) {
  const int nc = m.ncellP, nrt = T.nrt, nphi = T.nphi, deg = T.deg;
  const std::vector<double> pn = equispaced(deg);
  std::vector<double> rx(4 * nrt), ry(4 * nrt), ph(4 * nphi);
  for (int v = 0; v < 4; ++v) {
    for (int i = 0; i < nrt; ++i) {
      double dv;
      T.rt_eval(i, kCorner[v][0], kCorner[v][1], rx[v * nrt + i], ry[v * nrt + i], dv);
    }
    for (int b = 0; b <= deg; ++b)
      for (int a = 0; a <= deg; ++a) {
        double lx, dx, ly, dy;
        lagrange(pn, a, kCorner[v][0], lx, dx);
        lagrange(pn, b, kCorner[v][1], ly, dy);
        ph[v * nphi + a + (deg + 1) * b] = lx * ly;
      }
  }
  Field E{"Field", 3, std::vector<double>((size_t)12 * nc, 0.0)};
  Field P{"Potential", 1, std::vector<double>((size_t)4 * nc, 0.0)};
  for (int c = 0; c < nc; ++c) {
    const int* id = &L.idx[(size_t)c * L.nloc];
    const double* V = &m.pV[8 * (size_t)c];
    for (int v = 0; v < 4; ++v) {
      double hx = 0, hy = 0, phi = 0;
      for (int i = 0; i < nrt; ++i) {
        hx += pv[id[i]] * rx[v * nrt + i];
        hy += pv[id[i]] * ry[v * nrt + i];
      }
      for (int i = 0; i < nphi; ++i) phi += pv[id[nrt + i]] * ph[v * nphi + i];
      const Geo g = bilinear(V, kCorner[v][0], kCorner[v][1]);
      const size_t p = 4 * (size_t)c + v;
      E.vals[3 * p] = s.field * (g.J00 * hx + g.J01 * hy) / g.det;
      E.vals[3 * p + 1] = s.field * (g.J10 * hx + g.J11 * hy) / g.det;
      P.vals[p] = s.potential * phi;
    }
  }
  write_file(path, m.pV.data(), nc, {E, P});
}

}  // namespace

void write_state_vtu(const SolarCellMesh& m, const PoissonLayout& L, const RefTables& T, const double* const u[4],
                     const double* pv, int step, const std::string& dir, const VtuScales& s) {
  carrier_pair_file(m.semi, T, u[0], u[1], "Electrons", "Holes", file_name(dir, "Semiconductor-", step), s);
  carrier_pair_file(m.elec, T, u[2], u[3], "Reductants", "Oxidants", file_name(dir, "Electrolyte-", step), s);
  poisson_file(m, L, T, pv, file_name(dir, "Poisson-", step), s);
}

}  // namespace pecs
