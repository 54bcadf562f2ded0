// The Poisson vector [D | Phi] in deal.II's basis and numbering (host side).
//
// The reference's Poisson element is FESystem(FE_RaviartThomas(k-1),
// FE_DGQ(k-1)) (SolarCell.cpp:23-24), its DoFs distributed by
// DoFHandler::distribute_dofs and renumbered component-wise (Poisson.cpp:18-30),
// so a caller of the reference holds Poisson_object.solution and .system_rhs in
// that layout.  This library computes in its own nodal RT basis (fe.hpp); the
// functions here convert at the boundary.  Restated from deal.II's published
// definitions (deal.II 8.x):
//   * FE_RaviartThomas(deg) node functionals on the reference square.  Face f
//     (0: xi = 0, 1: xi = 1, 2: eta = 0, 3: eta = 1), DoF i = 0..deg:
//       l_{f,i}(v) = sum_q w_q P_i(s_q) v_d(face point s_q),
//     QGauss(deg+1) on the face, d the coordinate normal to the face (x for
//     faces 0, 1; y for 2, 3: the coordinate direction, not the outward
//     normal), P_i the Legendre polynomials on [0, 1] scaled orthonormal,
//     P_i(t) = sqrt(2i+1) L_i(2t-1).  Interior DoF 4(deg+1) + 2j + d:
//       l_{j,d}(v) = sum_q w_q Q_{d,j}(x_q) v_d(x_q),
//     QGauss(deg+1)^2 on the cell, Q_{d,j} the tensor products of Legendre
//     polynomials of degree <= deg-1 in direction d and <= deg in the other,
//     j = jx + nx jy (x fastest).  Shape functions are the dual basis of the
//     functionals, mapped by the contravariant Piola transform -- as this
//     library's basis is -- so the change of basis is one reference-cell
//     matrix N (deal.II coefficients = N x own coefficients), the same on
//     every cell.
//   * DoF numbering: active cells in order; on each cell the lines 0..3 not
//     numbered yet get deg+1 DoFs each, then the cell interior; component-wise
//     renumbering keeps that order within the RT block and puts Phi, per cell
//     (FE_DGQ(deg), lexicographic, the same basis as here), after it.
// Coefficient vectors (solutions) transform with N per cell; right-hand sides
// (b_i = a functional of basis function i) with N^{-T}, so that the pairing
// b . x is the same in both layouts.
#pragma once
#include <cstdint>
#include <vector>

#include "fe.hpp"
#include "mesh.hpp"

namespace pecs {

struct DealiiPoisson {
  int deg = 0, nf = 0, nint = 0, nrt = 0, nphi = 0, ncell = 0;
  int64_t n = 0, nD = 0;
  std::vector<double> N;     // nrt x nrt: rows deal.II local DoFs, columns this library's local DoFs
  std::vector<double> Ninv;  // its inverse (the same block structure, exact zeros kept)
  std::vector<int64_t> gd;   // per cell: deal.II global D index of each deal.II local DoF
  std::vector<int64_t> go;   // per cell: this library's global D index of each of its local DoFs
  std::vector<char> owns;    // per cell and face: the cell owns the face (left/bottom, or right/top on the boundary)
  std::vector<int> own_face; // per local DoF of this library: its face (0..3), -1 interior
  int64_t phi_off = 0;       // Phi starts at nD in both layouts
};

DealiiPoisson dealii_poisson(const SolarCellMesh& m, const PoissonLayout& L, const RefTables& T);
// coefficient vectors (the solution)
void poisson_to_dealii(const DealiiPoisson& X, const double* own, double* dii);
void poisson_from_dealii(const DealiiPoisson& X, const double* dii, double* own);
// right-hand sides
void poisson_rhs_to_dealii(const DealiiPoisson& X, const double* own, double* dii);
void poisson_rhs_from_dealii(const DealiiPoisson& X, const double* dii, double* own);

}  // namespace pecs
