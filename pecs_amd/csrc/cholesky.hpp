// Supernodal sparse Cholesky on a structured cell grid (host-side setup).
//
// Role in the reference: SparseDirectUMFPACK::initialize in Carrier::set_solver
// and PoissonData::set_solver (Carrier.cpp:26-32, Poisson.cpp:90-96), called
// once by SolarCellProblem::set_solvers (SolarCell.cpp:1733-1747).  Here the
// factored matrices are the SPD density Schur complements and the SPD Poisson
// multiplier matrix (ops.hpp), ordered by geometric nested dissection on the
// logically Cartesian cell grid (separators of `sep_width` cell columns/rows),
// factored by a multifrontal method.  The factor is stored per supernode as
// the explicit inverse of its diagonal block plus its off-diagonal panel, so
// every per-step triangular solve is a sequence of dense matrix-vector
// products scheduled level by level on the elimination tree.
#pragma once
#include <cstdint>
#include <vector>

#include "ops.hpp"

namespace pecs {

struct SNode {
  int p0 = 0, ns = 0, ms = 0, level = 0, parent = -1;
  int64_t linv_off = 0, l21_off = 0, rows_off = 0;
};

struct SupernodalFactor {
  int n = 0;
  std::vector<int> cell_base;      // per grid cell: permuted index of its first DoF
  std::vector<SNode> nodes;        // postorder
  std::vector<int> rows;           // per node: ms permuted row indices (struct)
  std::vector<double> Linv;        // per node: ns x ns row-major, lower triangular
  std::vector<double> L21;         // per node: ms x ns row-major
  int nlevels = 0;
  std::vector<std::vector<int>> level_nodes;
  // forward contributions: for each permuted row p the (node, panel row) pairs
  std::vector<int64_t> contrib_ptr;
  std::vector<int> contrib_node, contrib_row;
  double flops = 0.0;
  double t_assemble = 0.0, t_factor = 0.0, t_inverse = 0.0;
  int64_t nnz_L() const { return (int64_t)Linv.size() + (int64_t)L21.size(); }
};

SupernodalFactor cholesky_factor(const BlockSPD& S, int leaf_cells);

}  // namespace pecs
