// VTU output of the state in the layout of the reference's print_results
// (SolarCell.cpp:1826-1858): one file per carrier pair and one for Poisson,
// rescaled as PostProcessor does (PostProcessor.cpp:9-27, 110-140).
#pragma once
#include <string>

#include "fe.hpp"
#include "mesh.hpp"

namespace pecs {

struct VtuScales {
  double current = 1.0;    // 1.6e-19 * char_density * char_length / char_time
  double density = 1.0;    // densities are written unscaled (PostProcessor.cpp:122-123)
  double field = 1.0;      // 0.2585 / char_length
  double potential = 1.0;  // 0.02585
};
VtuScales vtu_scales(double char_length, double char_density, double char_time);

// Writes <dir>/Semiconductor-NNN.vtu, <dir>/Electrolyte-NNN.vtu and
// <dir>/Poisson-NNN.vtu (NNN = step, three digits as Utilities::int_to_string
// (step, 3)).  u[0..3] = electrons, holes, reductants, oxidants ([Jx|Jy|u]
// per region), pv = Poisson [D|Phi] (PoissonLayout).  Each cell is one patch
// with its four vertices (DataOut::build_patches() without subdivision);
// values are the cell's own polynomial at the vertex.
void write_state_vtu(const SolarCellMesh& m, const PoissonLayout& L, const RefTables& T, const double* const u[4],
                     const double* pv, int step, const std::string& dir, const VtuScales& s);

}  // namespace pecs
