// Structured meshes, boundary ids and maps of the PECS half cell (host side).
//
// Grid_Maker (Grid.cpp:44-461) builds one coarse quad per region and refines it
// globally r times; deal.II then iterates the 4^r cells in Morton (Z) order
// (level index = 4*parent + child, child 0..3 = (0,0),(1,0),(0,1),(1,1)).
// Boundary ids are set by exact comparisons of face centres with 0, L, H and
// r1 (make_Dirichlet/Neumann/Schottky_boundaries).  setup_mappings
// (SolarCell.cpp:156-372) pairs the interface faces and maps region cells to
// Poisson cells by an O(N^2) centre search; on these logically Cartesian grids
// both maps are index maps, computed here directly.
// Boundary-layer local refinement (Grid.cpp:11-44, 72-111) with one level:
// each region is two blocks side by side (bulk at level r, boundary layer at
// level r+1), listed level by level as deal.II's active cells are, with
// hanging faces between the blocks (make_solarcell_mesh with g.local = 1).
#pragma once
#include <cstdint>
#include <vector>

namespace pecs {

enum BoundaryId { INTERFACE = 0, DIRICHLET = 1, NEUMANN = 2, SCHOTTKY = 3, INTERIOR = -1 };
enum MaterialId { SEMI_ID = 0, SEMI_BL_ID = 1, ELEC_ID = 2, ELEC_BL_ID = 3 };

struct GeometrySpec {
  double r1 = 0.5, r2 = 0.5, H = 1.0, L = 1.0;
  bool insulated = false, schottky = false;
  double bl = 0.0;  // boundary-layer width (Grid.cpp:19-40)
  int local = 0;    // local refinements of the boundary layer (0 or 1)
};

// A hanging face: subface `sub` (0: the half at the face's first vertex) of
// face `face` of the level-r cell `coarse` is face `fine_face` of `fine`.
struct HangFace {
  int coarse, face, sub, fine, fine_face;
};

struct Region {
  int r = 0, N = 0, ncell = 0, material = 0;
  // grid positions of the cells: (i, j) of Morton cell c on a uniform mesh;
  // with local refinement, lower-left points on one fine lattice (2 per
  // bulk cell width, 1 per layer cell); the lattice is NXg x NYg
  std::vector<int> ci, cj;
  int NXg = 0, NYg = 0;
  std::vector<int> cell_of;     // uniform meshes: Morton cell of (i, j): cell_of[i + N*j]
  std::vector<double> V;        // 8 doubles per cell: v0..v3 (x, y)
  std::vector<int> nbr;         // 4 per cell, -1 at the boundary and on hanging faces
  std::vector<int> bid;         // 4 per cell, INTERIOR for interior faces (hanging ones too)
  std::vector<int> mat;         // material id per cell
  std::vector<double> pdiam;    // diameter of the level-r cell containing the cell
  std::vector<HangFace> hang;   // hanging faces (empty on uniform meshes)
};

struct SolarCellMesh {
  GeometrySpec geo;
  int r = 0, N = 0;
  Region semi, elec;
  // merged Poisson mesh: semiconductor cells (Morton) then electrolyte cells
  int NX = 0, NY = 0, ncellP = 0;
  std::vector<double> pV;       // 8 per Poisson cell
  std::vector<int> pI, pJ;      // global (I, J), I in [0, 2N)
  std::vector<int> pmat;        // material id
  std::vector<int> pbid;        // 4 per Poisson cell
  std::vector<int> pcell_of;    // Poisson cell of (I, J): pcell_of[I + NX*J]
  int local = 0;
  std::vector<int> pnbr;        // 4 per Poisson cell, -1 at the boundary and on hanging faces
  std::vector<HangFace> phang;  // hanging faces of the Poisson mesh
  // interface pairs, ordered by semiconductor cell index (SolarCell.cpp:163-255)
  std::vector<int> sigma_semi, sigma_elec;
  std::vector<int> semi_sigma_nb, elec_sigma_nb;  // per cell: partner or -1
};

int morton_i(int c, int r);
int morton_j(int c, int r);
int morton_index(int i, int j);
void vertex_grid(const double coarse[8], int r, std::vector<double>& G);
int classify_face(double cx, double cy, const GeometrySpec& g, bool insulated, bool schottky);
Region make_region(const double coarse[8], int r, int material, const GeometrySpec& g,
                   bool insulated, bool schottky);
SolarCellMesh make_solarcell_mesh(int r, const GeometrySpec& g);

// Poisson [D | Phi] layout (own layout, component-wise like Poisson.cpp:18-69):
// vertical faces j*(NX+1)+i, horizontal faces j*NX+i (deg+1 DoFs each), then
// per cell 2*deg*(deg+1) interior RT DoFs, then (deg+1)^2 Phi DoFs per cell.
// With hanging faces (m.local) the face blocks are numbered on first
// encounter instead (cells in order, faces 0..3; a conforming interior face
// shares its neighbour's block, the coarse and each fine side of a hanging
// face keep their own), nV = that count and nH = 0.
struct PoissonLayout {
  int deg = 0, nf = 0, nV = 0, nH = 0, nint = 0, nD = 0, nphi = 0, n = 0, nloc = 0;
  std::vector<int> idx;  // [ncellP * nloc]
  std::vector<int> face_base;  // [ncellP * 4] first DoF of each face block
};
PoissonLayout poisson_layout(const SolarCellMesh& m, int deg);

}  // namespace pecs
