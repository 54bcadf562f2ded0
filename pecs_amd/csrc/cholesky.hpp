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
  int64_t linv_size = 0, l21_size = 0;  // lengths of Linv / L21 when filled
  double flops = 0.0;
  double t_assemble = 0.0, t_factor = 0.0, t_inverse = 0.0;
  int64_t nnz_L() const { return (int64_t)Linv.size() + (int64_t)L21.size(); }
};

SupernodalFactor cholesky_factor(const BlockSPD& S, int leaf_cells);

// Index plan of the multifrontal factorisation for execution on the device:
// each node's dense nf x nf row-major front (nf = ns + ms; fronts of one level
// contiguous), the S blocks scattered into the fronts (mode 0: own-own block,
// lower triangle kept; mode 1: block of a panel row, stored transposed), and
// the extend-add rows (child c's update row r into its parent's front through
// the child's row map), in two passes per level (first, then second child).
struct FrontPlan {
  int nlevels = 0;
  std::vector<int64_t> front_off, level_front;  // per node; per level [begin, end)
  int64_t front_total = 0;
  std::vector<double> svals;                    // S blocks concatenated
  std::vector<int64_t> sblk_off;                // per (cell, offset): block offset or -1
  std::vector<int64_t> sc_src, sc_dst;          // scatter items
  std::vector<int> sc_nx, sc_ny, sc_nf, sc_mode, sc_rx, sc_ry, sc_level;
  std::vector<int> ea_child, ea_row, ea_parent, ea_level;  // extend-add rows; ea_level per (level, pass)
  std::vector<int> ea_map;                      // per child: its panel rows' positions in the parent front
  std::vector<int64_t> ea_map_off;
};
// The ordering, supernodes and row structure of cholesky_factor (no numeric
// values: Linv / L21 stay empty) and the device plan of its numeric phase.
SupernodalFactor cholesky_symbolic(const BlockSPD& S, int leaf_cells, FrontPlan& plan);
// Host execution of a plan (for checking the plan's indexing against
// cholesky_factor): fills Linv and L21 as cholesky_factor does.
void execute_plan_host(const FrontPlan& plan, const SupernodalFactor& F, std::vector<double>& Linv,
                       std::vector<double>& L21);

// The factor rearranged for level-scheduled dense matrix-vector sweeps (the
// layout the GPU solve kernels read).  Per supernode s:
//   fwd: packed lower triangle of Linv (row r: r+1 values at r(r+1)/2), then
//        W21 = L21 Linv (ms x ns row-major);
//   bwd: per own row j, [Linv^T(j, j:ns) | -W21^T(j, :)] (ns - j + ms values)
//        at offset j(ns+ms) - j(j-1)/2.
// Forward at a level: t_p = b_p - sum z[czidx[cptr[p]:cptr[p+1]]], then
// [y_own; z_s] = [Linv; W21] t_own.  Backward: x_own = bwd rows . [y; x_rows].
// The sweeps run over row tiles (node, first row, rows): W lanes per row,
// 256/W rows per pass, at most MAXR passes per tile, chosen per level and
// direction (a default rule here; the device library times the candidates).
inline int sweep_tile_rows(int w) { return w <= 16 ? 64 : (w == 32 ? 32 : 16); }

struct PackedFactor {
  int n = 0, nlevels = 0;
  int64_t zlen = 0;
  std::vector<int> cell_base, p0, ns, ms, rows;
  std::vector<int64_t> fwd_off, bwd_off, rows_off, z_off, cptr, czidx;
  std::vector<double> fwd, bwd;
  std::vector<int> ftiles, btiles;    // flattened (node, row0, nrows, 0) grouped by level
  std::vector<int> foff, boff;        // per level offsets (tiles)
  std::vector<int> lrows, goff;       // own rows of each level's nodes + offsets
  std::vector<int64_t> xidx;          // panel-row positions of each level's nodes
  std::vector<int> xoff;              // per level offsets into xidx
  std::vector<int> fw, bw;            // lanes per row of the sweeps per level
  std::vector<int> fm, bm;            // row passes per tile per level
  std::vector<int> frows, brows;      // tile height per level (<= (256 / W) * MAXR)
  std::vector<int> fks, bks;          // per level column-segment length of split tiles (0: no split)
  std::vector<int> fpf, bpf;          // per level: the scalar sweep's prefetching variant (0 / 1)
  std::vector<int> frtiles, brtiles;  // split levels: row tiles (node, row0, nrows, 64 * first tile + nseg)
  std::vector<int> froff, broff;      // per level offsets into frtiles / brtiles
  std::vector<int> lnodes, lnoff;     // nodes of each level + offsets
  int64_t nvals = 0;                  // entries of fwd (= of bwd)
  double tri_bytes = 0.0;             // bytes read by one forward + backward sweep
  // MFMA-blocked copies (see pack_blocked): per node, blocks of 16 rows; a
  // block holds its columns [clo, clo + 16 nsteps) in 16-column steps of four
  // 64-entry groups in MFMA A-operand lane order
  std::vector<double> fblk, bblk;
  std::vector<int64_t> fblk_off, bblk_off;  // per block: offset of its data
  std::vector<int> fblk0, bblk0;            // per node: index of its first block
};
PackedFactor pack_factor(const SupernodalFactor& F);
// pack_factor in two parts: every index array, offset and tile (no values;
// needs only the symbolic factor), then fwd / bwd from Linv and L21.
PackedFactor pack_layout(const SupernodalFactor& F);
void pack_values(const SupernodalFactor& F, PackedFactor& P);

// Build the MFMA-blocked copies from fwd / bwd and release the row-major
// arrays.  Block b of node s covers node rows [16b, 16b + 16) (forward: the
// ns + ms rows of [Linv; W21]; backward: the ns rows of [Linv^T | -W21^T]);
// its columns start at clo = 0 (forward) or 16b (backward, the first row's own
// column) and end at the block's longest row.  Entry (step st, group q, lane
// l = i + 16 k) is row 16b + i, column clo + 16 st + 4 k + q, zero outside the
// row.  A wave streaming one block reads 512 contiguous bytes per load.
void pack_blocked(PackedFactor& P);
// The block offsets of pack_blocked alone (fblk0 / fblk_off / bblk0 /
// bblk_off; the last offset is the total length).
void blocked_layout(PackedFactor& P);
inline int blk_clo(bool fwd, int b) { return fwd ? 0 : 16 * b; }
inline int blk_chi(bool fwd, int b, int ns, int ms) {
  if (!fwd) return ns + ms;
  const int last = std::min(16 * b + 15, ns + ms - 1);
  return last < ns ? last + 1 : ns;
}

// Row tiles of level l (forward: all ns + ms rows of each node; backward: the
// ns own rows) with `rows` rows per tile, appended as (node, row0, nrows, -1).
void level_tiles(const PackedFactor& P, int l, bool fwd, int rows, std::vector<int>& out);
// Split tiles: each row tile is cut into column segments of ks columns
// (multiple of 16), one tile (node, row0, nrows, segment) per non-empty
// segment, and one row tile (node, row0, nrows, 64 * first tile + nseg) per
// row tile for the deterministic reduction of the segments' partial sums.
// Returns false if a row tile would need more than 63 segments.
bool level_tiles_split(const PackedFactor& P, int l, bool fwd, int rows, int ks, std::vector<int>& tiles,
                       std::vector<int>& rtiles);
// Rebuild ftiles/btiles and their level offsets from fw/fm/frows, bw/bm/brows.
void rebuild_tiles(PackedFactor& P);

}  // namespace pecs
