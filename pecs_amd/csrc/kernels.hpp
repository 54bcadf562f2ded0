// Device-side data descriptors and kernel launchers of the PECS IMEX path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pecs {

struct DevTabs {
  const double *qx, *qy, *qw, *fw, *phi, *fphi, *phiP;
  const double *psix, *psiy;
  const double *fs, *fpsix, *fpsiy;  // face points and RT values there (manufactured Dirichlet data)
};

struct DevRegion {
  int N, ncell;
  const double* V;       // 8 per cell
  const double* h;       // cell diameter
  const int* nbr;        // 4 per cell
  const int* bid;        // 4 per cell
  const int* type;       // LDG operator type per cell
  const double* table;   // per type: Minv | Bself | Bf[4]
  int stride;
  const int* sigma_nb;   // interface partner cell in the other region, or -1
  const int* perm_base;  // permuted index of the cell's first density DoF in the Schur system
  // Cells grouped so that each group of 64 (one wave) shares the operator
  // types the Schur kernels read, groups padded with -1 (null: cells in
  // order, per-lane types): by the cell's own type (recovery), and by the
  // types of the cell and its 4 face neighbours (right-hand side).
  const int* ucell_own;
  const int* ucell_nbh;
  int nu_own, nu_nbh;
  // mixed precision (null: FP64): FP32 copy of the table; with the
  // wave-uniform lists the Schur kernels then contract in FP32
  const float* table_f;
  // manufactured-solution data of the right-hand side (null: the solar-cell
  // data), for both carriers of the region: a source at the cell quadrature
  // points (ncell x nq), the Dirichlet datum at the face points replacing the
  // equilibrium constants, and a prescribed normal flux h on interface and
  // Neumann faces, entering as -<v, h> (ncell x 4 x nfq each)
  const double* msrc;
  const double* mdir;
  const double* mflux;
  // hanging faces of boundary-layer refinement (nhx = nhm = 0: none).  Their
  // couplings B_{M,X} (2nb x nb blocks in hblk: cell M's current rows, cell
  // X's density columns) are not in nbr / table; they are applied after the
  // kernels above by target cell X (right-hand side: g_X += B^T W_M; entries
  // hx_ptr[t] .. hx_ptr[t+1] of target hx_cell[t] name M in hx_src and the
  // block in hx_blk) and by row cell M (recovery: q_M -= mu Minv_M B u_X).
  int nhx, nhm;
  const int *hx_cell, *hx_ptr, *hx_src, *hx_blk;
  const int *hm_cell, *hm_ptr, *hm_dst, *hm_blk;
  const double* hblk;
};

struct DevPoisson {
  int ncell, ncell_semi, nloc, nrt, nphi, NX, NY;
  const double* V;
  const int* idx;        // nloc global DoFs per cell
  const int* type;
  const double* table;   // per type: Ainv | U0 | U1 | U2 (unit Dirichlet data of phi_bi, phi_app, phi_sch)
  int stride;
  const int* pbid;       // 4 per cell
  const int* pI;
  const int* pJ;
  const int* mult_face;  // 4 per cell: first multiplier index of the face or -1
  const int* mult_perm;  // multiplier index -> permuted index in the H system
  int nmult;
  const int* mcell;      // 2 per multiplier: cells sharing it (-1)
  const int* mloc;       // 2 per multiplier: local RT DoF in that cell
  const double* mcoef;   // 2 per multiplier: sign * Gauss weight
  const double* gw;      // face Gauss weights (k)
  const int* ucell;      // cells grouped by type in whole waves, padded with -1 (null: cells in order)
  int nucell;
  const float* table_f;  // mixed precision (null: FP64): FP32 copy of the table (Ainv used)
  // manufactured-solution data (null: the solar-cell data): a source added
  // to the charge at the cell quadrature points (ncell x nq), and a Dirichlet
  // datum at the face points of Dirichlet faces (ncell x 4 x nfq), entering
  // as -<psi.n, g> in place of the bias data
  const double* msrc;
  const double* mdir;
  // per cell: bit f set if the cell writes face f's D DoFs
  const unsigned char* own;
  // hanging faces: the coarse side's D.n couples to the fine faces'
  // multipliers.  Per hanging pair h (nhang): coarse cell, coarse face,
  // subface and first multiplier of the fine face; the coupling of
  // multiplier (s, m) to the coarse face's local DoF j is
  // sign(face) * hE[(s k + m) k + j].  hc_cell / hc_ptr group the pairs by
  // coarse cell (nhc cells).
  int nhang, nhc;
  const int *h_cell, *h_face, *h_sub, *h_mbase;
  const int *hc_cell, *hc_ptr;
  const double* hE;
};

struct DevPhys {
  double dt, penalty, k_et, k_ht, v_rec_n, v_rec_p;
  double ne, pe, rinf, oinf, eps_s, eps_e, G0, alpha, H, ni, tau_n, tau_p;
  double phi_bi, phi_app, phi_sch;
  int gen_mode, srh;
};

// Supernodal Cholesky factor S = L L^T stored for level-scheduled dense
// matrix-vector sweeps.  Per supernode s (own DoFs [p0, p0+ns), panel rows
// `rows` of length ms):
//   fwd: packed lower triangle of Linv = L11^{-1} (row r has r+1 entries),
//        then W21 = L21 Linv (ms x ns), so one GEMV [Linv; W21] t_own gives
//        y_own and the updates z = L21 y_own owed to the ancestors;
//   bwd: per own row j the concatenation [Linv^T(j, j:ns) | -W21^T(j, :)],
//        so x_own = Linv^T y_own - W21^T x_rows is one GEMV per row.
struct DevFactor {
  int n, nnodes;
  int64_t zlen;             // sum of ms (length of one right-hand side's z)
  const int* p0;
  const int* ns;
  const int* ms;
  const int64_t* fwd_off;
  const int64_t* bwd_off;
  const int64_t* rows_off;
  const int64_t* z_off;
  const double* fwd;
  const double* bwd;
  const int* rows;
  const int64_t* cptr;      // per permuted row p: range in czidx
  const int64_t* czidx;     // z indices subtracted from row p
  const int* lrows;         // own rows of each level's nodes, grouped by level
  const int4* ftiles;       // forward row tiles (node, first row, rows, 0), grouped by level
  const int4* btiles;       // backward row tiles, grouped by level
  const int64_t* xidx;      // panel-row positions (into rows) of each level's nodes, grouped by level
  // MFMA-blocked factor copies (batched contexts; fwd / bwd are then null):
  // per node blocks of 16 rows in MFMA A-operand lane order (pack_blocked)
  const double* fblk;
  const double* bblk;
  const int64_t* fblk_off;  // per block: data offset (one extra entry at the end)
  const int64_t* bblk_off;
  const int* fblk0;         // per node: first block index
  const int* bblk0;
  const int4* frtiles;      // split row tiles of the MFMA sweeps, grouped by level
  const int4* brtiles;
  double* Part;             // partial sums of split tiles (MFMA sweeps)
  double* Z;                // nr * zlen scratch: forward updates (two right-hand sides pair-interleaved,
                            // Z[2 z + h]; MFMA sweeps NRP-interleaved), then backward gathered x
  double* X;                // 2 * n scratch: forward t, then backward output
};

struct FactorLevels {
  const int* foff;  // host: per level offsets into ftiles (nlevels+1)
  const int* boff;  // host: per level offsets into btiles
  const int* goff;  // host: per level offsets into lrows
  const int* xoff;  // host: per level offsets into xidx
  const int* fw;    // host: lanes per row of the forward sweep per level
  const int* bw;    // host: lanes per row of the backward sweep per level
  const int* fm;    // host: row passes per tile (tile rows <= (256 / fw) * fm) per level
  const int* bm;    // host: the same for the backward sweep
  const int* fks;   // host: column-segment length of split forward tiles per level (0: no split)
  const int* bks;   // host: the same for the backward sweep
  const int* froff; // host: per level offsets into frtiles (nlevels+1)
  const int* broff; // host: per level offsets into brtiles
  int nlevels;
  const int* fpf;   // host: scalar sweep, prefetching variant per forward level (null: none)
  const int* bpf;   // host: the same for the backward sweep
};

// launchers (return hipError_t as int).  The cell-local launchers run nb
// instances of a batch at once (grid dimension y); each vector argument of
// instance i is at its pointer + i * (its stride), physics at phs[i].
int launch_rhs_carrier(int k, bool semi, const DevRegion& R, const DevRegion& O, const DevPoisson& P,
                       const DevTabs& T, const DevPhys* phs, const double* u1, const double* u2, const double* ou1,
                       const double* ou2, const double* pvec, double* b1, double* b2, hipStream_t s, int nb = 1,
                       int64_t sR = 0, int64_t sO = 0, int64_t sP = 0);
// Schur transform g = f_v + B^T (mu A^{-1} f_p) and recovery
// q = mu A^{-1} f_p - mu A^{-1} B u of one region's carriers (b: the LDG RHS
// [f_p | f_v] of each carrier; g / xs in the factor's permuted order).
int launch_schur_rhs(int k, const DevRegion& R, const double* b1, const double* b2, double mu1, double mu2,
                     double* g1, double* g2, int nrhs, hipStream_t s, int nb = 1, int64_t sb = 0, int64_t sg = 0);
int launch_schur_recover(int k, const DevRegion& R, const double* b1, const double* b2, const double* xs1,
                         const double* xs2, double mu1, double mu2, double* out1, double* out2, int nrhs,
                         hipStream_t s, int nb = 1, int64_t sb = 0, int64_t sx = 0, int64_t so = 0);
// One level's sweep over `nt` row tiles.  w > 0: W = w lanes per row and
// `maxr` row passes per tile (W in {8,16,32,64} with maxr in {1,2,4}, or W in
// {32,64} with maxr 8; every tile has at most (256 / W) * maxr rows), for 1 or
// 2 right-hand sides.  w == 0: the FP64-MFMA sweep with maxr (1, 2 or 4)
// 16-row blocks per wave (tiles of at most 64 * maxr rows), for up to 32
// right-hand sides; ks > 0 splits its tiles into column segments of ks
// columns whose partial sums (F.Part) are added by a reduction over rtiles.
// Vectors are h-major: right-hand side h at t[h * n].
// src (scalar forward sweeps only): read t from src instead of X (the leaf
// level, whose rows are owed no contributions).
int launch_sweep(const DevFactor& F, const int4* tiles, int nt, int w, int maxr, bool fwd, int nrhs,
                 double* t, hipStream_t s, int ks = 0, const int4* rtiles = nullptr, int nrt = 0,
                 const double* src = nullptr, bool pf = false);
// Padded right-hand-side count of the MFMA sweep (16 or 32).
// Right-hand sides of the MFMA sweeps: padded to 16, 32, 64 or 128
// interleaved columns; at most MAX_MFMA_RHS.  mfma_shape_ok: whether the
// sweep is built for nrp padded columns with mb 16-row blocks per wave.
constexpr int MAX_MFMA_RHS = 128;
int mfma_rhs_pad(int nrhs);
bool mfma_shape_ok(int nrp, int mb, bool split = false);
// rows a workgroup of the MFMA sweep covers: 64 mb, or with the right-hand
// sides split over waves (sweep width -1) 64 mb / (waves per row block)
int mfma_tile_rows(int nrp, int mb, bool split);
// x = S^{-1} v for nrhs h-major right-hand sides in the factor's permuted
// order (v is overwritten with x).  t is scratch: nrhs * n doubles for a
// factor shaped for 1-2 right-hand sides, mfma_rhs_pad(nrhs) * n for an
// MFMA-shaped one (whose scratch X, Z and t hold RHS-interleaved vectors).
int launch_chol_solve(const DevFactor& F, const FactorLevels& L, double* v, double* t, int nrhs, hipStream_t s);
int launch_poisson_cell(int k, const DevPoisson& P, const DevTabs& T, const DevPhys* phs, const double* un,
                        const double* up, const double* ur, const double* uo, int ncs, int nce, double* y,
                        hipStream_t s, int nb = 1, int64_t ss = 0, int64_t se = 0, int64_t sy = 0);
int launch_poisson_cell_from_b(const DevPoisson& P, const double* b, double* y, hipStream_t s);
int launch_poisson_face(const DevPoisson& P, const double* y, double* r, hipStream_t s, int nb = 1, int64_t sy = 0,
                        int64_t sr = 0);
int launch_poisson_recover(int k, const DevPoisson& P, const double* lam, const double* y, double* out,
                           hipStream_t s, int nb = 1, int64_t sl = 0, int64_t sy = 0, int64_t so = 0);
int launch_poisson_rhs_assemble(int k, const DevPoisson& P, const DevTabs& T, const DevPhys* phs, const double* un,
                                const double* up, const double* ur, const double* uo, int ncs, int nce, double* b,
                                hipStream_t s);
// Steady residual: per vector, RES_PARTS partial pairs (|a-b|^2, |b|^2) into
// part[0 .. 2 RES_PARTS); then per instance the sum of nvec vectors' partials
// (stored consecutively, instance after instance) into acc[2 i], acc[2 i + 1].
constexpr int RES_PARTS = 1024;
int launch_diff_norms(const double* a, const double* b, int64_t n, double* part, hipStream_t s);
int launch_sum_partials(const double* part, int nvec, double* acc, int ninst, hipStream_t s);
// dst[0:n] = src[0:n] as a kernel (stays on the stream's compute queue)
int launch_copy(const double* src, double* dst, int64_t n, hipStream_t s);
// Density of a carrier vector (region refinement r) at the G x G logical
// cell centres of its region, and the density block from such a grid
// (nodal values of the grid's bilinear interpolant).
int launch_grid_sample(int k, const double* u, int r, int G, double* out, hipStream_t s);
int launch_grid_project(int k, const double* grid, int G, int r, double* u, hipStream_t s);

}  // namespace pecs
