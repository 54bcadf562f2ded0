// Constant local operators of the LDG and mixed systems (host side, setup).
//
// LDG (LDG.cpp:40-678).  Per carrier the reference assembles
//     [ A   B ] [q]   [f_p]     A = mu^{-1} (p, q)          (block diagonal)
//     [ C   D ] [u] = [f_v]     B = -(div p, u) + face terms
//                               C = -(grad v, q) + face terms
//                               D = (theta/dt)(v, u) + penalty terms
// and solves it with UMFPACK (Carrier.cpp:26-40).  Written per cell with its
// own outward normal n_M the interior-face terms of LDG.cpp:461-621 are
//     p-rows: 1/2 p.n_M (u_s + u_o) + (beta.p)(u_s - u_o)
//     v-rows: v [1/2 n_M.(q_s + q_o) - beta.(q_s - q_o) + sigma_F (u_s - u_o)]
// (the lower-index rule of LDG.cpp:390-391 does not change the matrix), and
// integrating the cell term of C by parts shows C = -B^T exactly.  With A SPD
// and D symmetric positive definite the density Schur complement
//     S = D + B^T A^{-1} B
// is SPD; it is factored once by sparse Cholesky and each step solves
//     S u = f_v + B^T A^{-1} f_p,   q = A^{-1}(f_p - B u),
// which is the same solution as the reference's direct solve.
//
// Mixed Poisson (MixedFEM.cpp:29-164, Poisson.cpp:18-105) is hybridised: the
// RT space is broken per cell and normal continuity (and D.n = 0 on Neumann
// faces) is imposed by face multipliers lambda.  Eliminating each cell,
//     H lambda = sum_K E_K A_K^{-1} f_K,   H = sum_K E_K A_K^{-1} E_K^T  (SPD)
// and x_K = A_K^{-1}(f_K - E_K^T lambda) recovers the same (D, Phi) as the
// reference's continuous mixed solve.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "fe.hpp"
#include "mesh.hpp"

namespace pecs {

// Scaled parameters of one problem instance (Parameters.hpp:25-80 roles).
struct Params {
  int k = 1;
  double dt = 1e-3, theta = 1.0, penalty = 1.0;
  double mu[4] = {1, 1, 1, 1};  // electrons, holes, reductants, oxidants
  double eps_s = 1, eps_e = 1, lam2 = 1;
  double k_et = 1e-2, k_ht = 1e-2, v_rec_n = 0, v_rec_p = 0;
  double ne = 1, pe = 0, rinf = 1, oinf = 1;
  double phi_bi = 0, phi_app = 0, phi_sch = 0, schottky_location = 0;
  int schottky_location_set = 0;
  int gen_mode = 0;  // 0: reference alpha*G0*exp(alpha*(y-H)); 1: G0*exp(-alpha*x); 2: off
  double G0 = 1, alpha = 1;
  int srh = 0;
  double ni = 0.1, tau_n = 1, tau_p = 1;
};

// Dense per-cell LDG operators (row-major).
struct LdgCellOps {
  std::vector<double> Minv;   // nb x nb, inverse of the density mass matrix
  std::vector<double> Bself;  // 2nb x nb
  std::vector<double> Bf;     // 4 x (2nb x nb), coupling to the face neighbour
  std::vector<double> Dself;  // nb x nb
  std::vector<double> Df;     // 4 x (nb x nb)
  std::vector<double> Mk;     // nb x nb mass
};

// Coupling of cell M's rows to cell X's density across a hanging face
// (LDG.cpp:338-383, written per cell like the blocks above): B (2nb x nb,
// M's current rows) and D (nb x nb, M's density rows).  Each hanging pair
// gives two: coarse -> fine and fine -> coarse.
struct LdgHangBlock {
  int M = -1, X = -1;
  std::vector<double> B, D;
};

struct LdgRegionOps {
  int nb = 0, ncell = 0;
  std::vector<LdgCellOps> cell;  // per cell (Bf / Df zero on hanging faces)
  std::vector<LdgHangBlock> hang;
  // de-duplicated device table: per type Minv | Bself | Bf[4]
  std::vector<int> type;
  std::vector<double> table;
  int ntype = 0, type_stride = 0;
};

LdgRegionOps build_ldg_ops(const Region& R, const RefTables& T, const Params& P);

// Block-sparse SPD matrix on a structured cell grid (input to the Cholesky).
// sep_width = 0: every nested-dissection separator is made as wide as the
// couplings across it require (meshes embedded in a lattice with empty
// positions, whose coupling offsets differ from place to place).
struct BlockSPD {
  int NXc = 0, NYc = 0;
  std::vector<int> ndof;          // per grid cell gid = i + NXc*j
  std::vector<int> offi, offj;    // coupling offsets
  std::vector<std::vector<double>> blk;  // [gid*noff + o]: ndof[gid] x ndof[nbr]
  int sep_width = 1;
  int noff() const { return (int)offi.size(); }
};

BlockSPD schur_matrix(const Region& R, const LdgRegionOps& ops, double mu);

// Full constant LDG matrix [[A,B],[C,D]] in CSR (reference layout) and the mass
// matrix, rebuilt from the local operators (for parity checks and callers that
// want the constant matrices, SolarCell.cpp:822-932).
void ldg_csr(const Region& R, const LdgRegionOps& ops, double mu, double dt, std::vector<int64_t>& rowptr,
             std::vector<int>& col, std::vector<double>& val, bool mass_only);

struct PoissonOps {
  int nloc = 0, nrt = 0, nphi = 0, k = 0;
  std::vector<double> Ainv;   // per cell nloc x nloc
  // constant Dirichlet / Schottky data of the D rows, split by the bias that
  // multiplies it (BiasValues.cpp): per cell 3 x nrt, the unit-datum parts of
  // phi_bi (points on x = 0), phi_app (Dirichlet points on x = 0) and phi_sch
  // (Schottky points at the Schottky location); the right-hand side is
  // phi_bi U0 - phi_app U1 - phi_sch U2, so the operators hold no bias value
  std::vector<double> ubias;
  std::vector<double> A;      // per cell nloc x nloc (kept for CSR export)
  // multipliers
  std::vector<int> mult_face;  // per cell, 4 faces: multiplier index base (global, unpermuted) or -1
  int nmult = 0;
  std::vector<int> ndof_owner;      // per grid cell: number of owned multiplier DoFs
  std::vector<int> owner_base;      // per grid cell: first owned multiplier index
  // de-duplicated device table: per type Ainv | U0 | U1 | U2
  std::vector<int> type;
  std::vector<double> table;
  int ntype = 0, type_stride = 0;
  std::vector<char> constrained;  // per global Poisson DoF: Neumann D.n constraint
  // per cell: bit f set if the cell writes face f's D DoFs (faces 0 and 2,
  // and faces 1 and 3 without a conforming neighbour)
  std::vector<unsigned char> own;
  // hanging faces: the coarse side has no multipliers of its own; its D.n is
  // tied to the multipliers of the fine faces.  Per hanging pair: coarse
  // cell, coarse face, subface, first multiplier of the fine face.  The
  // coupling of multiplier (s, m) to the coarse face's local DoF j is
  // sign(face) * hE[(s K + m) K + j], hE = w_m / 2 * G_j((s + x_m) / 2) with
  // the K-point Gauss rule (x, w) and its Lagrange basis G.
  struct HangMult {
    int coarse, face, sub, mbase;
  };
  std::vector<HangMult> hang;
  std::vector<double> hE;
};

PoissonOps build_poisson_ops(const SolarCellMesh& m, const PoissonLayout& L, const RefTables& T, const Params& P);
BlockSPD multiplier_matrix(const SolarCellMesh& m, const PoissonOps& ops, const RefTables& T);
void mixed_csr(const SolarCellMesh& m, const PoissonLayout& L, const PoissonOps& ops, std::vector<int64_t>& rowptr,
               std::vector<int>& col, std::vector<double>& val);

// small dense helpers
bool dense_inverse(int n, const double* A, double* Ainv);

}  // namespace pecs
