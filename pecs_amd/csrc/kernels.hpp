// Device-side data descriptors and kernel launchers of the PECS IMEX path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pecs {

struct DevTabs {
  const double *qx, *qy, *qw, *fw, *phi, *fphi, *phiP;
  const double *psix, *psiy;
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
};

struct DevPoisson {
  int ncell, ncell_semi, nloc, nrt, nphi, NX, NY;
  const double* V;
  const int* idx;        // nloc global DoFs per cell
  const int* type;
  const double* table;   // per type: Ainv | fpsi
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
};

struct DevPhys {
  double dt, penalty, k_et, k_ht, v_rec_n, v_rec_p;
  double ne, pe, rinf, oinf, eps_s, eps_e, G0, alpha, H, ni, tau_n, tau_p;
  int gen_mode, srh;
};

struct DevFactor {
  int n, nnodes;
  const int* p0;
  const int* ns;
  const int* ms;
  const int64_t* linv_off;
  const int64_t* l21_off;
  const int64_t* rows_off;
  const double* Linv;   // row-major lower
  const double* LinvT;  // row-major of the transpose (upper)
  const double* L21;    // ms x ns
  const double* L21T;   // ns x ms
  const int* rows;
  const int64_t* cptr;
  const int* cnode;
  const int* crow;
  const int2* tasks;    // (node, local row), grouped by level
};

// launchers (return hipError_t as int)
int launch_rhs_carrier(int k, bool semi, const DevRegion& R, const DevRegion& O, const DevPoisson& P,
                       const DevTabs& T, const DevPhys& ph, const double* u1, const double* u2, const double* ou1,
                       const double* ou2, const double* pvec, double* b1, double* b2, hipStream_t s);
int launch_schur_prep(int k, const DevRegion& R, const double* b, double mu, double* W, hipStream_t s);
int launch_schur_rhs(int k, const DevRegion& R, const double* b1, const double* b2, const double* W1,
                     const double* W2, double* g1, double* g2, int nrhs, hipStream_t s);
int launch_schur_recover(int k, const DevRegion& R, const double* W1, const double* W2, const double* xs1,
                         const double* xs2, double mu1, double mu2, double* out1, double* out2, const double* b1,
                         const double* b2, int nrhs, hipStream_t s);
int launch_chol_solve(const DevFactor& F, const int* level_task_off_host, int nlevels, double* v0, double* v1,
                      double* tmp0, double* tmp1, int nrhs, hipStream_t s);
int launch_poisson_cell(int k, const DevPoisson& P, const DevTabs& T, const DevPhys& ph, const double* un,
                        const double* up, const double* ur, const double* uo, int ncs, int nce, double* y,
                        hipStream_t s);
int launch_poisson_cell_from_b(const DevPoisson& P, const double* b, double* y, hipStream_t s);
int launch_poisson_face(const DevPoisson& P, const double* y, double* r, hipStream_t s);
int launch_poisson_recover(int k, const DevPoisson& P, const double* lam, const double* y, double* out,
                           hipStream_t s);
int launch_poisson_rhs_assemble(int k, const DevPoisson& P, const DevTabs& T, const DevPhys& ph, const double* un,
                                const double* up, const double* ur, const double* uo, int ncs, int nce, double* b,
                                hipStream_t s);
int launch_diff_norms(const double* a, const double* b, int64_t n, double* acc, hipStream_t s);

}  // namespace pecs
