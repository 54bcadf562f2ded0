/* C ABI of the MI355X-native PECS per-time-step IMEX path.
 *
 * One opaque context owns the three meshes, the constant matrices (as factored
 * SPD Schur / multiplier systems), the device state and the scratch buffers of
 * one problem instance -- the roles of SOLARCELL::SolarCellProblem's members
 * (include/SolarCell.hpp:327-444).  All functions return 0 on success and a
 * negative code on error; pecs_last_error() gives the message (the reference
 * reports failures by exceptions caught in main, source/main/main.cpp:28-51).
 * One host thread per context; calls are ordered on the context's HIP stream.
 *
 * Vector layouts are the reference's component-wise layouts:
 *   carrier vector (electrons, holes on the semiconductor mesh; reductants,
 *   oxidants on the electrolyte mesh): [Jx | Jy | u], each block cell-major in
 *   deal.II cell order (Morton order of the 2^r x 2^r cells), (k+1)^2
 *   lexicographic DoFs per cell (CarrierPair.cpp:23-64);
 *   Poisson vector: [D | Phi] (Poisson.cpp:18-69); D uses a nodal RT basis
 *   (face DoFs: vertical faces j*(NX+1)+i then horizontal faces j*NX+i, k DoFs
 *   each; then 2(k-1)k interior DoFs per cell), Phi is FE_DGQ(k-1) per cell;
 *   on a locally refined mesh the face blocks are numbered on first encounter
 *   instead (cells in order, faces 0..3; a conforming face shares its
 *   neighbour's block, the coarse face and each fine face of a hanging face
 *   have their own, the fine ones equal to the coarse flux restricted);
 *   pecs_set_dof_layout switches the host-pointer calls to deal.II's basis
 *   and numbering of D.
 * Host-pointer functions copy in and out; *_dev functions take device
 * pointers and do not synchronise.
 */
#ifndef PECS_ABI_H
#define PECS_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pecs_ctx pecs_ctx;

/* Geometry of the half cell (Grid_Maker::Grid, source/Grid.cpp:11-44,
 * 135-336): semiconductor trapezoid (0,0),(r2,0),(0,H),(r1,H); electrolyte
 * (r2,0),(L,0),(r1,H),(L,H); both refined n_refine times.
 * local_refine = 1 ("local refinements" of the parameter file,
 * Grid.cpp:72-111): each region is split at boundary_layer from the r line
 * into a bulk quad and a boundary-layer quad, and the layer cells are refined
 * once more, leaving hanging faces between bulk and layer (n_refine 1..7).
 * Cell order is then deal.II's active-cell order per region: the level-r
 * bulk cells (Morton), then the level-(r+1) layer cells (Morton); the Poisson
 * vector uses the face-block layout described at the top of this file. */
typedef struct {
  double r1, r2, H, L;
  int n_refine;
  int insulated; /* make_Neumann_boundaries, Grid.cpp:376-430 */
  int schottky;  /* make_Schottky_boundaries, Grid.cpp:432-461 */
  double boundary_layer; /* "boundary layer" width (Grid.cpp:19-40) */
  int local_refine;      /* 0 or 1 */
} pecs_mesh;

/* Scaled parameters (ParameterSpace::Parameters, include/Parameters.hpp:25-80,
 * after parse_and_scale_parameters). */
typedef struct {
  int degree;            /* LDG degree k (main.cpp:11); Poisson RT_{k-1} x DGQ_{k-1} */
  double dt, theta, penalty;
  double mu[4];          /* electrons, holes, reductants, oxidants */
  double eps_s, eps_e, lambda2;
  double k_et, k_ht, v_rec_n, v_rec_p;
  double n_e, p_e, r_inf, o_inf;   /* equilibria / doping (InitialConditions.cpp) */
  double phi_bi, phi_app, phi_sch, schottky_location;
  int schottky_location_set;
  int gen_mode;          /* 0: alpha*G0*exp(alpha*(y-H)) (Generation.cpp:29-44); 1: G0*exp(-alpha*x); 2: off */
  double G0, alpha;
  int srh;               /* 0: R = 0 (SolarCell.hpp:86-98); 1: SRH formula */
  double n_i, tau_n, tau_p;
  int leaf_cells;        /* nested-dissection leaf size (0: default) */
} pecs_params;

/* Physical inputs of input_file.prm (ParameterReader.cpp:29-208). */
typedef struct {
  double char_length, char_time, char_density;
  double electron_mobility, hole_mobility, reductant_mobility, oxidant_mobility;
  double electron_recombo_t, hole_recombo_t, electron_recombo_v, hole_recombo_v;
  double k_et, k_ht, intrinsic_density, photon_flux, absorption_coeff;
  double applied_bias, built_in_bias, schottky_bias;
} pecs_physical;

typedef struct {
  double mobility[4], recombo_t[2], recombo_v[2], k_et, k_ht, intrinsic_density;
  double photon_flux, absorption_coeff, debye_length, applied_bias, built_in_bias, schottky_bias;
  double rescaled_k_et, rescaled_k_ht, rescale_current;
} pecs_scaled;

enum { PECS_ELECTRONS = 0, PECS_HOLES = 1, PECS_REDUCTANTS = 2, PECS_OXIDANTS = 3, PECS_POISSON = 4, PECS_MASS_SEMI = 5,
       PECS_MASS_ELEC = 6 };

/* Layout of the Poisson vector [D | Phi] in the host-pointer calls.
 * PECS_LAYOUT_OWN: this library's nodal RT basis for D (the layout above).
 * PECS_LAYOUT_DEALII: the reference's own -- FE_RaviartThomas(k-1) node
 * functionals (Legendre moments of the normal component on each face,
 * interior moments) with deal.II's DoF distribution and component-wise
 * renumbering (SolarCell.cpp:23-24, Poisson.cpp:18-30), so a deal.II caller
 * passes Poisson_object.solution and .system_rhs unchanged.  Phi (FE_DGQ(k-1))
 * is the same in both. */
enum { PECS_LAYOUT_OWN = 0, PECS_LAYOUT_DEALII = 1 };

/* Parameters::parse_and_scale_parameters, include/Parameters.hpp:117-287. */
int pecs_scale_parameters(const pecs_physical* in, pecs_scaled* out);

/* SolarCellProblem constructor + run_full_system setup (SolarCell.cpp:15-92,
 * 1898-1970): meshes, maps, constant matrices, factorisation, device upload. */
int pecs_create(const pecs_mesh* mesh, const pecs_params* params, pecs_ctx** out);
/* A batch of `batch` (1..64) instances on one mesh that share every constant
 * matrix and factor: an I-V bias sweep, illumination or doping series.  The
 * reference runs one SolarCellProblem per bias point (main.cpp:22-23); here the
 * instances step together and every triangular sweep reads each factor once
 * for all of them.  params[i] may differ only in right-hand-side data (biases,
 * generation, recombination, transfer rates, equilibria); the matrix
 * parameters (degree, dt, theta, penalty, mobilities, permittivities, lambda2,
 * leaf size, Schottky location) must agree.  No counterpart in the reference. */
int pecs_create_batch(const pecs_mesh* mesh, const pecs_params* params, int batch, pecs_ctx** out);
int pecs_batch_size(pecs_ctx* ctx, int* batch);
int pecs_destroy(pecs_ctx* ctx);
const char* pecs_last_error(pecs_ctx* ctx);
/* Error text of the last failed pecs_create / stateless call on this thread. */
const char* pecs_global_error(void);
int pecs_set_stream(pecs_ctx* ctx, void* hip_stream);
/* Poisson layout of the host-pointer calls from now on: the Poisson vectors
 * of pecs_set_state / pecs_get_state (and _batch), the Poisson input of
 * pecs_rhs_semiconductor / pecs_rhs_electrolyte, the output of
 * pecs_rhs_poisson and both vectors of pecs_solve(PECS_POISSON).  Solutions
 * convert with the change of basis per cell, right-hand sides with its
 * inverse transpose.  Device-pointer (_dev) calls, pecs_write_vtu and
 * pecs_export_matrix keep PECS_LAYOUT_OWN.  The role of
 * PoissonData::setup_dofs's numbering (Poisson.cpp:18-30). */
int pecs_set_dof_layout(pecs_ctx* ctx, int layout);
/* Stateless conversion of one Poisson vector between the two layouts
 * (to_dealii: PECS_LAYOUT_OWN -> PECS_LAYOUT_DEALII, else back; rhs: the
 * vector is a right-hand side rather than a solution).  n_poisson doubles in
 * and out (pecs_sizes).  Host only; errors: pecs_global_error(). */
int pecs_poisson_layout_convert(const pecs_mesh* mesh, const pecs_params* params, int to_dealii, int rhs,
                                const double* in, double* out);
int pecs_sizes(pecs_ctx* ctx, int64_t* n_semi, int64_t* n_elec, int64_t* n_poisson);
/* Setup statistics, up to 12 values: factor sizes in doubles (semiconductor,
 * electrolyte, multiplier), setup seconds, factorisation seconds (symbolic +
 * numeric + sweep layout + tuning), supernode levels (3), Schur system size,
 * multiplier system size, seconds and dense flops of the numeric phase on the
 * device (0 when PECS_HOST_FACTOR=1 factors on the host). */
int pecs_stats(pecs_ctx* ctx, double* out, int n);

/* Constant matrices as CSR (rows, nnz) in the reference layouts, rebuilt on the
 * host from the local operators (SolarCell.cpp:377-417, 822-932): call with
 * rowptr == NULL to get sizes.  which: PECS_ELECTRONS..PECS_OXIDANTS (system
 * matrices), PECS_POISSON (unconstrained mixed matrix), PECS_MASS_*.
 * `constrained` (may be NULL, PECS_POISSON only) flags Neumann D.n DoFs; on
 * a locally refined mesh the hanging-node constraints (each fine face's D
 * DoFs equal to the coarse face's flux restricted to it) are not flagged. */
int pecs_export_matrix(const pecs_mesh* mesh, const pecs_params* params, int which, int64_t* n, int64_t* nnz,
                       int64_t* rowptr, int32_t* col, double* val, int8_t* constrained);

/* VTU output of a state, as print_results (SolarCell.cpp:1826-1858) with the
 * rescaling of PostProcessor (PostProcessor.cpp:9-27, 110-140):
 * <dir>/Semiconductor-NNN.vtu ("Electrons Current", "Electrons Density",
 * "Holes ..."), <dir>/Electrolyte-NNN.vtu ("Reductants ...", "Oxidants ...")
 * and <dir>/Poisson-NNN.vtu ("Field", "Potential"), NNN = step in three
 * digits.  Vectors as pecs_get_state returns them; char_* are the
 * "characteristic length / density / time" of the parameter file.  Stateless:
 * runs on the host. */
int pecs_write_vtu(const pecs_mesh* mesh, const pecs_params* params, const double* u_n, const double* u_p,
                   const double* u_r, const double* u_o, const double* poisson, int step, const char* dir,
                   double char_length, double char_density, double char_time);

/* Per-step right-hand sides.
 * assemble_semiconductor_rhs (SolarCell.cpp:1037-1071, worker 1073-1414):
 *   rhs_n, rhs_p = M u_old + cell drift/generation + Dirichlet, interface and
 *   Schottky face terms, with the lagged electrolyte traces u_r, u_o. */
int pecs_rhs_semiconductor(pecs_ctx* ctx, const double* u_n, const double* u_p, const double* poisson,
                           const double* u_r, const double* u_o, double* rhs_n, double* rhs_p);
/* assemble_electrolyte_rhs (SolarCell.cpp:1417-1449, worker 1451-1726). */
int pecs_rhs_electrolyte(pecs_ctx* ctx, const double* u_r, const double* u_o, const double* poisson,
                         const double* u_n, const double* u_p, double* rhs_r, double* rhs_o);
/* assemble_Poisson_rhs (SolarCell.cpp:430-485, workers 487-815). */
int pecs_rhs_poisson(pecs_ctx* ctx, const double* u_n, const double* u_p, const double* u_r, const double* u_o,
                     double* rhs);
/* Carrier::solve / PoissonData::solve (Carrier.cpp:34-40, Poisson.cpp:98-105):
 * x = A^{-1} rhs for which = PECS_ELECTRONS..PECS_POISSON. */
int pecs_solve(pecs_ctx* ctx, int which, const double* rhs, double* x);

/* State (Carrier::solution, PoissonData::solution). poisson may be NULL in
 * set_state (then it is recomputed from the densities as in SolarCell.cpp:2033).
 * pecs_set_state / pecs_get_state address instance 0 of a batch; the _batch
 * variants take the instance index.  The per-call functions above
 * (pecs_rhs_*, pecs_solve) use instance 0's right-hand-side data. */
int pecs_set_state(pecs_ctx* ctx, const double* u_n, const double* u_p, const double* u_r, const double* u_o,
                   const double* poisson);
int pecs_get_state(pecs_ctx* ctx, double* u_n, double* u_p, double* u_r, double* u_o, double* poisson);
int pecs_set_state_batch(pecs_ctx* ctx, int instance, const double* u_n, const double* u_p, const double* u_r,
                         const double* u_o, const double* poisson);
int pecs_get_state_batch(pecs_ctx* ctx, int instance, double* u_n, double* u_p, double* u_r, double* u_o,
                         double* poisson);
/* Equilibrium initial data + first Poisson solve (SolarCell.cpp:1996-2034), every instance. */
int pecs_initial_state(pecs_ctx* ctx);
/* nsteps of the IMEX loop body (SolarCell.cpp:2057-2075) for every instance, device resident. */
int pecs_step(pecs_ctx* ctx, int nsteps);
/* Relative steady residual ||x^{n+1}-x^n|| / (||x^n|| dt) over the five state
 * vectors after the last pecs_step call (the reference keeps dead diff_* vectors,
 * SolarCell.cpp:2042-2046). */
int pecs_last_residual(pecs_ctx* ctx, double* residual);       /* max over the instances */
int pecs_last_residual_batch(pecs_ctx* ctx, double* residual);  /* one per instance */

/* Uneven-mesh sweeps: contexts with different meshes (refinement, geometry)
 * stepped together.  With nstreams 1..4 each time step launches every active
 * member's captured step graph, members spread round-robin over the streams
 * so that their steps overlap; with nstreams 0 each time step is one graph
 * launch, of a graph holding every active member's step graph as a child
 * graph (rebuilt when the active set or a member's graphs change).
 * pecs_sweep_step
 * updates every active member's steady residual (pecs_last_residual[_batch]);
 * set_active takes one flag per member (0: left out of later steps, e.g. once
 * steady).  The members must not be used elsewhere during pecs_sweep_step and
 * must outlive the sweep (destroy the sweep first).
 * Errors: pecs_global_error().  No reference counterpart (the reference runs
 * one problem per process). */
typedef struct pecs_sweep pecs_sweep;
int pecs_sweep_create(pecs_ctx** ctxs, int n, int nstreams, pecs_sweep** out);
int pecs_sweep_set_active(pecs_sweep* sweep, const int* active);
int pecs_sweep_step(pecs_sweep* sweep, int nsteps);
int pecs_sweep_destroy(pecs_sweep* sweep);

/* Learned warm start (no reference counterpart): a fully connected network,
 * nin inputs -> nlayers hidden layers of nhidden ReLU units -> nout linear
 * outputs, trained in FP32 on the GPU (MSE loss, Adam) and evaluated with
 * BF16 weights and activations and FP32 accumulation (bf16 = 1) or in FP32;
 * every matrix product runs on the matrix cores (FP32 / BF16 MFMA).
 * Parameters are one FP32 array: W_1 (nin x nhidden, row-major), b_1, W_2,
 * b_2, ...  x: batch x nin, y: batch x nout, host arrays, row-major.
 * pecs_mlp_gradients returns the loss and the gradient of every parameter
 * (same layout) without updating them.  Errors: pecs_mlp_last_error().
 *
 * pecs_mlp_create_decoder: the same encoder, whose last layer (ReLU) has
 * c0 g0^2 units read as a c0 x g0 x g0 feature map, then nconv decoder
 * stages: nearest x2 upsampling if conv_up[j] (else none) and a 3x3
 * convolution with zero padding to conv_cout[j] channels, ReLU after all but
 * the last.  Outputs (pecs_mlp_num_outputs): cout_last x G x G per sample,
 * G = g0 2^(number of upsamplings), channel-major.  Convolution parameters
 * follow the dense ones: W (cout x cin x 3 x 3), b (cout) per stage. */
typedef struct pecs_mlp pecs_mlp;
int pecs_mlp_create(int nin, int nhidden, int nlayers, int nout, unsigned seed, pecs_mlp** out);
int pecs_mlp_create_decoder(int nin, int nhidden, int nlayers, int c0, int g0, int nconv, const int* conv_cout,
                            const int* conv_up, unsigned seed, pecs_mlp** out);
int64_t pecs_mlp_num_outputs(pecs_mlp* net);
int pecs_mlp_destroy(pecs_mlp* net);
int64_t pecs_mlp_num_params(pecs_mlp* net);
int pecs_mlp_get_params(pecs_mlp* net, float* params);
int pecs_mlp_set_params(pecs_mlp* net, const float* params);
int pecs_mlp_gradients(pecs_mlp* net, const float* x, const float* y, int batch, float* grads, float* loss);
int pecs_mlp_train_step(pecs_mlp* net, const float* x, const float* y, int batch, float lr, float* loss);
int pecs_mlp_predict(pecs_mlp* net, const float* x, int batch, float* y, int bf16);
const char* pecs_mlp_last_error(void);

/* pecs_step replays time steps captured as HIP graphs at its first call: one
 * step, and the last step of a call together with the copy of the state
 * before it and the steady-residual kernels (kernels and stream dependencies
 * only).  enable = 0 launches the kernels one by one instead; the results are
 * identical.  No reference counterpart (launch mechanics of this library). */
int pecs_use_graph(pecs_ctx* ctx, int enable);
/* Per-phase timing of pecs_step with hipEvents on the context stream
 * (the TimerOutput sections of SolarCell.cpp:1905-2087): phases 0 rhs semi,
 * 1 rhs elec, 2 Schur transforms of both regions, 3 carrier Cholesky solves of
 * both regions (concurrent on two streams), 4 recoveries of both regions,
 * 5 Poisson condensation, 6 Poisson multiplier Cholesky, 7 Poisson recovery.
 * ms[] accumulates since enable. */
int pecs_enable_timing(pecs_ctx* ctx, int enable);
int pecs_phase_times(pecs_ctx* ctx, double* ms, int n, int* steps);
/* Algorithmic bytes of one forward + backward triangular sweep with a factor
 * (0 semiconductor Schur, 1 electrolyte Schur, 2 Poisson multipliers). */
int pecs_factor_bytes(pecs_ctx* ctx, int which, double* bytes);

/* Tile shapes chosen at setup for the triangular sweeps of a factor (0
 * semiconductor Schur, 1 electrolyte Schur, 2 Poisson multipliers): per level
 * l, out[6l..6l+5] = forward lanes per row, row passes, rows per tile, then the
 * same for the backward sweep.  No counterpart in the reference (UMFPACK's
 * solve has no tunable shape). */
int pecs_sweep_config(pecs_ctx* ctx, int which, int* out, int n, int* nlevels);

/* Which cell-local kernels take wave-uniform cell lists (cells grouped so
 * that every 64-cell wave shares the operator types it reads): bit 0/1
 * semiconductor Schur recovery / right-hand side, bit 2/3 the same for the
 * electrolyte, bit 4 the Poisson condensation and recovery.  No counterpart
 * in the reference (a kernel launch shape). */
int pecs_uniform_lists(pecs_ctx* ctx, int* flags);

/* Mixed precision (enable != 0): the cell-local contractions with the
 * cell-type tables (Schur transform and recovery, Poisson condensation and
 * recovery) run in FP32 with FP32 accumulation on FP32 copies of the tables;
 * the state, the right-hand-side assembly and the triangular solves stay
 * FP64.  Needs the wave-uniform cell lists (all five flags of
 * pecs_uniform_lists; else an error).  Takes effect from the next step.
 * Accuracy depends on the mesh and the horizon (the FP32 rounding of the
 * contractions is amplified by the cancellation in the recovered currents):
 * relative L2 differences from the FP64 path of about 5e-8 in the densities
 * after 20 steps at refinement 4, 2e-4 (carrier vectors) after 200 steps at
 * refinement 6 and 8e-4 after 200 steps at refinement 8; the potential stays
 * near 3e-7.  No counterpart in the reference (FP64 throughout); it replaces
 * nothing. */
int pecs_set_mixed_precision(pecs_ctx* ctx, int enable);

/* Densities on a regular grid (the warm-start network's output space; no
 * counterpart in the reference).  Each region is read as its logical
 * G x G grid of cell centres ((a + 1/2) / G, (b + 1/2) / G) of the
 * region's coarse quadrilateral, the same points for every refinement and
 * cone angle.  pecs_sample_grid writes the densities of instance i's four
 * carriers there: out (4 x G x G: electrons, holes on the semiconductor;
 * reductants, oxidants on the electrolyte; row b, column a).
 * pecs_set_state_from_grid sets instance i's densities from such a grid --
 * every nodal DoF takes the grid's bilinear interpolant at its logical
 * position (constant continuation in the half-cell band along the edges) --
 * zeroes the currents (the next step recomputes them) and solves the
 * potential of the new densities, as pecs_set_state without a Poisson
 * vector does. */
int pecs_sample_grid(pecs_ctx* ctx, int instance, int G, double* out);
/* Per instance, ||u^{n+1} - u^n|| / ||u^n|| over the densities of the four
 * carriers across the last step of the last pecs_step / pecs_sweep_step call
 * (the warm start's predicted residual after one step from a prediction).
 * out: one value per instance.  No counterpart in the reference. */
int pecs_last_density_change(pecs_ctx* ctx, double* out);
int pecs_set_state_from_grid(pecs_ctx* ctx, int instance, int G, const double* grid);

/* Manufactured-solution data for the right-hand sides (test hooks).  target
 * 0 / 1: the carrier right-hand sides of the semiconductor / electrolyte
 * (both carriers of the region take the same data); src (ncell x nq) is
 * added to the source at the cell quadrature points, dir (ncell x 4 x nfq)
 * replaces the equilibrium constant as the Dirichlet datum at the face
 * points of Dirichlet faces, flux (ncell x 4 x nfq) is a prescribed normal
 * flux h on interface and Neumann faces, entering as -<v, h> (on top of the
 * interface terms, which vanish with k_et = k_ht = 0).  target 2: the Poisson
 * right-hand side; src is added to the charge at the cell quadrature points,
 * dir is the Dirichlet datum g of -<psi.n, g> on Dirichlet faces (on top of
 * the bias data, which vanishes with phi_bi = phi_app = phi_sch = 0); flux
 * must be NULL.  A NULL array switches that term off.  Every instance of a
 * batch takes the same data.  The arrays are copied; new contents take effect
 * at the next right-hand side or step.  Cells are numbered as in the region
 * vectors (Poisson: semiconductor cells, then electrolyte cells), faces in
 * deal.II order, points as pecs_quadrature_points returns them.  Not
 * combinable with pecs_set_mixed_precision.  Replaces the reference's test
 * assemblies of its manufactured studies: LDG.cpp:681-982
 * (assemble_local_test_rhs, assemble_local_test_transient_rhs),
 * MixedFEM.cpp:166-254, SolarCell.cpp:2105-2745 (test_DD_Poisson and
 * interface assemblies), test_functions.cpp:13-527 (the data). */
int pecs_set_manufactured(pecs_ctx* ctx, int target, const double* src, const double* dir, const double* flux);

/* Physical coordinates of the points pecs_set_manufactured takes data at:
 * cell_xy (ncell x nq x 2), face_xy (ncell x 4 x nfq x 2); any pointer may be
 * NULL (sizes only).  No counterpart in the reference (deal.II's FEValues
 * quadrature_point()). */
int pecs_quadrature_points(pecs_ctx* ctx, int target, int64_t* ncell, int* nq, int* nfq, double* cell_xy,
                           double* face_xy);

/* Device-pointer variants (no copies, no synchronisation). */
int pecs_rhs_semiconductor_dev(pecs_ctx* ctx, const double* u_n, const double* u_p, const double* poisson,
                               const double* u_r, const double* u_o, double* rhs_n, double* rhs_p);
int pecs_rhs_electrolyte_dev(pecs_ctx* ctx, const double* u_r, const double* u_o, const double* poisson,
                             const double* u_n, const double* u_p, double* rhs_r, double* rhs_o);
int pecs_rhs_poisson_dev(pecs_ctx* ctx, const double* u_n, const double* u_p, const double* u_r,
                         const double* u_o, double* rhs);
int pecs_solve_dev(pecs_ctx* ctx, int which, const double* rhs, double* x);
/* pecs_solve_dev without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory at the point of the call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result as pecs_solve_dev (same kernels, same order), but
 * its scratch (the Schur or multiplier right-hand side, the triangular
 * sweeps' update vectors, the condensed Poisson cell data) lives in one
 * static device array of 64 MiB per GPU, held until the process ends, instead
 * of the context's buffers.  Single-instance contexts only.  A system whose
 * scratch exceeds 64 MiB is rejected with an error (refinement 8, k = 2
 * needs 32.7 MB for a carrier and 42.2 MB for the Poisson system).  The
 * array is found on the device that is current at the call; contexts created
 * on that device set it up.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, two host threads) race. */
int pecs_solve_dev_noalloc(pecs_ctx* ctx, int which, const double* rhs, double* x);
/* Device pointers of the resident state (valid until pecs_destroy); a batch
 * stores instance i at offset i * (n_semi, n_elec or n_poisson). */
int pecs_state_dev(pecs_ctx* ctx, double** u_n, double** u_p, double** u_r, double** u_o, double** poisson);

#ifdef __cplusplus
}
#endif

#endif /* PECS_ABI_H */
