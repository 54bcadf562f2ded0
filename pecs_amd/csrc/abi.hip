// C-ABI of the PECS IMEX path: context, setup, device upload, step loop.
// Declarations and reference citations: include/pecs_abi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pecs_abi.h"
#include "cholesky.hpp"
#include "dealii_layout.hpp"
#include "devfactor.hpp"
#include "fe.hpp"
#include "kernels.hpp"
#include "mesh.hpp"
#include "ops.hpp"
#include "vtu.hpp"

using namespace pecs;

// Device scratch of pecs_solve_dev_noalloc, the entry point that allocates
// nothing: one array per GPU (the code object's own global), held until the
// process ends.  Host code never uses the symbol itself (its host-side name is
// not a device address): set-up looks up each device's address once with
// hipGetSymbolAddress and the solves take that pointer.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

thread_local std::string g_error;

struct FactorDev {
  DevFactor d{};
  std::vector<int> foff, boff, goff, xoff, fw, bw, fm, bm, frows, brows, fks, bks, froff, broff, fpf,
      bpf;  // per level
  int nlevels = 0, n = 0;
  double nnz = 0.0;
  double tri_bytes = 0.0;      // algorithmic bytes of a forward + backward sweep
  double device_seconds = 0.0, device_flops = 0.0;  // numeric factorisation on the device (0: on the host)
  std::vector<int> cell_base;  // host copy: grid cell -> permuted base
  FactorLevels lv() const {
    return FactorLevels{foff.data(), boff.data(), goff.data(),  xoff.data(),  fw.data(), bw.data(),
                        fm.data(),   bm.data(),   fks.data(),   bks.data(),   froff.data(), broff.data(),
                        nlevels,     fpf.empty() ? nullptr : fpf.data(), bpf.empty() ? nullptr : bpf.data()};
  }
};

}  // namespace

struct pecs_ctx {
  std::string err;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;            // second stream for the concurrent region solves
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipEvent_t rhs_ev[2] = {};  // a region's right-hand side done (it has read the other region's old state)
  // one time step captured as a graph (pecs_step replays it); cap is the
  // stream it is captured on (capture needs a stream other than the null stream)
  hipStream_t cap = nullptr;
  hipGraph_t step_graph = nullptr, last_graph = nullptr;
  hipGraphExec_t step_exec = nullptr, last_exec = nullptr;
  bool use_graph = true;
  Params P;                  // parameters of instance 0 (matrix parameters are common to all)
  int B = 1;                 // instances sharing mesh, matrices and factors (a bias / illumination / doping batch)
  std::vector<Params> Pb;    // per instance
  std::vector<DevPhys> phs;  // per instance
  DevPhys* phs_d = nullptr;  // the same on the device
  int k = 1;
  SolarCellMesh mesh;
  PoissonLayout L;
  std::unique_ptr<RefTables> T;
  DevTabs tabs{};
  DevRegion Rs{}, Re{};
  DevPoisson Pd{};
  // mixed precision: FP32 copies of the semiconductor, electrolyte and
  // Poisson cell-type tables (kept on the host when the wave-uniform lists
  // are in use; uploaded by pecs_set_mixed_precision)
  std::vector<float> tabf_host[3];
  float* tabf_dev[3] = {};
  DevPhys ph{};  // instance 0
  FactorDev Fs[2], Fe[2], FH;
  bool share_s = false, share_e = false;
  int64_t Ns = 0, Ne = 0, NP = 0;
  // device vectors, instance-major (instance i at offset i * per-instance size)
  double* u[4] = {};
  double* pv = nullptr;
  double* b[4] = {};

  // per region: the Schur right-hand sides / solutions of all carriers and
  // instances, h-major with h = carrier * B + instance (region factor order)
  double* g[2] = {};
  double* t[2] = {};
  double* yP = nullptr;
  int64_t ysz = 0;
  double* rH = nullptr;  // B multiplier right-hand sides, h-major
  double* tH = nullptr;
  double* old[5] = {};
  double* acc = nullptr;  // 2 per instance
  double* rpart = nullptr;  // residual partial sums: 5 vectors x RES_PARTS pairs per instance
  double* io[4] = {};     // staging for host-pointer calls (one instance)
  std::vector<double> last_res;
  double setup_seconds = 0.0, factor_seconds = 0.0;
  std::vector<void*> allocs;
  void* blas = nullptr;  // dense-library handle of the device factorisations (created at the first)
  // layout of the Poisson vectors of the host-pointer calls (PECS_LAYOUT_*)
  int layout = 0;
  std::unique_ptr<DealiiPoisson> dii;
  std::vector<double> ptmp;  // host conversion buffer
  // manufactured-solution data per target (semiconductor, electrolyte,
  // Poisson): source, Dirichlet datum, flux (allocated at the first set)
  double* man_dev[3][3] = {};
  // grid values of pecs_sample_grid / pecs_set_state_from_grid (4 G^2)
  double* grid_dev = nullptr;
  int64_t grid_cap = 0;
  double* dacc = nullptr;  // pecs_last_density_change: 2 per instance
  uint64_t graph_gen = 0;  // step graphs captured so far (sweeps rebuild their combined graphs on a change)
  // optional per-phase timing of pecs_step (hipEvents recorded between phases)
  int timing = 0;
  std::vector<hipEvent_t> events;  // NPHASE+1 per recorded step
  int ev_used = 0;
  std::vector<double> phase_ms;
  int timed_steps = 0;

  ~pecs_ctx() {
    if (step_exec) (void)hipGraphExecDestroy(step_exec);
    if (step_graph) (void)hipGraphDestroy(step_graph);
    if (last_exec) (void)hipGraphExecDestroy(last_exec);
    if (last_graph) (void)hipGraphDestroy(last_graph);
    if (cap) {
      (void)hipStreamSynchronize(cap);
      (void)hipStreamDestroy(cap);
    }
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    for (hipEvent_t e : rhs_ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    factor_handle_destroy(blas);
  }
};

namespace {

#define HIPCHECK(x)                                                                               \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
T* dalloc(pecs_ctx* c, size_t n) {
  void* p = nullptr;
  HIPCHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
  c->allocs.push_back(p);
  return (T*)p;
}

// instances of one batched context: two carriers sharing a Schur factor
// then solve 128 right-hand sides together, the widest MFMA sweep
constexpr int MAX_BATCH = MAX_MFMA_RHS / 2;

// per-device address of `scratch`, looked up once per device (with that
// device current) when the first context is created on it
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

void setup_noalloc_scratch() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return;  // pecs_solve_dev_noalloc then reports it
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev]) return;
  void* p = nullptr;
  HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_dev[dev] = static_cast<double*>(p);
}

template <class T>
T* upload(pecs_ctx* c, const std::vector<T>& v) {
  T* p = dalloc<T>(c, v.size());
  if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

Params to_params(const pecs_params* q) {
  Params P;
  P.k = q->degree;
  P.dt = q->dt;
  P.theta = q->theta;
  P.penalty = q->penalty;
  for (int i = 0; i < 4; ++i) P.mu[i] = q->mu[i];
  P.eps_s = q->eps_s;
  P.eps_e = q->eps_e;
  P.lam2 = q->lambda2;
  P.k_et = q->k_et;
  P.k_ht = q->k_ht;
  P.v_rec_n = q->v_rec_n;
  P.v_rec_p = q->v_rec_p;
  P.ne = q->n_e;
  P.pe = q->p_e;
  P.rinf = q->r_inf;
  P.oinf = q->o_inf;
  P.phi_bi = q->phi_bi;
  P.phi_app = q->phi_app;
  P.phi_sch = q->phi_sch;
  P.schottky_location = q->schottky_location;
  P.schottky_location_set = q->schottky_location_set;
  P.gen_mode = q->gen_mode;
  P.G0 = q->G0;
  P.alpha = q->alpha;
  P.srh = q->srh;
  P.ni = q->n_i;
  P.tau_n = q->tau_n;
  P.tau_p = q->tau_p;
  return P;
}

GeometrySpec to_geo(const pecs_mesh* m) {
  GeometrySpec g;
  g.r1 = m->r1;
  g.r2 = m->r2;
  g.H = m->H;
  g.L = m->L;
  g.insulated = m->insulated != 0;
  g.schottky = m->schottky != 0;
  g.bl = m->boundary_layer;
  g.local = m->local_refine;
  return g;
}

void validate(const pecs_mesh* m, const pecs_params* q) {
  if (!m || !q) throw std::runtime_error("null mesh or params");
  if (q->degree < 1 || q->degree > KMAX) throw std::runtime_error("degree must be 1..3");
  if (m->n_refine < 1 || m->n_refine > 10) throw std::runtime_error("n_refine must be 1..10");
  if (m->local_refine < 0 || m->local_refine > 1)
    throw std::runtime_error("local_refine must be 0 or 1 (one level of boundary-layer refinement)");
  if (m->local_refine) {
    if (m->n_refine > 7) throw std::runtime_error("with local refinement n_refine must be 1..7");
    if (!(m->boundary_layer > 0 && m->boundary_layer < std::min(m->r1, m->r2) &&
          std::max(m->r1, m->r2) + m->boundary_layer < m->L))
      throw std::runtime_error("boundary_layer must be positive and fit inside both regions");
  }
  if (!(q->dt > 0)) throw std::runtime_error("dt must be positive");
  if (!(m->r1 > 0 && m->r2 > 0 && m->L > m->r1 && m->L > m->r2 && m->H > 0)) throw std::runtime_error("bad geometry");
  for (int i = 0; i < 4; ++i)
    if (!(q->mu[i] > 0)) throw std::runtime_error("mobilities must be positive");
  if (!(q->eps_s > 0 && q->eps_e > 0 && q->lambda2 > 0)) throw std::runtime_error("permittivities/lambda2 must be positive");
}

// Pick, per level and sweep direction, the tile shape (lanes per row W, row
// passes MAXR, rows per tile) with the shortest measured sweep on this GPU
// (W 8..64 with MAXR 1, 2, 4; W 32 and 64 also with MAXR 8).
// Each candidate is timed on the real factor level with the number of
// right-hand sides it will be used with; a 512 MiB fill between timed runs
// keeps the level's factor out of the caches, as it is in a solve.  Inputs
// are zero vectors; outputs go to the factor's own scratch (X, Z) and to a
// temporary Y.  PECS_NO_TUNE=1 keeps the default shapes of pack_factor.
void tune_factor(pecs_ctx* c, PackedFactor& P, DevFactor& d, int nrhs) {
  const char* env = std::getenv("PECS_NO_TUNE");
  if (env && env[0] == '1') {
    if (nrhs > 2) {  // the default shapes are for 1-2 right-hand sides
      for (int l = 0; l < P.nlevels; ++l) {
        P.fw[l] = P.bw[l] = 0;
        P.fm[l] = P.bm[l] = 1;
        P.frows[l] = P.brows[l] = 64;
      }
      rebuild_tiles(P);
    }
    return;
  }
  // PECS_SWEEP_FORCE=m,rows,ks gives every level of a batched (MFMA-shaped)
  // factor one shape of the MFMA sweep instead of the tuned ones -- m 16-row
  // blocks per wave, rows per tile, column segments of ks (0: unsplit) -- so
  // that tests can run the shapes the tuner picks only on large meshes.
  if (const char* fs = std::getenv("PECS_SWEEP_FORCE"); fs && nrhs > 2) {
    int m = 0, rows = 0, ks = -1, split = 0;
    const int nf = std::sscanf(fs, "%d,%d,%d,%d", &m, &rows, &ks, &split);
    const int nrp = mfma_rhs_pad(nrhs);
    const bool lds = split == 2;  // the LDS-staged vector operand: m 1 or 2, at most 64 m rows
    if ((nf != 3 && nf != 4) || (m != 1 && m != 2 && m != 4) || split < 0 || split > 2 || rows < 16 ||
        rows > (lds ? 64 * m : mfma_tile_rows(nrp, m, split)) || rows % 16 != 0 || ks < 0 ||
        (ks > 0 && (ks < 16 || ks % 16 != 0)) || !mfma_shape_ok(nrp, m, split == 1) || (lds && m == 4))
      throw std::runtime_error("PECS_SWEEP_FORCE: expected m,rows,ks[,split] (m 1, 2 or 4 -- 1 or 2 above 64 "
                               "right-hand sides unless split -- rows a multiple of 16 up to the rows a workgroup "
                               "covers, 64 m or with split 1 fewer, ks 0 or a multiple of 16; split 2: the "
                               "LDS-staged sweep, m 1 or 2, rows up to 64 m)");
    // (both split forms need more than 16 right-hand sides; a factor with 16
    // or fewer keeps the plain MFMA sweep)
    const int w = nrp <= 16 ? 0 : (lds ? -2 : (split ? -1 : 0));
    for (int l = 0; l < P.nlevels; ++l) {
      P.fw[l] = P.bw[l] = w;
      P.fm[l] = P.bm[l] = m;
      P.frows[l] = P.brows[l] = rows;
    }
    P.fks.assign(P.nlevels, ks);
    P.bks.assign(P.nlevels, ks);
    rebuild_tiles(P);
    return;
  }
  const size_t scrub_bytes = (size_t)512 << 20;
  void* scrub = nullptr;
  double* Y = nullptr;
  int4* tbuf = nullptr;
  int4* rtbuf = nullptr;
  double* part = nullptr;
  size_t part_cap = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  size_t tcap = 0;
  for (int l = 0; l < P.nlevels; ++l) {
    size_t rows = 0;
    for (int q = P.lnoff[l]; q < P.lnoff[l + 1]; ++q) rows += P.ns[P.lnodes[q]] + P.ms[P.lnodes[q]];
    tcap = std::max(tcap, rows);  // at most one tile per row
  }
  auto release = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (scrub) (void)hipFree(scrub);
    if (Y) (void)hipFree(Y);
    if (tbuf) (void)hipFree(tbuf);
    if (rtbuf) (void)hipFree(rtbuf);
    if (part) (void)hipFree(part);
    d.Part = nullptr;
  };
  try {
    HIPCHECK(hipMalloc(&scrub, scrub_bytes));
    const int nrs = nrhs > 2 ? mfma_rhs_pad(nrhs) : nrhs;  // scratch right-hand sides
    HIPCHECK(hipMalloc((void**)&Y, (size_t)nrs * P.n * sizeof(double)));
    tcap *= 4;  // split tiles: at most 63 segments per row tile of >= 16 rows
    HIPCHECK(hipMalloc((void**)&tbuf, std::max<size_t>(tcap, 1) * sizeof(int4)));
    HIPCHECK(hipMalloc((void**)&rtbuf, std::max<size_t>(tcap, 1) * sizeof(int4)));
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    hipStream_t s = c->stream;
    HIPCHECK(hipMemsetAsync(Y, 0, (size_t)nrs * P.n * sizeof(double), s));
    HIPCHECK(hipMemsetAsync(d.X, 0, (size_t)nrs * P.n * sizeof(double), s));
    HIPCHECK(hipMemsetAsync(d.Z, 0, (size_t)nrs * std::max<int64_t>(P.zlen, 1) * sizeof(double), s));
    std::vector<int> tl, rtl;
    std::vector<int4> t4, r4;
    const char* tr = std::getenv("PECS_SETUP_TRACE");
    const bool trace = tr && std::string(tr) == "1";
    P.fks.assign(P.nlevels, 0);
    P.bks.assign(P.nlevels, 0);
    for (int l = 0; l < P.nlevels; ++l) {
      for (int dir = 0; dir < 2; ++dir) {
        const bool fwd = dir == 0;
        float best = 1e30f;
        int bw = fwd ? P.fw[l] : P.bw[l], bm = fwd ? P.fm[l] : P.bm[l], br = fwd ? P.frows[l] : P.brows[l];
        int bk = 0, bpf = 0;
        if (nrhs > 2) {
          bw = 0;
          bm = 1;
          br = 64;
        }
        // candidates: (w, m, rows, ks); w = 0 is the MFMA sweep (m 16-row
        // blocks per wave, rows <= 64 m, column segments of ks), used for more
        // than 2 right-hand sides
        std::vector<std::array<int, 5>> cand;  // (w, m, rows, ks, prefetching scalar variant)
        if (nrhs > 2) {
          const int nrp = mfma_rhs_pad(nrhs);
          for (int split = 0; split < (nrp > 16 ? 2 : 1); ++split)
            for (int m : {1, 2, 4}) {
              const int tr = mfma_tile_rows(nrp, m, split);
              for (int rows : {tr, tr / 2})
                for (int ks : {0, 4096, 2048, 1024, 512, 256, 128})
                  if (rows >= 16 && mfma_shape_ok(nrp, m, split)) cand.push_back({split ? -1 : 0, m, rows, ks, 0});
            }
          if (nrp >= 32)  // the LDS-staged vector operand (w = -2): m blocks per wave, 64 m rows per tile
            for (int m : {1, 2})
              for (int ks : {0, 4096, 2048, 1024, 512, 256, 128}) cand.push_back({-2, m, 64 * m, ks, 0});
        } else {
          // the prefetching variant is a candidate on levels with few nodes,
          // where each lane's chain of load latencies sets the time
          const bool few = P.lnoff[l + 1] - P.lnoff[l] <= 64;
          for (int pf = 0; pf < (few ? 2 : 1); ++pf)
            for (int w : {8, 16, 32, 64})
              for (int m : {1, 2, 4, 8}) {
                if (m == 8 && w < 32) continue;
                for (int half = 0; half < 2; ++half) {
                  const int R = 256 / w;
                  const int rows = (R * m) >> half;
                  if (rows >= R) cand.push_back({w, m, rows, 0, pf});
                }
              }
        }
        for (const auto& cd : cand) {
          const int w = cd[0], m = cd[1], rows = cd[2], ks = cd[3], pf = cd[4];
          tl.clear();
          rtl.clear();
          if (!level_tiles_split(P, l, fwd, rows, ks, tl, rtl)) continue;
          const int nt = (int)(tl.size() / 4), nrt = (int)(rtl.size() / 4);
          if (nt == 0) continue;
          if (ks > 0 && nt == nrt) continue;  // no row tile is split: same as ks = 0
          if ((size_t)nt > tcap) continue;
          t4.resize(nt);
          for (int t = 0; t < nt; ++t) t4[t] = make_int4(tl[4 * t], tl[4 * t + 1], tl[4 * t + 2], tl[4 * t + 3]);
          HIPCHECK(hipMemcpyAsync(tbuf, t4.data(), nt * sizeof(int4), hipMemcpyHostToDevice, s));
          if (ks > 0) {
            const size_t need = (size_t)nt * 64 * m * nrs;
            if (need > part_cap) {
              HIPCHECK(hipStreamSynchronize(s));
              if (part) HIPCHECK(hipFree(part));
              part = nullptr;
              HIPCHECK(hipMalloc((void**)&part, need * sizeof(double)));
              part_cap = need;
            }
            d.Part = part;
            r4.resize(nrt);
            for (int t = 0; t < nrt; ++t) r4[t] = make_int4(rtl[4 * t], rtl[4 * t + 1], rtl[4 * t + 2], rtl[4 * t + 3]);
            HIPCHECK(hipMemcpyAsync(rtbuf, r4.data(), nrt * sizeof(int4), hipMemcpyHostToDevice, s));
          }
          float tmin = 1e30f;
          for (int rep = 0; rep < 3; ++rep) {
            HIPCHECK(hipMemsetAsync(scrub, rep, scrub_bytes, s));
            HIPCHECK(hipEventRecord(e0, s));
            HIPCHECK((hipError_t)launch_sweep(d, tbuf, nt, w, m, fwd, nrhs, Y, s, ks, rtbuf, nrt, nullptr, pf != 0));
            HIPCHECK(hipEventRecord(e1, s));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            tmin = std::min(tmin, ms);
          }
          if (tmin < best) {
            best = tmin;
            bw = w;
            bm = m;
            br = rows;
            bk = ks;
            bpf = pf;
          }
        }
        if (trace) {
          // the level's factor bytes (MFMA-blocked copy when batched) and nodes
          double bytes = 0.0;
          int maxns = 0;
          for (int q = P.lnoff[l]; q < P.lnoff[l + 1]; ++q) {
            const int sn = P.lnodes[q];
            maxns = std::max(maxns, P.ns[sn]);
            if (nrhs > 2) {
              const std::vector<int>& b0 = fwd ? P.fblk0 : P.bblk0;
              const std::vector<int64_t>& bo = fwd ? P.fblk_off : P.bblk_off;
              bytes += 8.0 * (double)(bo[b0[sn + 1]] - bo[b0[sn]]);
            } else {  // packed triangle and panel of one copy
              const double ns_ = P.ns[sn], ms_ = P.ms[sn];
              bytes += 8.0 * (0.5 * ns_ * (ns_ + 1.0) + ms_ * ns_);
            }
          }
          std::fprintf(stderr,
                       "[setup] tune n=%d level %d %s: W %d MAXR %d rows %d segment %d%s (%.3f ms) nodes %d max ns %d"
                       " factor %.1f MB (%.2f TB/s)\n",
                       P.n, l, fwd ? "fwd" : "bwd", bw, bm, br, bk, bpf ? " prefetch" : "", best,
                       P.lnoff[l + 1] - P.lnoff[l], maxns, bytes / 1e6, bytes / (best * 1e-3) / 1e12);
        }
        if (fwd) {
          P.fw[l] = bw;
          P.fm[l] = bm;
          P.frows[l] = br;
          P.fks[l] = bk;
          P.fpf[l] = bpf;
        } else {
          P.bw[l] = bw;
          P.bm[l] = bm;
          P.brows[l] = br;
          P.bks[l] = bk;
          P.bpf[l] = bpf;
        }
      }
    }
    HIPCHECK(hipStreamSynchronize(s));
  } catch (...) {
    release();
    throw;
  }
  release();
  rebuild_tiles(P);
}

// Factor S and upload the sweep layout.  The numeric factorisation runs on
// the device (factor_on_device) after a host symbolic phase;
// PECS_HOST_FACTOR=1 runs it on the host instead (cholesky_factor +
// pack_factor), for comparison.
FactorDev make_factor(pecs_ctx* c, const BlockSPD& S, int leaf, int nrhs) {
  const char* hf = std::getenv("PECS_HOST_FACTOR");
  const bool on_host = hf && std::string(hf) == "1";
  PackedFactor P;
  FactorDev D;
  double* dfwd = nullptr;  // device-factor output (sweep layout), owned here until handed over
  double* dbwd = nullptr;
  auto free_tmp = [&]() {
    if (dfwd) (void)hipFree(dfwd);
    if (dbwd) (void)hipFree(dbwd);
    dfwd = dbwd = nullptr;
  };
  const char* tr = std::getenv("PECS_SETUP_TRACE");
  const bool trace = tr && std::string(tr) == "1";
  auto tnow = [] { return std::chrono::steady_clock::now(); };
  auto secs = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  const auto t0 = tnow();
  if (on_host) {
    SupernodalFactor F = cholesky_factor(S, leaf);
    P = pack_factor(F);
  } else {
    FrontPlan plan;
    SupernodalFactor F = cholesky_symbolic(S, leaf, plan);
    P = pack_layout(F);
    if (trace) std::fprintf(stderr, "[setup] n=%d: symbolic + layout %.3f s\n", P.n, secs(t0, tnow()));
    HIPCHECK(hipMalloc(&dfwd, std::max<int64_t>(P.nvals, 1) * sizeof(double)));
    if (hipMalloc(&dbwd, std::max<int64_t>(P.nvals, 1) * sizeof(double)) != hipSuccess) {
      free_tmp();
      throw std::runtime_error("make_factor: out of device memory for the factor");
    }
    try {
      if (!c->blas) c->blas = factor_handle_create();
      const DeviceFactorStats st = factor_on_device(plan, F, P, dfwd, dbwd, c->stream, c->blas);
      D.device_seconds = st.seconds;
      D.device_flops = st.flops;
    } catch (...) {
      free_tmp();
      throw;
    }
  }
  D.n = P.n;
  D.nlevels = P.nlevels;
  D.cell_base = P.cell_base;
  D.goff = P.goff;
  D.xoff = P.xoff;
  D.nnz = (double)P.nvals;
  D.tri_bytes = P.tri_bytes;
  const int nn = (int)P.p0.size();
  DevFactor& d = D.d;
  d.n = P.n;
  d.nnodes = nn;
  d.zlen = P.zlen;
  d.p0 = upload(c, P.p0);
  d.ns = upload(c, P.ns);
  d.ms = upload(c, P.ms);
  d.fwd_off = upload(c, P.fwd_off);
  d.bwd_off = upload(c, P.bwd_off);
  d.rows_off = upload(c, P.rows_off);
  d.z_off = upload(c, P.z_off);
  if (nrhs > 2) {  // MFMA sweeps: blocked copies only
    if (on_host) {
      pack_blocked(P);
      d.fblk = upload(c, P.fblk);
      d.bblk = upload(c, P.bblk);
      std::vector<double>().swap(P.fblk);
      std::vector<double>().swap(P.bblk);
    } else {
      blocked_layout(P);
      try {
        double* fb = dalloc<double>(c, P.fblk_off.back());
        double* bb = dalloc<double>(c, P.bblk_off.back());
        blocked_on_device(P, dfwd, dbwd, fb, bb, c->stream);
        d.fblk = fb;
        d.bblk = bb;
      } catch (...) {
        free_tmp();
        throw;
      }
      free_tmp();
    }
    d.fwd = nullptr;
    d.bwd = nullptr;
    d.fblk_off = upload(c, P.fblk_off);
    d.bblk_off = upload(c, P.bblk_off);
    d.fblk0 = upload(c, P.fblk0);
    d.bblk0 = upload(c, P.bblk0);
  } else {
    if (on_host) {
      d.fwd = upload(c, P.fwd);
      d.bwd = upload(c, P.bwd);
    } else {  // hand the device arrays over to the context
      c->allocs.push_back(dfwd);
      c->allocs.push_back(dbwd);
      d.fwd = dfwd;
      d.bwd = dbwd;
      dfwd = dbwd = nullptr;
    }
    d.fblk = d.bblk = nullptr;
    d.fblk_off = d.bblk_off = nullptr;
    d.fblk0 = d.bblk0 = nullptr;
  }
  d.rows = upload(c, P.rows);
  d.cptr = upload(c, P.cptr);
  d.czidx = upload(c, P.czidx);
  d.lrows = upload(c, P.lrows);
  d.xidx = upload(c, P.xidx);
  // scratch for up to nrhs right-hand sides (at least 2: single-carrier
  // calls; the MFMA sweep pads to 16 or 32 interleaved right-hand sides)
  const int nr = nrhs > 2 ? mfma_rhs_pad(nrhs) : 2;
  d.Z = dalloc<double>(c, (size_t)nr * std::max<int64_t>(P.zlen, 1));
  d.X = dalloc<double>(c, (size_t)nr * P.n);
  HIPCHECK(hipMemset(d.Z, 0, (size_t)nr * std::max<int64_t>(P.zlen, 1) * sizeof(double)));
  HIPCHECK(hipMemset(d.X, 0, (size_t)nr * P.n * sizeof(double)));
  const auto tt = tnow();
  if (trace) std::fprintf(stderr, "[setup] n=%d: numeric + upload %.3f s\n", P.n, secs(t0, tt));
  tune_factor(c, P, d, nrhs);
  if (trace) std::fprintf(stderr, "[setup] n=%d: tuning %.3f s\n", P.n, secs(tt, tnow()));
  auto to4 = [](const std::vector<int>& v) {
    std::vector<int4> o(v.size() / 4);
    for (size_t t = 0; t < o.size(); ++t) o[t] = make_int4(v[4 * t], v[4 * t + 1], v[4 * t + 2], v[4 * t + 3]);
    return o;
  };
  d.ftiles = upload(c, to4(P.ftiles));
  d.btiles = upload(c, to4(P.btiles));
  d.frtiles = upload(c, to4(P.frtiles));
  d.brtiles = upload(c, to4(P.brtiles));
  // partial sums of the split levels: one slot of 64 MAXR rows per tile
  size_t part = 0;
  const int nrs = nrhs > 2 ? mfma_rhs_pad(nrhs) : nrhs;
  for (int l = 0; l < P.nlevels; ++l) {
    if (P.fks[l] > 0) part = std::max(part, (size_t)(P.foff[l + 1] - P.foff[l]) * 64 * P.fm[l] * nrs);
    if (P.bks[l] > 0) part = std::max(part, (size_t)(P.boff[l + 1] - P.boff[l]) * 64 * P.bm[l] * nrs);
  }
  d.Part = part > 0 ? dalloc<double>(c, part) : nullptr;
  D.foff = P.foff;
  D.boff = P.boff;
  D.fw = P.fw;
  D.bw = P.bw;
  D.fm = P.fm;
  D.bm = P.bm;
  D.fpf = P.fpf;
  D.bpf = P.bpf;
  D.frows = P.frows;
  D.brows = P.brows;
  D.fks = P.fks;
  D.bks = P.bks;
  D.froff = P.froff;
  D.broff = P.broff;
  return D;
}

// Cells 0..n-1 grouped by a key of nkey ints (ascending cell order within a
// group), each group padded with -1 to a whole number of 64-cell waves.
template <class KeyFn>
std::vector<int> wave_groups(int n, int nkey, KeyFn key) {
  std::vector<int> keys((size_t)n * nkey);
  for (int i = 0; i < n; ++i) key(i, &keys[(size_t)i * nkey]);
  const auto kb = [&](int i) { return keys.begin() + (size_t)i * nkey; };
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return std::lexicographical_compare(kb(a), kb(a) + nkey, kb(b), kb(b) + nkey);
  });
  std::vector<int> out;
  for (int i = 0; i < n;) {
    int j = i;
    while (j < n && std::equal(kb(order[i]), kb(order[i]) + nkey, kb(order[j]))) ++j;
    out.insert(out.end(), order.begin() + i, order.begin() + j);
    while (out.size() % 64) out.push_back(-1);
    i = j;
  }
  return out;
}

DevRegion make_region_dev(pecs_ctx* c, int reg, const Region& R, const LdgRegionOps& O,
                          const std::vector<int>& sigma, const FactorDev& F) {
  DevRegion d{};
  d.N = R.N;
  d.ncell = R.ncell;
  d.V = upload(c, R.V);
  std::vector<double> h(R.ncell);
  for (int i = 0; i < R.ncell; ++i) h[i] = cell_diameter(&R.V[8 * i]);
  d.h = upload(c, h);
  d.nbr = upload(c, R.nbr);
  d.bid = upload(c, R.bid);
  d.type = upload(c, O.type);
  d.table = upload(c, O.table);
  d.stride = O.type_stride;
  d.sigma_nb = upload(c, sigma);
  std::vector<int> pb(R.ncell);
  for (int i = 0; i < R.ncell; ++i) pb[i] = F.cell_base[R.ci[i] + (size_t)R.NXg * R.cj[i]];
  d.perm_base = upload(c, pb);
  if (!O.hang.empty()) {  // couplings across hanging faces, by target cell and by row cell
    std::vector<double> blk;
    std::vector<std::vector<std::pair<int, int>>> byx(R.ncell), bym(R.ncell);
    for (size_t h = 0; h < O.hang.size(); ++h) {
      const LdgHangBlock& hb = O.hang[h];
      blk.insert(blk.end(), hb.B.begin(), hb.B.end());
      byx[hb.X].push_back({hb.M, (int)h});
      bym[hb.M].push_back({hb.X, (int)h});
    }
    auto csr = [&](const std::vector<std::vector<std::pair<int, int>>>& L, const int*& cell, const int*& ptr,
                   const int*& other, const int*& bi) {
      std::vector<int> vc, vp{0}, vo, vb;
      for (int x = 0; x < R.ncell; ++x) {
        if (L[x].empty()) continue;
        vc.push_back(x);
        for (const auto& pr : L[x]) {
          vo.push_back(pr.first);
          vb.push_back(pr.second);
        }
        vp.push_back((int)vo.size());
      }
      cell = upload(c, vc);
      ptr = upload(c, vp);
      other = upload(c, vo);
      bi = upload(c, vb);
      return (int)vc.size();
    };
    d.nhx = csr(byx, d.hx_cell, d.hx_ptr, d.hx_src, d.hx_blk);
    d.nhm = csr(bym, d.hm_cell, d.hm_ptr, d.hm_dst, d.hm_blk);
    d.hblk = upload(c, blk);
  }
  // Wave-uniform cell lists for the Schur kernels, kept when the padding to
  // whole waves adds at most max(ncell, 2048) cells (meshes with few
  // operator types; on conic meshes every cell has its own type and the
  // lists are not used).
  // PECS_SCHUR_UNIFORM=0 turns them off.
  const char* su = std::getenv("PECS_SCHUR_UNIFORM");
  if (!(su && su[0] == '0')) {
    std::vector<int> own = wave_groups(R.ncell, 1, [&](int M, int* key) { key[0] = O.type[M]; });
    std::vector<int> nbh = wave_groups(R.ncell, 5, [&](int X, int* key) {
      key[0] = O.type[X];
      for (int f = 0; f < 4; ++f) {
        const int M = R.nbr[4 * X + f];
        key[1 + f] = M < 0 ? -1 : O.type[M];
      }
    });
    const size_t lim = std::max<size_t>(2 * (size_t)R.ncell, (size_t)R.ncell + 2048);
    if (own.size() <= lim) {
      d.ucell_own = upload(c, own);
      d.nu_own = (int)own.size();
    }
    if (nbh.size() <= lim) {
      d.ucell_nbh = upload(c, nbh);
      d.nu_nbh = (int)nbh.size();
    }
    if (d.ucell_own && d.ucell_nbh)
      c->tabf_host[reg].assign(O.table.begin(), O.table.end());
  }
  return d;
}

// single-carrier solve with its own factor (instance 0's scratch)
void solve_carrier(pecs_ctx* c, int which, const double* b, double* x) {
  const int reg = which / 2, car = which % 2;
  const DevRegion& R = reg == 0 ? c->Rs : c->Re;
  FactorDev* F = reg == 0 ? c->Fs : c->Fe;
  const bool share = reg == 0 ? c->share_s : c->share_e;
  const FactorDev& Fc = (car == 1 && !share) ? F[1] : F[0];
  const double mu = c->P.mu[which];
  hipStream_t s = c->stream;
  HIPCHECK((hipError_t)launch_schur_rhs(c->k, R, b, b, mu, mu, c->g[reg], c->g[reg], 1, s));
  HIPCHECK((hipError_t)launch_chol_solve(Fc.d, Fc.lv(), c->g[reg], c->t[reg], 1, s));
  HIPCHECK((hipError_t)launch_schur_recover(c->k, R, b, b, c->g[reg], c->g[reg], mu, mu, x, x, 1, s));
}

// Poisson solves of instances [i0, i0 + n) from their condensed cell data yP:
// face assembly per instance, one multiplier solve with n right-hand sides,
// recovery per instance into out + i * NP
void poisson_solve_from_y(pecs_ctx* c, int i0, int n, double* out) {
  hipStream_t s = c->stream;
  const int64_t nH = c->FH.n;
  HIPCHECK((hipError_t)launch_poisson_face(c->Pd, c->yP + i0 * c->ysz, c->rH, s, n, c->ysz, nH));
  HIPCHECK((hipError_t)launch_chol_solve(c->FH.d, c->FH.lv(), c->rH, c->tH, n, s));
  HIPCHECK((hipError_t)launch_poisson_recover(c->k, c->Pd, c->rH, c->yP + i0 * c->ysz, out, s, n, nH, c->ysz, c->NP));
}

// Poisson condensation of instance i from its densities
void poisson_cells(pecs_ctx* c, int i, const double* un, const double* up, const double* ur, const double* uo) {
  HIPCHECK((hipError_t)launch_poisson_cell(c->k, c->Pd, c->tabs, c->phs_d + i, un, up, ur, uo, c->mesh.semi.ncell,
                                           c->mesh.elec.ncell, c->yP + i * c->ysz, c->stream));
}

// the Poisson states of all instances from their current densities
void poisson_all_cells(pecs_ctx* c) {
  HIPCHECK((hipError_t)launch_poisson_cell(c->k, c->Pd, c->tabs, c->phs_d, c->u[0], c->u[1], c->u[2], c->u[3],
                                           c->mesh.semi.ncell, c->mesh.elec.ncell, c->yP, c->stream, c->B, c->Ns,
                                           c->Ne, c->ysz));
}

void poisson_all_from_densities(pecs_ctx* c) {
  poisson_all_cells(c);
  poisson_solve_from_y(c, 0, c->B, c->pv);
}

// Phases of one step (timing mode records an event on the context stream
// after each): 0 rhs semiconductor, 1 rhs electrolyte, 2 Schur transforms of
// both regions, 3 the carrier Cholesky solves of both regions (run
// concurrently on the context stream and a second stream, joined before the
// mark), 4 recoveries of both regions, 5 Poisson condensation, 6 Poisson
// multiplier Cholesky, 7 Poisson recovery.  Every phase covers all instances.
constexpr int NPHASE = 8;

struct PhaseMarks {
  pecs_ctx* c;
  hipEvent_t* ev;
  int i = 0;
  void mark() {
    if (ev) HIPCHECK(hipEventRecord(ev[i++], c->stream));
  }
};

// right-hand side slot of carrier car (0, 1) of instance i in a region's g / t
inline int64_t gslot(const pecs_ctx* c, int reg, int car, int i) {
  const int64_t n = reg == 0 ? c->Fs[0].n : c->Fe[0].n;
  return ((int64_t)car * c->B + i) * n;
}

void region_schur(pecs_ctx* c, int reg, hipStream_t s) {
  const DevRegion& R = reg == 0 ? c->Rs : c->Re;
  const int64_t N = reg == 0 ? c->Ns : c->Ne;
  const int64_t n = reg == 0 ? c->Fs[0].n : c->Fe[0].n;
  const int c0 = 2 * reg, c1 = 2 * reg + 1, B = c->B;
  // the instance stride of the g slots is n (slot = carrier * B + instance)
  HIPCHECK((hipError_t)launch_schur_rhs(c->k, R, c->b[c0], c->b[c1], c->P.mu[c0], c->P.mu[c1],
                                        c->g[reg] + gslot(c, reg, 0, 0), c->g[reg] + gslot(c, reg, 1, 0), 2, s, B, N,
                                        n));
}

void region_chol(pecs_ctx* c, int reg, hipStream_t s) {
  FactorDev* F = reg == 0 ? c->Fs : c->Fe;
  const bool share = reg == 0 ? c->share_s : c->share_e;
  if (share) {  // both carriers of all instances against one factor read
    HIPCHECK((hipError_t)launch_chol_solve(F[0].d, F[0].lv(), c->g[reg], c->t[reg], 2 * c->B, s));
  } else {  // one factor per carrier, one after the other on the same stream (t is reused)
    const int64_t h1 = gslot(c, reg, 1, 0);
    HIPCHECK((hipError_t)launch_chol_solve(F[0].d, F[0].lv(), c->g[reg], c->t[reg], c->B, s));
    HIPCHECK((hipError_t)launch_chol_solve(F[1].d, F[1].lv(), c->g[reg] + h1, c->t[reg], c->B, s));
  }
}

void region_recover(pecs_ctx* c, int reg, hipStream_t s) {
  const DevRegion& R = reg == 0 ? c->Rs : c->Re;
  const int64_t N = reg == 0 ? c->Ns : c->Ne;
  const int64_t n = reg == 0 ? c->Fs[0].n : c->Fe[0].n;
  const int c0 = 2 * reg, c1 = 2 * reg + 1;
  HIPCHECK((hipError_t)launch_schur_recover(c->k, R, c->b[c0], c->b[c1], c->g[reg] + gslot(c, reg, 0, 0),
                                            c->g[reg] + gslot(c, reg, 1, 0), c->P.mu[c0], c->P.mu[c1], c->u[c0],
                                            c->u[c1], 2, s, c->B, N, n, N));
}

// The Poisson phases of a step on the context stream (cell condensation,
// face assembly and multiplier solve, recovery), with phase marks if timed.
void poisson_step(pecs_ctx* c, PhaseMarks* pm = nullptr) {
  hipStream_t s = c->stream;
  const int B = c->B;
  poisson_all_cells(c);
  if (pm) pm->mark();
  const int64_t nH = c->FH.n;
  HIPCHECK((hipError_t)launch_poisson_face(c->Pd, c->yP, c->rH, s, B, c->ysz, nH));
  HIPCHECK((hipError_t)launch_chol_solve(c->FH.d, c->FH.lv(), c->rH, c->tH, B, s));
  if (pm) pm->mark();
  HIPCHECK((hipError_t)launch_poisson_recover(c->k, c->Pd, c->rH, c->yP, c->pv, s, B, nH, c->ysz, c->NP));
  if (pm) pm->mark();
}

void step_once(pecs_ctx* c) {
  hipStream_t s = c->stream;
  PhaseMarks pm{c, nullptr};
  if (c->timing && c->ev_used + NPHASE + 1 <= (int)c->events.size()) {
    pm.ev = &c->events[c->ev_used];
    c->ev_used += NPHASE + 1;
    pm.mark();
  }
  const int64_t Ns = c->Ns, Ne = c->Ne, NP = c->NP;
  const int B = c->B;
  if (!pm.ev) {
    // Without per-phase timing each region's chain (right-hand side, Schur
    // transform, solve, recovery) runs on its own stream, joined before the
    // Poisson phase.  A region's recovery writes its new state, which the
    // other region's right-hand side reads as the old one (the lagged
    // interface traces): it waits for that right-hand side.
    hipStream_t d = c->side;
    HIPCHECK(hipEventRecord(c->fork_ev, s));
    HIPCHECK(hipStreamWaitEvent(d, c->fork_ev, 0));
    HIPCHECK((hipError_t)launch_rhs_carrier(c->k, true, c->Rs, c->Re, c->Pd, c->tabs, c->phs_d, c->u[0], c->u[1],
                                            c->u[2], c->u[3], c->pv, c->b[0], c->b[1], s, B, Ns, Ne, NP));
    HIPCHECK(hipEventRecord(c->rhs_ev[0], s));
    HIPCHECK((hipError_t)launch_rhs_carrier(c->k, false, c->Re, c->Rs, c->Pd, c->tabs, c->phs_d, c->u[2], c->u[3],
                                            c->u[0], c->u[1], c->pv, c->b[2], c->b[3], d, B, Ne, Ns, NP));
    HIPCHECK(hipEventRecord(c->rhs_ev[1], d));
    region_schur(c, 0, s);
    region_schur(c, 1, d);
    region_chol(c, 0, s);
    region_chol(c, 1, d);
    HIPCHECK(hipStreamWaitEvent(s, c->rhs_ev[1], 0));
    region_recover(c, 0, s);
    HIPCHECK(hipStreamWaitEvent(d, c->rhs_ev[0], 0));
    region_recover(c, 1, d);
    HIPCHECK(hipEventRecord(c->join_ev, d));
    HIPCHECK(hipStreamWaitEvent(s, c->join_ev, 0));
    poisson_step(c);
    return;
  }
  HIPCHECK((hipError_t)launch_rhs_carrier(c->k, true, c->Rs, c->Re, c->Pd, c->tabs, c->phs_d, c->u[0], c->u[1],
                                          c->u[2], c->u[3], c->pv, c->b[0], c->b[1], s, B, Ns, Ne, NP));
  pm.mark();
  HIPCHECK((hipError_t)launch_rhs_carrier(c->k, false, c->Re, c->Rs, c->Pd, c->tabs, c->phs_d, c->u[2], c->u[3],
                                          c->u[0], c->u[1], c->pv, c->b[2], c->b[3], s, B, Ne, Ns, NP));
  pm.mark();
  region_schur(c, 0, s);
  region_schur(c, 1, s);
  pm.mark();
  // the two regions' solves share no data: the electrolyte runs on the side
  // stream so its launches fill the GPU where the semiconductor's leave gaps
  HIPCHECK(hipEventRecord(c->fork_ev, s));
  HIPCHECK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
  region_chol(c, 0, s);
  region_chol(c, 1, c->side);
  HIPCHECK(hipEventRecord(c->join_ev, c->side));
  HIPCHECK(hipStreamWaitEvent(s, c->join_ev, 0));
  pm.mark();
  region_recover(c, 0, s);
  region_recover(c, 1, s);
  pm.mark();
  poisson_step(c, &pm);
}

// The state before the last step of a pecs_step call (for the steady residual).
// Copy kernels, not memcpys, so that they become kernel nodes of the
// last-step graph.  (Work queued on a non-blocking stream right behind a
// graph launch was seen to overlap the graph's tail; inside the graph the
// copy, the step and the residual are ordered by the graph's edges.)
void save_old(pecs_ctx* c, hipStream_t s) {
  const int64_t nsv[5] = {c->Ns, c->Ns, c->Ne, c->Ne, c->NP};
  for (int j = 0; j < 4; ++j) HIPCHECK((hipError_t)launch_copy(c->u[j], c->old[j], c->B * nsv[j], s));
  HIPCHECK((hipError_t)launch_copy(c->pv, c->old[4], c->B * nsv[4], s));
}

// ||x^{n+1} - x^n||^2 and ||x^n||^2 over the five state vectors, per
// instance, into c->acc (2 per instance).
void launch_residual(pecs_ctx* c, hipStream_t s) {
  const int64_t nsv[5] = {c->Ns, c->Ns, c->Ne, c->Ne, c->NP};
  const int B = c->B;
  for (int b = 0; b < B; ++b) {
    for (int j = 0; j < 5; ++j) {
      const double* cur = j < 4 ? c->u[j] : c->pv;
      HIPCHECK((hipError_t)launch_diff_norms(cur + b * nsv[j], c->old[j] + b * nsv[j], nsv[j],
                                             c->rpart + (size_t)(5 * b + j) * 2 * RES_PARTS, s));
    }
  }
  HIPCHECK((hipError_t)launch_sum_partials(c->rpart, 5, c->acc, B, s));
}

// c->acc to the host (after the stream has drained), then the residuals.
void finish_residual(pecs_ctx* c) {
  std::vector<double> h(2 * c->B);
  HIPCHECK(hipMemcpy(h.data(), c->acc, 2 * c->B * sizeof(double), hipMemcpyDeviceToHost));
  for (int b = 0; b < c->B; ++b)
    c->last_res[b] = (h[2 * b + 1] > 0) ? std::sqrt(h[2 * b] / h[2 * b + 1]) / c->P.dt : 0.0;
}


// Capture into a graph (once per context), on c->cap with the context's own
// stream swapped for it; the side-stream fork/join of the region solves
// becomes part of the graph.  Two graphs: one time step, and the last step
// of a pecs_step call with the copy of the state before it and the steady-
// residual kernels after it, so that these are ordered by the graph's own
// edges.  Neither holds a memset or host work: kernels and event
// dependencies only.
void capture(pecs_ctx* c, bool last, hipGraph_t& graph, hipGraphExec_t& exec) {
  hipStream_t keep = c->stream;
  const int timing = c->timing;
  c->stream = c->cap;
  c->timing = 0;
  hipGraph_t g = nullptr;
  try {
    HIPCHECK(hipStreamBeginCapture(c->cap, hipStreamCaptureModeThreadLocal));
    try {
      if (last) save_old(c, c->cap);
      step_once(c);
      if (last) launch_residual(c, c->cap);
    } catch (...) {
      (void)hipStreamEndCapture(c->cap, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(c->cap, &g));
    hipGraphExec_t e = nullptr;
    HIPCHECK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    graph = g;
    exec = e;
    ++c->graph_gen;
  } catch (...) {
    c->stream = keep;
    c->timing = timing;
    throw;
  }
  c->stream = keep;
  c->timing = timing;
}

// Forget the captured step graphs (they are captured again at the next step),
// after a change of the kernels' arguments.  The stream must be idle.
void drop_graphs(pecs_ctx* c) {
  if (c->step_exec) HIPCHECK(hipGraphExecDestroy(c->step_exec));
  if (c->step_graph) HIPCHECK(hipGraphDestroy(c->step_graph));
  if (c->last_exec) HIPCHECK(hipGraphExecDestroy(c->last_exec));
  if (c->last_graph) HIPCHECK(hipGraphDestroy(c->last_graph));
  c->step_exec = c->last_exec = nullptr;
  c->step_graph = c->last_graph = nullptr;
}

void ensure_step_graph(pecs_ctx* c) {
  if (!c->step_exec) capture(c, false, c->step_graph, c->step_exec);
  if (!c->last_exec) capture(c, true, c->last_graph, c->last_exec);
}

// One time step, launched on stream s: the captured graph (the last-step
// variant when last), or the launches themselves when graphs are off or
// per-phase timing is on (the timing events are host-recorded).
void run_step(pecs_ctx* c, bool last) {
  if (c->use_graph && !c->timing) {
    ensure_step_graph(c);
    HIPCHECK(hipGraphLaunch(last ? c->last_exec : c->step_exec, c->stream));
  } else {
    if (last) save_old(c, c->stream);
    step_once(c);
    if (last) launch_residual(c, c->stream);
  }
}

void collect_timing(pecs_ctx* c) {
  if (!c->timing || c->ev_used == 0) return;
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int b = 0; b + NPHASE + 1 <= c->ev_used; b += NPHASE + 1) {
    for (int p = 0; p < NPHASE; ++p) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, c->events[b + p], c->events[b + p + 1]));
      c->phase_ms[p] += ms;
    }
    c->timed_steps++;
  }
  c->ev_used = 0;
}

int fail(pecs_ctx* c, const std::exception& e) {
  if (c) c->err = e.what();
  g_error = e.what();
  return -1;
}

}  // namespace

namespace {

DevPhys make_phys(const Params& P, double H) {
  DevPhys ph{};
  ph.dt = P.dt;
  ph.penalty = P.penalty;
  ph.k_et = P.k_et;
  ph.k_ht = P.k_ht;
  ph.v_rec_n = P.v_rec_n;
  ph.v_rec_p = P.v_rec_p;
  ph.ne = P.ne;
  ph.pe = P.pe;
  ph.rinf = P.rinf;
  ph.oinf = P.oinf;
  ph.eps_s = P.eps_s;
  ph.eps_e = P.eps_e;
  ph.G0 = P.G0;
  ph.alpha = P.alpha;
  ph.H = H;
  ph.ni = P.ni;
  ph.tau_n = P.tau_n;
  ph.tau_p = P.tau_p;
  ph.phi_bi = P.phi_bi;
  ph.phi_app = P.phi_app;
  ph.phi_sch = P.phi_sch;
  ph.gen_mode = P.gen_mode;
  ph.srh = P.srh;
  return ph;
}

// members of a batch share the mesh and every constant matrix, so the
// parameters that enter the matrices (or the Schottky data's location) must agree
void validate_batch(const pecs_params* q, int B) {
  for (int i = 1; i < B; ++i) {
    const pecs_params& a = q[0];
    const pecs_params& b = q[i];
    bool same = a.degree == b.degree && a.dt == b.dt && a.theta == b.theta && a.penalty == b.penalty &&
                a.eps_s == b.eps_s && a.eps_e == b.eps_e && a.lambda2 == b.lambda2 && a.leaf_cells == b.leaf_cells &&
                a.schottky_location_set == b.schottky_location_set &&
                (!a.schottky_location_set || a.schottky_location == b.schottky_location);
    for (int j = 0; j < 4; ++j) same = same && a.mu[j] == b.mu[j];
    if (!same)
      throw std::runtime_error("batch member " + std::to_string(i) +
                               " differs from member 0 in a parameter that enters the matrices (degree, dt, theta, "
                               "penalty, mobilities, permittivities, lambda2, leaf size, Schottky location)");
  }
}

int create_impl(const pecs_mesh* mesh, const pecs_params* params, int B, pecs_ctx** out) {
  if (!out) return -1;
  *out = nullptr;
  auto c = std::make_unique<pecs_ctx>();
  try {
    if (B < 1 || B > MAX_BATCH) throw std::runtime_error("batch size must be 1..64");
    if (!params) throw std::runtime_error("null params");
    for (int i = 0; i < B; ++i) validate(mesh, &params[i]);
    validate_batch(params, B);
    c->B = B;
    auto t0 = std::chrono::steady_clock::now();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      throw std::runtime_error("no HIP device: the PECS path runs on the GPU only");
    setup_noalloc_scratch();
    HIPCHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
    for (hipEvent_t& e : c->rhs_ev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->P = to_params(params);
    for (int i = 0; i < B; ++i) c->Pb.push_back(to_params(&params[i]));
    c->k = params->degree;
    const int leaf = params->leaf_cells > 0 ? params->leaf_cells : 16;
    c->mesh = make_solarcell_mesh(mesh->n_refine, to_geo(mesh));
    c->T = std::make_unique<RefTables>(c->k);
    const RefTables& T = *c->T;
    c->L = poisson_layout(c->mesh, c->k - 1);
    const int nb = T.nb;
    c->Ns = 3LL * nb * c->mesh.semi.ncell;
    c->Ne = 3LL * nb * c->mesh.elec.ncell;
    c->NP = c->L.n;
    LdgRegionOps os = build_ldg_ops(c->mesh.semi, T, c->P);
    LdgRegionOps oe = build_ldg_ops(c->mesh.elec, T, c->P);
    PoissonOps op = build_poisson_ops(c->mesh, c->L, T, c->P);
    auto tf = std::chrono::steady_clock::now();
    const char* tr = std::getenv("PECS_SETUP_TRACE");
    if (tr && std::string(tr) == "1")
      std::fprintf(stderr, "[setup] mesh + local operators %.3f s\n",
                   std::chrono::duration<double>(tf - t0).count());
    c->share_s = c->P.mu[0] == c->P.mu[1];
    c->share_e = c->P.mu[2] == c->P.mu[3];
    auto timed = [&](const char* what, auto&& build) {
      const auto ta = std::chrono::steady_clock::now();
      BlockSPD S = build();
      if (tr && std::string(tr) == "1")
        std::fprintf(stderr, "[setup] %s matrix %.3f s\n", what,
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - ta).count());
      return S;
    };
    auto schur = [&](const Region& R, const LdgRegionOps& o, double mu) {
      return timed("Schur", [&] { return schur_matrix(R, o, mu); });
    };
    c->Fs[0] = make_factor(c.get(), schur(c->mesh.semi, os, c->P.mu[0]), leaf, c->share_s ? 2 * B : B);
    if (!c->share_s) c->Fs[1] = make_factor(c.get(), schur(c->mesh.semi, os, c->P.mu[1]), leaf, B);
    c->Fe[0] = make_factor(c.get(), schur(c->mesh.elec, oe, c->P.mu[2]), leaf, c->share_e ? 2 * B : B);
    if (!c->share_e) c->Fe[1] = make_factor(c.get(), schur(c->mesh.elec, oe, c->P.mu[3]), leaf, B);
    c->FH = make_factor(c.get(), timed("multiplier", [&] { return multiplier_matrix(c->mesh, op, T); }), leaf, B);
    c->factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tf).count();
    // reference tables
    DevTabs& tb = c->tabs;
    tb.qx = upload(c.get(), T.qx);
    tb.qy = upload(c.get(), T.qy);
    tb.qw = upload(c.get(), T.qw);
    tb.fw = upload(c.get(), T.fw);
    tb.phi = upload(c.get(), T.phi);
    tb.fphi = upload(c.get(), T.fphi);
    tb.phiP = upload(c.get(), T.phiP);
    tb.psix = upload(c.get(), T.psix);
    tb.psiy = upload(c.get(), T.psiy);
    tb.fs = upload(c.get(), T.fs);
    tb.fpsix = upload(c.get(), T.fpsix);
    tb.fpsiy = upload(c.get(), T.fpsiy);
    // regions
    c->Rs = make_region_dev(c.get(), 0, c->mesh.semi, os, c->mesh.semi_sigma_nb, c->Fs[0]);
    c->Re = make_region_dev(c.get(), 1, c->mesh.elec, oe, c->mesh.elec_sigma_nb, c->Fe[0]);
    // Poisson
    DevPoisson& P = c->Pd;
    const SolarCellMesh& m = c->mesh;
    P.ncell = m.ncellP;
    P.ncell_semi = m.semi.ncell;
    P.nloc = op.nloc;
    P.nrt = op.nrt;
    P.nphi = op.nphi;
    P.NX = m.NX;
    P.NY = m.NY;
    P.V = upload(c.get(), m.pV);
    P.idx = upload(c.get(), c->L.idx);
    P.type = upload(c.get(), op.type);
    P.table = upload(c.get(), op.table);
    P.stride = op.type_stride;
    {  // wave-uniform cell list for the recovery (as for the Schur kernels)
      const char* su = std::getenv("PECS_SCHUR_UNIFORM");
      std::vector<int> own = wave_groups(m.ncellP, 1, [&](int K, int* key) { key[0] = op.type[K]; });
      if (!(su && su[0] == '0') && own.size() <= std::max<size_t>(2 * (size_t)m.ncellP, (size_t)m.ncellP + 2048)) {
        P.ucell = upload(c.get(), own);
        P.nucell = (int)own.size();
        c->tabf_host[2].assign(op.table.begin(), op.table.end());
      }
    }
    P.pbid = upload(c.get(), m.pbid);
    P.pI = upload(c.get(), m.pI);
    P.pJ = upload(c.get(), m.pJ);
    P.mult_face = upload(c.get(), op.mult_face);
    P.nmult = op.nmult;
    std::vector<int> owner_of(op.nmult), mperm(op.nmult), mcell(2 * (size_t)op.nmult, -1), mloc(2 * (size_t)op.nmult, 0);
    std::vector<double> mcoef(2 * (size_t)op.nmult, 0.0), gx, gw;
    for (int G = 0; G < m.NX * m.NY; ++G)
      for (int t = 0; t < op.ndof_owner[G]; ++t) {
        owner_of[op.owner_base[G] + t] = G;
        mperm[op.owner_base[G] + t] = c->FH.cell_base[G] + t;
      }
    gauss01(c->k, gx, gw);
    for (int K = 0; K < m.ncellP; ++K)
      for (int f = 0; f < 4; ++f) {
        const int base = op.mult_face[4 * K + f];
        if (base < 0) continue;
        const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
        for (int mm = 0; mm < c->k; ++mm) {
          const int id = base + mm;
          const int e = (mcell[2 * id] < 0) ? 0 : 1;
          mcell[2 * id + e] = K;
          mloc[2 * id + e] = T.rt_face_dof(f, mm);
          mcoef[2 * id + e] = sf * gw[mm];
        }
      }
    P.own = upload(c.get(), op.own);
    if (!op.hang.empty()) {  // hanging faces, grouped by coarse cell
      std::vector<PoissonOps::HangMult> hs = op.hang;
      std::stable_sort(hs.begin(), hs.end(), [](const PoissonOps::HangMult& a, const PoissonOps::HangMult& b) {
        return a.coarse < b.coarse;
      });
      std::vector<int> hc, hf, hsub, hmb, cells, ptr{0};
      for (size_t q = 0; q < hs.size(); ++q) {
        hc.push_back(hs[q].coarse);
        hf.push_back(hs[q].face);
        hsub.push_back(hs[q].sub);
        hmb.push_back(hs[q].mbase);
        if (q + 1 == hs.size() || hs[q + 1].coarse != hs[q].coarse) {
          cells.push_back(hs[q].coarse);
          ptr.push_back((int)q + 1);
        }
      }
      P.nhang = (int)hs.size();
      P.nhc = (int)cells.size();
      P.h_cell = upload(c.get(), hc);
      P.h_face = upload(c.get(), hf);
      P.h_sub = upload(c.get(), hsub);
      P.h_mbase = upload(c.get(), hmb);
      P.hc_cell = upload(c.get(), cells);
      P.hc_ptr = upload(c.get(), ptr);
      P.hE = upload(c.get(), op.hE);
    }
    P.mult_perm = upload(c.get(), mperm);
    P.mcell = upload(c.get(), mcell);
    P.mloc = upload(c.get(), mloc);
    P.mcoef = upload(c.get(), mcoef);
    P.gw = upload(c.get(), gw);
    // physics, per instance
    for (int i = 0; i < B; ++i) c->phs.push_back(make_phys(c->Pb[i], m.geo.H));
    c->ph = c->phs[0];
    c->phs_d = upload(c.get(), c->phs);
    // state and scratch (instance-major)
    const int64_t nsv[4] = {c->Ns, c->Ns, c->Ne, c->Ne};
    for (int i = 0; i < 4; ++i) {
      c->u[i] = dalloc<double>(c.get(), B * nsv[i]);
      c->b[i] = dalloc<double>(c.get(), B * nsv[i]);
      c->old[i] = dalloc<double>(c.get(), B * nsv[i]);
      c->io[i] = dalloc<double>(c.get(), std::max(nsv[i], c->NP));
      HIPCHECK(hipMemset(c->u[i], 0, B * nsv[i] * sizeof(double)));
    }
    c->pv = dalloc<double>(c.get(), B * c->NP);
    c->old[4] = dalloc<double>(c.get(), B * c->NP);
    HIPCHECK(hipMemset(c->pv, 0, B * c->NP * sizeof(double)));
    // scratch right-hand sides of a solve: 2B (or B) h-major, or the padded
    // interleaved block of the MFMA sweep
    auto scratch_rhs = [](int nrhs) { return nrhs > 2 ? mfma_rhs_pad(nrhs) : nrhs; };
    for (int r = 0; r < 2; ++r) {
      const int64_t n = r == 0 ? c->Fs[0].n : c->Fe[0].n;
      c->g[r] = dalloc<double>(c.get(), 2 * B * n);
      c->t[r] = dalloc<double>(c.get(), std::max(2 * B, scratch_rhs(2 * B)) * n);
      HIPCHECK(hipMemset(c->t[r], 0, std::max(2 * B, scratch_rhs(2 * B)) * n * sizeof(double)));
    }
    c->ysz = (int64_t)m.ncellP * op.nloc;
    c->yP = dalloc<double>(c.get(), B * c->ysz);
    c->rH = dalloc<double>(c.get(), B * (size_t)c->FH.n);
    c->tH = dalloc<double>(c.get(), std::max(B, scratch_rhs(B)) * (size_t)c->FH.n);
    c->acc = dalloc<double>(c.get(), 2 * B);
    c->rpart = dalloc<double>(c.get(), (size_t)10 * RES_PARTS * B);
    c->last_res.assign(B, -1.0);
    HIPCHECK(hipDeviceSynchronize());
    c->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
  *out = c.release();
  return 0;
}

}  // namespace

extern "C" {

int pecs_scale_parameters(const pecs_physical* in, pecs_scaled* out) {
  // Parameters.hpp:164-251 (constants Parameters.hpp:7-12)
  if (!in || !out) return -1;
  const double UT = 0.02585, q = 1.62e-19, eps0 = 8.85e-14;
  const double L = in->char_length, t = in->char_time, C = in->char_density;
  const double mob = (t * UT) / (L * L);
  out->mobility[0] = in->electron_mobility * mob;
  out->mobility[1] = in->hole_mobility * mob;
  out->mobility[2] = in->reductant_mobility * mob;
  out->mobility[3] = in->oxidant_mobility * mob;
  out->recombo_t[0] = in->electron_recombo_t / t;
  out->recombo_t[1] = in->hole_recombo_t / t;
  out->recombo_v[0] = in->electron_recombo_v * (t / L);
  out->recombo_v[1] = in->hole_recombo_v * (t / L);
  out->intrinsic_density = in->intrinsic_density / C;
  out->photon_flux = in->photon_flux * (t / C);
  out->absorption_coeff = in->absorption_coeff * L;
  out->debye_length = (UT * eps0) / (q * C * L * L);
  out->rescaled_k_et = q * in->k_et * C * C;
  out->rescaled_k_ht = q * in->k_ht * C * C;
  out->k_et = in->k_et * (t * C / L);
  out->k_ht = in->k_ht * (t * C / L);
  out->applied_bias = in->applied_bias / UT;
  out->built_in_bias = in->built_in_bias / UT;
  out->schottky_bias = in->schottky_bias / UT;
  out->rescale_current = (q * C * L) / t;
  return 0;
}

const char* pecs_global_error(void) { return g_error.c_str(); }


int pecs_create(const pecs_mesh* mesh, const pecs_params* params, pecs_ctx** out) {
  return create_impl(mesh, params, 1, out);
}

int pecs_create_batch(const pecs_mesh* mesh, const pecs_params* params, int batch, pecs_ctx** out) {
  return create_impl(mesh, params, batch, out);
}

int pecs_batch_size(pecs_ctx* ctx, int* batch) {
  if (!ctx || !batch) return -1;
  *batch = ctx->B;
  return 0;
}

int pecs_destroy(pecs_ctx* ctx) {
  if (!ctx) return -1;
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return 0;
}

const char* pecs_last_error(pecs_ctx* ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

int pecs_set_stream(pecs_ctx* ctx, void* s) {
  if (!ctx) return -1;
  ctx->stream = (hipStream_t)s;
  return 0;
}

int pecs_set_dof_layout(pecs_ctx* ctx, int layout) {
  if (!ctx) return -1;
  try {
    if (layout != PECS_LAYOUT_OWN && layout != PECS_LAYOUT_DEALII)
      throw std::runtime_error("unknown DoF layout (PECS_LAYOUT_OWN or PECS_LAYOUT_DEALII)");
    if (layout == PECS_LAYOUT_DEALII && ctx->mesh.local)
      throw std::runtime_error("the deal.II layout is not available on locally refined meshes");
    if (layout == PECS_LAYOUT_DEALII && !ctx->dii)
      ctx->dii = std::make_unique<DealiiPoisson>(dealii_poisson(ctx->mesh, ctx->L, *ctx->T));
    ctx->layout = layout;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return -1;
  }
  return 0;
}

int pecs_poisson_layout_convert(const pecs_mesh* mesh, const pecs_params* params, int to_dealii, int rhs,
                                const double* in, double* out) {
  try {
    if (!mesh || !params || !in || !out) throw std::runtime_error("null argument");
    validate(mesh, params);
    if (mesh->local_refine) throw std::runtime_error("the deal.II layout is not available on locally refined meshes");
    const SolarCellMesh m = make_solarcell_mesh(mesh->n_refine, to_geo(mesh));
    const RefTables T(params->degree);
    const PoissonLayout L = poisson_layout(m, params->degree - 1);
    const DealiiPoisson X = dealii_poisson(m, L, T);
    if (to_dealii) {
      if (rhs) poisson_rhs_to_dealii(X, in, out);
      else poisson_to_dealii(X, in, out);
    } else {
      if (rhs) poisson_rhs_from_dealii(X, in, out);
      else poisson_from_dealii(X, in, out);
    }
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
  return 0;
}

int pecs_sizes(pecs_ctx* ctx, int64_t* ns, int64_t* ne, int64_t* np) {
  if (!ctx) return -1;
  if (ns) *ns = ctx->Ns;
  if (ne) *ne = ctx->Ne;
  if (np) *np = ctx->NP;
  return 0;
}

int pecs_stats(pecs_ctx* ctx, double* o, int n) {
  if (!ctx || !o) return -1;
  const FactorDev* fs[5] = {&ctx->Fs[0], &ctx->Fs[1], &ctx->Fe[0], &ctx->Fe[1], &ctx->FH};
  double dev_s = 0.0, dev_flops = 0.0;
  for (const FactorDev* f : fs) {
    dev_s += f->device_seconds;
    dev_flops += f->device_flops;
  }
  const double v[12] = {ctx->Fs[0].nnz + ctx->Fs[1].nnz, ctx->Fe[0].nnz + ctx->Fe[1].nnz, ctx->FH.nnz,
                        ctx->setup_seconds, ctx->factor_seconds, (double)ctx->Fs[0].nlevels,
                        (double)ctx->Fe[0].nlevels, (double)ctx->FH.nlevels, (double)ctx->Fs[0].n,
                        (double)ctx->FH.n, dev_s, dev_flops};
  for (int i = 0; i < n && i < 12; ++i) o[i] = v[i];
  return 0;
}

int pecs_write_vtu(const pecs_mesh* mesh, const pecs_params* params, const double* u_n, const double* u_p,
                   const double* u_r, const double* u_o, const double* poisson, int step, const char* dir,
                   double char_length, double char_density, double char_time) {
  try {
    validate(mesh, params);
    if (!u_n || !u_p || !u_r || !u_o || !poisson) throw std::runtime_error("pecs_write_vtu: null state vector");
    if (!(char_length > 0 && char_density > 0 && char_time > 0))
      throw std::runtime_error("pecs_write_vtu: characteristic scales must be positive");
    SolarCellMesh m = make_solarcell_mesh(mesh->n_refine, to_geo(mesh));
    RefTables T(params->degree);
    PoissonLayout L = poisson_layout(m, params->degree - 1);
    const double* u[4] = {u_n, u_p, u_r, u_o};
    write_state_vtu(m, L, T, u, poisson, step, dir ? dir : ".", vtu_scales(char_length, char_density, char_time));
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
  return 0;
}

int pecs_export_matrix(const pecs_mesh* mesh, const pecs_params* params, int which, int64_t* n, int64_t* nnz,
                       int64_t* rowptr, int32_t* col, double* val, int8_t* constrained) {
  try {
    validate(mesh, params);
    Params P = to_params(params);
    SolarCellMesh m = make_solarcell_mesh(mesh->n_refine, to_geo(mesh));
    RefTables T(params->degree);
    std::vector<int64_t> rp;
    std::vector<int> cl;
    std::vector<double> vl;
    std::vector<char> cons;
    if (which >= PECS_ELECTRONS && which <= PECS_OXIDANTS) {
      const Region& R = which < 2 ? m.semi : m.elec;
      LdgRegionOps O = build_ldg_ops(R, T, P);
      ldg_csr(R, O, P.mu[which], P.dt, rp, cl, vl, false);
    } else if (which == PECS_MASS_SEMI || which == PECS_MASS_ELEC) {
      const Region& R = which == PECS_MASS_SEMI ? m.semi : m.elec;
      LdgRegionOps O = build_ldg_ops(R, T, P);
      ldg_csr(R, O, 1.0, P.dt, rp, cl, vl, true);
    } else if (which == PECS_POISSON) {
      PoissonLayout L = poisson_layout(m, params->degree - 1);
      PoissonOps O = build_poisson_ops(m, L, T, P);
      mixed_csr(m, L, O, rp, cl, vl);
      cons = O.constrained;
    } else {
      throw std::runtime_error("unknown matrix id");
    }
    if (n) *n = (int64_t)rp.size() - 1;
    if (nnz) *nnz = (int64_t)cl.size();
    if (rowptr) {
      std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int64_t));
      std::memcpy(col, cl.data(), cl.size() * sizeof(int32_t));
      std::memcpy(val, vl.data(), vl.size() * sizeof(double));
      if (constrained && !cons.empty())
        for (size_t i = 0; i < cons.size(); ++i) constrained[i] = (int8_t)cons[i];
    }
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
  return 0;
}

#define PECS_TRY(ctx, ...)                      \
  if (!ctx) return -1;                          \
  try {                                         \
    __VA_ARGS__                                 \
  } catch (const std::exception& e) {           \
    return fail(ctx, e);                        \
  }                                             \
  return 0;

// A host Poisson vector handed in by the caller, in this library's layout
// (converted from deal.II's when the context uses it).  rhs: a right-hand
// side (dual vector) rather than a coefficient vector.
static const double* poisson_in(pecs_ctx* c, const double* v, bool rhs) {
  if (c->layout != PECS_LAYOUT_DEALII || !v) return v;
  c->ptmp.resize(c->NP);
  if (rhs) poisson_rhs_from_dealii(*c->dii, v, c->ptmp.data());
  else poisson_from_dealii(*c->dii, v, c->ptmp.data());
  return c->ptmp.data();
}

// A host Poisson vector about to be returned, converted in place to the
// caller's layout.
static void poisson_out(pecs_ctx* c, double* v, bool rhs) {
  if (c->layout != PECS_LAYOUT_DEALII || !v) return;
  c->ptmp.assign(v, v + c->NP);
  if (rhs) poisson_rhs_to_dealii(*c->dii, c->ptmp.data(), v);
  else poisson_to_dealii(*c->dii, c->ptmp.data(), v);
}

static void h2d(pecs_ctx* c, double* d, const double* h, int64_t n) {
  HIPCHECK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
}
static void d2h(pecs_ctx* c, double* h, const double* d, int64_t n) {
  HIPCHECK(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
}

int pecs_rhs_semiconductor_dev(pecs_ctx* ctx, const double* un, const double* up, const double* pv,
                               const double* ur, const double* uo, double* bn, double* bp) {
  PECS_TRY(ctx, {
    HIPCHECK((hipError_t)launch_rhs_carrier(ctx->k, true, ctx->Rs, ctx->Re, ctx->Pd, ctx->tabs, ctx->phs_d, un, up, ur,
                                            uo, pv, bn, bp, ctx->stream));
  })
}

int pecs_rhs_electrolyte_dev(pecs_ctx* ctx, const double* ur, const double* uo, const double* pv,
                             const double* un, const double* up, double* br, double* bo) {
  PECS_TRY(ctx, {
    HIPCHECK((hipError_t)launch_rhs_carrier(ctx->k, false, ctx->Re, ctx->Rs, ctx->Pd, ctx->tabs, ctx->phs_d, ur, uo,
                                            un, up, pv, br, bo, ctx->stream));
  })
}

int pecs_rhs_poisson_dev(pecs_ctx* ctx, const double* un, const double* up, const double* ur, const double* uo,
                         double* rhs) {
  PECS_TRY(ctx, {
    HIPCHECK(hipMemsetAsync(rhs, 0, ctx->NP * sizeof(double), ctx->stream));
    HIPCHECK((hipError_t)launch_poisson_rhs_assemble(ctx->k, ctx->Pd, ctx->tabs, ctx->phs_d, un, up, ur, uo,
                                                     ctx->mesh.semi.ncell, ctx->mesh.elec.ncell, rhs, ctx->stream));
  })
}

int pecs_solve_dev(pecs_ctx* ctx, int which, const double* rhs, double* x) {
  PECS_TRY(ctx, {
    if (which >= PECS_ELECTRONS && which <= PECS_OXIDANTS) {
      solve_carrier(ctx, which, rhs, x);
    } else if (which == PECS_POISSON) {
      HIPCHECK((hipError_t)launch_poisson_cell_from_b(ctx->Pd, rhs, ctx->yP, ctx->stream));
      poisson_solve_from_y(ctx, 0, 1, x);
    } else {
      throw std::runtime_error("unknown system id");
    }
  })
}

int pecs_solve_dev_noalloc(pecs_ctx* ctx, int which, const double* rhs, double* x) {
  PECS_TRY(ctx, {
    if (ctx->B != 1)
      throw std::runtime_error("pecs_solve_dev_noalloc: single-instance contexts only (a batch's MFMA-shaped "
                               "factors need right-hand-side-interleaved scratch)");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    double* base = (dev >= 0 && dev < NOALLOC_MAX_DEVICES) ? g_scratch_dev[dev] : nullptr;
    if (!base)
      throw std::runtime_error("pecs_solve_dev_noalloc: no device scratch on the current device (create a "
                               "context on it first)");
    size_t cap = NOALLOC_SCRATCH_BYTES / sizeof(double);
    // PECS_NOALLOC_LIMIT_BYTES may only lower the limit (to exercise the size check)
    if (const char* lim = std::getenv("PECS_NOALLOC_LIMIT_BYTES")) {
      const size_t l = (size_t)std::strtoull(lim, nullptr, 10) / sizeof(double);
      if (l < cap) cap = l;
    }
    hipStream_t s = ctx->stream;
    if (which >= PECS_ELECTRONS && which <= PECS_OXIDANTS) {
      // scratch: X (n), Z (zlen) of the factor, then the Schur right-hand side
      // g (n) and the forward output t (n); every entry is written before it is read
      const int reg = which / 2, car = which % 2;
      const DevRegion& R = reg == 0 ? ctx->Rs : ctx->Re;
      FactorDev* F = reg == 0 ? ctx->Fs : ctx->Fe;
      const bool share = reg == 0 ? ctx->share_s : ctx->share_e;
      const FactorDev& Fc = (car == 1 && !share) ? F[1] : F[0];
      const size_t n = (size_t)Fc.n, zl = (size_t)std::max<int64_t>(Fc.d.zlen, 1);
      const size_t need = 3 * n + zl;
      if (need > cap)
        throw std::runtime_error("pecs_solve_dev_noalloc: carrier solve needs " + std::to_string(need * 8) +
                                 " bytes of scratch, above the limit of " + std::to_string(cap * 8));
      DevFactor d = Fc.d;
      d.X = base;
      d.Z = base + n;
      d.Part = nullptr;
      double* g = base + n + zl;
      double* t = g + n;
      const double mu = ctx->P.mu[which];
      HIPCHECK((hipError_t)launch_schur_rhs(ctx->k, R, rhs, rhs, mu, mu, g, g, 1, s));
      HIPCHECK((hipError_t)launch_chol_solve(d, Fc.lv(), g, t, 1, s));
      HIPCHECK((hipError_t)launch_schur_recover(ctx->k, R, rhs, rhs, g, g, mu, mu, x, x, 1, s));
    } else if (which == PECS_POISSON) {
      // scratch: X (nH), Z (zlen) of the multiplier factor, the condensed
      // cell data y (ysz), the multiplier right-hand side r (nH) and the
      // forward output t (nH)
      const FactorDev& FH = ctx->FH;
      const size_t nH = (size_t)FH.n, zl = (size_t)std::max<int64_t>(FH.d.zlen, 1), ys = (size_t)ctx->ysz;
      const size_t need = 3 * nH + zl + ys;
      if (need > cap)
        throw std::runtime_error("pecs_solve_dev_noalloc: Poisson solve needs " + std::to_string(need * 8) +
                                 " bytes of scratch, above the limit of " + std::to_string(cap * 8));
      DevFactor d = FH.d;
      d.X = base;
      d.Z = base + nH;
      d.Part = nullptr;
      double* y = base + nH + zl;
      double* r = y + ys;
      double* t = r + nH;
      HIPCHECK((hipError_t)launch_poisson_cell_from_b(ctx->Pd, rhs, y, s));
      HIPCHECK((hipError_t)launch_poisson_face(ctx->Pd, y, r, s, 1, ctx->ysz, (int64_t)nH));
      HIPCHECK((hipError_t)launch_chol_solve(d, FH.lv(), r, t, 1, s));
      HIPCHECK((hipError_t)launch_poisson_recover(ctx->k, ctx->Pd, r, y, x, s, 1, (int64_t)nH, ctx->ysz, ctx->NP));
    } else {
      throw std::runtime_error("unknown system id");
    }
  })
}

int pecs_rhs_semiconductor(pecs_ctx* ctx, const double* un, const double* up, const double* pv, const double* ur,
                           const double* uo, double* bn, double* bp) {
  PECS_TRY(ctx, {
    h2d(ctx, ctx->io[0], un, ctx->Ns);
    h2d(ctx, ctx->io[1], up, ctx->Ns);
    h2d(ctx, ctx->io[2], ur, ctx->Ne);
    h2d(ctx, ctx->io[3], uo, ctx->Ne);
    h2d(ctx, ctx->old[4], poisson_in(ctx, pv, false), ctx->NP);
    HIPCHECK((hipError_t)launch_rhs_carrier(ctx->k, true, ctx->Rs, ctx->Re, ctx->Pd, ctx->tabs, ctx->phs_d, ctx->io[0],
                                            ctx->io[1], ctx->io[2], ctx->io[3], ctx->old[4], ctx->b[0], ctx->b[1],
                                            ctx->stream));
    d2h(ctx, bn, ctx->b[0], ctx->Ns);
    d2h(ctx, bp, ctx->b[1], ctx->Ns);
  })
}

int pecs_rhs_electrolyte(pecs_ctx* ctx, const double* ur, const double* uo, const double* pv, const double* un,
                         const double* up, double* br, double* bo) {
  PECS_TRY(ctx, {
    h2d(ctx, ctx->io[0], un, ctx->Ns);
    h2d(ctx, ctx->io[1], up, ctx->Ns);
    h2d(ctx, ctx->io[2], ur, ctx->Ne);
    h2d(ctx, ctx->io[3], uo, ctx->Ne);
    h2d(ctx, ctx->old[4], poisson_in(ctx, pv, false), ctx->NP);
    HIPCHECK((hipError_t)launch_rhs_carrier(ctx->k, false, ctx->Re, ctx->Rs, ctx->Pd, ctx->tabs, ctx->phs_d, ctx->io[2],
                                            ctx->io[3], ctx->io[0], ctx->io[1], ctx->old[4], ctx->b[2], ctx->b[3],
                                            ctx->stream));
    d2h(ctx, br, ctx->b[2], ctx->Ne);
    d2h(ctx, bo, ctx->b[3], ctx->Ne);
  })
}

int pecs_rhs_poisson(pecs_ctx* ctx, const double* un, const double* up, const double* ur, const double* uo,
                     double* rhs) {
  PECS_TRY(ctx, {
    h2d(ctx, ctx->io[0], un, ctx->Ns);
    h2d(ctx, ctx->io[1], up, ctx->Ns);
    h2d(ctx, ctx->io[2], ur, ctx->Ne);
    h2d(ctx, ctx->io[3], uo, ctx->Ne);
    HIPCHECK(hipMemsetAsync(ctx->old[4], 0, ctx->NP * sizeof(double), ctx->stream));
    HIPCHECK((hipError_t)launch_poisson_rhs_assemble(ctx->k, ctx->Pd, ctx->tabs, ctx->phs_d, ctx->io[0], ctx->io[1],
                                                     ctx->io[2], ctx->io[3], ctx->mesh.semi.ncell,
                                                     ctx->mesh.elec.ncell, ctx->old[4], ctx->stream));
    d2h(ctx, rhs, ctx->old[4], ctx->NP);
    poisson_out(ctx, rhs, true);
  })
}

int pecs_solve(pecs_ctx* ctx, int which, const double* rhs, double* x) {
  PECS_TRY(ctx, {
    int64_t n = 0;
    if (which == PECS_ELECTRONS || which == PECS_HOLES) n = ctx->Ns;
    else if (which == PECS_REDUCTANTS || which == PECS_OXIDANTS) n = ctx->Ne;
    else if (which == PECS_POISSON) n = ctx->NP;
    else throw std::runtime_error("unknown system id");
    h2d(ctx, ctx->io[0], which == PECS_POISSON ? poisson_in(ctx, rhs, true) : rhs, n);
    if (which == PECS_POISSON) {
      HIPCHECK((hipError_t)launch_poisson_cell_from_b(ctx->Pd, ctx->io[0], ctx->yP, ctx->stream));
      poisson_solve_from_y(ctx, 0, 1, ctx->io[1]);
    } else {
      solve_carrier(ctx, which, ctx->io[0], ctx->io[1]);
    }
    d2h(ctx, x, ctx->io[1], n);
    if (which == PECS_POISSON) poisson_out(ctx, x, false);
  })
}

int pecs_set_state_batch(pecs_ctx* ctx, int i, const double* un, const double* up, const double* ur,
                         const double* uo, const double* pv) {
  PECS_TRY(ctx, {
    if (i < 0 || i >= ctx->B) throw std::runtime_error("instance index out of range");
    const int64_t Ns = ctx->Ns, Ne = ctx->Ne, NP = ctx->NP;
    h2d(ctx, ctx->u[0] + i * Ns, un, Ns);
    h2d(ctx, ctx->u[1] + i * Ns, up, Ns);
    h2d(ctx, ctx->u[2] + i * Ne, ur, Ne);
    h2d(ctx, ctx->u[3] + i * Ne, uo, Ne);
    if (pv) {
      h2d(ctx, ctx->pv + i * NP, poisson_in(ctx, pv, false), NP);
    } else {
      poisson_cells(ctx, i, ctx->u[0] + i * Ns, ctx->u[1] + i * Ns, ctx->u[2] + i * Ne, ctx->u[3] + i * Ne);
      poisson_solve_from_y(ctx, i, 1, ctx->pv + i * NP);
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  })
}

int pecs_set_state(pecs_ctx* ctx, const double* un, const double* up, const double* ur, const double* uo,
                   const double* pv) {
  return pecs_set_state_batch(ctx, 0, un, up, ur, uo, pv);
}

int pecs_get_state_batch(pecs_ctx* ctx, int i, double* un, double* up, double* ur, double* uo, double* pv) {
  PECS_TRY(ctx, {
    if (i < 0 || i >= ctx->B) throw std::runtime_error("instance index out of range");
    const int64_t Ns = ctx->Ns, Ne = ctx->Ne, NP = ctx->NP;
    if (un) d2h(ctx, un, ctx->u[0] + i * Ns, Ns);
    if (up) d2h(ctx, up, ctx->u[1] + i * Ns, Ns);
    if (ur) d2h(ctx, ur, ctx->u[2] + i * Ne, Ne);
    if (uo) d2h(ctx, uo, ctx->u[3] + i * Ne, Ne);
    if (pv) {
      d2h(ctx, pv, ctx->pv + i * NP, NP);
      poisson_out(ctx, pv, false);
    }
  })
}

int pecs_get_state(pecs_ctx* ctx, double* un, double* up, double* ur, double* uo, double* pv) {
  return pecs_get_state_batch(ctx, 0, un, up, ur, uo, pv);
}

int pecs_last_density_change(pecs_ctx* ctx, double* out) {
  PECS_TRY(ctx, {
    if (!out) throw std::runtime_error("pecs_last_density_change: null output");
    const int B = ctx->B;
    if (!ctx->dacc) ctx->dacc = dalloc<double>(ctx, 2 * (size_t)B);
    const int64_t nsv[4] = {ctx->Ns, ctx->Ns, ctx->Ne, ctx->Ne};
    for (int b = 0; b < B; ++b)
      for (int j = 0; j < 4; ++j) {
        const int64_t n = nsv[j], off = b * n + 2 * n / 3;  // the density block of [Jx | Jy | u]
        HIPCHECK((hipError_t)launch_diff_norms(ctx->u[j] + off, ctx->old[j] + off, n / 3,
                                               ctx->rpart + (size_t)(4 * b + j) * 2 * RES_PARTS, ctx->stream));
      }
    HIPCHECK((hipError_t)launch_sum_partials(ctx->rpart, 4, ctx->dacc, B, ctx->stream));
    std::vector<double> h(2 * (size_t)B);
    d2h(ctx, h.data(), ctx->dacc, 2 * (int64_t)B);
    for (int b = 0; b < B; ++b) out[b] = h[2 * b + 1] > 0 ? std::sqrt(h[2 * b] / h[2 * b + 1]) : 0.0;
  })
}

namespace {
double* grid_buffer(pecs_ctx* c, int G) {
  if (G < 1 || G > 4096) throw std::runtime_error("grid side must be 1..4096");
  const int64_t n = 4 * (int64_t)G * G;
  if (n > c->grid_cap) {
    c->grid_dev = dalloc<double>(c, n);
    c->grid_cap = n;
  }
  return c->grid_dev;
}
}  // namespace

int pecs_sample_grid(pecs_ctx* ctx, int i, int G, double* out) {
  PECS_TRY(ctx, {
    if (ctx->mesh.local) throw std::runtime_error("the density grid needs a uniformly refined mesh");
    if (i < 0 || i >= ctx->B) throw std::runtime_error("instance index out of range");
    if (!out) throw std::runtime_error("pecs_sample_grid: null output");
    double* g = grid_buffer(ctx, G);
    const int64_t G2 = (int64_t)G * G;
    for (int car = 0; car < 4; ++car) {
      const int64_t n = car < 2 ? ctx->Ns : ctx->Ne;
      const int rc = launch_grid_sample(ctx->k, ctx->u[car] + i * n, ctx->mesh.r, G, g + car * G2, ctx->stream);
      if (rc) throw std::runtime_error("launch_grid_sample failed");
    }
    d2h(ctx, out, g, 4 * G2);
  })
}

int pecs_set_state_from_grid(pecs_ctx* ctx, int i, int G, const double* grid) {
  PECS_TRY(ctx, {
    if (ctx->mesh.local) throw std::runtime_error("the density grid needs a uniformly refined mesh");
    if (i < 0 || i >= ctx->B) throw std::runtime_error("instance index out of range");
    if (!grid) throw std::runtime_error("pecs_set_state_from_grid: null grid");
    double* g = grid_buffer(ctx, G);
    const int64_t G2 = (int64_t)G * G;
    h2d(ctx, g, grid, 4 * G2);
    double* u[4];
    for (int car = 0; car < 4; ++car) {
      const int64_t n = car < 2 ? ctx->Ns : ctx->Ne;
      u[car] = ctx->u[car] + i * n;
      HIPCHECK(hipMemsetAsync(u[car], 0, n * sizeof(double), ctx->stream));  // currents 0
      const int rc = launch_grid_project(ctx->k, g + car * G2, G, ctx->mesh.r, u[car], ctx->stream);
      if (rc) throw std::runtime_error("launch_grid_project failed");
    }
    // the potential of these densities, as pecs_set_state without one
    poisson_cells(ctx, i, u[0], u[1], u[2], u[3]);
    poisson_solve_from_y(ctx, i, 1, ctx->pv + i * ctx->NP);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  })
}

int pecs_initial_state(pecs_ctx* ctx) {
  PECS_TRY(ctx, {
    // VectorTools::project of the constant equilibria (SolarCell.cpp:1999-2021):
    // the L2 projection of a constant onto FE_DGQ is that constant, J = 0.
    const int64_t nsv[4] = {ctx->Ns, ctx->Ns, ctx->Ne, ctx->Ne};
    for (int i = 0; i < ctx->B; ++i) {
      const Params& P = ctx->Pb[i];
      const double vals[4] = {P.ne, P.pe, P.rinf, P.oinf};
      for (int c = 0; c < 4; ++c) {
        std::vector<double> h(nsv[c], 0.0);
        const int64_t n3 = nsv[c] / 3;
        for (int64_t j = 2 * n3; j < nsv[c]; ++j) h[j] = vals[c];
        HIPCHECK(hipMemcpy(ctx->u[c] + i * nsv[c], h.data(), nsv[c] * sizeof(double), hipMemcpyHostToDevice));
      }
    }
    poisson_all_from_densities(ctx);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  })
}


int pecs_step(pecs_ctx* ctx, int nsteps) {
  PECS_TRY(ctx, {
    if (nsteps < 0) throw std::runtime_error("nsteps must be >= 0");
    for (int i = 0; i < nsteps; ++i) {
      run_step(ctx, i == nsteps - 1);
      if (ctx->timing && ctx->ev_used + NPHASE + 1 > (int)ctx->events.size()) collect_timing(ctx);
    }
    collect_timing(ctx);
    if (nsteps > 0) {
      HIPCHECK(hipStreamSynchronize(ctx->stream));
      finish_residual(ctx);
    }
  })
}

// ---------------------------------------------------------------------------
// Sweeps over contexts with different meshes (uneven-mesh I-V / geometry
// sweeps): a time step of all active members is one launch of each member's
// captured step graph, the members spread round-robin over up to 4 streams so
// that the small members' launch-latency-bound steps overlap each other and
// the large members' bandwidth-bound ones.  (One parent graph holding the
// members' graphs as child nodes measured slower than these launches.)

struct pecs_sweep {
  std::vector<pecs_ctx*> m;
  std::vector<char> active;
  std::vector<hipStream_t> st;
  // one launch per step (nstreams 0 at creation): the active members' step
  // graphs (and last-step graphs) as child graphs of one graph each, rebuilt
  // when the active set or a member's captured graphs change
  bool one_launch = false;
  hipGraph_t pg[2] = {};
  hipGraphExec_t pe[2] = {};
  std::vector<char> built_active;
  std::vector<uint64_t> built_gen;
  void drop() {
    for (int j = 0; j < 2; ++j) {
      if (pe[j]) (void)hipGraphExecDestroy(pe[j]);
      if (pg[j]) (void)hipGraphDestroy(pg[j]);
      pe[j] = nullptr;
      pg[j] = nullptr;
    }
  }
  ~pecs_sweep() {
    for (hipStream_t s : st) (void)hipStreamSynchronize(s);
    drop();
    for (hipStream_t s : st) (void)hipStreamDestroy(s);
  }
};

namespace {
void ensure_sweep_graphs(pecs_sweep* w) {
  std::vector<uint64_t> gen(w->m.size());
  for (size_t i = 0; i < w->m.size(); ++i) gen[i] = w->active[i] ? w->m[i]->graph_gen : 0;
  if (w->pe[0] && w->built_active == w->active && w->built_gen == gen) return;
  w->drop();
  for (int j = 0; j < 2; ++j) {
    HIPCHECK(hipGraphCreate(&w->pg[j], 0));
    for (size_t i = 0; i < w->m.size(); ++i) {
      if (!w->active[i]) continue;
      hipGraphNode_t node;
      HIPCHECK(hipGraphAddChildGraphNode(&node, w->pg[j], nullptr, 0, j ? w->m[i]->last_graph : w->m[i]->step_graph));
    }
    HIPCHECK(hipGraphInstantiate(&w->pe[j], w->pg[j], nullptr, nullptr, 0));
  }
  w->built_active = w->active;
  w->built_gen = gen;
}
}  // namespace

int pecs_sweep_create(pecs_ctx** ctxs, int n, int nstreams, pecs_sweep** out) {
  if (!out) return -1;
  *out = nullptr;
  try {
    if (!ctxs || n < 1) throw std::runtime_error("pecs_sweep_create: need at least one context");
    if (nstreams < 0 || nstreams > 4) throw std::runtime_error("pecs_sweep_create: nstreams must be 0..4");
    for (int i = 0; i < n; ++i) {
      if (!ctxs[i]) throw std::runtime_error("pecs_sweep_create: null context");
      for (int j = 0; j < i; ++j)
        if (ctxs[j] == ctxs[i]) throw std::runtime_error("pecs_sweep_create: a context appears twice");
    }
    auto w = std::make_unique<pecs_sweep>();
    w->m.assign(ctxs, ctxs + n);
    w->active.assign(n, 1);
    w->one_launch = nstreams == 0;
    w->st.assign(std::max(nstreams, 1), nullptr);
    for (hipStream_t& s : w->st) HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = w.release();
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
  return 0;
}

int pecs_sweep_set_active(pecs_sweep* w, const int* active) {
  if (!w || !active) return -1;
  for (size_t i = 0; i < w->m.size(); ++i) w->active[i] = active[i] ? 1 : 0;
  return 0;
}

int pecs_sweep_step(pecs_sweep* w, int nsteps) {
  if (!w) return -1;
  try {
    if (nsteps < 0) throw std::runtime_error("nsteps must be >= 0");
    const int ns = (int)w->st.size();
    std::vector<int> q_of(w->m.size(), -1);
    int slot = 0;
    for (size_t i = 0; i < w->m.size(); ++i) {
      if (!w->active[i]) continue;
      // the member's own stream drained: its state is complete before the sweep reads it
      HIPCHECK(hipStreamSynchronize(w->m[i]->stream));
      ensure_step_graph(w->m[i]);
      q_of[i] = slot++ % ns;
    }
    if (w->one_launch) {
      if (slot > 0 && nsteps > 0) {
        ensure_sweep_graphs(w);
        for (int t = 0; t < nsteps; ++t) HIPCHECK(hipGraphLaunch(w->pe[t == nsteps - 1 ? 1 : 0], w->st[0]));
      }
    } else {
      for (int t = 0; t < nsteps; ++t)
        for (size_t i = 0; i < w->m.size(); ++i)
          if (q_of[i] >= 0)
            HIPCHECK(hipGraphLaunch(t == nsteps - 1 ? w->m[i]->last_exec : w->m[i]->step_exec, w->st[q_of[i]]));
    }
    if (nsteps > 0) {
      for (hipStream_t s : w->st) HIPCHECK(hipStreamSynchronize(s));
      for (size_t i = 0; i < w->m.size(); ++i)
        if (q_of[i] >= 0) finish_residual(w->m[i]);
    }
  } catch (const std::exception& e) {
    for (hipStream_t s : w->st) (void)hipStreamSynchronize(s);
    g_error = e.what();
    return -1;
  }
  return 0;
}

int pecs_sweep_destroy(pecs_sweep* w) {
  if (!w) return -1;
  delete w;
  return 0;
}

int pecs_set_mixed_precision(pecs_ctx* ctx, int enable) {
  PECS_TRY(ctx, {
    if (enable && (ctx->Rs.msrc || ctx->Rs.mdir || ctx->Rs.mflux || ctx->Re.msrc || ctx->Re.mdir ||
                   ctx->Re.mflux || ctx->Pd.msrc || ctx->Pd.mdir))
      throw std::runtime_error("mixed precision and manufactured data do not combine");
    if (enable) {
      if (!(ctx->Rs.ucell_own && ctx->Rs.ucell_nbh && ctx->Re.ucell_own && ctx->Re.ucell_nbh && ctx->Pd.ucell) ||
          ctx->tabf_host[0].empty() || ctx->tabf_host[1].empty() || ctx->tabf_host[2].empty())
        throw std::runtime_error(
            "mixed precision needs the wave-uniform cell lists (a planar mesh with few cell types)");
      for (int i = 0; i < 3; ++i)
        if (!ctx->tabf_dev[i]) ctx->tabf_dev[i] = upload(ctx, ctx->tabf_host[i]);
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->Rs.table_f = enable ? ctx->tabf_dev[0] : nullptr;
    ctx->Re.table_f = enable ? ctx->tabf_dev[1] : nullptr;
    ctx->Pd.table_f = enable ? ctx->tabf_dev[2] : nullptr;
    // the captured step graphs hold the old kernel arguments: capture again
    drop_graphs(ctx);
  })
}

int pecs_set_manufactured(pecs_ctx* ctx, int target, const double* src, const double* dir, const double* flux) {
  PECS_TRY(ctx, {
    if (target < 0 || target > 2) throw std::runtime_error("target must be 0 (semiconductor), 1 (electrolyte) or 2 (Poisson)");
    if (target == 2 && flux) throw std::runtime_error("the Poisson right-hand side takes no manufactured flux");
    if (ctx->Rs.table_f) throw std::runtime_error("mixed precision and manufactured data do not combine");
    const RefTables& T = *ctx->T;
    const int64_t ncell = target == 0 ? ctx->mesh.semi.ncell : target == 1 ? ctx->mesh.elec.ncell : ctx->mesh.ncellP;
    const int64_t len[3] = {ncell * T.nq, ncell * 4 * T.nfq, ncell * 4 * T.nfq};
    const double* in[3] = {src, dir, flux};
    double* dev[3] = {};
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int j = 0; j < 3; ++j) {
      if (!in[j]) continue;
      if (!ctx->man_dev[target][j]) ctx->man_dev[target][j] = dalloc<double>(ctx, len[j]);
      dev[j] = ctx->man_dev[target][j];
      HIPCHECK(hipMemcpy(dev[j], in[j], len[j] * sizeof(double), hipMemcpyHostToDevice));
    }
    bool changed;
    if (target < 2) {
      DevRegion& R = target == 0 ? ctx->Rs : ctx->Re;
      changed = R.msrc != dev[0] || R.mdir != dev[1] || R.mflux != dev[2];
      R.msrc = dev[0];
      R.mdir = dev[1];
      R.mflux = dev[2];
    } else {
      changed = ctx->Pd.msrc != dev[0] || ctx->Pd.mdir != dev[1];
      ctx->Pd.msrc = dev[0];
      ctx->Pd.mdir = dev[1];
    }
    // captured step graphs hold the kernel arguments: capture again when the
    // pointers change (new contents in the same buffers need no capture)
    if (changed) drop_graphs(ctx);
  })
}

int pecs_quadrature_points(pecs_ctx* ctx, int target, int64_t* ncell, int* nq, int* nfq, double* cell_xy,
                           double* face_xy) {
  PECS_TRY(ctx, {
    if (target < 0 || target > 2) throw std::runtime_error("target must be 0 (semiconductor), 1 (electrolyte) or 2 (Poisson)");
    const RefTables& T = *ctx->T;
    const std::vector<double>& V = target == 0 ? ctx->mesh.semi.V : target == 1 ? ctx->mesh.elec.V : ctx->mesh.pV;
    const int64_t nc = (int64_t)V.size() / 8;
    if (ncell) *ncell = nc;
    if (nq) *nq = T.nq;
    if (nfq) *nfq = T.nfq;
    for (int64_t c = 0; c < nc; ++c) {
      const double* v = &V[8 * c];
      if (cell_xy)
        for (int q = 0; q < T.nq; ++q) {
          const Geo g = bilinear(v, T.qx[q], T.qy[q]);
          cell_xy[2 * (c * T.nq + q)] = g.x;
          cell_xy[2 * (c * T.nq + q) + 1] = g.y;
        }
      if (face_xy)
        for (int f = 0; f < 4; ++f)
          for (int q = 0; q < T.nfq; ++q) {
            double xi, eta;
            face_point(f, T.fs[q], xi, eta);
            const Geo g = bilinear(v, xi, eta);
            const int64_t o = 2 * ((c * 4 + f) * T.nfq + q);
            face_xy[o] = g.x;
            face_xy[o + 1] = g.y;
          }
    }
  })
}

int pecs_uniform_lists(pecs_ctx* ctx, int* flags) {
  if (!ctx || !flags) return -1;
  *flags = (ctx->Rs.ucell_own ? 1 : 0) | (ctx->Rs.ucell_nbh ? 2 : 0) | (ctx->Re.ucell_own ? 4 : 0) |
           (ctx->Re.ucell_nbh ? 8 : 0) | (ctx->Pd.ucell ? 16 : 0);
  return 0;
}

int pecs_sweep_config(pecs_ctx* ctx, int which, int* out, int n, int* nlevels) {
  PECS_TRY(ctx, {
    const FactorDev* F = which == 0 ? &ctx->Fs[0] : which == 1 ? &ctx->Fe[0] : which == 2 ? &ctx->FH : nullptr;
    if (!F) throw std::runtime_error("which must be 0 (semiconductor), 1 (electrolyte) or 2 (Poisson multipliers)");
    if (nlevels) *nlevels = F->nlevels;
    for (int l = 0; l < F->nlevels && 6 * l + 5 < n; ++l) {
      const int v[6] = {F->fw[l], F->fm[l], F->frows[l], F->bw[l], F->bm[l], F->brows[l]};
      for (int j = 0; j < 6; ++j) out[6 * l + j] = v[j];
    }
  })
}

int pecs_use_graph(pecs_ctx* ctx, int enable) {
  PECS_TRY(ctx, { ctx->use_graph = enable != 0; })
}

int pecs_enable_timing(pecs_ctx* ctx, int enable) {
  PECS_TRY(ctx, {
    ctx->timing = enable ? 1 : 0;
    if (ctx->timing && ctx->events.empty()) {
      ctx->events.resize(64 * (NPHASE + 1));
      for (auto& e : ctx->events) HIPCHECK(hipEventCreate(&e));
    }
    ctx->phase_ms.assign(NPHASE, 0.0);
    ctx->timed_steps = 0;
    ctx->ev_used = 0;
  })
}

int pecs_phase_times(pecs_ctx* ctx, double* ms, int n, int* steps) {
  if (!ctx || !ms) return -1;
  for (int i = 0; i < n && i < NPHASE; ++i) ms[i] = i < (int)ctx->phase_ms.size() ? ctx->phase_ms[i] : 0.0;
  if (steps) *steps = ctx->timed_steps;
  return 0;
}

int pecs_factor_bytes(pecs_ctx* ctx, int which, double* bytes) {
  // Algorithmic bytes of one triangular solve pair (forward + backward) with the
  // factor of system `which` (0: semiconductor Schur, 1: electrolyte Schur,
  // 2: Poisson multipliers): the lower triangles of the inverse diagonal
  // blocks and the off-diagonal panels, read once per sweep, 8 B per value.
  if (!ctx || !bytes) return -1;
  const FactorDev* F = which == 0 ? &ctx->Fs[0] : (which == 1 ? &ctx->Fe[0] : &ctx->FH);
  *bytes = F->tri_bytes;
  return 0;
}

int pecs_last_residual(pecs_ctx* ctx, double* r) {
  if (!ctx || !r) return -1;
  double m = -1.0;
  for (double v : ctx->last_res) m = std::max(m, v);
  *r = m;
  return 0;
}

int pecs_last_residual_batch(pecs_ctx* ctx, double* r) {
  if (!ctx || !r) return -1;
  for (int i = 0; i < ctx->B; ++i) r[i] = ctx->last_res[i];
  return 0;
}

int pecs_state_dev(pecs_ctx* ctx, double** un, double** up, double** ur, double** uo, double** pv) {
  if (!ctx) return -1;
  if (un) *un = ctx->u[0];
  if (up) *up = ctx->u[1];
  if (ur) *ur = ctx->u[2];
  if (uo) *uo = ctx->u[3];
  if (pv) *pv = ctx->pv;
  return 0;
}

}  // extern "C"
