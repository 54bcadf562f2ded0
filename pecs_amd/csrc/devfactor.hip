// Numeric multifrontal Cholesky factorisation on the device (gfx950): front
// assembly kernels, rocSOLVER / rocBLAS batched dense kernels per front shape,
// and the kernels that write the sweep layouts.
//
// Front storage: the front of supernode s (nf = ns + ms) is an nf x nf
// row-major array whose lower triangle holds the symmetric matrix.  Read
// column-major with leading dimension nf, the same memory is the upper
// triangle, so the dense calls below work on Upper / column-major operands:
//   potrf(Upper):            leading block = U11 (U11^T U11 = F11), U11 = L11^T
//   trsm(Left, Upper, T):    panel (columns ns..nf of rows 0..ns) = L21^T
//   gemm(T, N, -1, 1):       trailing block -= L21 L21^T (both triangles)
//   trtri(Upper):            leading block = U11^-1 = Linv^T
//   trmm(Left, Upper, N):    W21^T = Linv^T L21^T, column-major ns x ms
//                            = W21 row-major ms x ns, straight into fwd.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "devfactor.hpp"

namespace pecs {
namespace {

#define DFCHECK(x)                                                                                \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define RBCHECK(x)                                                                                 \
  do {                                                                                             \
    rocblas_status s_ = (x);                                                                       \
    if (s_ != rocblas_status_success)                                                              \
      throw std::runtime_error(std::string(#x) + ": " + rocblas_status_to_string(s_));             \
  } while (0)

struct ScItem {
  int64_t src, dst;
  int nx, ny, nf, mode, rx, ry;
};
struct EaItem {
  int child, row, parent;
};
struct PkItem {
  int node, j0;
};

// Device allocations of one call, freed together once the stream has drained.
struct Pool {
  hipStream_t st;
  std::vector<void*> p;
  explicit Pool(hipStream_t s) : st(s) {}
  template <class T>
  T* alloc(size_t n) {
    void* q = nullptr;
    DFCHECK(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
    p.push_back(q);
    return (T*)q;
  }
  template <class T>
  T* up(const std::vector<T>& v, hipStream_t s) {
    T* q = alloc<T>(v.size());
    if (!v.empty()) DFCHECK(hipMemcpyAsync(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return q;
  }
  ~Pool() {
    (void)hipStreamSynchronize(st);
    for (void* q : p) (void)hipFree(q);
  }
};

// S blocks into the fronts of one level: one wave per (cell, neighbour) block.
// Mode 0 (both cells own): the lower-triangle part of the block at rows rx..,
// columns ry..; mode 1 (neighbour in the struct rows): the block transposed.
// Items write disjoint positions of zeroed fronts.
__global__ __launch_bounds__(64) void k_front_scatter(const double* __restrict__ sv, const ScItem* __restrict__ items,
                                                      int i0, double* __restrict__ fr) {
  const ScItem I = items[i0 + blockIdx.x];
  const int n = I.nx * I.ny;
  for (int e = threadIdx.x; e < n; e += 64) {
    const int a = e / I.ny, b = e - a * I.ny;
    const double v = sv[I.src + e];
    if (I.mode == 0) {
      if (I.rx + a >= I.ry + b) fr[I.dst + (int64_t)a * I.nf + b] = v;
    } else {
      fr[I.dst + (int64_t)b * I.nf + a] = v;
    }
  }
}

// Extend-add of one struct row r of a child's updated trailing block into
// its parent's front (lower triangle: child columns 0..r).  The child's row
// map is injective, so the rows of one pass write disjoint positions.
__global__ __launch_bounds__(64) void k_extend_add(const EaItem* __restrict__ items, int i0, const int* __restrict__ ns,
                                                   const int* __restrict__ ms, const int64_t* __restrict__ front_off,
                                                   const int* __restrict__ map, const int64_t* __restrict__ map_off,
                                                   double* __restrict__ fr) {
  const EaItem I = items[i0 + blockIdx.x];
  const int nsc = ns[I.child], nfc = nsc + ms[I.child];
  const int nfp = ns[I.parent] + ms[I.parent];
  const int* m = map + map_off[I.child];
  const int r = I.row;
  double* dst = fr + front_off[I.parent] + (int64_t)m[r] * nfp;
  const double* src = fr + front_off[I.child] + (int64_t)(nsc + r) * nfc + nsc;
  for (int q = threadIdx.x; q <= r; q += 64) dst[m[q]] += src[q];
}

// Sweep layout of 16 own rows j0.. of a node from its factored front:
//   fwd row r < ns: Linv(r, 0..r)                      (row-major copy)
//   bwd row j:      [Linv(j..ns-1, j) | -W21(0..ms-1, j)]  (a column: staged
//                   through LDS in 16 x 16 tiles so loads and stores both
//                   run over 16 consecutive doubles)
// W21 (ms x ns row-major) is read from fwd, where trmm wrote it.
__global__ __launch_bounds__(256) void k_pack_rows(const PkItem* __restrict__ items, const int* __restrict__ ns_,
                                                   const int* __restrict__ ms_, const int64_t* __restrict__ front_off,
                                                   const int64_t* __restrict__ fwd_off, const int64_t* __restrict__ bwd_off,
                                                   const double* __restrict__ fr, double* __restrict__ fwd,
                                                   double* __restrict__ bwd) {
  __shared__ double t[16][17];
  const PkItem I = items[blockIdx.x];
  const int s = I.node, j0 = I.j0, ns = ns_[s], ms = ms_[s], nf = ns + ms;
  const double* Fr = fr + front_off[s];
  double* f = fwd + fwd_off[s];
  const double* W = f + (int64_t)ns * (ns + 1) / 2;
  double* bw = bwd + bwd_off[s];
  const int jend = min(j0 + 16, ns);
  for (int r = j0; r < jend; ++r)
    for (int c = threadIdx.x; c <= r; c += 256) f[(int64_t)r * (r + 1) / 2 + c] = Fr[(int64_t)r * nf + c];
  const int a = threadIdx.x & 15, b = threadIdx.x >> 4;
  for (int i0 = j0; i0 < nf; i0 += 16) {
    {
      const int j = j0 + a, i = i0 + b;
      double v = 0.0;
      if (j < ns && i >= j && i < nf) v = i < ns ? Fr[(int64_t)i * nf + j] : -W[(int64_t)(i - ns) * ns + j];
      t[a][b] = v;
    }
    __syncthreads();
    {
      const int j = j0 + b, i = i0 + a;
      if (j < ns && i >= j && i < nf) bw[(int64_t)j * nf - (int64_t)j * (j - 1) / 2 + (i - j)] = t[b][a];
    }
    __syncthreads();
  }
}

// One 16-row block of the MFMA-blocked copy (pack_blocked's layout): entry
// e = (16-column step st, group q, lane l = i + 16 k) holds row 16 b + i,
// column clo + 16 st + 4 k + q of the node's fwd (or bwd) rows, zero outside
// the row.  Every entry of the block is written.
__global__ __launch_bounds__(256) void k_pack_blocked(int fwd, const int* __restrict__ blk_node,
                                                      const int* __restrict__ blk0, const int64_t* __restrict__ boff,
                                                      const int* __restrict__ ns_, const int* __restrict__ ms_,
                                                      const int64_t* __restrict__ src_off, const double* __restrict__ src,
                                                      double* __restrict__ dst) {
  const int gb = blockIdx.x;
  const int s = blk_node[gb], b = gb - blk0[s];
  const int ns = ns_[s], ms = ms_[s];
  const int clo = fwd ? 0 : 16 * b;
  int chi;
  if (fwd) {
    const int last = min(16 * b + 15, ns + ms - 1);
    chi = last < ns ? last + 1 : ns;
  } else {
    chi = ns + ms;
  }
  const int nsteps = (chi - clo + 15) / 16;
  const int rows = fwd ? ns + ms : ns;
  const double* sr = src + src_off[s];
  double* d = dst + boff[gb];
  for (int e = threadIdx.x; e < nsteps * 256; e += 256) {
    const int st = e >> 8, q = (e >> 6) & 3, lane = e & 63, i = lane & 15, k = lane >> 4;
    const int lr = 16 * b + i, col = clo + 16 * st + 4 * k + q;
    double v = 0.0;
    if (lr < rows) {
      int64_t off;
      int rs, re;
      if (fwd) {
        rs = 0;
        if (lr < ns) {
          off = (int64_t)lr * (lr + 1) / 2;
          re = lr + 1;
        } else {
          off = (int64_t)ns * (ns + 1) / 2 + (int64_t)(lr - ns) * ns;
          re = ns;
        }
      } else {
        off = (int64_t)lr * (ns + ms) - (int64_t)lr * (lr - 1) / 2 - lr;
        rs = lr;
        re = ns + ms;
      }
      if (col >= rs && col < re) v = sr[off + col];
    }
    d[e] = v;
  }
}

// Groups of at most kFewFronts fronts with ns >= kLargeFront use one
// non-batched call per front.
constexpr int kLargeFront = 512, kFewFronts = 8;

struct Group {
  int level, ns, ms, count;
  size_t ptr0;  // first of 4 * count pointers: A11[count] | A12[count] | A22[count] | W[count]
  size_t info0;
};

}  // namespace

void* factor_handle_create() {
  rocblas_handle h = nullptr;
  RBCHECK(rocblas_create_handle(&h));
  return h;
}

void factor_handle_destroy(void* handle) {
  if (handle) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(handle));
}

DeviceFactorStats factor_on_device(const FrontPlan& plan, const SupernodalFactor& F, const PackedFactor& P,
                                   double* fwd, double* bwd, hipStream_t st, void* handle) {
  const auto t0 = std::chrono::steady_clock::now();
  DeviceFactorStats stats;
  const int nn = (int)F.nodes.size();
  if (!handle) throw std::runtime_error("device factorisation: no dense-library handle");
  {  // the fronts are the large allocation: fail with a clear message, not inside the dense calls
    size_t free_b = 0, total_b = 0;
    DFCHECK(hipMemGetInfo(&free_b, &total_b));
    const double need = 8.0 * (double)plan.front_total + 8.0 * (double)plan.svals.size() + (256.0 * (1 << 20));
    if (need > (double)free_b)
      throw std::runtime_error("device factorisation: the fronts need " + std::to_string(need * 1e-9) +
                               " GB, the device has " + std::to_string(free_b * 1e-9) + " GB free");
  }
  Pool pool(st);
  // plan arrays
  std::vector<ScItem> sc(plan.sc_src.size());
  for (size_t i = 0; i < sc.size(); ++i)
    sc[i] = ScItem{plan.sc_src[i], plan.sc_dst[i], plan.sc_nx[i], plan.sc_ny[i],
                   plan.sc_nf[i],  plan.sc_mode[i], plan.sc_rx[i], plan.sc_ry[i]};
  std::vector<EaItem> ea(plan.ea_child.size());
  for (size_t i = 0; i < ea.size(); ++i) ea[i] = EaItem{plan.ea_child[i], plan.ea_row[i], plan.ea_parent[i]};
  std::vector<PkItem> pk;
  for (int s = 0; s < nn; ++s)
    for (int j0 = 0; j0 < P.ns[s]; j0 += 16) pk.push_back(PkItem{s, j0});
  double* sv = pool.up(plan.svals, st);
  ScItem* d_sc = pool.up(sc, st);
  EaItem* d_ea = pool.up(ea, st);
  PkItem* d_pk = pool.up(pk, st);
  int* d_ns = pool.up(P.ns, st);
  int* d_ms = pool.up(P.ms, st);
  int64_t* d_front_off = pool.up(plan.front_off, st);
  int64_t* d_fwd_off = pool.up(P.fwd_off, st);
  int64_t* d_bwd_off = pool.up(P.bwd_off, st);
  int* d_map = pool.up(plan.ea_map, st);
  int64_t* d_map_off = pool.up(plan.ea_map_off, st);
  double* fr = pool.alloc<double>(plan.front_total);
  stats.front_bytes = 8.0 * (double)plan.front_total;
  DFCHECK(hipMemsetAsync(fr, 0, (size_t)plan.front_total * sizeof(double), st));
  // fronts of equal shape per level, and their operand pointers
  std::vector<Group> groups;
  std::vector<double*> ptrs;
  size_t ninfo = 0;
  for (int l = 0; l < F.nlevels; ++l) {
    std::map<std::pair<int, int>, std::vector<int>> by;
    for (int s : F.level_nodes[l]) by[{F.nodes[s].ns, F.nodes[s].ms}].push_back(s);
    for (auto& kv : by) {
      const int ns = kv.first.first, ms = kv.first.second, nf = ns + ms;
      if (ns == 0) continue;
      Group g{l, ns, ms, (int)kv.second.size(), ptrs.size(), ninfo};
      for (int s : kv.second) ptrs.push_back(fr + plan.front_off[s]);
      for (int s : kv.second) ptrs.push_back(fr + plan.front_off[s] + (int64_t)ns * nf);
      for (int s : kv.second) ptrs.push_back(fr + plan.front_off[s] + (int64_t)ns * nf + ns);
      for (int s : kv.second) ptrs.push_back(fwd + P.fwd_off[s] + (int64_t)ns * (ns + 1) / 2);
      ninfo += 2 * (size_t)g.count;
      groups.push_back(g);
      const double n_ = ns, m_ = ms;
      stats.flops += g.count * (2.0 * n_ * n_ * n_ / 3.0 + 2.0 * n_ * n_ * m_ + n_ * m_ * m_);
    }
  }
  double** d_ptrs = pool.up(ptrs, st);
  rocblas_int* d_info = pool.alloc<rocblas_int>(ninfo);
  DFCHECK(hipMemsetAsync(d_info, 0, std::max<size_t>(ninfo, 1) * sizeof(rocblas_int), st));
  // PECS_SETUP_TRACE=1: per-level times of the assembly kernels and the dense calls on stderr
  const char* tr = std::getenv("PECS_SETUP_TRACE");
  const bool trace = tr && std::string(tr) == "1";
  std::vector<hipEvent_t> ev;
  std::vector<std::pair<int, hipEvent_t>> opev;
  if (trace) {
    ev.resize(2 * F.nlevels + 2);
    for (auto& e : ev) DFCHECK(hipEventCreate(&e));
  }
  rocblas_handle h = static_cast<rocblas_handle>(handle);
  try {
    RBCHECK(rocblas_set_stream(h, st));
    const double one = 1.0, minus_one = -1.0;
    size_t gi = 0;
    if (trace) DFCHECK(hipEventRecord(ev[0], st));
    for (int l = 0; l < F.nlevels; ++l) {
      const int s0 = plan.sc_level[l], s1 = plan.sc_level[l + 1];
      if (s1 > s0) k_front_scatter<<<s1 - s0, 64, 0, st>>>(sv, d_sc, s0, fr);
      for (int pass = 0; pass < 2; ++pass) {
        const int e0 = plan.ea_level[2 * l + pass], e1 = plan.ea_level[2 * l + pass + 1];
        if (e1 > e0) k_extend_add<<<e1 - e0, 64, 0, st>>>(d_ea, e0, d_ns, d_ms, d_front_off, d_map, d_map_off, fr);
      }
      DFCHECK(hipGetLastError());
      if (trace) DFCHECK(hipEventRecord(ev[2 * l + 1], st));
      for (; gi < groups.size() && groups[gi].level == l; ++gi) {
        const Group& g = groups[gi];
        auto mark = [&](int op) {  // trace: per-call events, op 0..4 = potrf trsm gemm trtri trmm, 5 = end
          if (!trace) return;
          hipEvent_t e;
          DFCHECK(hipEventCreate(&e));
          DFCHECK(hipEventRecord(e, st));
          opev.push_back({op, e});
        };
        const int ns = g.ns, ms = g.ms, nf = ns + ms, cnt = g.count;
        double* const* A11 = d_ptrs + g.ptr0;
        double* const* A12 = A11 + cnt;
        double* const* A22 = A12 + cnt;
        double* const* Wo = A22 + cnt;
        rocblas_int* info = d_info + g.info0;
        if (ns >= kLargeFront && cnt <= kFewFronts) {
          // a few large fronts: one non-batched (blocked, GEMM-rich) call per
          // front runs faster than the batched routines
          for (int f = 0; f < cnt; ++f) {
            double* a11 = ptrs[g.ptr0 + f];
            double* a12 = ptrs[g.ptr0 + cnt + f];
            double* a22 = ptrs[g.ptr0 + 2 * cnt + f];
            double* wo = ptrs[g.ptr0 + 3 * cnt + f];
            mark(0);
            RBCHECK(rocsolver_dpotrf(h, rocblas_fill_upper, ns, a11, nf, info + f));
            if (ms > 0) {
              mark(1);
              RBCHECK(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                    rocblas_diagonal_non_unit, ns, ms, &one, a11, nf, a12, nf));
              mark(2);
              RBCHECK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, ms, ms, ns, &minus_one,
                                    a12, nf, a12, nf, &one, a22, nf));
            }
            mark(3);
            RBCHECK(rocsolver_dtrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, ns, a11, nf, info + cnt + f));
            if (ms > 0) {
              mark(4);
              RBCHECK(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                    rocblas_diagonal_non_unit, ns, ms, &one, a11, nf, a12, nf, wo, ns));
            }
            mark(5);
          }
          continue;
        }
        mark(0);
        RBCHECK(rocsolver_dpotrf_batched(h, rocblas_fill_upper, ns, A11, nf, info, cnt));
        if (ms > 0) {
          mark(1);
          RBCHECK(rocblas_dtrsm_batched(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                        rocblas_diagonal_non_unit, ns, ms, &one, (const double* const*)A11, nf, A12,
                                        nf, cnt));
          mark(2);
          // trailing update A22 -= A12^T A12 as a full GEMM: twice the flops of
          // SYRK, but rocBLAS's GEMM kernels run it several times faster; the
          // extra (strictly lower, column-major) triangle lands in the part of
          // the front that nothing reads
          RBCHECK(rocblas_dgemm_batched(h, rocblas_operation_transpose, rocblas_operation_none, ms, ms, ns,
                                        &minus_one, (const double* const*)A12, nf, (const double* const*)A12, nf,
                                        &one, A22, nf, cnt));
        }
        mark(3);
        RBCHECK(rocsolver_dtrtri_batched(h, rocblas_fill_upper, rocblas_diagonal_non_unit, ns, A11, nf, info + cnt,
                                         cnt));
        if (ms > 0) {
          mark(4);
          RBCHECK(rocblas_dtrmm_batched(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                        rocblas_diagonal_non_unit, ns, ms, &one, (const double* const*)A11, nf,
                                        (const double* const*)A12, nf, Wo, ns, cnt));
        }
        mark(5);
      }
      if (trace) DFCHECK(hipEventRecord(ev[2 * l + 2], st));
    }
    if (!pk.empty())
      k_pack_rows<<<(unsigned)pk.size(), 256, 0, st>>>(d_pk, d_ns, d_ms, d_front_off, d_fwd_off, d_bwd_off, fr, fwd,
                                                       bwd);
    DFCHECK(hipGetLastError());
    if (trace) DFCHECK(hipEventRecord(ev[2 * F.nlevels + 1], st));
    std::vector<rocblas_int> info(ninfo);
    if (ninfo) DFCHECK(hipMemcpyAsync(info.data(), d_info, ninfo * sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
    DFCHECK(hipStreamSynchronize(st));
    for (const Group& g : groups)
      for (int i = 0; i < 2 * g.count; ++i)
        if (info[g.info0 + i] != 0)
          throw std::runtime_error("device factorisation: front of shape (" + std::to_string(g.ns) + ", " +
                                   std::to_string(g.ms) + ") at level " + std::to_string(g.level) +
                                   (i < g.count ? " is not positive definite" : " has a singular factor"));
    if (trace) {
      float a = 0, d = 0, ms = 0;
      size_t g2 = 0;
      for (int l = 0; l < F.nlevels; ++l) {
        DFCHECK(hipEventElapsedTime(&a, ev[2 * l], ev[2 * l + 1]));
        DFCHECK(hipEventElapsedTime(&d, ev[2 * l + 1], ev[2 * l + 2]));
        int ng = 0, mxn = 0, mxm = 0;
        for (; g2 < groups.size() && groups[g2].level == l; ++g2, ++ng) {
          mxn = std::max(mxn, groups[g2].ns);
          mxm = std::max(mxm, groups[g2].ms);
        }
        std::fprintf(stderr, "[factor] level %2d: %5zu fronts, %3d shapes (max ns %d, ms %d): assembly %.2f ms, dense %.2f ms\n",
                     l, F.level_nodes[l].size(), ng, mxn, mxm, a, d);
      }
      double opms[5] = {0, 0, 0, 0, 0};
      for (size_t i = 0; i + 1 < opev.size(); ++i) {
        if (opev[i].first == 5) continue;
        float t = 0;
        DFCHECK(hipEventElapsedTime(&t, opev[i].second, opev[i + 1].second));
        opms[opev[i].first] += t;
      }
      std::fprintf(stderr, "[factor] dense calls: potrf %.1f ms, trsm %.1f ms, gemm %.1f ms, trtri %.1f ms, trmm %.1f ms\n",
                   opms[0], opms[1], opms[2], opms[3], opms[4]);
      for (auto& pe : opev) (void)hipEventDestroy(pe.second);
      opev.clear();
      DFCHECK(hipEventElapsedTime(&ms, ev[2 * F.nlevels], ev[2 * F.nlevels + 1]));
      std::fprintf(stderr, "[factor] sweep layout %.2f ms; fronts %.2f GB; %.3g dense flops\n", ms,
                   stats.front_bytes * 1e-9, stats.flops);
      for (auto& e : ev) (void)hipEventDestroy(e);
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (auto& pe : opev) (void)hipEventDestroy(pe.second);
    throw;
  }
  stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return stats;
}

void blocked_on_device(const PackedFactor& P, const double* fwd, const double* bwd, double* fblk, double* bblk,
                       hipStream_t st) {
  Pool pool(st);
  const int nn = (int)P.p0.size();
  int* d_ns = pool.up(P.ns, st);
  int* d_ms = pool.up(P.ms, st);
  for (int dir = 0; dir < 2; ++dir) {
    const bool fw = dir == 0;
    const std::vector<int>& b0 = fw ? P.fblk0 : P.bblk0;
    const std::vector<int64_t>& boff = fw ? P.fblk_off : P.bblk_off;
    const int nblk = b0[nn];
    if (nblk == 0) continue;
    std::vector<int> node(nblk);
    for (int s = 0; s < nn; ++s)
      for (int b = b0[s]; b < b0[s + 1]; ++b) node[b] = s;
    int* d_node = pool.up(node, st);
    int* d_b0 = pool.up(b0, st);
    int64_t* d_boff = pool.up(boff, st);
    int64_t* d_src_off = pool.up(fw ? P.fwd_off : P.bwd_off, st);
    k_pack_blocked<<<nblk, 256, 0, st>>>(fw ? 1 : 0, d_node, d_b0, d_boff, d_ns, d_ms, d_src_off, fw ? fwd : bwd,
                                         fw ? fblk : bblk);
    DFCHECK(hipGetLastError());
    DFCHECK(hipStreamSynchronize(st));  // node is a host vector of this iteration
  }
}

}  // namespace pecs
