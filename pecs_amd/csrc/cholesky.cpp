// Supernodal multifrontal Cholesky with geometric nested dissection (setup).
#include "cholesky.hpp"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>

namespace pecs {

namespace {

struct NDNode {
  std::vector<int> cells;  // grid cell ids
  std::vector<int> children;
  int parent = -1;
};

struct NDBuilder {
  const BlockSPD& S;
  int leaf;
  std::vector<NDNode> nodes;  // postorder
  NDBuilder(const BlockSPD& s, int l) : S(s), leaf(l) {}

  // Width of a separator starting at column (row) `mid` of the rectangle
  // that leaves no coupling between the part before it and the part after
  // it: the farthest reach past mid of any coupling from a cell before mid.
  int reach(int i0, int i1, int j0, int j1, int mid, bool vertical) const {
    const int no = S.noff();
    int maxo = 0;
    for (int o = 0; o < no; ++o) maxo = std::max(maxo, std::abs(vertical ? S.offi[o] : S.offj[o]));
    int w = 1;
    const int lo = std::max(vertical ? i0 : j0, mid - maxo);
    for (int t = lo; t < mid; ++t) {
      const int u0 = vertical ? j0 : i0, u1 = vertical ? j1 : i1;
      for (int u = u0; u < u1; ++u) {
        const int xi = vertical ? t : u, xj = vertical ? u : t;
        const int X = xi + S.NXc * xj;
        if (S.ndof[X] == 0) continue;
        for (int o = 0; o < no; ++o) {
          const int yi = xi + S.offi[o], yj = xj + S.offj[o];
          if (yi < i0 || yi >= i1 || yj < j0 || yj >= j1) continue;
          if (S.blk[(size_t)X * no + o].empty() || S.ndof[yi + S.NXc * yj] == 0) continue;
          const int p = vertical ? yi : yj;
          if (p >= mid) w = std::max(w, p - mid + 1);
        }
      }
    }
    return w;
  }

  int build(int i0, int i1, int j0, int j1) {
    const int ni = i1 - i0, nj = j1 - j0, w0 = S.sep_width > 0 ? S.sep_width : 2;
    if (ni <= 0 || nj <= 0) return -1;
    if (ni * nj <= leaf || (ni <= w0 && nj <= w0)) {
      NDNode n;
      for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) n.cells.push_back(i + S.NXc * j);
      nodes.push_back(n);
      return (int)nodes.size() - 1;
    }
    std::vector<int> kids;
    NDNode sep;
    if (ni >= nj && ni > w0) {
      const int mid = i0 + (ni - w0) / 2;
      const int w = S.sep_width > 0 ? w0 : reach(i0, i1, j0, j1, mid, true);
      int a = build(i0, mid, j0, j1);
      int b = build(std::min(mid + w, i1), i1, j0, j1);
      if (a >= 0) kids.push_back(a);
      if (b >= 0) kids.push_back(b);
      for (int j = j0; j < j1; ++j)
        for (int i = mid; i < std::min(mid + w, i1); ++i) sep.cells.push_back(i + S.NXc * j);
    } else {
      const int mid = j0 + (nj - w0) / 2;
      const int w = S.sep_width > 0 ? w0 : reach(i0, i1, j0, j1, mid, false);
      int a = build(i0, i1, j0, mid);
      int b = build(i0, i1, std::min(mid + w, j1), j1);
      if (a >= 0) kids.push_back(a);
      if (b >= 0) kids.push_back(b);
      for (int j = mid; j < std::min(mid + w, j1); ++j)
        for (int i = i0; i < i1; ++i) sep.cells.push_back(i + S.NXc * j);
    }
    sep.children = kids;
    nodes.push_back(sep);
    const int id = (int)nodes.size() - 1;
    for (int k : kids) nodes[k].parent = id;
    return id;
  }
};

// F[i][j] -= sum_t P[i][t] P[j][t] for k1 <= j <= i < nf, P = F[:, k0:k0+pw].
// Register-blocked 4 x 8 micro-kernel over a packed copy of the panel and of
// its transpose.
void trailing_update(double* F, int nf, int k0, int k1, int pw) {
  const int nt = nf - k1;
  const int ntp = (nt + 7) & ~7;
  std::vector<double> P((size_t)(nt + 4) * pw, 0.0), PT((size_t)pw * (ntp + 8), 0.0);
  const int ld = ntp + 8;
#pragma omp parallel for schedule(static) if (nt > 256)
  for (int i = 0; i < nt; ++i)
    for (int t = 0; t < pw; ++t) {
      const double v = F[(size_t)(k1 + i) * nf + k0 + t];
      P[(size_t)i * pw + t] = v;
      PT[(size_t)t * ld + i] = v;
    }
  const int nib = (nt + 3) / 4;
#pragma omp parallel for schedule(dynamic, 4) if (nt > 64)
  for (int ib = nib - 1; ib >= 0; --ib) {
    const int i0 = 4 * ib;
    const double* a0 = &P[(size_t)(i0 + 0) * pw];
    const double* a1 = &P[(size_t)(i0 + 1) * pw];
    const double* a2 = &P[(size_t)(i0 + 2) * pw];
    const double* a3 = &P[(size_t)(i0 + 3) * pw];
    for (int j0 = 0; j0 <= i0 + 3 && j0 < nt; j0 += 8) {
      double acc[4][8];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) acc[r][c] = 0.0;
      const double* bt = &PT[j0];
      for (int t = 0; t < pw; ++t) {
        const double* b = bt + (size_t)t * ld;
        const double x0 = a0[t], x1 = a1[t], x2 = a2[t], x3 = a3[t];
#pragma omp simd
        for (int c = 0; c < 8; ++c) {
          acc[0][c] += x0 * b[c];
          acc[1][c] += x1 * b[c];
          acc[2][c] += x2 * b[c];
          acc[3][c] += x3 * b[c];
        }
      }
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + r;
        if (i >= nt) break;
        double* Fi = &F[(size_t)(k1 + i) * nf + k1];
        const int jmax = std::min(j0 + 8, i + 1);
        for (int j = j0; j < jmax; ++j) Fi[j] -= acc[r][j - j0];
      }
    }
  }
}

// Partial Cholesky of the leading ns columns of the lower triangle of the
// nf x nf row-major front F (upper triangle ignored).  On return F holds L11,
// L21 and the Schur update in its lower triangle.
void partial_cholesky(double* F, int nf, int ns, double& flops) {
  const int BS = 64;
  for (int k0 = 0; k0 < ns; k0 += BS) {
    const int k1 = std::min(ns, k0 + BS);
    // factor diagonal block and the panel below it (rows k0..nf-1, cols k0..k1-1)
    for (int k = k0; k < k1; ++k) {
      double d = F[(size_t)k * nf + k];
      for (int t = k0; t < k; ++t) d -= F[(size_t)k * nf + t] * F[(size_t)k * nf + t];
      if (!(d > 0.0)) throw std::runtime_error("Cholesky: matrix not positive definite (pivot " + std::to_string(d) + ")");
      d = std::sqrt(d);
      F[(size_t)k * nf + k] = d;
      const double inv = 1.0 / d;
#pragma omp parallel for schedule(static) if (nf - k > 2048)
      for (int i = k + 1; i < nf; ++i) {
        double s = F[(size_t)i * nf + k];
        const double* Li = &F[(size_t)i * nf];
        const double* Lk = &F[(size_t)k * nf];
        for (int t = k0; t < k; ++t) s -= Li[t] * Lk[t];
        F[(size_t)i * nf + k] = s * inv;
      }
    }
    // trailing update of rows/cols >= k1 with the panel columns [k0, k1)
    const int nt = nf - k1;
    if (nt <= 0) continue;
    const int pw = k1 - k0;
    flops += (double)nt * nt * pw;
    trailing_update(F, nf, k0, k1, pw);
  }
}

// Inverse of a lower-triangular n x n block stored with leading dimension ld.
// Blocked: each column block J of X = L^{-1} is independent; row block I of it
// is L_II^{-1} (delta_IJ - sum_{J<=K<I} L_IK X_KJ), computed with contiguous
// row segments.
void lower_inverse(const double* L, int n, int ld, double* X) {
  const int B = 32;
  const int nbk = (n + B - 1) / B;
#pragma omp parallel for schedule(static) if (n > 256)
  for (size_t t = 0; t < (size_t)n * n; ++t) X[t] = 0.0;
#pragma omp parallel for schedule(dynamic, 1) if (n > 128)
  for (int jb = 0; jb < nbk; ++jb) {
    const int j0 = jb * B, j1 = std::min(n, j0 + B), w = j1 - j0;
    double T[32];
    for (int i = j0; i < n; ++i) {
      // T = delta_i - sum_{t in [j0, i)} L[i][t] X[t][j0:j1]
      for (int c = 0; c < w; ++c) T[c] = (i == j0 + c) ? 1.0 : 0.0;
      const double* Li = &L[(size_t)i * ld];
      for (int t = j0; t < i; ++t) {
        const double l = Li[t];
        if (l == 0.0) continue;
        const double* Xt = &X[(size_t)t * n + j0];
#pragma omp simd
        for (int c = 0; c < 32; ++c)
          if (c < w) T[c] -= l * Xt[c];
      }
      const double inv = 1.0 / Li[i];
      double* Xi = &X[(size_t)i * n + j0];
      for (int c = 0; c < w; ++c) Xi[c] = (j0 + c <= i) ? T[c] * inv : 0.0;
    }
  }
}

}  // namespace

namespace {

// Ordering, supernodes, row structure and levels (no numeric values).
struct Symbolic {
  SupernodalFactor F;
  std::vector<NDNode> nodes;
  std::vector<int> pos, node_pos0, node_pos1;
  std::vector<std::vector<int>> structc;
};

Symbolic symbolic(const BlockSPD& S, int leaf_cells) {
  const int NG = S.NXc * S.NYc;
  const int no = S.noff();
  NDBuilder nd(S, leaf_cells);
  nd.build(0, S.NXc, 0, S.NYc);
  const int nn = (int)nd.nodes.size();
  Symbolic Y;
  SupernodalFactor& F = Y.F;
  // cell positions and DoF bases
  std::vector<int> pos(NG, -1), node_of(NG, -1);
  std::vector<int> node_pos0(nn), node_pos1(nn);
  F.cell_base.assign(NG, 0);
  int p = 0, d = 0;
  F.nodes.resize(nn);
  for (int s = 0; s < nn; ++s) {
    node_pos0[s] = p;
    F.nodes[s].p0 = d;
    for (int c : nd.nodes[s].cells) {
      pos[c] = p++;
      node_of[c] = s;
      F.cell_base[c] = d;
      d += S.ndof[c];
    }
    node_pos1[s] = p;
    F.nodes[s].ns = d - F.nodes[s].p0;
    F.nodes[s].parent = nd.nodes[s].parent;
  }
  F.n = d;
  // symbolic: struct cells per node (sorted by position)
  std::vector<std::vector<int>> structc(nn);
  std::vector<int> mark(NG, -1);
  for (int s = 0; s < nn; ++s) {
    std::vector<int> st;
    auto consider = [&](int c) {
      if (pos[c] >= node_pos1[s] && mark[c] != s) {
        mark[c] = s;
        st.push_back(c);
      }
    };
    for (int X : nd.nodes[s].cells) {
      const int xi = X % S.NXc, xj = X / S.NXc;
      for (int o = 0; o < no; ++o) {
        const int yi = xi + S.offi[o], yj = xj + S.offj[o];
        if (yi < 0 || yi >= S.NXc || yj < 0 || yj >= S.NYc) continue;
        const int Y = yi + S.NXc * yj;
        if (S.blk[(size_t)X * no + o].empty() || S.ndof[Y] == 0) continue;
        consider(Y);
      }
    }
    for (int c : nd.nodes[s].children)
      for (int Y : structc[c]) consider(Y);
    std::sort(st.begin(), st.end(), [&](int a, int b) { return pos[a] < pos[b]; });
    structc[s] = st;
  }
  // levels (height in the elimination tree)
  for (int s = 0; s < nn; ++s) {
    int lv = 0;
    for (int c : nd.nodes[s].children) lv = std::max(lv, F.nodes[c].level + 1);
    F.nodes[s].level = lv;
    F.nlevels = std::max(F.nlevels, lv + 1);
  }
  F.level_nodes.assign(F.nlevels, {});
  for (int s = 0; s < nn; ++s) F.level_nodes[F.nodes[s].level].push_back(s);
  // offsets
  int64_t linv = 0, l21 = 0, rows = 0;
  for (int s = 0; s < nn; ++s) {
    SNode& N = F.nodes[s];
    int ms = 0;
    for (int c : structc[s]) ms += S.ndof[c];
    N.ms = ms;
    N.linv_off = linv;
    N.l21_off = l21;
    N.rows_off = rows;
    linv += (int64_t)N.ns * N.ns;
    l21 += (int64_t)ms * N.ns;
    rows += ms;
  }
  F.rows.assign(rows, 0);
  for (int s = 0; s < nn; ++s) {
    int t = 0;
    for (int c : structc[s])
      for (int a = 0; a < S.ndof[c]; ++a) F.rows[F.nodes[s].rows_off + t++] = F.cell_base[c] + a;
  }
  F.linv_size = linv;
  F.l21_size = l21;
  // forward contribution lists
  std::vector<int64_t> cnt(F.n + 1, 0);
  for (int s = 0; s < nn; ++s)
    for (int r = 0; r < F.nodes[s].ms; ++r) cnt[F.rows[F.nodes[s].rows_off + r] + 1]++;
  for (int i = 0; i < F.n; ++i) cnt[i + 1] += cnt[i];
  F.contrib_ptr = cnt;
  F.contrib_node.assign(cnt[F.n], 0);
  F.contrib_row.assign(cnt[F.n], 0);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (int s = 0; s < nn; ++s)
    for (int r = 0; r < F.nodes[s].ms; ++r) {
      const int pp = F.rows[F.nodes[s].rows_off + r];
      F.contrib_node[fill[pp]] = s;
      F.contrib_row[fill[pp]] = r;
      fill[pp]++;
    }
  Y.nodes = nd.nodes;
  Y.pos = pos;
  Y.node_pos0 = node_pos0;
  Y.node_pos1 = node_pos1;
  Y.structc = structc;
  return Y;
}

}  // namespace

SupernodalFactor cholesky_factor(const BlockSPD& S, int leaf_cells) {
  const int NG = S.NXc * S.NYc;
  const int no = S.noff();
  Symbolic Y = symbolic(S, leaf_cells);
  SupernodalFactor& F = Y.F;
  const int nn = (int)F.nodes.size();
  const auto& pos = Y.pos;
  const auto& node_pos0 = Y.node_pos0;
  const auto& node_pos1 = Y.node_pos1;
  const auto& structc = Y.structc;
  struct {
    const std::vector<NDNode>& nodes;
  } nd{Y.nodes};
  F.Linv.assign(F.linv_size, 0.0);
  F.L21.assign(F.l21_size, 0.0);
  // numeric multifrontal factorisation, level by level on the elimination tree:
  // nodes of one level are independent (their children are in lower levels).
  std::vector<std::vector<double>> upd(nn);
  const int nthreads = omp_get_max_threads();
  std::vector<std::vector<int>> front_off_t(nthreads, std::vector<int>(NG, -1));
  std::vector<double> flops_t(nthreads, 0.0), ta_t(nthreads, 0.0), tf_t(nthreads, 0.0), ti_t(nthreads, 0.0);
  auto process = [&](int s, int tid) {
    std::vector<int>& front_off = front_off_t[tid];
    const SNode& N = F.nodes[s];
    const int ns = N.ns, ms = N.ms, nf = ns + ms;
    const bool big = nf > 1024;
    auto tA = std::chrono::steady_clock::now();
    std::vector<double> Fr((size_t)nf * nf);
#pragma omp parallel for schedule(static) if (big)
    for (int r = 0; r < nf; ++r) std::fill(&Fr[(size_t)r * nf], &Fr[(size_t)r * nf] + r + 1, 0.0);
    int off = 0;
    for (int c : nd.nodes[s].cells) {
      front_off[c] = off;
      off += S.ndof[c];
    }
    for (int c : structc[s]) {
      front_off[c] = off;
      off += S.ndof[c];
    }
    // original entries: each symmetric block pair once, by the earlier cell's node
    for (int X : nd.nodes[s].cells) {
      const int nx = S.ndof[X];
      if (nx == 0) continue;
      const int xi = X % S.NXc, xj = X / S.NXc;
      const int rx = front_off[X];
      for (int o = 0; o < no; ++o) {
        const int yi = xi + S.offi[o], yj = xj + S.offj[o];
        if (yi < 0 || yi >= S.NXc || yj < 0 || yj >= S.NYc) continue;
        const int Y = yi + S.NXc * yj;
        const auto& B = S.blk[(size_t)X * no + o];
        if (B.empty() || S.ndof[Y] == 0) continue;
        const int ny = S.ndof[Y];
        if (pos[Y] < node_pos0[s]) continue;
        const int ry = front_off[Y];
        if (pos[Y] < node_pos1[s]) {
          for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
              if (rx + a >= ry + b) Fr[(size_t)(rx + a) * nf + ry + b] += B[(size_t)a * ny + b];
        } else {
          for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b) Fr[(size_t)(ry + b) * nf + rx + a] += B[(size_t)a * ny + b];
        }
      }
    }
    // extend-add children
    for (int c : nd.nodes[s].children) {
      const int mc = F.nodes[c].ms;
      std::vector<int> map(mc);
      int t = 0;
      for (int Y : structc[c])
        for (int a = 0; a < S.ndof[Y]; ++a) map[t++] = front_off[Y] + a;
      const std::vector<double>& U = upd[c];
#pragma omp parallel for schedule(dynamic, 64) if (big)
      for (int r = 0; r < mc; ++r) {
        double* row = &Fr[(size_t)map[r] * nf];
        const double* ur = &U[(size_t)r * mc];
        for (int q = 0; q <= r; ++q) row[map[q]] += ur[q];
      }
      std::vector<double>().swap(upd[c]);
    }
    for (int c : nd.nodes[s].cells) front_off[c] = -1;
    for (int c : structc[s]) front_off[c] = -1;
    auto tB = std::chrono::steady_clock::now();
    if (ns > 0) partial_cholesky(Fr.data(), nf, ns, flops_t[tid]);
    flops_t[tid] += (double)ns * ns * ns / 3.0;
    auto tC = std::chrono::steady_clock::now();
    // store Linv, L21 and the update matrix
    if (ns > 0) lower_inverse(Fr.data(), ns, nf, &F.Linv[N.linv_off]);
    auto tD = std::chrono::steady_clock::now();
    ta_t[tid] += std::chrono::duration<double>(tB - tA).count();
    tf_t[tid] += std::chrono::duration<double>(tC - tB).count();
    ti_t[tid] += std::chrono::duration<double>(tD - tC).count();
#pragma omp parallel for schedule(static) if (big)
    for (int r = 0; r < ms; ++r)
      for (int c2 = 0; c2 < ns; ++c2) F.L21[N.l21_off + (int64_t)r * ns + c2] = Fr[(size_t)(ns + r) * nf + c2];
    if (ms > 0) {
      upd[s].resize((size_t)ms * ms);
#pragma omp parallel for schedule(static) if (big)
      for (int r = 0; r < ms; ++r)
        for (int q = 0; q <= r; ++q) upd[s][(size_t)r * ms + q] = Fr[(size_t)(ns + r) * nf + ns + q];
    }
  };
  for (int l = 0; l < F.nlevels; ++l) {
    const std::vector<int>& lv = F.level_nodes[l];
    if ((int)lv.size() >= 2 * nthreads) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int i = 0; i < (int)lv.size(); ++i) process(lv[i], omp_get_thread_num());
    } else {
      for (int s : lv) process(s, 0);
    }
  }
  for (int t = 0; t < nthreads; ++t) {
    F.flops += flops_t[t];
    F.t_assemble += ta_t[t];
    F.t_factor += tf_t[t];
    F.t_inverse += ti_t[t];
  }
  return F;
}

SupernodalFactor cholesky_symbolic(const BlockSPD& S, int leaf_cells, FrontPlan& plan) {
  const int no = S.noff();
  Symbolic Y = symbolic(S, leaf_cells);
  SupernodalFactor& F = Y.F;
  const int nn = (int)F.nodes.size();
  const int NG = S.NXc * S.NYc;
  plan = FrontPlan{};
  plan.nlevels = F.nlevels;
  // S blocks, concatenated
  plan.sblk_off.assign((size_t)NG * no, -1);
  int64_t sv = 0;
  for (int X = 0; X < NG; ++X)
    for (int o = 0; o < no; ++o) {
      const auto& B = S.blk[(size_t)X * no + o];
      if (B.empty()) continue;
      plan.sblk_off[(size_t)X * no + o] = sv;
      sv += (int64_t)B.size();
    }
  plan.svals.resize(sv);
  for (int X = 0; X < NG; ++X)
    for (int o = 0; o < no; ++o) {
      const int64_t off = plan.sblk_off[(size_t)X * no + o];
      if (off >= 0) std::copy(S.blk[(size_t)X * no + o].begin(), S.blk[(size_t)X * no + o].end(), &plan.svals[off]);
    }
  // fronts, level by level (each level's fronts contiguous)
  plan.front_off.assign(nn, 0);
  plan.level_front.assign(F.nlevels + 1, 0);
  int64_t fo = 0;
  for (int l = 0; l < F.nlevels; ++l) {
    plan.level_front[l] = fo;
    for (int s : F.level_nodes[l]) {
      const int64_t nf = F.nodes[s].ns + F.nodes[s].ms;
      plan.front_off[s] = fo;
      fo += nf * nf;
    }
  }
  plan.level_front[F.nlevels] = fo;
  plan.front_total = fo;
  // local front position of each cell of a node's front (own cells, then struct)
  std::vector<int> loc(NG, -1);
  plan.sc_level.assign(F.nlevels + 1, 0);
  plan.ea_level.assign(2 * F.nlevels + 1, 0);
  plan.ea_map_off.assign(nn + 1, 0);
  // child row maps into the parent's front, per child (ms entries)
  std::vector<std::vector<int>> cmap(nn);
  for (int l = 0; l < F.nlevels; ++l) {
    plan.sc_level[l] = (int)plan.sc_src.size();
    for (int s : F.level_nodes[l]) {
      const int nf = F.nodes[s].ns + F.nodes[s].ms;
      int off = 0;
      for (int c : Y.nodes[s].cells) {
        loc[c] = off;
        off += S.ndof[c];
      }
      for (int c : Y.structc[s]) {
        loc[c] = off;
        off += S.ndof[c];
      }
      for (int X : Y.nodes[s].cells) {
        const int nx = S.ndof[X];
        if (nx == 0) continue;
        const int xi = X % S.NXc, xj = X / S.NXc;
        for (int o = 0; o < no; ++o) {
          const int yi = xi + S.offi[o], yj = xj + S.offj[o];
          if (yi < 0 || yi >= S.NXc || yj < 0 || yj >= S.NYc) continue;
          const int Yc = yi + S.NXc * yj;
          const int64_t so = plan.sblk_off[(size_t)X * no + o];
          if (so < 0 || S.ndof[Yc] == 0) continue;
          if (Y.pos[Yc] < Y.node_pos0[s]) continue;
          const bool own = Y.pos[Yc] < Y.node_pos1[s];
          const int rx = loc[X], ry = loc[Yc];
          plan.sc_src.push_back(so);
          plan.sc_dst.push_back(plan.front_off[s] + (own ? (int64_t)rx * nf + ry : (int64_t)ry * nf + rx));
          plan.sc_nx.push_back(nx);
          plan.sc_ny.push_back(S.ndof[Yc]);
          plan.sc_nf.push_back(nf);
          plan.sc_mode.push_back(own ? 0 : 1);
          plan.sc_rx.push_back(rx);
          plan.sc_ry.push_back(ry);
        }
      }
      for (int c : Y.nodes[s].children) {
        std::vector<int>& m = cmap[c];
        m.clear();
        for (int Yc : Y.structc[c])
          for (int a = 0; a < S.ndof[Yc]; ++a) m.push_back(loc[Yc] + a);
      }
      for (int c : Y.nodes[s].cells) loc[c] = -1;
      for (int c : Y.structc[s]) loc[c] = -1;
    }
  }
  plan.sc_level[F.nlevels] = (int)plan.sc_src.size();
  for (int s = 0; s < nn; ++s) plan.ea_map_off[s + 1] = plan.ea_map_off[s] + (int64_t)cmap[s].size();
  plan.ea_map.resize(plan.ea_map_off[nn]);
  for (int s = 0; s < nn; ++s) std::copy(cmap[s].begin(), cmap[s].end(), &plan.ea_map[plan.ea_map_off[s]]);
  // extend-add rows, per parent level and pass (child 0, then child 1 of every parent)
  for (int l = 0; l < F.nlevels; ++l)
    for (int pass = 0; pass < 2; ++pass) {
      plan.ea_level[2 * l + pass] = (int)plan.ea_child.size();
      for (int s : F.level_nodes[l]) {
        if ((int)Y.nodes[s].children.size() <= pass) continue;
        const int c = Y.nodes[s].children[pass];
        for (int r = 0; r < F.nodes[c].ms; ++r) {
          plan.ea_child.push_back(c);
          plan.ea_row.push_back(r);
          plan.ea_parent.push_back(s);
        }
      }
    }
  plan.ea_level[2 * F.nlevels] = (int)plan.ea_child.size();
  for (int s = 0; s < nn; ++s)
    if (Y.nodes[s].children.size() > 2) throw std::runtime_error("cholesky_symbolic: more than two children");
  return F;
}

void execute_plan_host(const FrontPlan& plan, const SupernodalFactor& F, std::vector<double>& Linv,
                       std::vector<double>& L21) {
  const int nn = (int)F.nodes.size();
  std::vector<double> fr(plan.front_total, 0.0);
  Linv.assign(F.linv_size, 0.0);
  L21.assign(F.l21_size, 0.0);
  double flops = 0.0;
  for (int l = 0; l < F.nlevels; ++l) {
    for (int i = plan.sc_level[l]; i < plan.sc_level[l + 1]; ++i) {
      const int nx = plan.sc_nx[i], ny = plan.sc_ny[i], nf = plan.sc_nf[i];
      for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) {
          const double v = plan.svals.at(plan.sc_src[i] + (int64_t)a * ny + b);
          if (plan.sc_mode[i] == 0) {
            if (plan.sc_rx[i] + a < plan.sc_ry[i] + b) continue;
            fr.at(plan.sc_dst[i] + (int64_t)a * nf + b) += v;
          } else {
            fr.at(plan.sc_dst[i] + (int64_t)b * nf + a) += v;
          }
        }
    }
    for (int pass = 0; pass < 2; ++pass)
      for (int i = plan.ea_level[2 * l + pass]; i < plan.ea_level[2 * l + pass + 1]; ++i) {
        const int c = plan.ea_child[i], r = plan.ea_row[i], s = plan.ea_parent[i];
        const int nsc = F.nodes[c].ns, nfc = nsc + F.nodes[c].ms, nfp = F.nodes[s].ns + F.nodes[s].ms;
        const int* m = &plan.ea_map[plan.ea_map_off[c]];
        for (int q = 0; q <= r; ++q)
          fr.at(plan.front_off[s] + (int64_t)m[r] * nfp + m[q]) +=
              fr.at(plan.front_off[c] + (int64_t)(nsc + r) * nfc + nsc + q);
      }
    for (int s : F.level_nodes[l]) {
      const SNode& N = F.nodes[s];
      const int nf = N.ns + N.ms;
      double* Fr = &fr[plan.front_off[s]];
      if (N.ns > 0) {
        partial_cholesky(Fr, nf, N.ns, flops);
        lower_inverse(Fr, N.ns, nf, &Linv[N.linv_off]);
      }
      for (int r = 0; r < N.ms; ++r)
        for (int c2 = 0; c2 < N.ns; ++c2) L21[N.l21_off + (int64_t)r * N.ns + c2] = Fr[(size_t)(N.ns + r) * nf + c2];
    }
  }
  (void)nn;
}

PackedFactor pack_layout(const SupernodalFactor& F) {
  PackedFactor P;
  const int nn = (int)F.nodes.size();
  P.n = F.n;
  P.nlevels = F.nlevels;
  P.cell_base = F.cell_base;
  P.p0.resize(nn);
  P.ns.resize(nn);
  P.ms.resize(nn);
  P.fwd_off.resize(nn);
  P.bwd_off.resize(nn);
  P.rows_off.resize(nn);
  P.z_off.resize(nn);
  int64_t nf = 0, zl = 0;
  for (int s = 0; s < nn; ++s) {
    P.p0[s] = F.nodes[s].p0;
    P.ns[s] = F.nodes[s].ns;
    P.ms[s] = F.nodes[s].ms;
    P.rows_off[s] = F.nodes[s].rows_off;
    const int64_t n_ = P.ns[s], m_ = P.ms[s];
    P.fwd_off[s] = nf;
    P.bwd_off[s] = nf;
    P.z_off[s] = zl;
    nf += n_ * (n_ + 1) / 2 + m_ * n_;
    zl += m_;
    P.tri_bytes += 2.0 * 8.0 * (0.5 * n_ * (n_ + 1.0) + (double)m_ * n_);
  }
  P.zlen = zl;
  P.nvals = nf;
  // lanes per row: the smallest power of two >= an eighth of the mean row length
  auto lanes = [](double mean) {
    int w = 8;
    while (w < 64 && w * 8 < mean) w *= 2;
    return w;
  };
  P.goff.assign(F.nlevels + 1, 0);
  P.xoff.assign(F.nlevels + 1, 0);
  P.fw.assign(F.nlevels, 64);
  P.bw.assign(F.nlevels, 64);
  P.fm.assign(F.nlevels, 1);
  P.bm.assign(F.nlevels, 1);
  P.fpf.assign(F.nlevels, 0);
  P.bpf.assign(F.nlevels, 0);
  P.frows.assign(F.nlevels, 4);
  P.brows.assign(F.nlevels, 4);
  P.lnoff.assign(F.nlevels + 1, 0);
  for (int l = 0; l < F.nlevels; ++l) {
    double fl = 0, fr = 0, bl = 0, br = 0;
    for (int s : F.level_nodes[l]) {
      const double n_ = P.ns[s], m_ = P.ms[s];
      fl += 0.5 * n_ * (n_ + 1) + m_ * n_;
      fr += n_ + m_;
      bl += 0.5 * n_ * (n_ + 1) + m_ * n_;
      br += n_;
    }
    P.fw[l] = lanes(fr > 0 ? fl / fr : 1);
    P.bw[l] = lanes(br > 0 ? bl / br : 1);
    // tile height: at most sweep_tile_rows(W), smaller on levels with few rows
    // so that each level still spreads over many workgroups
    auto height = [](int w, double nrows) {
      int h = sweep_tile_rows(w);
      const int pass = 256 / w;
      while (h > pass && nrows / h < 8192) h /= 2;
      return h;
    };
    P.frows[l] = height(P.fw[l], fr);
    P.brows[l] = height(P.bw[l], br);
    P.fm[l] = std::max(1, P.frows[l] / (256 / P.fw[l]));
    P.bm[l] = std::max(1, P.brows[l] / (256 / P.bw[l]));
    P.goff[l] = (int)P.lrows.size();
    P.xoff[l] = (int)P.xidx.size();
    P.lnoff[l] = (int)P.lnodes.size();
    for (int s : F.level_nodes[l]) {
      P.lnodes.push_back(s);
      for (int r = 0; r < P.ns[s]; ++r) P.lrows.push_back(P.p0[s] + r);
      for (int i = 0; i < P.ms[s]; ++i) P.xidx.push_back(P.rows_off[s] + i);
    }
  }
  P.goff[F.nlevels] = (int)P.lrows.size();
  P.xoff[F.nlevels] = (int)P.xidx.size();
  P.lnoff[F.nlevels] = (int)P.lnodes.size();
  rebuild_tiles(P);
  P.rows = F.rows;
  P.cptr = F.contrib_ptr;
  // the solves read the leaf level's right-hand side without a gather
  for (int q = P.goff[0]; q < P.goff[1]; ++q)
    if (P.cptr[P.lrows[q] + 1] != P.cptr[P.lrows[q]])
      throw std::runtime_error("pack_layout: a leaf-level row is owed contributions");
  P.czidx.resize(F.contrib_node.size());
  for (size_t t = 0; t < P.czidx.size(); ++t) P.czidx[t] = P.z_off[F.contrib_node[t]] + F.contrib_row[t];
  return P;
}

void pack_values(const SupernodalFactor& F, PackedFactor& P) {
  const int nn = (int)P.p0.size();
  P.fwd.assign(P.nvals, 0.0);
  P.bwd.assign(P.nvals, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nn; ++s) {
    const int n_ = P.ns[s], m_ = P.ms[s];
    const double* Li = &F.Linv[F.nodes[s].linv_off];
    const double* L21 = &F.L21[F.nodes[s].l21_off];
    double* f = &P.fwd[P.fwd_off[s]];
    for (int r = 0; r < n_; ++r)
      for (int j = 0; j <= r; ++j) f[(int64_t)r * (r + 1) / 2 + j] = Li[(int64_t)r * n_ + j];
    // W21 = L21 Linv, 8 panel rows at a time so each Linv row is reused
    double* W = f + (int64_t)n_ * (n_ + 1) / 2;
    for (int i0 = 0; i0 < m_; i0 += 8) {
      const int i1 = std::min(m_, i0 + 8);
      for (int k = 0; k < n_; ++k) {
        const double* Lk = Li + (int64_t)k * n_;
        for (int i = i0; i < i1; ++i) {
          const double a = L21[(int64_t)i * n_ + k];
          if (a == 0.0) continue;
          double* Wi = W + (int64_t)i * n_;
#pragma omp simd
          for (int j = 0; j <= k; ++j) Wi[j] += a * Lk[j];
        }
      }
    }
    double* bw = &P.bwd[P.bwd_off[s]];
    for (int j = 0; j < n_; ++j) {
      double* row = bw + (int64_t)j * (n_ + m_) - (int64_t)j * (j - 1) / 2;
      for (int i = j; i < n_; ++i) row[i - j] = Li[(int64_t)i * n_ + j];
      for (int i = 0; i < m_; ++i) row[n_ - j + i] = -W[(int64_t)i * n_ + j];
    }
  }
}

PackedFactor pack_factor(const SupernodalFactor& F) {
  PackedFactor P = pack_layout(F);
  pack_values(F, P);
  return P;
}

void blocked_layout(PackedFactor& P) {
  const int nn = (int)P.p0.size();
  for (int dir = 0; dir < 2; ++dir) {
    const bool fwd = dir == 0;
    std::vector<int>& b0 = fwd ? P.fblk0 : P.bblk0;
    std::vector<int64_t>& boff = fwd ? P.fblk_off : P.bblk_off;
    b0.assign(nn + 1, 0);
    boff.clear();
    int64_t total = 0;
    for (int s = 0; s < nn; ++s) {
      b0[s] = (int)boff.size();
      const int rows = fwd ? P.ns[s] + P.ms[s] : P.ns[s];
      for (int b = 0; b * 16 < rows; ++b) {
        boff.push_back(total);
        const int nsteps = (blk_chi(fwd, b, P.ns[s], P.ms[s]) - blk_clo(fwd, b) + 15) / 16;
        total += (int64_t)nsteps * 256;
      }
    }
    b0[nn] = (int)boff.size();
    boff.push_back(total);
  }
}

void pack_blocked(PackedFactor& P) {
  const int nn = (int)P.p0.size();
  blocked_layout(P);
  for (int dir = 0; dir < 2; ++dir) {
    const bool fwd = dir == 0;
    const std::vector<int>& b0 = fwd ? P.fblk0 : P.bblk0;
    const std::vector<int64_t>& boff = fwd ? P.fblk_off : P.bblk_off;
    std::vector<double>& dst = fwd ? P.fblk : P.bblk;
    dst.assign(boff.back(), 0.0);
    const std::vector<double>& src = fwd ? P.fwd : P.bwd;
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nn; ++s) {
      const int ns = P.ns[s], ms = P.ms[s];
      const int rows = fwd ? ns + ms : ns;
      const int64_t base = fwd ? P.fwd_off[s] : P.bwd_off[s];
      for (int b = 0; b * 16 < rows; ++b) {
        const int clo = blk_clo(fwd, b), chi = blk_chi(fwd, b, ns, ms);
        const int nsteps = (chi - clo + 15) / 16;
        double* d = &dst[boff[b0[s] + b]];
        for (int i = 0; i < 16; ++i) {
          const int lr = 16 * b + i;
          if (lr >= rows) continue;
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
          for (int st = 0; st < nsteps; ++st)
            for (int k = 0; k < 4; ++k)
              for (int q = 0; q < 4; ++q) {
                const int col = clo + 16 * st + 4 * k + q;
                if (col >= rs && col < re) d[((int64_t)st * 4 + q) * 64 + i + 16 * k] = src[base + off + col];
              }
        }
      }
    }
  }
  std::vector<double>().swap(P.fwd);
  std::vector<double>().swap(P.bwd);
}

void level_tiles(const PackedFactor& P, int l, bool fwd, int rows, std::vector<int>& out) {
  if (rows < 1) throw std::runtime_error("level_tiles: rows must be >= 1");
  for (int q = P.lnoff[l]; q < P.lnoff[l + 1]; ++q) {
    const int s = P.lnodes[q];
    const int nr = fwd ? P.ns[s] + P.ms[s] : P.ns[s];
    for (int r0 = 0; r0 < nr; r0 += rows) out.insert(out.end(), {s, r0, std::min(rows, nr - r0), -1});
  }
}

bool level_tiles_split(const PackedFactor& P, int l, bool fwd, int rows, int ks, std::vector<int>& tiles,
                       std::vector<int>& rtiles) {
  if (ks <= 0) {
    level_tiles(P, l, fwd, rows, tiles);
    return true;
  }
  if (ks % 16 != 0) throw std::runtime_error("level_tiles_split: segment length must be a multiple of 16");
  const int t_base = (int)(tiles.size() / 4);
  for (int q = P.lnoff[l]; q < P.lnoff[l + 1]; ++q) {
    const int s = P.lnodes[q];
    const int ns = P.ns[s], ms = P.ms[s];
    const int nr = fwd ? ns + ms : ns;
    for (int r0 = 0; r0 < nr; r0 += rows) {
      const int n = std::min(rows, nr - r0);
      const int last = r0 + n - 1;
      // columns the row tile touches: forward [0, longest row), backward [r0, ns + ms)
      const int cbeg = fwd ? 0 : r0;
      const int cend = fwd ? (last < ns ? last + 1 : ns) : ns + ms;
      const int k0 = cbeg / ks, k1 = (cend + ks - 1) / ks;
      if (k1 - k0 > 63) return false;
      const int t0 = (int)(tiles.size() / 4) - t_base;
      for (int k = k0; k < k1; ++k) tiles.insert(tiles.end(), {s, r0, n, k});
      rtiles.insert(rtiles.end(), {s, r0, n, 64 * t0 + (k1 - k0)});
    }
  }
  return true;
}

void rebuild_tiles(PackedFactor& P) {
  P.ftiles.clear();
  P.btiles.clear();
  P.frtiles.clear();
  P.brtiles.clear();
  P.foff.assign(P.nlevels + 1, 0);
  P.boff.assign(P.nlevels + 1, 0);
  P.froff.assign(P.nlevels + 1, 0);
  P.broff.assign(P.nlevels + 1, 0);
  if ((int)P.fks.size() != P.nlevels) P.fks.assign(P.nlevels, 0);
  if ((int)P.bks.size() != P.nlevels) P.bks.assign(P.nlevels, 0);
  for (int l = 0; l < P.nlevels; ++l) {
    // lanes per row W > 0: (256 / W) * MAXR rows per tile; W <= 0 (the MFMA
    // sweep): at most 64 * MAXR rows per tile (fewer with its right-hand
    // sides split over waves, W == -1); only the MFMA sweep splits columns
    auto cap = [](int w, int m) { return w > 0 ? (256 / w) * m : 64 * m; };
    if (P.frows[l] > cap(P.fw[l], P.fm[l]) || P.brows[l] > cap(P.bw[l], P.bm[l]))
      throw std::runtime_error("rebuild_tiles: tile height exceeds the rows one workgroup covers");
    if ((P.fks[l] && P.fw[l] > 0) || (P.bks[l] && P.bw[l] > 0))
      throw std::runtime_error("rebuild_tiles: only the MFMA sweep splits columns");
    P.foff[l] = (int)(P.ftiles.size() / 4);
    P.boff[l] = (int)(P.btiles.size() / 4);
    P.froff[l] = (int)(P.frtiles.size() / 4);
    P.broff[l] = (int)(P.brtiles.size() / 4);
    if (!level_tiles_split(P, l, true, P.frows[l], P.fks[l], P.ftiles, P.frtiles) ||
        !level_tiles_split(P, l, false, P.brows[l], P.bks[l], P.btiles, P.brtiles))
      throw std::runtime_error("rebuild_tiles: too many column segments");
  }
  P.foff[P.nlevels] = (int)(P.ftiles.size() / 4);
  P.boff[P.nlevels] = (int)(P.btiles.size() / 4);
  P.froff[P.nlevels] = (int)(P.frtiles.size() / 4);
  P.broff[P.nlevels] = (int)(P.brtiles.size() / 4);
}

}  // namespace pecs
