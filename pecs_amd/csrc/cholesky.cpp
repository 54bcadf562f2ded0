// Supernodal multifrontal Cholesky with geometric nested dissection (setup).
#include "cholesky.hpp"

#include <omp.h>

#include <algorithm>
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

  int build(int i0, int i1, int j0, int j1) {
    const int ni = i1 - i0, nj = j1 - j0, w = S.sep_width;
    if (ni <= 0 || nj <= 0) return -1;
    if (ni * nj <= leaf || (ni <= w && nj <= w)) {
      NDNode n;
      for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) n.cells.push_back(i + S.NXc * j);
      nodes.push_back(n);
      return (int)nodes.size() - 1;
    }
    std::vector<int> kids;
    NDNode sep;
    if (ni >= nj && ni > w) {
      const int mid = i0 + (ni - w) / 2;
      int a = build(i0, mid, j0, j1);
      int b = build(mid + w, i1, j0, j1);
      if (a >= 0) kids.push_back(a);
      if (b >= 0) kids.push_back(b);
      for (int j = j0; j < j1; ++j)
        for (int i = mid; i < mid + w; ++i) sep.cells.push_back(i + S.NXc * j);
    } else {
      const int mid = j0 + (nj - w) / 2;
      int a = build(i0, i1, j0, mid);
      int b = build(i0, i1, mid + w, j1);
      if (a >= 0) kids.push_back(a);
      if (b >= 0) kids.push_back(b);
      for (int j = mid; j < mid + w; ++j)
        for (int i = i0; i < i1; ++i) sep.cells.push_back(i + S.NXc * j);
    }
    sep.children = kids;
    nodes.push_back(sep);
    const int id = (int)nodes.size() - 1;
    for (int k : kids) nodes[k].parent = id;
    return id;
  }
};

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
#pragma omp parallel for schedule(dynamic, 16) if (nt > 128)
    for (int i = k1; i < nf; ++i) {
      const double* Li = &F[(size_t)i * nf + k0];
      double* Fi = &F[(size_t)i * nf];
      for (int j = k1; j <= i; ++j) {
        const double* Lj = &F[(size_t)j * nf + k0];
        double s = 0.0;
#pragma omp simd reduction(+ : s)
        for (int t = 0; t < pw; ++t) s += Li[t] * Lj[t];
        Fi[j] -= s;
      }
    }
  }
}

// Inverse of a lower-triangular n x n block stored with leading dimension ld.
void lower_inverse(const double* L, int n, int ld, double* X) {
  for (size_t t = 0; t < (size_t)n * n; ++t) X[t] = 0.0;
#pragma omp parallel for schedule(dynamic, 8) if (n > 256)
  for (int j = 0; j < n; ++j) {
    X[(size_t)j * n + j] = 1.0 / L[(size_t)j * ld + j];
    for (int i = j + 1; i < n; ++i) {
      double s = 0.0;
      const double* Li = &L[(size_t)i * ld];
      for (int t = j; t < i; ++t) s += Li[t] * X[(size_t)t * n + j];
      X[(size_t)i * n + j] = -s / Li[i];
    }
  }
}

}  // namespace

SupernodalFactor cholesky_factor(const BlockSPD& S, int leaf_cells) {
  const int NG = S.NXc * S.NYc;
  const int no = S.noff();
  NDBuilder nd(S, leaf_cells);
  nd.build(0, S.NXc, 0, S.NYc);
  const int nn = (int)nd.nodes.size();
  SupernodalFactor F;
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
  F.Linv.assign(linv, 0.0);
  F.L21.assign(l21, 0.0);
  F.rows.assign(rows, 0);
  for (int s = 0; s < nn; ++s) {
    int t = 0;
    for (int c : structc[s])
      for (int a = 0; a < S.ndof[c]; ++a) F.rows[F.nodes[s].rows_off + t++] = F.cell_base[c] + a;
  }
  // numeric multifrontal factorisation
  std::vector<std::vector<double>> upd(nn);
  std::vector<int> front_off(NG, -1);
  for (int s = 0; s < nn; ++s) {
    const SNode& N = F.nodes[s];
    const int ns = N.ns, ms = N.ms, nf = ns + ms;
    if (ns == 0) {
      // a supernode without DoFs only forwards its children's updates
    }
    std::vector<double> Fr((size_t)nf * nf, 0.0);
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
      for (int r = 0; r < mc; ++r) {
        double* row = &Fr[(size_t)map[r] * nf];
        const double* ur = &U[(size_t)r * mc];
        for (int q = 0; q <= r; ++q) row[map[q]] += ur[q];
      }
      std::vector<double>().swap(upd[c]);
    }
    for (int c : nd.nodes[s].cells) front_off[c] = -1;
    for (int c : structc[s]) front_off[c] = -1;
    if (ns > 0) partial_cholesky(Fr.data(), nf, ns, F.flops);
    F.flops += (double)ns * ns * ns / 3.0;
    // store Linv, L21 and the update matrix
    if (ns > 0) lower_inverse(Fr.data(), ns, nf, &F.Linv[N.linv_off]);
    for (int r = 0; r < ms; ++r)
      for (int c2 = 0; c2 < ns; ++c2) F.L21[N.l21_off + (int64_t)r * ns + c2] = Fr[(size_t)(ns + r) * nf + c2];
    if (ms > 0) {
      upd[s].assign((size_t)ms * ms, 0.0);
      for (int r = 0; r < ms; ++r)
        for (int q = 0; q <= r; ++q) upd[s][(size_t)r * ms + q] = Fr[(size_t)(ns + r) * nf + ns + q];
    }
  }
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
  return F;
}

}  // namespace pecs
