// Native self-test of the setup path (CPU only, not part of the library):
// builds the meshes, local operators, the SPD Schur / multiplier matrices,
// factors them and checks ||S x - b|| / ||b|| of the stored factor through a
// plain host triangular solve written only for this test.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "cholesky.hpp"
#include "mesh.hpp"
#include "ops.hpp"

using namespace pecs;

static void block_matvec(const BlockSPD& S, const std::vector<int>& base, const double* x, double* y) {
  const int no = S.noff();
  for (int X = 0; X < S.NXc * S.NYc; ++X) {
    const int nx = S.ndof[X];
    for (int a = 0; a < nx; ++a) y[base[X] + a] = 0.0;
    const int xi = X % S.NXc, xj = X / S.NXc;
    for (int o = 0; o < no; ++o) {
      const int yi = xi + S.offi[o], yj = xj + S.offj[o];
      if (yi < 0 || yi >= S.NXc || yj < 0 || yj >= S.NYc) continue;
      const int Y = yi + S.NXc * yj;
      const auto& B = S.blk[(size_t)X * no + o];
      if (B.empty()) continue;
      const int ny = S.ndof[Y];
      for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) y[base[X] + a] += B[(size_t)a * ny + b] * x[base[Y] + b];
    }
  }
}

static void factor_solve_host(const SupernodalFactor& F, const double* b, double* x) {
  std::vector<double> y(b, b + F.n);
  for (size_t s = 0; s < F.nodes.size(); ++s) {  // forward, postorder
    const SNode& N = F.nodes[s];
    std::vector<double> t(N.ns);
    for (int i = 0; i < N.ns; ++i) {
      double v = 0;
      for (int j = 0; j <= i; ++j) v += F.Linv[N.linv_off + (int64_t)i * N.ns + j] * y[N.p0 + j];
      t[i] = v;
    }
    for (int i = 0; i < N.ns; ++i) y[N.p0 + i] = t[i];
    for (int r = 0; r < N.ms; ++r) {
      double v = 0;
      for (int j = 0; j < N.ns; ++j) v += F.L21[N.l21_off + (int64_t)r * N.ns + j] * t[j];
      y[F.rows[N.rows_off + r]] -= v;
    }
  }
  for (int i = 0; i < F.n; ++i) x[i] = y[i];
  for (int s = (int)F.nodes.size() - 1; s >= 0; --s) {  // backward
    const SNode& N = F.nodes[s];
    std::vector<double> t(N.ns);
    for (int j = 0; j < N.ns; ++j) {
      double v = x[N.p0 + j];
      for (int r = 0; r < N.ms; ++r) v -= F.L21[N.l21_off + (int64_t)r * N.ns + j] * x[F.rows[N.rows_off + r]];
      t[j] = v;
    }
    for (int j = 0; j < N.ns; ++j) {
      double v = 0;
      for (int i = j; i < N.ns; ++i) v += F.Linv[N.linv_off + (int64_t)i * N.ns + j] * t[i];
      x[N.p0 + j] = v;
    }
  }
}

// Host emulation of the GPU sweep algorithm on the packed factor: the same row
// tiles, the same staged vectors, the same storage offsets.
static void packed_solve_host(const PackedFactor& P, const double* b, double* x) {
  std::vector<double> Y(P.n), z(P.zlen), X(P.n), vec;
  for (int l = 0; l < P.nlevels; ++l)
    for (int q = P.foff[l]; q < P.foff[l + 1]; ++q) {
      const int s = P.ftiles[4 * q], r0 = P.ftiles[4 * q + 1], nr = P.ftiles[4 * q + 2];
      const int ns = P.ns[s], p0 = P.p0[s];
      vec.assign(ns, 0.0);
      for (int pos = 0; pos < ns; ++pos) {
        double t = b[p0 + pos];
        for (int64_t c = P.cptr[p0 + pos]; c < P.cptr[p0 + pos + 1]; ++c) t -= z[P.czidx[c]];
        vec[pos] = t;
      }
      for (int lr = r0; lr < r0 + nr; ++lr) {
        const double* ptr;
        int end;
        if (lr < ns) {
          ptr = &P.fwd[P.fwd_off[s] + (int64_t)lr * (lr + 1) / 2];
          end = lr + 1;
        } else {
          ptr = &P.fwd[P.fwd_off[s] + (int64_t)ns * (ns + 1) / 2 + (int64_t)(lr - ns) * ns];
          end = ns;
        }
        double a = 0;
        for (int pos = 0; pos < end; ++pos) a += ptr[pos] * vec[pos];
        if (lr < ns) Y[p0 + lr] = a;
        else z[P.z_off[s] + lr - ns] = a;
      }
    }
  for (int l = P.nlevels - 1; l >= 0; --l)
    for (int q = P.boff[l]; q < P.boff[l + 1]; ++q) {
      const int s = P.btiles[4 * q], r0 = P.btiles[4 * q + 1], nr = P.btiles[4 * q + 2];
      const int ns = P.ns[s], ms = P.ms[s], p0 = P.p0[s];
      vec.assign(ns + ms, 0.0);
      for (int pos = 0; pos < ns; ++pos) vec[pos] = Y[p0 + pos];
      for (int i = 0; i < ms; ++i) vec[ns + i] = X[P.rows[P.rows_off[s] + i]];
      for (int lr = r0; lr < r0 + nr; ++lr) {
        const double* ptr = &P.bwd[P.bwd_off[s] + (int64_t)lr * (ns + ms) - (int64_t)lr * (lr - 1) / 2 - lr];
        double a = 0;
        for (int pos = lr; pos < ns + ms; ++pos) a += ptr[pos] * vec[pos];
        X[p0 + lr] = a;
      }
    }
  for (int i = 0; i < P.n; ++i) x[i] = X[i];
}

// Host emulation of the FP64-MFMA solve (k_pack_il, k_fwd_gather_il,
// k_sweep_mfma, k_bwd_gather_il, k_unpack_il) lane by lane: the same tiles, the
// RHS-interleaved vectors ([p * NRP + h]), the same per-lane loads and
// v_mfma_f64_16x16x4f64 operand / result layouts (A: lane l = row l%16, k
// l/16; B: lane l = k l/16, column l%16; D: lane l register r = row l/16 + 4r,
// column l%16), with every global index checked against its array.  nrhs
// h-major right-hand sides in b, h-major solutions in x.
#define EMU_CHECK(cond, what)                                              \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "mfma emulation: out of bounds: %s\n", what);  \
      std::exit(3);                                                        \
    }                                                                      \
  } while (0)

static void mfma_solve_emulate(const PackedFactor& P, int nrhs, int MB, int KS, const std::vector<double>& b,
                               std::vector<double>& x) {
  const int NRP = nrhs <= 16 ? 16 : 32, NH = NRP / 16;
  const int64_t n = P.n, zl = P.zlen;
  std::vector<double> X(NRP * n, 0.0), Y(NRP * n, 0.0), Z(NRP * std::max<int64_t>(zl, 1), 0.0);
  std::vector<int> tl;
  for (int dir = 0; dir < 2; ++dir) {
    const bool FWD = dir == 0;
    for (int li = 0; li < P.nlevels; ++li) {
      const int l = FWD ? li : P.nlevels - 1 - li;
      if (FWD) {
        if (l == 0)  // k_pack_il: the right-hand sides interleaved, padding zero
          for (int64_t p = 0; p < n; ++p)
            for (int h = 0; h < NRP; ++h) X[p * NRP + h] = h < nrhs ? b[h * n + p] : 0.0;
        for (int i = P.goff[l]; i < P.goff[l + 1]; ++i)
          for (int h = 0; h < nrhs && l > 0; ++h) {
            const int p = P.lrows[i];
            for (int64_t c = P.cptr[p]; c < P.cptr[p + 1]; ++c) {
              EMU_CHECK(P.czidx[c] >= 0 && P.czidx[c] < zl, "gathered z");
              X[(int64_t)p * NRP + h] -= Z[P.czidx[c] * NRP + h];
            }
          }
      } else {
        for (int i = P.xoff[l]; i < P.xoff[l + 1]; ++i)
          for (int h = 0; h < NRP; ++h) {
            const int64_t q = P.xidx[i];
            EMU_CHECK(q >= 0 && q < zl, "backward gather");
            Z[q * NRP + h] = X[(int64_t)P.rows[q] * NRP + h];
          }
      }
      tl.clear();
      std::vector<int> rtl;
      if (!level_tiles_split(P, l, FWD, 64 * MB, KS, tl, rtl)) {
        std::fprintf(stderr, "mfma emulation: too many segments\n");
        std::exit(3);
      }
      const int cap = 64 * MB;
      std::vector<double> part(KS > 0 ? tl.size() / 4 * cap * NRP : 0, 0.0);
      for (size_t t = 0; t < tl.size() / 4; ++t) {
        const int s = tl[4 * t], row0 = tl[4 * t + 1], nr = tl[4 * t + 2], kseg = tl[4 * t + 3];
        EMU_CHECK(nr <= 64 * MB, "tile rows");
        const int ns = P.ns[s], ms = P.ms[s], p0 = P.p0[s];
        const int VL = FWD ? ns : ns + ms;
        const std::vector<double>& blk = FWD ? P.fblk : P.bblk;
        const std::vector<int64_t>& boffs = FWD ? P.fblk_off : P.bblk_off;
        const int bnode = FWD ? P.fblk0[s] : P.bblk0[s], bnext = FWD ? P.fblk0[s + 1] : P.bblk0[s + 1];
        auto vecv = [&](int col, int h) -> double {
          if (col >= VL) return 0.0;
          if (FWD) {
            EMU_CHECK(p0 + col < n, "B operand X");
            return X[(int64_t)(p0 + col) * NRP + h];
          }
          if (col < ns) return Y[(int64_t)(p0 + col) * NRP + h];
          const int64_t q = P.rows_off[s] + (col - ns);
          EMU_CHECK(q >= 0 && q < zl, "B operand Z");
          return Z[q * NRP + h];
        };
        for (int wave = 0; wave < 4; ++wave) {
          const int first = row0 + wave * MB * 16;
          if (first >= row0 + nr) continue;
          const int last = std::min(row0 + nr, first + MB * 16) - 1;
          const int lo = FWD ? 0 : first, hi = FWD ? (last < ns ? last + 1 : ns) : ns + ms;
          int c0 = (lo >> 4) << 4, cend = hi;
          if (KS > 0) {
            EMU_CHECK(kseg >= 0, "segment index");
            c0 = std::max(c0, kseg * KS);
            cend = std::min(cend, (kseg + 1) * KS);
          }
          const int nstep = cend > c0 ? (cend - c0 + 15) >> 4 : 0;
          EMU_CHECK(row0 % 16 == 0, "tile start on a 16-row block");
          std::vector<double> C((size_t)MB * NH * 256, 0.0);
          for (int st = 0; st < nstep; ++st) {
            const int cc = c0 + 16 * st;
            for (int q = 0; q < 4; ++q)
              for (int mb = 0; mb < MB; ++mb) {
                double A[16][4] = {}, Bm[NH][4][16] = {};
                const int b = first / 16 + mb;
                const bool exists = first + mb * 16 < row0 + nr;
                int64_t bbase = 0;
                int bclo = 0, bsteps = 0;
                if (exists) {
                  EMU_CHECK(bnode + b < bnext, "block of the node");
                  bbase = boffs[bnode + b];
                  bclo = FWD ? 0 : 16 * b;
                  bsteps = (int)((boffs[bnode + b + 1] - bbase) >> 8);
                }
                const int idx = (cc - bclo) >> 4;
                const bool in = exists && cc >= bclo && idx < bsteps;
                for (int lane = 0; lane < 64; ++lane) {
                  const int ri = lane & 15, kq = lane >> 4;
                  const int col = cc + 4 * kq + q;
                  double L = 0.0;
                  if (in) {
                    const int64_t e = bbase + ((int64_t)idx * 4 + q) * 64 + lane;
                    EMU_CHECK(e >= 0 && e < (int64_t)blk.size(), "blocked factor entry");
                    L = blk[e];
                  }
                  A[ri][kq] = L;
                  for (int hg = 0; hg < NH; ++hg) Bm[hg][kq][ri] = vecv(col, hg * 16 + ri);
                }
                for (int hg = 0; hg < NH; ++hg) {
                  double* Cm = &C[(size_t)(mb * NH + hg) * 256];
                  for (int i = 0; i < 16; ++i)
                    for (int j = 0; j < 16; ++j) {
                      double v = 0;
                      for (int k = 0; k < 4; ++k) v += A[i][k] * Bm[hg][k][j];
                      Cm[i * 16 + j] += v;
                    }
                }
              }
          }
          for (int mb = 0; mb < MB; ++mb)
            for (int hg = 0; hg < NH; ++hg)
              for (int lane = 0; lane < 64; ++lane)
                for (int r = 0; r < 4; ++r) {
                  const int ri = lane & 15, kq = lane >> 4;
                  const int row = kq + 4 * r, lr = first + mb * 16 + row, h = hg * 16 + ri;
                  if (lr >= row0 + nr) continue;
                  const double v = C[(size_t)(mb * NH + hg) * 256 + row * 16 + ri];
                  if (KS > 0) {
                    const size_t e = ((size_t)t * cap + (lr - row0)) * NRP + h;
                    EMU_CHECK(e < part.size(), "partial slot");
                    part[e] = v;
                  } else if (FWD) {
                    if (lr < ns) Y[(int64_t)(p0 + lr) * NRP + h] = v;
                    else {
                      const int64_t z = P.z_off[s] + (lr - ns);
                      EMU_CHECK(z >= 0 && z < zl, "Z output");
                      Z[z * NRP + h] = v;
                    }
                  } else {
                    EMU_CHECK(p0 + lr < n, "X output");
                    X[(int64_t)(p0 + lr) * NRP + h] = v;
                  }
                }
        }
      }
      // reduction of split row tiles, segments in order (k_sweep_reduce)
      for (size_t q = 0; q < rtl.size() / 4; ++q) {
        const int s = rtl[4 * q], row0 = rtl[4 * q + 1], nr = rtl[4 * q + 2];
        const int t0 = rtl[4 * q + 3] >> 6, nseg = rtl[4 * q + 3] & 63;
        const int ns = P.ns[s], p0 = P.p0[s];
        for (int r = 0; r < nr; ++r)
          for (int h = 0; h < NRP; ++h) {
            double v = 0.0;
            for (int k = 0; k < nseg; ++k) {
              const size_t e = ((size_t)(t0 + k) * cap + r) * NRP + h;
              EMU_CHECK(e < part.size(), "reduce slot");
              v += part[e];
            }
            const int lr = row0 + r;
            if (FWD) {
              if (lr < ns) Y[(int64_t)(p0 + lr) * NRP + h] = v;
              else Z[(P.z_off[s] + (lr - ns)) * NRP + h] = v;
            } else {
              X[(int64_t)(p0 + lr) * NRP + h] = v;
            }
          }
      }
    }
  }
  x.assign((size_t)nrhs * n, 0.0);
  for (int64_t p = 0; p < n; ++p)
    for (int h = 0; h < nrhs; ++h) x[h * n + p] = X[p * NRP + h];
}

// CPU emulation of factor_on_device / blocked_on_device: the kernels' index
// arithmetic thread by thread (bounds-checked) and the dense calls in their
// column-major Upper meaning, written as plain loops.
#define DEV_CHECK(cond, what)                                                  \
  do {                                                                         \
    if (!(cond)) {                                                             \
      std::fprintf(stderr, "device factor emulation: out of bounds: %s\n", what); \
      std::exit(3);                                                            \
    }                                                                          \
  } while (0)

static void device_factor_emulate(const FrontPlan& plan, const SupernodalFactor& F, const PackedFactor& P,
                                  std::vector<double>& fwd, std::vector<double>& bwd) {
  const int nn = (int)F.nodes.size();
  const int64_t FT = plan.front_total;
  std::vector<double> fr(FT, 0.0);
  fwd.assign(P.nvals, 0.0);
  bwd.assign(P.nvals, 0.0);
  auto FR = [&](int64_t i) -> double& {
    DEV_CHECK(i >= 0 && i < FT, "front");
    return fr[i];
  };
  auto FW = [&](int64_t i) -> double& {
    DEV_CHECK(i >= 0 && i < P.nvals, "fwd");
    return fwd[i];
  };
  auto BW = [&](int64_t i) -> double& {
    DEV_CHECK(i >= 0 && i < P.nvals, "bwd");
    return bwd[i];
  };
  for (int l = 0; l < F.nlevels; ++l) {
    // k_front_scatter: 64 lanes per item
    for (int it = plan.sc_level[l]; it < plan.sc_level[l + 1]; ++it) {
      const int nx = plan.sc_nx[it], ny = plan.sc_ny[it], nf = plan.sc_nf[it];
      for (int t = 0; t < 64; ++t)
        for (int e = t; e < nx * ny; e += 64) {
          const int a = e / ny, b = e - a * ny;
          DEV_CHECK(plan.sc_src[it] + e < (int64_t)plan.svals.size(), "svals");
          const double v = plan.svals[plan.sc_src[it] + e];
          if (plan.sc_mode[it] == 0) {
            if (plan.sc_rx[it] + a >= plan.sc_ry[it] + b) FR(plan.sc_dst[it] + (int64_t)a * nf + b) = v;
          } else {
            FR(plan.sc_dst[it] + (int64_t)b * nf + a) = v;
          }
        }
    }
    // k_extend_add, two passes
    for (int pass = 0; pass < 2; ++pass)
      for (int it = plan.ea_level[2 * l + pass]; it < plan.ea_level[2 * l + pass + 1]; ++it) {
        const int c = plan.ea_child[it], r = plan.ea_row[it], s = plan.ea_parent[it];
        const int nsc = P.ns[c], nfc = nsc + P.ms[c], nfp = P.ns[s] + P.ms[s];
        DEV_CHECK(r < P.ms[c], "child row");
        const int* m = &plan.ea_map[plan.ea_map_off[c]];
        DEV_CHECK(plan.ea_map_off[c] + P.ms[c] <= plan.ea_map_off[c + 1], "map");
        for (int t = 0; t < 64; ++t)
          for (int q = t; q <= r; q += 64) {
            DEV_CHECK(m[r] >= m[q] && m[r] < nfp, "map lower triangle");
            FR(plan.front_off[s] + (int64_t)m[r] * nfp + m[q]) += FR(plan.front_off[c] + (int64_t)(nsc + r) * nfc + nsc + q);
          }
      }
    // dense calls, column-major with leading dimension nf
    for (int s : F.level_nodes[l]) {
      const int ns = P.ns[s], ms = P.ms[s], nf = ns + ms;
      if (ns == 0) continue;
      const int64_t o = plan.front_off[s];
      DEV_CHECK(o + (int64_t)nf * nf <= FT, "front extent");
      auto A = [&](int i, int j) -> double& { return FR(o + i + (int64_t)j * nf); };
      // potrf(Upper): A11 = U^T U
      for (int j = 0; j < ns; ++j) {
        double d = A(j, j);
        for (int k = 0; k < j; ++k) d -= A(k, j) * A(k, j);
        if (!(d > 0)) {
          std::printf("emulated potrf: not positive definite\n");
          std::exit(2);
        }
        A(j, j) = std::sqrt(d);
        for (int i = j + 1; i < ns; ++i) {
          double v = A(j, i);
          for (int k = 0; k < j; ++k) v -= A(k, j) * A(k, i);
          A(j, i) = v / A(j, j);
        }
      }
      // trsm(Left, Upper, Transpose): U^T X = A12
      for (int jc = 0; jc < ms; ++jc)
        for (int i = 0; i < ns; ++i) {
          double v = A(i, ns + jc);
          for (int k = 0; k < i; ++k) v -= A(k, i) * A(k, ns + jc);
          A(i, ns + jc) = v / A(i, i);
        }
      // gemm(Transpose, None, -1, 1): A22 -= A12^T A12 (both triangles)
      for (int j = 0; j < ms; ++j)
        for (int i = 0; i < ms; ++i) {
          double v = 0;
          for (int k = 0; k < ns; ++k) v += A(k, ns + i) * A(k, ns + j);
          A(ns + i, ns + j) -= v;
        }
      // trtri(Upper): A11 = U^-1 (column by column, back substitution)
      std::vector<double> Ui((size_t)ns * ns, 0.0);
      for (int j = 0; j < ns; ++j) {
        Ui[j + (size_t)j * ns] = 1.0 / A(j, j);
        for (int i = j - 1; i >= 0; --i) {
          double v = 0;
          for (int k = i + 1; k <= j; ++k) v += A(i, k) * Ui[k + (size_t)j * ns];
          Ui[i + (size_t)j * ns] = -v / A(i, i);
        }
      }
      for (int j = 0; j < ns; ++j)
        for (int i = 0; i <= j; ++i) A(i, j) = Ui[i + (size_t)j * ns];
      // trmm(Left, Upper, None): W = U^-1 A12 into fwd, ldc = ns
      const int64_t wo = P.fwd_off[s] + (int64_t)ns * (ns + 1) / 2;
      for (int jc = 0; jc < ms; ++jc)
        for (int i = 0; i < ns; ++i) {
          double v = 0;
          for (int k = i; k < ns; ++k) v += A(i, k) * A(k, ns + jc);
          FW(wo + i + (int64_t)jc * ns) = v;
        }
    }
  }
  // k_pack_rows: items (node, 16-row chunk), 256 threads, LDS tile t[16][17]
  for (int s = 0; s < nn; ++s)
    for (int j0 = 0; j0 < P.ns[s]; j0 += 16) {
      const int ns = P.ns[s], ms = P.ms[s], nf = ns + ms;
      const int64_t Fo = plan.front_off[s], fo = P.fwd_off[s], Wo = fo + (int64_t)ns * (ns + 1) / 2, bo = P.bwd_off[s];
      const int jend = std::min(j0 + 16, ns);
      for (int r = j0; r < jend; ++r)
        for (int t = 0; t < 256; ++t)
          for (int c = t; c <= r; c += 256) FW(fo + (int64_t)r * (r + 1) / 2 + c) = FR(Fo + (int64_t)r * nf + c);
      double tile[16][17];
      for (int i0 = j0; i0 < nf; i0 += 16) {
        for (int t = 0; t < 256; ++t) {
          const int a = t & 15, b = t >> 4, j = j0 + a, i = i0 + b;
          double v = 0.0;
          if (j < ns && i >= j && i < nf) v = i < ns ? FR(Fo + (int64_t)i * nf + j) : -FW(Wo + (int64_t)(i - ns) * ns + j);
          tile[a][b] = v;
        }
        for (int t = 0; t < 256; ++t) {
          const int a = t & 15, b = t >> 4, j = j0 + b, i = i0 + a;
          if (j < ns && i >= j && i < nf) {
            DEV_CHECK((int64_t)j * nf - (int64_t)j * (j - 1) / 2 + (i - j) < (int64_t)ns * (ns + 1) / 2 + (int64_t)ns * ms,
                      "bwd row within node");
            BW(bo + (int64_t)j * nf - (int64_t)j * (j - 1) / 2 + (i - j)) = tile[b][a];
          }
        }
      }
    }
}

// k_pack_blocked, thread by thread.
static void blocked_emulate(const PackedFactor& P, const std::vector<double>& fwd, const std::vector<double>& bwd,
                            std::vector<double>& fblk, std::vector<double>& bblk) {
  const int nn = (int)P.p0.size();
  for (int dir = 0; dir < 2; ++dir) {
    const bool fw = dir == 0;
    const std::vector<int>& b0 = fw ? P.fblk0 : P.bblk0;
    const std::vector<int64_t>& boff = fw ? P.fblk_off : P.bblk_off;
    std::vector<double>& dst = fw ? fblk : bblk;
    const std::vector<double>& src = fw ? fwd : bwd;
    dst.assign(boff.back(), -1.0);  // every entry must be written
    for (int s = 0; s < nn; ++s)
      for (int gb = b0[s]; gb < b0[s + 1]; ++gb) {
        const int b = gb - b0[s], ns = P.ns[s], ms = P.ms[s];
        const int clo = fw ? 0 : 16 * b;
        int chi;
        if (fw) {
          const int last = std::min(16 * b + 15, ns + ms - 1);
          chi = last < ns ? last + 1 : ns;
        } else {
          chi = ns + ms;
        }
        const int nsteps = (chi - clo + 15) / 16;
        DEV_CHECK(boff[gb] + (int64_t)nsteps * 256 == boff[gb + 1], "block extent");
        const int rows = fw ? ns + ms : ns;
        const int64_t so = fw ? P.fwd_off[s] : P.bwd_off[s];
        for (int t = 0; t < 256; ++t)
          for (int e = t; e < nsteps * 256; e += 256) {
            const int st = e >> 8, q = (e >> 6) & 3, lane = e & 63, i = lane & 15, k = lane >> 4;
            const int lr = 16 * b + i, col = clo + 16 * st + 4 * k + q;
            double v = 0.0;
            if (lr < rows) {
              int64_t off;
              int rs, re;
              if (fw) {
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
              if (col >= rs && col < re) {
                DEV_CHECK(so + off + col >= 0 && so + off + col < (int64_t)src.size(), "blocked source");
                v = src[so + off + col];
              }
            }
            dst[boff[gb] + e] = v;
          }
      }
  }
}

static double check(const char* name, const BlockSPD& S, int leaf) {
  auto t0 = std::chrono::steady_clock::now();
  SupernodalFactor F = cholesky_factor(S, leaf);
  {  // the device plan, executed on the host, must reproduce the factor
    FrontPlan plan;
    SupernodalFactor G = cholesky_symbolic(S, leaf, plan);
    std::vector<double> Li, L2;
    execute_plan_host(plan, G, Li, L2);
    double d = 0;
    for (size_t i = 0; i < Li.size(); ++i) d = std::max(d, std::fabs(Li[i] - F.Linv[i]));
    for (size_t i = 0; i < L2.size(); ++i) d = std::max(d, std::fabs(L2[i] - F.L21[i]));
    std::printf("   front plan on host: max |diff| to cholesky_factor %.3e (fronts %.1f MB, %zu scatter items, %zu "
                "extend-add rows)\n",
                d, plan.front_total * 8e-6, plan.sc_src.size(), plan.ea_child.size());
    if (d != 0.0 || Li.size() != F.Linv.size() || L2.size() != F.L21.size()) {
      std::printf("front plan mismatch\n");
      std::exit(2);
    }
    // the device numeric phase and its sweep-layout kernels, emulated
    PackedFactor PL = pack_layout(G);
    std::vector<double> ef, eb;
    device_factor_emulate(plan, G, PL, ef, eb);
    PackedFactor PR = pack_factor(F);
    double df = 0, mx = 0;
    for (int64_t i = 0; i < PL.nvals; ++i) {
      df = std::max(df, std::max(std::fabs(ef[i] - PR.fwd[i]), std::fabs(eb[i] - PR.bwd[i])));
      mx = std::max(mx, std::fabs(PR.fwd[i]));
    }
    PL.fwd = ef;
    PL.bwd = eb;
    PackedFactor PB2 = PL;
    pack_blocked(PB2);
    blocked_layout(PL);
    std::vector<double> fb, bb;
    blocked_emulate(PL, ef, eb, fb, bb);
    const bool blk_same = fb == PB2.fblk && bb == PB2.bblk && PL.fblk_off == PB2.fblk_off && PL.bblk_off == PB2.bblk_off;
    std::printf("   device factor emulation: max |diff| of the sweep layout to pack_factor %.3e (max |entry| %.3e); "
                "blocked copies %s\n",
                df, mx, blk_same ? "identical" : "DIFFER");
    if (!(df <= 1e-10 * mx) || !blk_same) {
      std::printf("device factor emulation mismatch\n");
      std::exit(2);
    }
  }
  double ts = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::mt19937_64 rng(20261202);
  std::uniform_real_distribution<double> U(-1, 1);
  std::vector<double> x(F.n), b(F.n), xs(F.n), r(F.n);
  for (auto& v : x) v = U(rng);
  block_matvec(S, F.cell_base, x.data(), b.data());
  factor_solve_host(F, b.data(), xs.data());
  block_matvec(S, F.cell_base, xs.data(), r.data());
  double nr = 0, nb = 0, ne = 0, nx = 0;
  for (int i = 0; i < F.n; ++i) {
    nr += (r[i] - b[i]) * (r[i] - b[i]);
    nb += b[i] * b[i];
    ne += (xs[i] - x[i]) * (xs[i] - x[i]);
    nx += x[i] * x[i];
  }
  double rel = std::sqrt(nr / nb);
  {
    PackedFactor P = pack_factor(F);
    std::vector<double> xp(F.n), rp(F.n);
    packed_solve_host(P, b.data(), xp.data());
    block_matvec(S, F.cell_base, xp.data(), rp.data());
    double nr2 = 0;
    for (int i = 0; i < F.n; ++i) nr2 += (rp[i] - b[i]) * (rp[i] - b[i]);
    double rel2 = std::sqrt(nr2 / nb);
    std::printf("   packed sweeps: resid=%.3e\n", rel2);
    // MFMA sweep emulation, several right-hand sides (RHS h = b scaled by h+1)
    PackedFactor PB = P;
    pack_blocked(PB);
    for (int nrhs : {3, 17}) {
      for (int MB : {1, 2, 4})
      for (int KS : {0, 32, 256}) {
        std::vector<double> bb((size_t)nrhs * F.n), xm;
        for (int h = 0; h < nrhs; ++h)
          for (int i = 0; i < F.n; ++i) bb[(size_t)h * F.n + i] = (h + 1) * b[i];
        mfma_solve_emulate(PB, nrhs, MB, KS, bb, xm);
        double e2 = 0, n2 = 0;
        for (int h = 0; h < nrhs; ++h)
          for (int i = 0; i < F.n; ++i) {
            const double d = xm[(size_t)h * F.n + i] - (h + 1) * xs[i];
            e2 += d * d;
            n2 += (h + 1) * (h + 1) * xs[i] * xs[i];
          }
        const double relm = std::sqrt(e2 / n2);
        std::printf("   mfma emulation nrhs=%d MB=%d KS=%d: rel diff to host solve %.3e\n", nrhs, MB, KS, relm);
        rel2 = std::max(rel2, relm);
      }
    }
    if (std::getenv("PECS_LEVELS"))
      for (int l = 0; l < P.nlevels; ++l) {
        double fb = 0, bb = 0;
        int nodes = 0, rowsf = 0, rowsb = 0, maxns = 0, maxms = 0;
        for (int q = P.foff[l]; q < P.foff[l + 1]; ++q) {
          const int s = P.ftiles[4 * q], r0 = P.ftiles[4 * q + 1], nr = P.ftiles[4 * q + 2];
          for (int r = r0; r < r0 + nr; ++r) fb += 8.0 * (r < P.ns[s] ? r + 1 : P.ns[s]);
          if (r0 == 0) { nodes++; maxns = std::max(maxns, P.ns[s]); maxms = std::max(maxms, P.ms[s]); }
          rowsf += nr;
        }
        for (int q = P.boff[l]; q < P.boff[l + 1]; ++q) {
          const int s = P.btiles[4 * q], r0 = P.btiles[4 * q + 1], nr = P.btiles[4 * q + 2];
          for (int j = r0; j < r0 + nr; ++j) bb += 8.0 * (P.ns[s] - j + P.ms[s]);
          rowsb += nr;
        }
        std::printf("   level %2d nodes %5d maxns %5d maxms %5d fwd rows %8d %8.1f MB W%d  bwd rows %8d %8.1f MB W%d\n",
                    l, nodes, maxns, maxms, rowsf, fb / 1e6, P.fw[l], rowsb, bb / 1e6, P.bw[l]);
      }
    rel = std::max(rel, rel2);
  }
  std::printf("%s: n=%d nodes=%zu levels=%d nnzL=%lld (%.1f MB) flops=%.3g factor %.2fs  resid=%.3e err=%.3e\n", name,
              F.n, F.nodes.size(), F.nlevels, (long long)F.nnz_L(), F.nnz_L() * 8e-6, F.flops, ts, rel,
              std::sqrt(ne / nx));
  std::printf("   time: assemble %.2fs factor %.2fs inverse %.2fs\n", F.t_assemble, F.t_factor, F.t_inverse);
  return rel;
}

int main(int argc, char** argv) {
  int r = argc > 1 ? std::atoi(argv[1]) : 4;
  int k = argc > 2 ? std::atoi(argv[2]) : 1;
  int leaf = argc > 3 ? std::atoi(argv[3]) : 16;
  int conic = argc > 4 ? std::atoi(argv[4]) : 0;
  int local = argc > 5 ? std::atoi(argv[5]) : 0;
  GeometrySpec g;
  if (conic) {
    g.r1 = 0.5;
    g.r2 = 0.5 + std::tan(10.0 * M_PI / 180.0);
    g.L = 1.5;
  }
  if (local) {  // boundary-layer refinement: lattice-embedded matrices, adaptive separators
    g.local = 1;
    g.bl = 0.1;
  }
  Params P;
  P.k = k;
  P.dt = (k == 1) ? 1e-3 : 5e-4;
  auto t0 = std::chrono::steady_clock::now();
  SolarCellMesh m = make_solarcell_mesh(r, g);
  RefTables T(k);
  LdgRegionOps os = build_ldg_ops(m.semi, T, P);
  LdgRegionOps oe = build_ldg_ops(m.elec, T, P);
  PoissonLayout L = poisson_layout(m, k - 1);
  PoissonOps op = build_poisson_ops(m, L, T, P);
  double ts = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("setup ops %.2fs: ldg types semi=%d elec=%d poisson types=%d nmult=%d\n", ts, os.ntype, oe.ntype,
              op.ntype, op.nmult);
  t0 = std::chrono::steady_clock::now();
  BlockSPD Ss = schur_matrix(m.semi, os, 1.0);
  BlockSPD Se = schur_matrix(m.elec, oe, 1.0);
  BlockSPD H = multiplier_matrix(m, op, T);
  std::printf("assemble S/H %.2fs\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  double worst = 0;
  worst = std::max(worst, check("S_semi", Ss, leaf));
  worst = std::max(worst, check("S_elec", Se, leaf));
  worst = std::max(worst, check("H_poisson", H, leaf));
  std::printf("%s\n", worst < 1e-12 ? "SELFTEST OK" : "SELFTEST FAIL");
  return worst < 1e-12 ? 0 : 1;
}
