// HIP kernels of the PECS per-time-step IMEX path for gfx950 (CDNA4).
//
// Per step (SolarCell.cpp:2057-2075):
//   k_rhs_carrier   assemble_{semiconductor,electrolyte}_rhs: one thread per
//                   cell, geometry rebuilt from the 4 vertices (MappingQ1),
//                   old densities, Poisson field (Piola) and lagged traces of
//                   the other region evaluated at QGauss(k+2) points.
//   k_schur_rhs     g = f_v + B^T W (cell + 4 face neighbours), with
//                   W = mu A^{-1} f_p recomputed per cell from the RHS.
//   k_fwd_gather / k_sweep  supernodal triangular solves
//                   with S = L L^T, level by level on the elimination tree:
//                   row tiles of dense [Linv; L21 Linv] panels, vector chunks
//                   staged in LDS, W lanes per row (deterministic, no atomics).
//   k_schur_recover q = W - A^{-1} B u; writes [Jx | Jy | u].
//   k_poisson_*     hybridised mixed Poisson: cell condensation, face
//                   assembly, multiplier solve, cell recovery.
// Batches of instances sharing the factors: k_sweep_mfma (FP64 MFMA) with
// RHS-interleaved vectors and their gathers.  The Schur and Poisson cell
// kernels take wave-uniform cell lists where the mesh has few cell types
// (table entries as scalar operands), and FP32 variants (*_f32) serve the
// mixed-precision mode.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace pecs {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
  return x;
}

struct GeoD {
  double x, y, J00, J01, J10, J11, det;
};

__device__ __forceinline__ GeoD bilin(const double* v, double xi, double eta) {
  GeoD g;
  const double a = (1 - xi) * (1 - eta), b = xi * (1 - eta), c = (1 - xi) * eta, d = xi * eta;
  g.x = v[0] * a + v[2] * b + v[4] * c + v[6] * d;
  g.y = v[1] * a + v[3] * b + v[5] * c + v[7] * d;
  g.J00 = (v[2] - v[0]) * (1 - eta) + (v[6] - v[4]) * eta;
  g.J10 = (v[3] - v[1]) * (1 - eta) + (v[7] - v[5]) * eta;
  g.J01 = (v[4] - v[0]) * (1 - xi) + (v[6] - v[2]) * xi;
  g.J11 = (v[5] - v[1]) * (1 - xi) + (v[7] - v[3]) * xi;
  g.det = g.J00 * g.J11 - g.J01 * g.J10;
  return g;
}

__device__ __forceinline__ void face_nrm(const double* v, int f, double& nx, double& ny, double& len) {
  int a, b;
  if (f == 0) { a = 0; b = 2; }
  else if (f == 1) { a = 1; b = 3; }
  else if (f == 2) { a = 0; b = 1; }
  else { a = 2; b = 3; }
  const double tx = v[2 * b] - v[2 * a], ty = v[2 * b + 1] - v[2 * a + 1];
  len = sqrt(tx * tx + ty * ty);
  if (f == 1 || f == 2) { nx = ty / len; ny = -tx / len; }
  else { nx = -ty / len; ny = tx / len; }
}

template <int K>
struct Sz {
  static constexpr int NB = (K + 1) * (K + 1);
  static constexpr int NQ = (K + 2) * (K + 2);
  static constexpr int NF = K + 2;
  static constexpr int NRT = 2 * K * (K + 1);
  static constexpr int NRTX = K * (K + 1);
  static constexpr int NPHI = K * K;
  static constexpr int NL = NRT + NPHI;
};

// local RT index of face node m of face f (RT_{K-1} nodal basis)
template <int K>
__device__ __forceinline__ int rt_face(int f, int m) {
  if (f == 0) return (K + 1) * m;
  if (f == 1) return K + (K + 1) * m;
  if (f == 2) return Sz<K>::NRTX + m;
  return Sz<K>::NRTX + m + K * K;
}

// ------------------------------------------------------------------------
// Carrier right-hand sides (SolarCell.cpp:1073-1414 and 1451-1726)
// ------------------------------------------------------------------------
template <int K, bool SEMI>
__global__ __launch_bounds__(64) void k_rhs_carrier(DevRegion R, DevRegion O, DevPoisson P, DevTabs T,
                                                    const DevPhys* __restrict__ phs, const double* __restrict__ u1,
                                                    const double* __restrict__ u2, const double* __restrict__ ou1,
                                                    const double* __restrict__ ou2, const double* __restrict__ pvec,
                                                    double* __restrict__ b1, double* __restrict__ b2, int64_t sR,
                                                    int64_t sO, int64_t sP) {
  using S = Sz<K>;
  constexpr int NB = S::NB, NQ = S::NQ, NF = S::NF, NRT = S::NRT;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R.ncell) return;
  // instance blockIdx.y of a batch: its physics and vectors
  const DevPhys ph = phs[blockIdx.y];
  {
    const int64_t i = blockIdx.y;
    u1 += i * sR;
    u2 += i * sR;
    b1 += i * sR;
    b2 += i * sR;
    ou1 += i * sO;
    ou2 += i * sO;
    pvec += i * sP;
  }
  const int nc = R.ncell;
  double v[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) v[t] = R.V[8 * (size_t)c + t];
  double d1[NB], d2[NB];
  const size_t ub = 2 * (size_t)nc * NB + (size_t)c * NB;
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    d1[a] = u1[ub + a];
    d2[a] = u2[ub + a];
  }
  const int pc = SEMI ? c : P.ncell_semi + c;
  double dp[NRT];
#pragma unroll
  for (int i = 0; i < NRT; ++i) dp[i] = pvec[P.idx[(size_t)pc * P.nloc + i]];
  double a1p[2 * NB], a1v[NB], a2p[2 * NB], a2v[NB];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    a1p[a] = a1p[NB + a] = a1v[a] = 0.0;
    a2p[a] = a2p[NB + a] = a2v[a] = 0.0;
  }
  const double inv_perm = 1.0 / (SEMI ? ph.eps_s : ph.eps_e);
  const double idt = 1.0 / ph.dt;
  for (int q = 0; q < NQ; ++q) {
    const GeoD g = bilin(v, T.qx[q], T.qy[q]);
    const double jxw = T.qw[q] * g.det;
    double n1 = 0.0, n2 = 0.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      const double ph_a = T.phi[a * NQ + q];
      n1 += d1[a] * ph_a;
      n2 += d2[a] * ph_a;
    }
    double hx = 0.0, hy = 0.0;
#pragma unroll
    for (int i = 0; i < NRT; ++i) {
      hx += dp[i] * T.psix[i * NQ + q];
      hy += dp[i] * T.psiy[i * NQ + q];
    }
    const double Ex = (g.J00 * hx + g.J01 * hy) / g.det;
    const double Ey = (g.J10 * hx + g.J11 * hy) / g.det;
    double src = 0.0;
    if (SEMI) {
      if (ph.gen_mode == 0) src = ph.alpha * ph.G0 * exp(ph.alpha * (g.y - ph.H));
      else if (ph.gen_mode == 1) src = ph.G0 * exp(-ph.alpha * g.x);
      if (ph.srh)
        src += (ph.ni * ph.ni - n1 * n2) / (ph.tau_n * (n1 + ph.ni) + ph.tau_p * (n2 + ph.ni));
    }
    if (R.msrc) src += R.msrc[(size_t)c * NQ + q];
    const double sv1 = jxw * (n1 * idt + src), sv2 = jxw * (n2 * idt + src);
    const double cp1 = -1.0 * inv_perm * n1 * jxw, cp2 = 1.0 * inv_perm * n2 * jxw;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      const double ph_a = T.phi[a * NQ + q];
      a1v[a] += sv1 * ph_a;
      a2v[a] += sv2 * ph_a;
      a1p[a] += cp1 * Ex * ph_a;
      a1p[NB + a] += cp1 * Ey * ph_a;
      a2p[a] += cp2 * Ex * ph_a;
      a2p[NB + a] += cp2 * Ey * ph_a;
    }
  }
  const double hc = R.h[c];
  for (int f = 0; f < 4; ++f) {
    const int bid = R.bid[4 * (size_t)c + f];
    if (bid < 0) continue;
    double nx, ny, len;
    face_nrm(v, f, nx, ny, len);
    const double* fph = T.fphi + (size_t)f * NB * NF;
    if (R.mflux && (bid == 0 || bid == 2)) {  // manufactured data: prescribed normal flux, -<v, h>
      for (int q = 0; q < NF; ++q) {
        const double w = -R.mflux[(4 * (size_t)c + f) * NF + q] * T.fw[q] * len;
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          a1v[a] += fph[a * NF + q] * w;
          a2v[a] += fph[a * NF + q] * w;
        }
      }
    }
    if (bid == 2) continue;  // Neumann: nothing (SolarCell.cpp:1403-1406)
    if (!SEMI && bid == 3) continue;
    if (bid == 1) {  // Dirichlet
      const double c1 = SEMI ? ph.ne : ph.rinf, c2 = SEMI ? ph.pe : ph.oinf;
      const double pen = ph.penalty / hc;
      for (int q = 0; q < NF; ++q) {
        const double jw = T.fw[q] * len;
        const double e1 = R.mdir ? R.mdir[(4 * (size_t)c + f) * NF + q] : c1;
        const double e2 = R.mdir ? e1 : c2;
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          const double fa = fph[a * NF + q];
          a1p[a] -= fa * nx * jw * e1;
          a1p[NB + a] -= fa * ny * jw * e1;
          a1v[a] += pen * fa * jw * e1;
          a2p[a] -= fa * nx * jw * e2;
          a2p[NB + a] -= fa * ny * jw * e2;
          a2v[a] += pen * fa * jw * e2;
        }
      }
      continue;
    }
    // interface (0) and Schottky (3) use the own traces
    double o1[NB], o2[NB];
    const double* foph = T.fphi + (size_t)(f ^ 1) * NB * NF;
    if (bid == 0) {
      const int e = R.sigma_nb[c];
      if (e < 0) continue;
      const size_t ob = 2 * (size_t)O.ncell * NB + (size_t)e * NB;
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        o1[a] = ou1[ob + a];
        o2[a] = ou2[ob + a];
      }
    }
    for (int q = 0; q < NF; ++q) {
      const double jw = T.fw[q] * len;
      double t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        t1 += d1[a] * fph[a * NF + q];
        t2 += d2[a] * fph[a * NF + q];
      }
      double w1, w2;
      if (bid == 0) {
        double r1 = 0.0, r2 = 0.0;
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          r1 += o1[a] * foph[a * NF + q];
          r2 += o2[a] * foph[a * NF + q];
        }
        if (SEMI) {
          w1 = -ph.k_et * (t1 - ph.ne) * r2 * jw;  // -v k_et (n - n^e) o
          w2 = ph.k_ht * (t2 - ph.pe) * r1 * jw;   // +v k_ht (p - p^e) r
        } else {
          const double cur = (-ph.k_et * (r1 - ph.ne) * t2 + ph.k_ht * (r2 - ph.pe) * t1) * jw;
          w1 = cur;
          w2 = -cur;
        }
      } else {  // Schottky (SolarCell.cpp:1348-1402)
        w1 = -ph.v_rec_n * (t1 - ph.ne) * jw;
        w2 = ph.v_rec_p * (t2 - ph.pe) * jw;
      }
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        const double fa = fph[a * NF + q];
        a1v[a] += fa * w1;
        a2v[a] += fa * w2;
      }
    }
  }
  const size_t jxo = (size_t)c * NB, jyo = (size_t)nc * NB + (size_t)c * NB, uo = 2 * (size_t)nc * NB + (size_t)c * NB;
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    b1[jxo + a] = a1p[a];
    b1[jyo + a] = a1p[NB + a];
    b1[uo + a] = a1v[a];
    b2[jxo + a] = a2p[a];
    b2[jyo + a] = a2p[NB + a];
    b2[uo + a] = a2v[a];
  }
}

// ------------------------------------------------------------------------
// Schur complement transforms
// ------------------------------------------------------------------------
// W_M = mu A_M^{-1} f_p,M of cell M (the current part of carrier RHS b):
// the cell-local inverse mass block times the cell's two current components.
template <int K>
__device__ __forceinline__ void schur_w(const DevRegion& R, const double* __restrict__ b, double mu, int M,
                                        int ty, double* w) {
  constexpr int NB = Sz<K>::NB;
  const size_t nc = R.ncell;
  const double* Minv = R.table + (size_t)ty * R.stride;
  double fx[NB], fy[NB];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    fx[a] = b[(size_t)M * NB + a];
    fy[a] = b[nc * NB + (size_t)M * NB + a];
  }
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const double m = Minv[a * NB + t];
      sx += m * fx[t];
      sy += m * fy[t];
    }
    w[a] = mu * sx;
    w[NB + a] = mu * sy;
  }
}

// Cell of thread t.  UNI: cells come from a list grouped so that every
// wave's cells share the operator types the kernel reads (groups padded to
// whole waves with -1), and the table offsets are made wave-uniform: the
// table entries are then fetched once per wave through the scalar cache and
// used as scalar operands, instead of one vector load per lane and entry.
template <bool UNI>
__device__ __forceinline__ int schur_cell(const int* __restrict__ list, int n, int t) {
  if (UNI) return t < n ? list[t] : -1;
  return t < n ? t : -1;
}

template <bool UNI>
__device__ __forceinline__ int wave_type(const DevRegion& R, int M) {
  const int t = R.type[M];
  return UNI ? __builtin_amdgcn_readfirstlane(t) : t;
}

// One right-hand side of k_schur_rhs on a wave-uniform cell list (the table
// reads are scalar, so a pass per right-hand side costs no extra vector
// loads; it halves the live registers, 165 -> 97 VGPRs, which pays off when
// a batch of instances fills the chip and not for a single instance).
template <int K>
__device__ __forceinline__ void schur_rhs_one(const DevRegion& R, const double* __restrict__ b, double mu,
                                              double* __restrict__ g, int X) {
  constexpr int NB = Sz<K>::NB;
  const size_t nc = R.ncell;
  double a1[NB];
#pragma unroll
  for (int a = 0; a < NB; ++a) a1[a] = b[2 * nc * NB + (size_t)X * NB + a];
  for (int f = -1; f < 4; ++f) {
    const int M = f < 0 ? X : R.nbr[4 * (size_t)X + f];
    if (M < 0) continue;
    const int ty = wave_type<true>(R, M);
    const double* B = R.table + (size_t)ty * R.stride + NB * NB +
                      (f < 0 ? 0 : 2 * NB * NB + (f ^ 1) * 2 * NB * NB);
    double w1[2 * NB];
    schur_w<K>(R, b, mu, M, ty, w1);
    for (int i = 0; i < 2 * NB; ++i) {
#pragma unroll
      for (int a = 0; a < NB; ++a) a1[a] += B[i * NB + a] * w1[i];
    }
  }
  const int pb = R.perm_base[X];
#pragma unroll
  for (int a = 0; a < NB; ++a) g[pb + a] = a1[a];
}

// g = f_v + B^T W of cell X: its own block and its 4 face neighbours' blocks,
// each neighbour's W recomputed from its right-hand side (no W in memory)
template <int K, int NRHS, bool UNI, bool SPLIT>
__global__ __launch_bounds__(64) void k_schur_rhs(DevRegion R, const double* __restrict__ b1,
                                                  const double* __restrict__ b2, double mu1, double mu2,
                                                  double* __restrict__ g1, double* __restrict__ g2, int64_t sb,
                                                  int64_t sg) {
  constexpr int NB = Sz<K>::NB;
  const int X = schur_cell<UNI>(R.ucell_nbh, UNI ? R.nu_nbh : R.ncell, blockIdx.x * blockDim.x + threadIdx.x);
  if (X < 0) return;
  {
    const int64_t i = blockIdx.y;
    b1 += i * sb;
    b2 += i * sb;
    g1 += i * sg;
    g2 += i * sg;
  }
  if constexpr (UNI && SPLIT) {
    schur_rhs_one<K>(R, b1, mu1, g1, X);
    if (NRHS > 1) schur_rhs_one<K>(R, b2, mu2, g2, X);
    return;
  }
  const size_t nc = R.ncell;
  double a1[NB], a2[NB];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    a1[a] = b1[2 * nc * NB + (size_t)X * NB + a];
    a2[a] = (NRHS > 1) ? b2[2 * nc * NB + (size_t)X * NB + a] : 0.0;
  }
  for (int f = -1; f < 4; ++f) {
    const int M = f < 0 ? X : R.nbr[4 * (size_t)X + f];
    if (M < 0) continue;
    const int ty = wave_type<UNI>(R, M);
    const double* B = R.table + (size_t)ty * R.stride + NB * NB +
                      (f < 0 ? 0 : 2 * NB * NB + (f ^ 1) * 2 * NB * NB);
    double w1[2 * NB], w2[2 * NB];
    schur_w<K>(R, b1, mu1, M, ty, w1);
    if (NRHS > 1) schur_w<K>(R, b2, mu2, M, ty, w2);
    for (int i = 0; i < 2 * NB; ++i) {
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        a1[a] += B[i * NB + a] * w1[i];
        if (NRHS > 1) a2[a] += B[i * NB + a] * w2[i];
      }
    }
  }
  const int pb = R.perm_base[X];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    g1[pb + a] = a1[a];
    if (NRHS > 1) g2[pb + a] = a2[a];
  }
}

// One right-hand side of k_schur_recover on a wave-uniform cell list (the
// same operations as the kernel's per carrier; half the live registers, for
// batches that fill the chip).
template <int K>
__device__ __forceinline__ void schur_recover_one(const DevRegion& R, const double* __restrict__ b,
                                                  const double* __restrict__ x, double mu, double* __restrict__ o,
                                                  int M) {
  constexpr int NB = Sz<K>::NB;
  const size_t nc = R.ncell;
  const int ty = wave_type<true>(R, M);
  const double* tb = R.table + (size_t)ty * R.stride;
  double sv[2 * NB], u[NB];
  const int pb = R.perm_base[M];
#pragma unroll
  for (int a = 0; a < NB; ++a) u[a] = x[pb + a];
  for (int i = 0; i < 2 * NB; ++i) {
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) t += tb[NB * NB + i * NB + a] * u[a];
    sv[i] = t;
  }
  // u is final: stored now, so that it is not live through the rest
#pragma unroll
  for (int a = 0; a < NB; ++a) o[2 * nc * NB + (size_t)M * NB + a] = u[a];
  for (int f = 0; f < 4; ++f) {
    const int X = R.nbr[4 * (size_t)M + f];
    if (X < 0) continue;
    const int px = R.perm_base[X];
    double v[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) v[a] = x[px + a];
    const double* B = tb + NB * NB + 2 * NB * NB + f * 2 * NB * NB;
    for (int i = 0; i < 2 * NB; ++i) {
      double t = 0.0;
#pragma unroll
      for (int a = 0; a < NB; ++a) t += B[i * NB + a] * v[a];
      sv[i] += t;
    }
  }
  double w[2 * NB];
  schur_w<K>(R, b, mu, M, ty, w);
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      double t = 0.0;
#pragma unroll
      for (int q = 0; q < NB; ++q) t += tb[a * NB + q] * sv[h * NB + q];
      o[(size_t)h * nc * NB + (size_t)M * NB + a] = w[h * NB + a] - mu * t;
    }
}

// q = W - A^{-1} B u; writes [Jx | Jy | u] (W recomputed from the RHS b)
template <int K, int NRHS, bool UNI, bool SPLIT>
__global__ __launch_bounds__(64) void k_schur_recover(DevRegion R, const double* __restrict__ b1,
                                                      const double* __restrict__ b2, const double* __restrict__ x1,
                                                      const double* __restrict__ x2, double mu1, double mu2,
                                                      double* __restrict__ o1, double* __restrict__ o2, int64_t sb,
                                                      int64_t sx, int64_t so) {
  constexpr int NB = Sz<K>::NB;
  const int M = schur_cell<UNI>(R.ucell_own, UNI ? R.nu_own : R.ncell, blockIdx.x * blockDim.x + threadIdx.x);
  if (M < 0) return;
  {
    const int64_t i = blockIdx.y;
    b1 += i * sb;
    b2 += i * sb;
    x1 += i * sx;
    x2 += i * sx;
    o1 += i * so;
    o2 += i * so;
  }
  if constexpr (UNI && SPLIT) {
    schur_recover_one<K>(R, b1, x1, mu1, o1, M);
    if (NRHS > 1) schur_recover_one<K>(R, b2, x2, mu2, o2, M);
    return;
  }
  const size_t nc = R.ncell;
  const int ty = wave_type<UNI>(R, M);
  const double* tb = R.table + (size_t)ty * R.stride;
  double s1[2 * NB], s2[2 * NB], u1[NB], u2[NB];
  const int pb = R.perm_base[M];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    u1[a] = x1[pb + a];
    u2[a] = (NRHS > 1) ? x2[pb + a] : 0.0;
  }
  for (int i = 0; i < 2 * NB; ++i) {
    double t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      t1 += tb[NB * NB + i * NB + a] * u1[a];
      if (NRHS > 1) t2 += tb[NB * NB + i * NB + a] * u2[a];
    }
    s1[i] = t1;
    s2[i] = t2;
  }
  for (int f = 0; f < 4; ++f) {
    const int X = R.nbr[4 * (size_t)M + f];
    if (X < 0) continue;
    const int px = R.perm_base[X];
    double v1[NB], v2[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      v1[a] = x1[px + a];
      v2[a] = (NRHS > 1) ? x2[px + a] : 0.0;
    }
    const double* B = tb + NB * NB + 2 * NB * NB + f * 2 * NB * NB;
    for (int i = 0; i < 2 * NB; ++i) {
      double t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        t1 += B[i * NB + a] * v1[a];
        if (NRHS > 1) t2 += B[i * NB + a] * v2[a];
      }
      s1[i] += t1;
      s2[i] += t2;
    }
  }
  double w1[2 * NB], w2[2 * NB];
  schur_w<K>(R, b1, mu1, M, ty, w1);
  if (NRHS > 1) schur_w<K>(R, b2, mu2, M, ty, w2);
  const double* Minv = tb;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      double t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        t1 += Minv[a * NB + t] * s1[h * NB + t];
        if (NRHS > 1) t2 += Minv[a * NB + t] * s2[h * NB + t];
      }
      const size_t off = (size_t)h * nc * NB + (size_t)M * NB + a;
      o1[off] = w1[h * NB + a] - mu1 * t1;
      if (NRHS > 1) o2[off] = w2[h * NB + a] - mu2 * t2;
    }
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    o1[2 * nc * NB + (size_t)M * NB + a] = u1[a];
    if (NRHS > 1) o2[2 * nc * NB + (size_t)M * NB + a] = u2[a];
  }
}

// ------------------------------------------------------------------------
// Mixed precision (FP32 element contractions, FP64 vectors and solves)
// ------------------------------------------------------------------------
// The cell-local contractions with the cell-type tables run in FP32 on an
// FP32 copy of the table, with FP32 accumulation; inputs are read and
// results written in FP64, and the terms that are not contractions (the
// density part f_v of the Schur right-hand side, the difference f - B u
// before the recovery's inverse mass) stay FP64.  Wave-uniform cell lists
// only (planar meshes): the table entries are scalar operands.
template <int K>
__device__ __forceinline__ void schur_rhs_one_f32(const DevRegion& R, const double* __restrict__ b, double mu,
                                                  double* __restrict__ g, int X) {
  constexpr int NB = Sz<K>::NB;
  const size_t nc = R.ncell;
  const float muf = (float)mu;
  float acc[NB];
#pragma unroll
  for (int a = 0; a < NB; ++a) acc[a] = 0.0f;
  for (int f = -1; f < 4; ++f) {
    const int M = f < 0 ? X : R.nbr[4 * (size_t)X + f];
    if (M < 0) continue;
    const int ty = wave_type<true>(R, M);
    const float* Minv = R.table_f + (size_t)ty * R.stride;
    const float* B = Minv + NB * NB + (f < 0 ? 0 : 2 * NB * NB + (f ^ 1) * 2 * NB * NB);
    float fx[NB], fy[NB], w[2 * NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      fx[a] = (float)b[(size_t)M * NB + a];
      fy[a] = (float)b[nc * NB + (size_t)M * NB + a];
    }
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      float sx = 0.0f, sy = 0.0f;
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        sx += Minv[a * NB + t] * fx[t];
        sy += Minv[a * NB + t] * fy[t];
      }
      w[a] = muf * sx;
      w[NB + a] = muf * sy;
    }
    for (int i = 0; i < 2 * NB; ++i) {
#pragma unroll
      for (int a = 0; a < NB; ++a) acc[a] += B[i * NB + a] * w[i];
    }
  }
  const int pb = R.perm_base[X];
#pragma unroll
  for (int a = 0; a < NB; ++a) g[pb + a] = b[2 * nc * NB + (size_t)X * NB + a] + (double)acc[a];
}

template <int K, int NRHS>
__global__ __launch_bounds__(64) void k_schur_rhs_f32(DevRegion R, const double* __restrict__ b1,
                                                      const double* __restrict__ b2, double mu1, double mu2,
                                                      double* __restrict__ g1, double* __restrict__ g2, int64_t sb,
                                                      int64_t sg) {
  const int X = schur_cell<true>(R.ucell_nbh, R.nu_nbh, blockIdx.x * blockDim.x + threadIdx.x);
  if (X < 0) return;
  const int64_t i = blockIdx.y;
  schur_rhs_one_f32<K>(R, b1 + i * sb, mu1, g1 + i * sg, X);
  if (NRHS > 1) schur_rhs_one_f32<K>(R, b2 + i * sb, mu2, g2 + i * sg, X);
}

template <int K>
__device__ __forceinline__ void schur_recover_one_f32(const DevRegion& R, const float* tb,
                                                      const double* __restrict__ b, const double* __restrict__ x,
                                                      double mu, double* __restrict__ o, int M) {
  constexpr int NB = Sz<K>::NB;
  const size_t nc = R.ncell;
  float s[2 * NB], u[NB];
  double ud[NB];
  const int pb = R.perm_base[M];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    ud[a] = x[pb + a];
    u[a] = (float)ud[a];
  }
  for (int i = 0; i < 2 * NB; ++i) {
    float t = 0.0f;
#pragma unroll
    for (int a = 0; a < NB; ++a) t += tb[NB * NB + i * NB + a] * u[a];
    s[i] = t;
  }
  for (int f = 0; f < 4; ++f) {
    const int X = R.nbr[4 * (size_t)M + f];
    if (X < 0) continue;
    const int px = R.perm_base[X];
    float v[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) v[a] = (float)x[px + a];
    const float* B = tb + NB * NB + 2 * NB * NB + f * 2 * NB * NB;
    for (int i = 0; i < 2 * NB; ++i) {
      float t = 0.0f;
#pragma unroll
      for (int a = 0; a < NB; ++a) t += B[i * NB + a] * v[a];
      s[i] += t;
    }
  }
  // q = mu A^{-1} (f_q - B u): the difference in FP64, the inverse mass in FP32
  const float muf = (float)mu;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    float d[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) d[t] = (float)(b[(size_t)h * nc * NB + (size_t)M * NB + t] - (double)s[h * NB + t]);
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      float t1 = 0.0f;
#pragma unroll
      for (int t = 0; t < NB; ++t) t1 += tb[a * NB + t] * d[t];
      o[(size_t)h * nc * NB + (size_t)M * NB + a] = (double)(muf * t1);
    }
  }
#pragma unroll
  for (int a = 0; a < NB; ++a) o[2 * nc * NB + (size_t)M * NB + a] = ud[a];
}

template <int K, int NRHS>
__global__ __launch_bounds__(64) void k_schur_recover_f32(DevRegion R, const double* __restrict__ b1,
                                                          const double* __restrict__ b2,
                                                          const double* __restrict__ x1,
                                                          const double* __restrict__ x2, double mu1, double mu2,
                                                          double* __restrict__ o1, double* __restrict__ o2,
                                                          int64_t sb, int64_t sx, int64_t so) {
  const int M = schur_cell<true>(R.ucell_own, R.nu_own, blockIdx.x * blockDim.x + threadIdx.x);
  if (M < 0) return;
  const int64_t i = blockIdx.y;
  const float* tb = R.table_f + (size_t)wave_type<true>(R, M) * R.stride;
  schur_recover_one_f32<K>(R, tb, b1 + i * sb, x1 + i * sx, mu1, o1 + i * so, M);
  if (NRHS > 1) schur_recover_one_f32<K>(R, tb, b2 + i * sb, x2 + i * sx, mu2, o2 + i * so, M);
}

// ------------------------------------------------------------------------
// Supernodal triangular solves (S = L L^T), level by level
// ------------------------------------------------------------------------
// One workgroup per tile of rows of one supernode.  The vector the rows are
// dotted with is staged in LDS in chunks of CW entries per right-hand side;
// W lanes share a row (256/W rows per pass), partial sums stay in registers
// across chunks.
//   forward  (FWD): vector t_own (= b_own - updates z owed by descendants,
//                   from k_fwd_gather), rows of [Linv; W21]; outputs
//                   y_own -> Y, z -> Z.
//   backward:       vector [y_own; x at the panel rows], rows
//                   [Linv^T(j, j:) | -W21^T(j, :)]; outputs x_own -> X.
constexpr int CW = 1024;

template <int W>
__device__ __forceinline__ double group_sum(double x) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, W);
  return x;
}

// W lanes per row, 256/W rows per pass, MAXR passes: a tile holds at most
// (256/W) * MAXR rows of one supernode.
template <int W, int MAXR_>
struct SweepShape {
  static constexpr int R = 256 / W;
  static constexpr int MAXR = MAXR_;
  static constexpr int RB = R * MAXR;
};

// forward, part 1: t_p = b_p - sum of the updates z owed to row p, into X.
// Vectors are h-major (b[h n + p]); with two right-hand sides the updates are
// pair-interleaved (Z[2 z + h]), so one 16-byte load fetches both.  L lanes
// per row (L = 1, or 8 on levels with few rows: their rows collect many
// descendants' contributions); the lanes' partial sums are combined by
// shuffles in a fixed order.
template <int L, int NRHS>
__global__ __launch_bounds__(256) void k_fwd_gather(DevFactor F, const int* __restrict__ lrows, int nrow,
                                                    const double* __restrict__ b) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / L, lane = t % L;
  if (i >= nrow) return;  // nrow * L threads: whole L-groups exit together
  const int p = lrows[i];
  const int64_t c1 = F.cptr[p + 1];
  double z0 = 0.0, z1 = 0.0;
  for (int64_t c = F.cptr[p] + lane; c < c1; c += L) {
    if (NRHS == 2) {
      const double2 z = *reinterpret_cast<const double2*>(F.Z + 2 * F.czidx[c]);
      z0 += z.x;
      z1 += z.y;
    } else {
      z0 += F.Z[F.czidx[c]];
    }
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) {
    z0 += __shfl_xor(z0, o, L);
    if (NRHS == 2) z1 += __shfl_xor(z1, o, L);
  }
  if (lane == 0) {
    F.X[p] = b[p] - z0;
    if (NRHS == 2) F.X[(int64_t)F.n + p] = b[(int64_t)F.n + p] - z1;
  }
}

// PF: the factor loads of the next pair of positions are issued before this
// pair's products (unconditional loads from clamped positions, masked at use,
// two register sets in turn).  Costlier in registers and issue slots, it
// pays on the levels with few tiles, whose time is the latency of each
// lane's chain of loads; the tuner picks it per level.
template <int NRHS, int W, int MAXR_, bool FWD, bool PF = false>
__global__ __launch_bounds__(256) void k_sweep(DevFactor F, const int4* __restrict__ tiles,
                                               const double* __restrict__ b0, const double* __restrict__ b1,
                                               double* __restrict__ Y0, double* __restrict__ Y1) {
  using SS = SweepShape<W, MAXR_>;
  constexpr int R = SS::R, MAXR = SS::MAXR;
  __shared__ double vec[NRHS][CW];
  const int4 tl = tiles[blockIdx.x];
  const int s = tl.x, row0 = tl.y, nr = tl.z;
  const int ns = F.ns[s], ms = F.ms[s], p0 = F.p0[s];
  const int VL = FWD ? ns : ns + ms;
  const int g = threadIdx.x / W, sl = threadIdx.x % W;
  const double* __restrict__ base = FWD ? F.fwd + F.fwd_off[s] : F.bwd + F.bwd_off[s];
  // The group's MAXR rows are walked together: every column step issues one
  // load per row, so each lane keeps 2*MAXR independent loads in flight and
  // reads each staged vector entry from LDS once for all its rows.  Row q
  // holds vector positions [rs[q], re[q]), its element at position p at
  // base[off[q] + p].
  int64_t off[MAXR];
  int rs[MAXR], re[MAXR];
  int lo = VL, hi = 0;
#pragma unroll
  for (int q = 0; q < MAXR; ++q) {
    const int lr = row0 + g + q * R;
    if (lr < row0 + nr) {
      if (FWD) {
        rs[q] = 0;
        if (lr < ns) {
          off[q] = (int64_t)lr * (lr + 1) / 2;
          re[q] = lr + 1;
        } else {
          off[q] = (int64_t)ns * (ns + 1) / 2 + (int64_t)(lr - ns) * ns;
          re[q] = ns;
        }
      } else {
        off[q] = (int64_t)lr * (ns + ms) - (int64_t)lr * (lr - 1) / 2 - lr;
        rs[q] = lr;
        re[q] = ns + ms;
      }
      lo = min(lo, rs[q]);
      hi = max(hi, re[q]);
    } else {
      off[q] = 0;
      rs[q] = 0;
      re[q] = 0;
    }
  }
  double acc[MAXR][NRHS];
#pragma unroll
  for (int q = 0; q < MAXR; ++q)
#pragma unroll
    for (int h = 0; h < NRHS; ++h) acc[q][h] = 0.0;
  // backward rows start at their own index, so no row of the tile uses
  // positions below row0
  const int tlo = FWD ? 0 : row0;
  for (int c0 = tlo; c0 < VL; c0 += CW) {
    const int cl = min(CW, VL - c0);
    __syncthreads();
    for (int i = threadIdx.x; i < cl; i += 256) {
      const int pos = c0 + i;
      if (FWD) {  // t from X, or straight from the right-hand side b0 (leaf level: nothing to gather)
        const double* __restrict__ xs = b0 ? b0 : F.X;
        vec[0][i] = xs[p0 + pos];
        if (NRHS > 1) vec[NRHS - 1][i] = xs[F.n + p0 + pos];
      } else if (pos < ns) {
        vec[0][i] = Y0[p0 + pos];
        if (NRHS > 1) vec[NRHS - 1][i] = Y1[p0 + pos];
      } else {  // x at the panel rows: rows of ancestors, solved by earlier launches
        const int64_t p = F.rows[F.rows_off[s] + (pos - ns)];
        vec[0][i] = F.X[p];
        if (NRHS > 1) vec[NRHS - 1][i] = F.X[F.n + p];
      }
    }
    __syncthreads();
    const int e = min(hi, c0 + cl);
    if constexpr (PF) {
      auto ld = [&](int p, double (&L)[MAXR]) {
#pragma unroll
        for (int q = 0; q < MAXR; ++q) {
          const int pc = re[q] > rs[q] ? min(max(p, rs[q]), re[q] - 1) : 0;
          L[q] = __builtin_nontemporal_load(base + off[q] + pc);
        }
      };
      auto fma_pair = [&](int pa, const double (&La)[MAXR], const double (&Lb)[MAXR]) {
        const int pb = pa + W;
        const bool okb = pb < e;
        const int ib = okb ? pb - c0 : 0;
        const double va0 = vec[0][pa - c0];
        const double vb0 = okb ? vec[0][ib] : 0.0;
        const double va1 = NRHS > 1 ? vec[NRHS - 1][pa - c0] : 0.0;
        const double vb1 = NRHS > 1 && okb ? vec[NRHS - 1][ib] : 0.0;
#pragma unroll
        for (int q = 0; q < MAXR; ++q) {
          const double a = (pa >= rs[q] && pa < re[q]) ? La[q] : 0.0;
          const double b = (pb >= rs[q] && pb < re[q]) ? Lb[q] : 0.0;
          acc[q][0] += a * va0 + b * vb0;
          if (NRHS > 1) acc[q][NRHS - 1] += a * va1 + b * vb1;
        }
      };
      int pa = max(lo, c0) + sl;
      if (pa < e) {
        double La[MAXR], Lb[MAXR], Ma[MAXR], Mb[MAXR];
        ld(pa, La);
        ld(pa + W, Lb);
        for (; pa < e; pa += 4 * W) {
          ld(pa + 2 * W, Ma);
          ld(pa + 3 * W, Mb);
          __builtin_amdgcn_sched_barrier(0);
          fma_pair(pa, La, Lb);
          __builtin_amdgcn_sched_barrier(0);
          ld(pa + 4 * W, La);
          ld(pa + 5 * W, Lb);
          __builtin_amdgcn_sched_barrier(0);
          if (pa + 2 * W < e) fma_pair(pa + 2 * W, Ma, Mb);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      continue;
    }
    for (int pa = max(lo, c0) + sl; pa < e; pa += 2 * W) {
      const int pb = pa + W;
      const bool okb = pb < e;
      double La[MAXR], Lb[MAXR];
#pragma unroll
      for (int q = 0; q < MAXR; ++q) {
        La[q] = (pa >= rs[q] && pa < re[q]) ? __builtin_nontemporal_load(base + off[q] + pa) : 0.0;
        Lb[q] = (okb && pb >= rs[q] && pb < re[q]) ? __builtin_nontemporal_load(base + off[q] + pb) : 0.0;
      }
      const double va0 = vec[0][pa - c0];
      const double vb0 = okb ? vec[0][pb - c0] : 0.0;
#pragma unroll
      for (int q = 0; q < MAXR; ++q) acc[q][0] += La[q] * va0 + Lb[q] * vb0;
      if (NRHS > 1) {
        const double va1 = vec[NRHS - 1][pa - c0];
        const double vb1 = okb ? vec[NRHS - 1][pb - c0] : 0.0;
#pragma unroll
        for (int q = 0; q < MAXR; ++q) acc[q][NRHS - 1] += La[q] * va1 + Lb[q] * vb1;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < MAXR; ++q) {
    const int lr = row0 + g + q * R;
    double r0 = group_sum<W>(acc[q][0]);
    double r1 = (NRHS > 1) ? group_sum<W>(acc[q][NRHS - 1]) : 0.0;
    if (lr >= row0 + nr || sl != 0) continue;
    if (FWD) {
      if (lr < ns) {
        Y0[p0 + lr] = r0;
        if (NRHS > 1) Y1[p0 + lr] = r1;
      } else {
        const int64_t z = F.z_off[s] + (lr - ns);
        if (NRHS > 1) {
          *reinterpret_cast<double2*>(F.Z + 2 * z) = double2{r0, r1};
        } else {
          F.Z[z] = r0;
        }
      }
    } else {
      F.X[p0 + lr] = r0;
      if (NRHS > 1) F.X[F.n + p0 + lr] = r1;
    }
  }
}

// Many right-hand sides (a batch of instances sharing the factor): the row
// tile times the vector block is a skinny GEMM, done with FP64 MFMA
// (v_mfma_f64_16x16x4f64).  These solves keep their vectors RHS-interleaved
// (entry p of right-hand side h at [p * NRP + h], NRP = 16 or 32 padded
// right-hand sides, the padding zero), so the B operand of a 16-column step is
// four 128-byte rows straight from L2: no LDS staging and no barriers.  Each
// wave owns MB blocks of 16 rows; per step a lane loads 4 consecutive entries
// of one row (lane l: row l%16, columns 4(l/16)..4(l/16)+3) and issues 4 MFMAs
// whose B operand holds the vector at the same columns (lane l: column
// 4(l/16)+e, right-hand side l%16), so the k order of A and B agree.
// Accumulator lane l, register r holds row l/16 + 4r, right-hand side l%16.
typedef double v4d __attribute__((ext_vector_type(4)));

// The interleaved vectors from and to one vector per right-hand side (entry
// p of right-hand side h at [h * n + p]): a tile of IL_TP rows passes through
// LDS, so that the per-right-hand-side side moves IL_TP consecutive rows
// (256 bytes) per wave access and the interleaved side whole rows of NRP
// entries.  The padding columns (h >= nrhs) are written as zeros.
constexpr int IL_TP = 32;

template <int NRP>
__global__ __launch_bounds__(256) void k_pack_il(DevFactor F, const double* __restrict__ v, int nrhs) {
  __shared__ double s[NRP * (IL_TP + 1)];
  const int64_t n = F.n, p0 = (int64_t)blockIdx.x * IL_TP;
  for (int e = threadIdx.x; e < NRP * IL_TP; e += 256) {
    const int h = e / IL_TP, pp = e % IL_TP;
    s[h * (IL_TP + 1) + pp] = (h < nrhs && p0 + pp < n) ? v[(int64_t)h * n + p0 + pp] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < IL_TP * NRP / 2; e += 256) {
    const int pp = e / (NRP / 2), h = 2 * (e % (NRP / 2));
    if (p0 + pp < n)
      *reinterpret_cast<double2*>(F.X + (p0 + pp) * NRP + h) =
          double2{s[h * (IL_TP + 1) + pp], s[(h + 1) * (IL_TP + 1) + pp]};
  }
}

template <int NRP>
__global__ __launch_bounds__(256) void k_unpack_il(DevFactor F, double* __restrict__ v, int nrhs) {
  __shared__ double s[NRP * (IL_TP + 1)];
  const int64_t n = F.n, p0 = (int64_t)blockIdx.x * IL_TP;
  for (int e = threadIdx.x; e < IL_TP * NRP / 2; e += 256) {
    const int pp = e / (NRP / 2), h = 2 * (e % (NRP / 2));
    if (p0 + pp < n) {
      const double2 x = *reinterpret_cast<const double2*>(F.X + (p0 + pp) * NRP + h);
      s[h * (IL_TP + 1) + pp] = x.x;
      s[(h + 1) * (IL_TP + 1) + pp] = x.y;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nrhs * IL_TP; e += 256) {
    const int h = e / IL_TP, pp = e % IL_TP;
    if (p0 + pp < n) v[(int64_t)h * n + p0 + pp] = s[h * (IL_TP + 1) + pp];
  }
}

// t_p = b_p - sum of the updates z owed to row p, in place on the packed
// right-hand sides (X holds b on entry; rows owed nothing are left alone).
// A thread takes two adjacent right-hand sides (16-byte loads and stores).
template <int NRP>
__global__ __launch_bounds__(256) void k_fwd_gather_il(DevFactor F, const int* __restrict__ lrows, int nrow,
                                                       int nrhs) {
  constexpr int NP = NRP / 2;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)nrow * NP) return;
  const int i = (int)(e / NP), h = 2 * (int)(e % NP);
  const int p = lrows[i];
  const int64_t c0 = F.cptr[p], c1 = F.cptr[p + 1];
  if (c0 == c1 || h >= nrhs) return;
  double2* x = reinterpret_cast<double2*>(F.X + (int64_t)p * NRP + h);
  double2 a = *x;
  // four contributions' loads in flight at a time, subtracted in order
  int64_t c = c0;
  for (; c + 4 <= c1; c += 4) {
    double2 z[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) z[u] = *reinterpret_cast<const double2*>(F.Z + F.czidx[c + u] * NRP + h);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a.x -= z[u].x;
      a.y -= z[u].y;
    }
  }
  for (; c < c1; ++c) {
    const double2 z = *reinterpret_cast<const double2*>(F.Z + F.czidx[c] * NRP + h);
    a.x -= z.x;
    a.y -= z.y;
  }
  if (h + 1 >= nrhs) a.y = 0.0;
  *x = a;
}

template <int NRP>
__global__ __launch_bounds__(256) void k_bwd_gather_il(DevFactor F, const int64_t* __restrict__ idx, int nidx) {
  constexpr int NP = NRP / 2;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)nidx * NP) return;
  const int i = (int)(e / NP), h = 2 * (int)(e % NP);
  const int64_t q = idx[i];
  *reinterpret_cast<double2*>(F.Z + q * NRP + h) =
      *reinterpret_cast<const double2*>(F.X + (int64_t)F.rows[q] * NRP + h);
}

// ks > 0: the tile is column segment tl.w (columns [tl.w ks, (tl.w + 1) ks))
// of its row tile and writes its partial sums to Part (slot blockIdx.x, rows
// relative to row0, 64 MB rows per slot); k_sweep_reduce adds the segments.
// Single-block tiles (MB = 1) with up to 32 right-hand sides ask for 4 waves
// per SIMD: the forward sweep then fits its accumulators in the VGPRs (111,
// no spills) instead of taking 130 registers with AGPRs, which allowed only 3
// waves; the backward sweep goes from 4 to 5 waves the same way.  With 64 or
// 128 right-hand sides the accumulators (NRP / 4 doubles per block) and the
// double-buffered vector operands need the full register file.
// HPW < NRP / 16 splits the right-hand-side groups of a row block over
// NRP / 16 / HPW waves (each loads the row block's factor and its own groups'
// vector entries): a workgroup then covers 64 MB HPW 16 / NRP rows, with
// fewer registers per wave and so more waves in flight.
constexpr int mfma_waves_per_eu(int nrp, int mb, int hpw) {
  if (hpw == nrp / 16) return mb == 1 && nrp <= 32 ? 4 : 1;
  return hpw == 1 ? (mb == 1 ? 6 : mb == 2 ? 4 : 2) : (mb == 4 ? 2 : 1);
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b % 8), each XCD with its own L2.  Tile of workgroup b such that every XCD
// takes a contiguous run of the tile list: the tiles of one node (which
// stage the same vector rows) then share an L2.  A bijection on [0, n).
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n / 8, r = n % 8, x = b % 8;
  return x * q + min(x, r) + b / 8;
}

template <int NRP, int MB, bool FWD, int HPW = NRP / 16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mfma_waves_per_eu(NRP, MB, HPW)))) void k_sweep_mfma(
    DevFactor F, const int4* __restrict__ tiles, double* __restrict__ Y, int ks, double* __restrict__ Part) {
  constexpr int NH = HPW;                  // right-hand-side groups of this wave
  constexpr int WPR = NRP / 16 / HPW;      // waves per row block
  static_assert(WPR >= 1 && 4 % WPR == 0, "a row block's waves must divide the workgroup");
  const int ti = xcd_tile(blockIdx.x, gridDim.x);
  const int4 tl = tiles[ti];
  const int s = tl.x, row0 = tl.y, nr = tl.z;  // row0 is a multiple of 16
  const int ns = F.ns[s], ms = F.ms[s], p0 = F.p0[s];
  const int VL = FWD ? ns : ns + ms;
  // the wave index through readfirstlane: everything derived from it (rows,
  // column range, step count) is then scalar and the step loop's control flow
  // is uniform, so the compiler counts outstanding loads exactly
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ri = lane & 15, kq = lane >> 4;
  const int hg0 = (wave % WPR) * HPW;                    // this wave's first right-hand-side group
  const int first = row0 + (wave / WPR) * MB * 16;       // wave-uniform
  if (first >= row0 + nr) return;                        // no rows for this wave (no barriers below)
  const int last = min(row0 + nr, first + MB * 16) - 1;
  const int lo = FWD ? 0 : first;
  const int hi = FWD ? (last < ns ? last + 1 : ns) : ns + ms;
  // the wave's blocks in the MFMA-blocked copy: data base, first column, steps
  const double* __restrict__ blk = FWD ? F.fblk : F.bblk;
  const int64_t* __restrict__ boff = FWD ? F.fblk_off : F.bblk_off;
  const int bnode = FWD ? F.fblk0[s] : F.bblk0[s];
  int64_t bbase[MB];
  int bclo[MB], bsteps[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    const int b = first / 16 + mb;
    bbase[mb] = 0;
    bclo[mb] = 0;
    bsteps[mb] = 0;
    if (first + mb * 16 < row0 + nr) {
      bbase[mb] = boff[bnode + b];
      bclo[mb] = FWD ? 0 : 16 * b;
      bsteps[mb] = (int)((boff[bnode + b + 1] - bbase[mb]) >> 8);
    }
  }
  v4d acc[MB][NH];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int hg = 0; hg < NH; ++hg) acc[mb][hg] = v4d{0.0, 0.0, 0.0, 0.0};
  // vector entry at column col of right-hand side h (interleaved storage).
  // Columns past the rows' last one (col >= VL, in the last step) read the
  // last column's entry instead: the factor's entries there are zero and the
  // vector entries are finite, so the products vanish.  The loads then carry
  // no per-lane branch, and the compiler issues them back to back instead of
  // waiting on each one to select between it and zero.
  const int64_t zrow = FWD ? 0 : F.rows_off[s] - ns;
  auto vecv = [&](int col, int h) -> double {
    const int c = min(col, VL - 1);
    if (FWD) return F.X[(int64_t)(p0 + c) * NRP + h];
    const double* src = c < ns ? Y + (int64_t)(p0 + c) * NRP : F.Z + (zrow + c) * NRP;
    return src[h];
  };
  int c0 = (lo >> 4) << 4, cend = hi;
  if (ks > 0) {  // this tile's column segment (ks and the segment start are multiples of 16)
    c0 = max(c0, tl.w * ks);
    cend = min(cend, (tl.w + 1) * ks);
  }
  const int nstep = cend > c0 ? (cend - c0 + 15) >> 4 : 0;
  // Step st's operands.  Every load is unconditional: a block outside its
  // stored columns at this step (only with MB > 1) loads its nearest stored
  // step and is masked to zero where the MFMA uses it (inb), so no load waits
  // behind a select.
  auto load = [&](int st, double (&L)[MB][4], double (&Bv)[4][NH], bool (&inb)[MB]) {
    const int cc = c0 + 16 * st;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int idx = (cc - bclo[mb]) >> 4;  // the block's own step (wave-uniform)
      inb[mb] = cc >= bclo[mb] && idx < bsteps[mb];
      // a block past the tile's rows (bsteps 0) reads block 0's first step,
      // which exists: the wave has rows, so block 0 has at least one step
      const int ic = min(max(idx, 0), max(bsteps[mb] - 1, 0));
      const double* src = blk + (bsteps[mb] > 0 ? bbase[mb] + (int64_t)ic * 256 : bbase[0]) + lane;
#pragma unroll
      for (int q = 0; q < 4; ++q) L[mb][q] = __builtin_nontemporal_load(src + q * 64);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int col = cc + 4 * kq + q;
#pragma unroll
      for (int hg = 0; hg < NH; ++hg) Bv[q][hg] = vecv(col, (hg0 + hg) * 16 + ri);
    }
  };
  auto mma = [&](const double (&L)[MB][4], const double (&Bv)[4][NH], const bool (&inb)[MB]) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int hg = 0; hg < NH; ++hg)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          acc[mb][hg] = __builtin_amdgcn_mfma_f64_16x16x4f64(MB == 1 || inb[mb] ? L[mb][q] : 0.0, Bv[q][hg],
                                                             acc[mb][hg], 0, 0, 0);
  };
  // loads run one step ahead of the MFMAs that use them (two register sets);
  // past the last step they repeat it (unused), so the loop body has no
  // conditional loads
  double LA[MB][4], LB[MB][4], BA[4][NH], BB[4][NH];
  bool IA[MB], IB[MB];
  // (sched_barrier keeps each step's loads ahead of the previous step's MFMAs:
  // left alone, the scheduler sinks them next to their use)
  if (nstep > 0) {
    load(0, LA, BA, IA);
    for (int st = 0; st < nstep; st += 2) {
      load(min(st + 1, nstep - 1), LB, BB, IB);
      __builtin_amdgcn_sched_barrier(0);
      mma(LA, BA, IA);
      __builtin_amdgcn_sched_barrier(0);
      load(min(st + 2, nstep - 1), LA, BA, IA);
      __builtin_amdgcn_sched_barrier(0);
      if (st + 1 < nstep) mma(LB, BB, IB);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int hg = 0; hg < NH; ++hg)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = first + mb * 16 + kq + 4 * r;
        const int h = (hg0 + hg) * 16 + ri;
        if (lr >= row0 + nr) continue;
        const double v = acc[mb][hg][r];
        if (ks > 0) {
          Part[((int64_t)ti * (64 * MB) + (lr - row0)) * NRP + h] = v;
        } else if (FWD) {
          if (lr < ns) Y[(int64_t)(p0 + lr) * NRP + h] = v;
          else F.Z[(F.z_off[s] + (lr - ns)) * NRP + h] = v;
        } else {
          F.X[(int64_t)(p0 + lr) * NRP + h] = v;
        }
      }
}

// The MFMA sweep with the vector operand staged in LDS once per workgroup:
// its 4 waves each own MBL consecutive 16-row blocks of the tile (64 MBL
// rows) and all right-hand-side groups, and step through the union of their
// column ranges together.  Per 16-column step the workgroup loads the step's
// 16 x NRP vector entries once (NRP / 16 per thread, 16-byte pieces), stores them
// in LDS and every wave reads its B operands there, each B operand serving
// the wave's MBL blocks: the vector bytes leaving L2 drop to 1 / (4 MBL) of
// per-block loads.  The next step's vector entries (and the factor operands
// of the step after it) are loaded into registers while this step's MFMAs
// run, with one barrier per step (double-buffered LDS).  The tile order is
// XCD-aware (xcd_tile).  A
// block contributes zero on steps outside its stored columns (masked factor
// operand), and a wave skips the steps none of its blocks stores (wave-uniform
// through readfirstlane); a wave without rows still loads, stores and meets
// every barrier.
template <int NRP, bool FWD, int MBL>
__global__ __launch_bounds__(256) void k_sweep_mfma_lds(DevFactor F, const int4* __restrict__ tiles,
                                                        double* __restrict__ Y, int ks,
                                                        double* __restrict__ Part) {
  constexpr int NH = NRP / 16;
  constexpr int S = NRP + 4;   // LDS row stride (doubles): the 4 column groups of a read fall in 2 bank halves
  constexpr int K = NRP / 16;  // staged entries per thread per step
  __shared__ __align__(16) double sB[2][16 * S];
  const int ti = xcd_tile(blockIdx.x, gridDim.x);
  const int4 tl = tiles[ti];
  const int s = tl.x, row0 = tl.y, nr = tl.z;  // row0 is a multiple of 16
  const int ns = F.ns[s], ms = F.ms[s], p0 = F.p0[s];
  const int VL = FWD ? ns : ns + ms;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ri = lane & 15, kq = lane >> 4;
  const int first = row0 + wave * 16 * MBL;
  const int last_t = row0 + nr - 1;  // the tile's last row
  // the workgroup's column range (the union of its blocks')
  const int lo = FWD ? 0 : row0;
  const int hi = FWD ? (last_t < ns ? last_t + 1 : ns) : ns + ms;
  int c0 = (lo >> 4) << 4, cend = hi;
  if (ks > 0) {
    c0 = max(c0, tl.w * ks);
    cend = min(cend, (tl.w + 1) * ks);
  }
  const int nstep = cend > c0 ? (cend - c0 + 15) >> 4 : 0;
  // this wave's blocks in the MFMA-blocked copy
  const double* __restrict__ blk = FWD ? F.fblk : F.bblk;
  const int64_t* __restrict__ boff = FWD ? F.fblk_off : F.bblk_off;
  const int bnode = FWD ? F.fblk0[s] : F.bblk0[s];
  bool rows[MBL];
  int64_t bbase[MBL];
  int bclo[MBL], bsteps[MBL];
#pragma unroll
  for (int j = 0; j < MBL; ++j) {
    const int b = first / 16 + j;
    rows[j] = first + 16 * j < row0 + nr;  // wave-uniform
    bbase[j] = rows[j] ? boff[bnode + b] : boff[bnode];
    bclo[j] = rows[j] ? (FWD ? 0 : 16 * b) : 0;
    bsteps[j] = rows[j] ? (int)((boff[bnode + b + 1] - bbase[j]) >> 8) : 0;
  }
  const int64_t zrow = FWD ? 0 : F.rows_off[s] - ns;
  // staged entries of thread t at step st: column c = t / 16 of the step,
  // right-hand-side pairs 2 (t % 16) + 32 jj (jj < K / 2), so that each
  // 16-lane group moves 256 contiguous bytes per instruction, from the
  // interleaved vector and into LDS (no bank conflicts on the stores); past
  // the rows' last column the last column's entries (finite; the factor is
  // zero there).  The backward sweep reads its panel columns (the
  // ancestors' rows, solved by earlier launches) straight from x through the
  // row index: no gathered copy.
  const int sc = threadIdx.x >> 4, sl = 2 * (threadIdx.x & 15);
  auto stage_load = [&](int st, double (&R)[K]) {
    const int col = min(c0 + 16 * st + sc, VL - 1);
    const double* src;
    if (FWD) src = F.X + (int64_t)(p0 + col) * NRP + sl;
    else src = col < ns ? Y + (int64_t)(p0 + col) * NRP + sl : F.X + (int64_t)F.rows[zrow + col] * NRP + sl;
#pragma unroll
    for (int j = 0; j < K; j += 2) {
      const double2 v = *reinterpret_cast<const double2*>(src + 16 * j);
      R[j] = v.x;
      R[j + 1] = v.y;
    }
  };
  auto stage_store = [&](int buf, const double (&R)[K]) {
#pragma unroll
    for (int j = 0; j < K; j += 2)
      *reinterpret_cast<double2*>(&sB[buf][sc * S + sl + 16 * j]) = double2{R[j], R[j + 1]};
  };
  // the factor operands of step st (clamped into each block's stored steps;
  // in[j] = whether the step is one of block j's, wave-uniform)
  auto a_load = [&](int st, double (&L)[MBL][4], int (&in)[MBL]) {
#pragma unroll
    for (int j = 0; j < MBL; ++j) {
      const int idx = (c0 + 16 * st - bclo[j]) >> 4;
      in[j] = __builtin_amdgcn_readfirstlane((int)(rows[j] && idx >= 0 && idx < bsteps[j]));
      const int ic = min(max(idx, 0), max(bsteps[j] - 1, 0));
      const double* src = blk + bbase[j] + (int64_t)ic * 256 + lane;
      // (a block without rows reads its node's first block: a valid
      // address, skipped by in)
#pragma unroll
      for (int q = 0; q < 4; ++q) L[j][q] = __builtin_nontemporal_load(src + q * 64);
    }
  };
  v4d acc[MBL][NH];
#pragma unroll
  for (int j = 0; j < MBL; ++j)
#pragma unroll
    for (int hg = 0; hg < NH; ++hg) acc[j][hg] = v4d{0.0, 0.0, 0.0, 0.0};
  auto mma = [&](int buf, const double (&L)[MBL][4], const int (&in)[MBL]) {
    int any = 0;
#pragma unroll
    for (int j = 0; j < MBL; ++j) any |= in[j];
    if (!any) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double* row = &sB[buf][(4 * kq + q) * S + ri];
#pragma unroll
      for (int hg = 0; hg < NH; ++hg) {
        const double bv = row[hg * 16];
#pragma unroll
        for (int j = 0; j < MBL; ++j)
          acc[j][hg] = __builtin_amdgcn_mfma_f64_16x16x4f64(in[j] ? L[j][q] : 0.0, bv, acc[j][hg], 0, 0, 0);
      }
    }
  };
  if (nstep > 0) {
    // the factor operands are loaded two steps ahead, into three register
    // sets that rotate (step s uses set s % 3 and loads step s + 2 into the
    // set step s - 1 used); the staged vector one step ahead, through the
    // two LDS buffers (step s reads buffer s % 2)
    double R[K], LA[MBL][4], LB[MBL][4], LC[MBL][4];
    int IA[MBL], IB[MBL], IC[MBL];
    auto step = [&](int st, int buf, const double (&Lu)[MBL][4], const int (&Iu)[MBL], double (&Ln)[MBL][4],
                    int (&In)[MBL]) {
      stage_load(min(st + 1, nstep - 1), R);
      a_load(min(st + 2, nstep - 1), Ln, In);
      __builtin_amdgcn_sched_barrier(0);
      mma(buf, Lu, Iu);
      __builtin_amdgcn_sched_barrier(0);
      if (st + 1 < nstep) stage_store(buf ^ 1, R);
      __syncthreads();
      return st + 1 < nstep;
    };
    stage_load(0, R);
    a_load(0, LA, IA);
    a_load(min(1, nstep - 1), LB, IB);
    stage_store(0, R);
    __syncthreads();
    for (int st = 0;; st += 6) {
      if (!step(st, 0, LA, IA, LC, IC)) break;
      if (!step(st + 1, 1, LB, IB, LA, IA)) break;
      if (!step(st + 2, 0, LC, IC, LB, IB)) break;
      if (!step(st + 3, 1, LA, IA, LC, IC)) break;
      if (!step(st + 4, 0, LB, IB, LA, IA)) break;
      if (!step(st + 5, 1, LC, IC, LB, IB)) break;
    }
  }
#pragma unroll
  for (int j = 0; j < MBL; ++j) {
    if (!rows[j]) continue;  // after the last barrier
#pragma unroll
    for (int hg = 0; hg < NH; ++hg)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = first + 16 * j + kq + 4 * r;
        const int h = hg * 16 + ri;
        if (lr >= row0 + nr) continue;
        const double v = acc[j][hg][r];
        if (ks > 0) {
          Part[((int64_t)ti * (64 * MBL) + (lr - row0)) * NRP + h] = v;
        } else if (FWD) {
          if (lr < ns) Y[(int64_t)(p0 + lr) * NRP + h] = v;
          else F.Z[(F.z_off[s] + (lr - ns)) * NRP + h] = v;
        } else {
          F.X[(int64_t)(p0 + lr) * NRP + h] = v;
        }
      }
  }
}

// workgroups per row tile in k_sweep_reduce (a level has a few hundred split
// row tiles: one workgroup each would leave most SIMDs idle)
constexpr int REDUCE_SPLIT = 8;

// Sum of the column segments of split row tiles, in segment order: row tile
// rt = (node, row0, nrows, 64 * first tile + nseg); partial sums of tile t at
// Part[(t * cap + r) * NRP + h] (cap = 64 MB rows per slot).
template <int NRP, bool FWD>
__global__ __launch_bounds__(256) void k_sweep_reduce(DevFactor F, const int4* __restrict__ rtiles, int cap,
                                                      const double* __restrict__ Part, double* __restrict__ Y) {
  const int4 rt = rtiles[blockIdx.x];
  const int s = rt.x, row0 = rt.y, nr = rt.z, t0 = rt.w >> 6, nseg = rt.w & 63;
  const int ns = F.ns[s], p0 = F.p0[s];
  // gridDim.y workgroups share a row tile's entries; each entry's segments
  // are added in segment order, four loads in flight at a time
  for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < nr * NRP; e += gridDim.y * blockDim.x) {
    const int r = e / NRP, h = e % NRP;
    const double* pp = Part + ((int64_t)t0 * cap + r) * NRP + h;
    const int64_t seg = (int64_t)cap * NRP;
    double v = 0.0;
    int k = 0;
    for (; k + 4 <= nseg; k += 4) {
      double q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = pp[(k + u) * seg];
#pragma unroll
      for (int u = 0; u < 4; ++u) v += q[u];
    }
    for (; k < nseg; ++k) v += pp[k * seg];
    const int lr = row0 + r;
    if (FWD) {
      if (lr < ns) Y[(int64_t)(p0 + lr) * NRP + h] = v;
      else F.Z[(F.z_off[s] + (lr - ns)) * NRP + h] = v;
    } else {
      F.X[(int64_t)(p0 + lr) * NRP + h] = v;
    }
  }
}

// ------------------------------------------------------------------------
// Hybridised mixed Poisson
// ------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void poisson_fphi(const DevPoisson& P, const DevTabs& T, const DevPhys& ph, int Kc,
                                             const double* un, const double* up, const double* ur,
                                             const double* uo, int ncs, int nce, double* fphi) {
  using S = Sz<K>;
  constexpr int NB = S::NB, NQ = S::NQ, NPHI = S::NPHI;
  const bool semi = Kc < ncs;
  const int c = semi ? Kc : Kc - ncs;
  const int nc = semi ? ncs : nce;
  const double* u1 = semi ? un : ur;
  const double* u2 = semi ? up : uo;
  const double rho0 = semi ? (ph.ne - ph.pe) : 0.0;  // N_D - N_A (SolarCell.cpp:542-548)
  double v[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) v[t] = P.V[8 * (size_t)Kc + t];
  double d1[NB], d2[NB];
  const size_t ub = 2 * (size_t)nc * NB + (size_t)c * NB;
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    d1[a] = u1[ub + a];
    d2[a] = u2[ub + a];
  }
#pragma unroll
  for (int l = 0; l < NPHI; ++l) fphi[l] = 0.0;
  for (int q = 0; q < NQ; ++q) {
    const GeoD g = bilin(v, T.qx[q], T.qy[q]);
    const double jxw = T.qw[q] * g.det;
    double n1 = 0.0, n2 = 0.0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      n1 += d1[a] * T.phi[a * NQ + q];
      n2 += d2[a] * T.phi[a * NQ + q];
    }
    double rho = rho0 + (-1.0 * n1 + 1.0 * n2);
    if (P.msrc) rho += P.msrc[(size_t)Kc * NQ + q];
#pragma unroll
    for (int l = 0; l < NPHI; ++l) fphi[l] -= T.phiP[l * NQ + q] * rho * jxw;
  }
}

// constant Dirichlet / Schottky data of D row i from the cell's unit parts
template <int K>
__device__ __forceinline__ double poisson_bias(const double* tb, const DevPhys& ph, int i) {
  constexpr int NL = Sz<K>::NL, NRT = Sz<K>::NRT;
  const double* U = tb + NL * NL;
  return ph.phi_bi * U[i] - ph.phi_app * U[NRT + i] - ph.phi_sch * U[2 * NRT + i];
}

// D rows of the cell's right-hand side from manufactured Dirichlet data:
// -<psi_i . n, g> over its Dirichlet faces (the same integral as the unit
// bias parts, with g given at the face points)
template <int K>
__device__ __forceinline__ void poisson_mdir(const DevPoisson& P, const DevTabs& T, int Kc, double* f) {
  constexpr int NRT = Sz<K>::NRT, NF = Sz<K>::NF;
  double v[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) v[t] = P.V[8 * (size_t)Kc + t];
  for (int fc = 0; fc < 4; ++fc) {
    if (P.pbid[4 * (size_t)Kc + fc] != 1) continue;  // Dirichlet faces only
    double nx, ny, len;
    face_nrm(v, fc, nx, ny, len);
    for (int q = 0; q < NF; ++q) {
      const double s = T.fs[q];
      const double xi = fc == 0 ? 0.0 : fc == 1 ? 1.0 : s, eta = fc == 2 ? 0.0 : fc == 3 ? 1.0 : s;
      const GeoD g = bilin(v, xi, eta);
      const double w = -P.mdir[(4 * (size_t)Kc + fc) * NF + q] * T.fw[q] * len;
#pragma unroll
      for (int i = 0; i < NRT; ++i) {
        const double hx = T.fpsix[(fc * NRT + i) * NF + q], hy = T.fpsiy[(fc * NRT + i) * NF + q];
        f[i] += ((g.J00 * hx + g.J01 * hy) * nx + (g.J10 * hx + g.J11 * hy) * ny) / g.det * w;
      }
    }
  }
}

// UNI: cells from P.ucell (waves share the cell type: scalar table reads)
template <int K, bool UNI>
__global__ __launch_bounds__(64) void k_poisson_cell(DevPoisson P, DevTabs T, const DevPhys* __restrict__ phs,
                                                     const double* __restrict__ un, const double* __restrict__ up,
                                                     const double* __restrict__ ur, const double* __restrict__ uo,
                                                     int ncs, int nce, double* __restrict__ y, int64_t ss, int64_t se,
                                                     int64_t sy) {
  const DevPhys ph = phs[blockIdx.y];
  {
    const int64_t i = blockIdx.y;
    un += i * ss;
    up += i * ss;
    ur += i * se;
    uo += i * se;
    y += i * sy;
  }
  using S = Sz<K>;
  constexpr int NRT = S::NRT, NL = S::NL, NPHI = S::NPHI;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int Kc = UNI ? (t0 < P.nucell ? P.ucell[t0] : -1) : (t0 < P.ncell ? t0 : -1);
  if (Kc < 0) return;
  double f[NL];
  const int ty = UNI ? __builtin_amdgcn_readfirstlane(P.type[Kc]) : P.type[Kc];
  const double* tb = P.table + (size_t)ty * P.stride;
#pragma unroll
  for (int i = 0; i < NRT; ++i) f[i] = poisson_bias<K>(tb, ph, i);
  if (P.mdir) poisson_mdir<K>(P, T, Kc, f);
  poisson_fphi<K>(P, T, ph, Kc, un, up, ur, uo, ncs, nce, f + NRT);
  for (int i = 0; i < NL; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NL; ++j) s += tb[i * NL + j] * f[j];
    y[(size_t)Kc * NL + i] = s;
  }
  (void)NPHI;
}

// face DoF ownership for the continuous D layout: left/bottom faces belong to
// the cell; right/top faces only on the right/top boundary.
template <int K>
__device__ __forceinline__ int face_of_local(int i, int& m) {
  constexpr int NRTX = Sz<K>::NRTX;
  if (i < NRTX) {
    const int a = i % (K + 1), b = i / (K + 1);
    m = b;
    if (a == 0) return 0;
    if (a == K) return 1;
    return -1;
  }
  const int ii = i - NRTX, a = ii % K, b = ii / K;
  m = a;
  if (b == 0) return 2;
  if (b == K) return 3;
  return -1;
}

template <int K>
__device__ __forceinline__ bool owns_face(const DevPoisson& P, int Kc, int f) {
  return (P.own[Kc] >> f) & 1;
}

template <int K>
__global__ __launch_bounds__(64) void k_poisson_cell_from_b(DevPoisson P, const double* __restrict__ b,
                                                            double* __restrict__ y) {
  constexpr int NL = Sz<K>::NL, NRT = Sz<K>::NRT;
  const int Kc = blockIdx.x * blockDim.x + threadIdx.x;
  if (Kc >= P.ncell) return;
  const int* id = P.idx + (size_t)Kc * NL;
  double f[NL];
  for (int i = 0; i < NL; ++i) {
    int m;
    const int fc = (i < NRT) ? face_of_local<K>(i, m) : -1;
    f[i] = (fc < 0 || owns_face<K>(P, Kc, fc)) ? b[id[i]] : 0.0;
  }
  const double* tb = P.table + (size_t)P.type[Kc] * P.stride;
  for (int i = 0; i < NL; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NL; ++j) s += tb[i * NL + j] * f[j];
    y[(size_t)Kc * NL + i] = s;
  }
}

__global__ __launch_bounds__(256) void k_poisson_face(DevPoisson P, const double* __restrict__ y,
                                                      double* __restrict__ r, int64_t sy, int64_t sr) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= P.nmult) return;
  y += blockIdx.y * sy;
  r += blockIdx.y * sr;
  double s = 0.0;
  for (int e = 0; e < 2; ++e) {
    const int c = P.mcell[2 * m + e];
    if (c >= 0) s += P.mcoef[2 * m + e] * y[(size_t)c * P.nloc + P.mloc[2 * m + e]];
  }
  r[P.mult_perm[m]] = s;
}

// UNI: cells from P.ucell, whose waves share the cell type, so the table
// offsets are wave-uniform (scalar-cache reads).  Faces without multipliers
// contribute exact zeros (s - 0 = s), so the sum equals the one over the
// faces that have them.
template <int K, bool UNI>
__global__ __launch_bounds__(64) void k_poisson_recover(DevPoisson P, const double* __restrict__ lam,
                                                        const double* __restrict__ y, double* __restrict__ out,
                                                        int64_t sl, int64_t sy, int64_t so) {
  constexpr int NL = Sz<K>::NL, NRT = Sz<K>::NRT;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int Kc = UNI ? (t0 < P.nucell ? P.ucell[t0] : -1) : (t0 < P.ncell ? t0 : -1);
  if (Kc < 0) return;
  lam += blockIdx.y * sl;
  y += blockIdx.y * sy;
  out += blockIdx.y * so;
  double e[4 * K];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int base = P.mult_face[4 * (size_t)Kc + f];
    const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
#pragma unroll
    for (int m = 0; m < K; ++m) e[f * K + m] = base < 0 ? 0.0 : sf * P.gw[m] * lam[P.mult_perm[base + m]];
  }
  const int ty = UNI ? __builtin_amdgcn_readfirstlane(P.type[Kc]) : P.type[Kc];
  const double* tb = P.table + (size_t)ty * P.stride;
  const int* id = P.idx + (size_t)Kc * NL;
  for (int i = 0; i < NL; ++i) {
    double s = y[(size_t)Kc * NL + i];
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int m = 0; m < K; ++m) s -= tb[i * NL + rt_face<K>(f, m)] * e[f * K + m];
    if (i < NRT) {
      int m;
      const int fc = face_of_local<K>(i, m);
      if (fc >= 0) {
        if (!owns_face<K>(P, Kc, fc)) continue;
        if (P.pbid[4 * (size_t)Kc + fc] == 2) s = 0.0;  // Neumann D.n = 0 (Poisson.cpp:41-47)
      }
    }
    out[id[i]] = s;
  }
}

// Mixed precision: the cell condensation y = A^{-1} f and the recovery's
// correction A^{-1} E^T lambda contract in FP32 with the FP32 table (the
// right-hand side f, y and the results stay FP64); wave-uniform lists only.
template <int K>
__global__ __launch_bounds__(64) void k_poisson_cell_f32(DevPoisson P, DevTabs T, const DevPhys* __restrict__ phs,
                                                         const double* __restrict__ un,
                                                         const double* __restrict__ up,
                                                         const double* __restrict__ ur,
                                                         const double* __restrict__ uo, int ncs, int nce,
                                                         double* __restrict__ y, int64_t ss, int64_t se,
                                                         int64_t sy) {
  const DevPhys ph = phs[blockIdx.y];
  {
    const int64_t i = blockIdx.y;
    un += i * ss;
    up += i * ss;
    ur += i * se;
    uo += i * se;
    y += i * sy;
  }
  using S = Sz<K>;
  constexpr int NRT = S::NRT, NL = S::NL;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int Kc = t0 < P.nucell ? P.ucell[t0] : -1;
  if (Kc < 0) return;
  const int ty = __builtin_amdgcn_readfirstlane(P.type[Kc]);
  const double* tb = P.table + (size_t)ty * P.stride;
  const float* tf = P.table_f + (size_t)ty * P.stride;
  double f[NL];
#pragma unroll
  for (int i = 0; i < NRT; ++i) f[i] = poisson_bias<K>(tb, ph, i);
  if (P.mdir) poisson_mdir<K>(P, T, Kc, f);
  poisson_fphi<K>(P, T, ph, Kc, un, up, ur, uo, ncs, nce, f + NRT);
  float ff[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) ff[j] = (float)f[j];
  for (int i = 0; i < NL; ++i) {
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < NL; ++j) s += tf[i * NL + j] * ff[j];
    y[(size_t)Kc * NL + i] = (double)s;
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_poisson_recover_f32(DevPoisson P, const double* __restrict__ lam,
                                                            const double* __restrict__ y, double* __restrict__ out,
                                                            int64_t sl, int64_t sy, int64_t so) {
  constexpr int NL = Sz<K>::NL, NRT = Sz<K>::NRT;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int Kc = t0 < P.nucell ? P.ucell[t0] : -1;
  if (Kc < 0) return;
  lam += blockIdx.y * sl;
  y += blockIdx.y * sy;
  out += blockIdx.y * so;
  float e[4 * K];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int base = P.mult_face[4 * (size_t)Kc + f];
    const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
#pragma unroll
    for (int m = 0; m < K; ++m)
      e[f * K + m] = base < 0 ? 0.0f : (float)(sf * P.gw[m] * lam[P.mult_perm[base + m]]);
  }
  const float* tf = P.table_f + (size_t)__builtin_amdgcn_readfirstlane(P.type[Kc]) * P.stride;
  const int* id = P.idx + (size_t)Kc * NL;
  for (int i = 0; i < NL; ++i) {
    float c = 0.0f;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int m = 0; m < K; ++m) c += tf[i * NL + rt_face<K>(f, m)] * e[f * K + m];
    double s = y[(size_t)Kc * NL + i] - (double)c;
    if (i < NRT) {
      int m;
      const int fc = face_of_local<K>(i, m);
      if (fc >= 0) {
        if (!owns_face<K>(P, Kc, fc)) continue;
        if (P.pbid[4 * (size_t)Kc + fc] == 2) s = 0.0;  // Neumann D.n = 0 (Poisson.cpp:41-47)
      }
    }
    out[id[i]] = s;
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_poisson_rhs_assemble(DevPoisson P, DevTabs T, const DevPhys* __restrict__ phs,
                                                             const double* __restrict__ un,
                                                             const double* __restrict__ up,
                                                             const double* __restrict__ ur,
                                                             const double* __restrict__ uo, int ncs, int nce,
                                                             double* __restrict__ b) {
  constexpr int NL = Sz<K>::NL, NRT = Sz<K>::NRT, NPHI = Sz<K>::NPHI;
  const int Kc = blockIdx.x * blockDim.x + threadIdx.x;
  if (Kc >= P.ncell) return;
  const DevPhys ph = phs[0];
  double fphi[NPHI];
  poisson_fphi<K>(P, T, ph, Kc, un, up, ur, uo, ncs, nce, fphi);
  const int* id = P.idx + (size_t)Kc * NL;
  const double* tb = P.table + (size_t)P.type[Kc] * P.stride;
  double fd[NRT];
#pragma unroll
  for (int i = 0; i < NRT; ++i) fd[i] = poisson_bias<K>(tb, ph, i);
  if (P.mdir) poisson_mdir<K>(P, T, Kc, fd);
  for (int i = 0; i < NRT; ++i) {
    int m;
    const int fc = face_of_local<K>(i, m);
    if (fc >= 0 && !owns_face<K>(P, Kc, fc)) continue;
    b[id[i]] = fd[i];
  }
#pragma unroll
  for (int l = 0; l < NPHI; ++l) b[id[NRT + l]] = fphi[l];
}

__global__ __launch_bounds__(256) void k_copy(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

// Steady-residual norms without atomics (the sums are bitwise reproducible):
// block g of k_diff_norms writes its partial sums of |a - b|^2 and |b|^2 to
// part[2 g], part[2 g + 1] (RES_BLOCKS blocks per vector); k_sum_partials then
// adds each instance's partials in a fixed order.
constexpr int RES_BLOCKS = RES_PARTS;

__global__ __launch_bounds__(256) void k_diff_norms(const double* __restrict__ a, const double* __restrict__ b,
                                                    int64_t n, double* __restrict__ part) {
  __shared__ double s0[256], s1[256];
  double d = 0.0, o = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double t = a[i] - b[i];
    d += t * t;
    o += b[i] * b[i];
  }
  s0[threadIdx.x] = d;
  s1[threadIdx.x] = o;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

// One block per instance: acc[2 i], acc[2 i + 1] = sums of the instance's
// nparts partial pairs at part + 2 nparts i.
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int nparts,
                                                      double* __restrict__ acc) {
  __shared__ double s0[256], s1[256];
  const double* p = part + 2 * (int64_t)nparts * blockIdx.x;
  double d = 0.0, o = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    d += p[2 * i];
    o += p[2 * i + 1];
  }
  s0[threadIdx.x] = d;
  s1[threadIdx.x] = o;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    acc[2 * blockIdx.x] = s0[0];
    acc[2 * blockIdx.x + 1] = s1[0];
  }
}

// ------------------------------------------------------------------------
// Density fields on a regular grid (the warm-start network's output space)
// ------------------------------------------------------------------------
// A region's cells are the logical grid (i, j) = Morton decode of c, cell
// (i, j) covering logical [i, i+1] / N x [j, j+1] / N with its reference
// coordinates along i and j (the mesh is a uniform logical subdivision of
// the region's coarse quadrilateral).
__device__ __forceinline__ int morton_code(int i, int j, int r) {
  int c = 0;
  for (int b = 0; b < r; ++b) c |= (((i >> b) & 1) << (2 * b)) | (((j >> b) & 1) << (2 * b + 1));
  return c;
}

// Lagrange basis of the K+1 equispaced nodes on [0, 1], node a, at x
template <int K>
__device__ __forceinline__ double lagrange_eq(int a, double x) {
  double v = 1.0;
#pragma unroll
  for (int m = 0; m <= K; ++m)
    if (m != a) v *= (x - (double)m / K) / ((double)(a - m) / K);
  return v;
}

// out[b G + a] = density of carrier vector u at logical ((a + 1/2) / G, (b + 1/2) / G)
template <int K>
__global__ __launch_bounds__(256) void k_grid_sample(const double* __restrict__ u, int r, int G,
                                                     double* __restrict__ out) {
  constexpr int NB = (K + 1) * (K + 1);
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= G * G) return;
  const int N = 1 << r, nc = N * N;
  const int a = e % G, b = e / G;
  const double s = (a + 0.5) / G * N, t = (b + 0.5) / G * N;
  const int i = min((int)s, N - 1), j = min((int)t, N - 1);
  const double xi = s - i, eta = t - j;
  const double* d = u + 2 * (size_t)nc * NB + (size_t)morton_code(i, j, r) * NB;
  double v = 0.0;
#pragma unroll
  for (int q = 0; q <= K; ++q) {
    const double ly = lagrange_eq<K>(q, eta);
#pragma unroll
    for (int p = 0; p <= K; ++p) v += d[p + (K + 1) * q] * lagrange_eq<K>(p, xi) * ly;
  }
  out[e] = v;
}

// Density block of carrier vector u from a G x G grid of values at the
// logical cell centres ((a + 1/2) / G, (b + 1/2) / G): each nodal DoF takes
// the bilinear interpolant of the grid at its logical position (constant
// continuation in the half-cell band along the region's edges).
template <int K>
__global__ __launch_bounds__(256) void k_grid_project(const double* __restrict__ grid, int G, int r,
                                                      double* __restrict__ u) {
  constexpr int NB = (K + 1) * (K + 1);
  const int N = 1 << r, nc = N * N;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nc * NB) return;
  const int c = e / NB, m = e % NB;
  int i = 0, j = 0;
  for (int b = 0; b < r; ++b) {
    i |= ((c >> (2 * b)) & 1) << b;
    j |= ((c >> (2 * b + 1)) & 1) << b;
  }
  const double s = ((double)i + (double)(m % (K + 1)) / K) / N, t = ((double)j + (double)(m / (K + 1)) / K) / N;
  const double gx = fmin(fmax(s * G - 0.5, 0.0), (double)(G - 1)), gy = fmin(fmax(t * G - 0.5, 0.0), (double)(G - 1));
  const int x0 = min((int)gx, G - 2 < 0 ? 0 : G - 2), y0 = min((int)gy, G - 2 < 0 ? 0 : G - 2);
  const int x1 = min(x0 + 1, G - 1), y1 = min(y0 + 1, G - 1);
  const double fx = gx - x0, fy = gy - y0;
  const double v = (1.0 - fy) * ((1.0 - fx) * grid[y0 * G + x0] + fx * grid[y0 * G + x1]) +
                   fy * ((1.0 - fx) * grid[y1 * G + x0] + fx * grid[y1 * G + x1]);
  u[2 * (size_t)nc * NB + e] = v;
}

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

#define PECS_DISPATCH_K(k, CALL) \
  switch (k) {                   \
    case 1: { constexpr int KK = 1; CALL; } break; \
    case 2: { constexpr int KK = 2; CALL; } break; \
    case 3: { constexpr int KK = 3; CALL; } break; \
    default: return (int)hipErrorInvalidValue;      \
  }

// ------------------------------------------------------------------------
// Hanging faces of boundary-layer refinement
// ------------------------------------------------------------------------
// The Schur kernels leave out the couplings across hanging faces (those faces
// are absent from the neighbour lists); these kernels add them afterwards, in
// FP64, one thread per cell that has such a coupling.  Both are linear in the
// terms they add, so the sum equals the one over all couplings.
template <int K, int NRHS>
__global__ __launch_bounds__(64) void k_schur_rhs_hang(DevRegion R, const double* __restrict__ b1,
                                                       const double* __restrict__ b2, double mu1, double mu2,
                                                       double* __restrict__ g1, double* __restrict__ g2, int64_t sb,
                                                       int64_t sg) {
  constexpr int NB = Sz<K>::NB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.nhx) return;
  {
    const int64_t i = blockIdx.y;
    b1 += i * sb;
    b2 += i * sb;
    g1 += i * sg;
    g2 += i * sg;
  }
  double a1[NB], a2[NB];
#pragma unroll
  for (int a = 0; a < NB; ++a) a1[a] = a2[a] = 0.0;
  for (int e = R.hx_ptr[t]; e < R.hx_ptr[t + 1]; ++e) {
    const int M = R.hx_src[e];
    const double* B = R.hblk + (size_t)R.hx_blk[e] * 2 * NB * NB;
    const int ty = R.type[M];
    double w1[2 * NB], w2[2 * NB];
    schur_w<K>(R, b1, mu1, M, ty, w1);
    if (NRHS > 1) schur_w<K>(R, b2, mu2, M, ty, w2);
    for (int i = 0; i < 2 * NB; ++i) {
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        a1[a] += B[i * NB + a] * w1[i];
        if (NRHS > 1) a2[a] += B[i * NB + a] * w2[i];
      }
    }
  }
  const int pb = R.perm_base[R.hx_cell[t]];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    g1[pb + a] += a1[a];
    if (NRHS > 1) g2[pb + a] += a2[a];
  }
}

template <int K, int NRHS>
__global__ __launch_bounds__(64) void k_schur_recover_hang(DevRegion R, const double* __restrict__ x1,
                                                           const double* __restrict__ x2, double mu1, double mu2,
                                                           double* __restrict__ o1, double* __restrict__ o2,
                                                           int64_t sx, int64_t so) {
  constexpr int NB = Sz<K>::NB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.nhm) return;
  {
    const int64_t i = blockIdx.y;
    x1 += i * sx;
    x2 += i * sx;
    o1 += i * so;
    o2 += i * so;
  }
  const int M = R.hm_cell[t];
  const size_t nc = R.ncell;
  double s1[2 * NB], s2[2 * NB];
#pragma unroll
  for (int i = 0; i < 2 * NB; ++i) s1[i] = s2[i] = 0.0;
  for (int e = R.hm_ptr[t]; e < R.hm_ptr[t + 1]; ++e) {
    const double* B = R.hblk + (size_t)R.hm_blk[e] * 2 * NB * NB;
    const int px = R.perm_base[R.hm_dst[e]];
    double v1[NB], v2[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      v1[a] = x1[px + a];
      v2[a] = (NRHS > 1) ? x2[px + a] : 0.0;
    }
    for (int i = 0; i < 2 * NB; ++i) {
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        s1[i] += B[i * NB + a] * v1[a];
        if (NRHS > 1) s2[i] += B[i * NB + a] * v2[a];
      }
    }
  }
  const double* Minv = R.table + (size_t)R.type[M] * R.stride;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      double t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        t1 += Minv[a * NB + q] * s1[h * NB + q];
        if (NRHS > 1) t2 += Minv[a * NB + q] * s2[h * NB + q];
      }
      const size_t off = (size_t)h * nc * NB + (size_t)M * NB + a;
      o1[off] -= mu1 * t1;
      if (NRHS > 1) o2[off] -= mu2 * t2;
    }
}

// H right-hand side: the coarse side's contribution to the fine faces'
// multipliers, one thread per (hanging pair, multiplier DoF)
template <int K>
__global__ __launch_bounds__(64) void k_poisson_face_hang(DevPoisson P, const double* __restrict__ y,
                                                          double* __restrict__ r, int64_t sy, int64_t sr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.nhang * K) return;
  y += blockIdx.y * sy;
  r += blockIdx.y * sr;
  const int h = t / K, m = t % K;
  const int Kc = P.h_cell[h], f = P.h_face[h], sub = P.h_sub[h];
  const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < K; ++j) s += P.hE[(sub * K + m) * K + j] * y[(size_t)Kc * P.nloc + rt_face<K>(f, j)];
  r[P.mult_perm[P.h_mbase[h] + m]] += sf * s;
}

// Recovery: x_K -= A_K^{-1} E_K^T lambda over the coarse cell's hanging faces
// (with the same write rules as k_poisson_recover)
template <int K>
__global__ __launch_bounds__(64) void k_poisson_recover_hang(DevPoisson P, const double* __restrict__ lam,
                                                             double* __restrict__ out, int64_t sl, int64_t so) {
  constexpr int NL = Sz<K>::NL, NRT = Sz<K>::NRT;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.nhc) return;
  lam += blockIdx.y * sl;
  out += blockIdx.y * so;
  const int Kc = P.hc_cell[t];
  double e[4 * K];
#pragma unroll
  for (int i = 0; i < 4 * K; ++i) e[i] = 0.0;
  for (int q = P.hc_ptr[t]; q < P.hc_ptr[t + 1]; ++q) {
    const int f = P.h_face[q], sub = P.h_sub[q], mb = P.h_mbase[q];
    const double sf = (f == 1 || f == 3) ? 1.0 : -1.0;
    for (int m = 0; m < K; ++m) {
      const double l = sf * lam[P.mult_perm[mb + m]];
#pragma unroll
      for (int j = 0; j < K; ++j) e[f * K + j] += P.hE[(sub * K + m) * K + j] * l;
    }
  }
  const double* tb = P.table + (size_t)P.type[Kc] * P.stride;
  const int* id = P.idx + (size_t)Kc * NL;
  for (int i = 0; i < NL; ++i) {
    if (i < NRT) {
      int m;
      const int fc = face_of_local<K>(i, m);
      if (fc >= 0 && (!owns_face<K>(P, Kc, fc) || P.pbid[4 * (size_t)Kc + fc] == 2)) continue;
    }
    double c = 0.0;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int j = 0; j < K; ++j) c += tb[i * NL + rt_face<K>(f, j)] * e[f * K + j];
    out[id[i]] -= c;
  }
}

int launch_rhs_carrier(int k, bool semi, const DevRegion& R, const DevRegion& O, const DevPoisson& P,
                       const DevTabs& T, const DevPhys* phs, const double* u1, const double* u2, const double* ou1,
                       const double* ou2, const double* pvec, double* b1, double* b2, hipStream_t s, int nb,
                       int64_t sR, int64_t sO, int64_t sP) {
  if (R.ncell == 0) return 0;
  const dim3 g(nblk(R.ncell, 64), nb);
  if (semi) {
    PECS_DISPATCH_K(k, (k_rhs_carrier<KK, true><<<g, 64, 0, s>>>(R, O, P, T, phs, u1, u2, ou1, ou2, pvec, b1, b2, sR,
                                                                 sO, sP)));
  } else {
    PECS_DISPATCH_K(k, (k_rhs_carrier<KK, false><<<g, 64, 0, s>>>(R, O, P, T, phs, u1, u2, ou1, ou2, pvec, b1, b2, sR,
                                                                  sO, sP)));
  }
  return (int)hipGetLastError();
}

static int launch_schur_rhs_main(int k, const DevRegion& R, const double* b1, const double* b2, double mu1, double mu2,
                     double* g1, double* g2, int nrhs, hipStream_t s, int nb, int64_t sb, int64_t sg) {
#define PECS_SCHUR_RHS(UU, SS, NCELL)                                                                       \
  {                                                                                                         \
    const dim3 g(nblk(NCELL, 64), nb);                                                                      \
    if (nrhs == 2) {                                                                                        \
      PECS_DISPATCH_K(k, (k_schur_rhs<KK, 2, UU, SS><<<g, 64, 0, s>>>(R, b1, b2, mu1, mu2, g1, g2, sb, sg))); \
    } else {                                                                                                \
      PECS_DISPATCH_K(k, (k_schur_rhs<KK, 1, UU, SS><<<g, 64, 0, s>>>(R, b1, b1, mu1, mu1, g1, g1, sb, sg))); \
    }                                                                                                       \
    return (int)hipGetLastError();                                                                          \
  }
  if (R.table_f && R.ucell_nbh) {  // mixed precision
    const dim3 g(nblk(R.nu_nbh, 64), nb);
    if (nrhs == 2) {
      PECS_DISPATCH_K(k, (k_schur_rhs_f32<KK, 2><<<g, 64, 0, s>>>(R, b1, b2, mu1, mu2, g1, g2, sb, sg)));
    } else {
      PECS_DISPATCH_K(k, (k_schur_rhs_f32<KK, 1><<<g, 64, 0, s>>>(R, b1, b1, mu1, mu1, g1, g1, sb, sg)));
    }
    return (int)hipGetLastError();
  }
  // a pass per right-hand side once a batch fills the chip (4 or more instances)
  if (R.ucell_nbh && nb >= 4) PECS_SCHUR_RHS(true, true, R.nu_nbh)
  if (R.ucell_nbh) PECS_SCHUR_RHS(true, false, R.nu_nbh)
  PECS_SCHUR_RHS(false, false, R.ncell)
#undef PECS_SCHUR_RHS
}

static int launch_schur_recover_main(int k, const DevRegion& R, const double* b1, const double* b2, const double* xs1,
                         const double* xs2, double mu1, double mu2, double* out1, double* out2, int nrhs,
                         hipStream_t s, int nb, int64_t sb, int64_t sx, int64_t so) {
#define PECS_RECOVER(UU, SS, NCELL)                                                                             \
  {                                                                                                             \
    const dim3 g(nblk(NCELL, 64), nb);                                                                          \
    if (nrhs == 2) {                                                                                            \
      PECS_DISPATCH_K(k, (k_schur_recover<KK, 2, UU, SS><<<g, 64, 0, s>>>(R, b1, b2, xs1, xs2, mu1, mu2, out1,   \
                                                                          out2, sb, sx, so)));                  \
    } else {                                                                                                    \
      PECS_DISPATCH_K(k, (k_schur_recover<KK, 1, UU, SS><<<g, 64, 0, s>>>(R, b1, b1, xs1, xs1, mu1, mu1, out1,   \
                                                                          out1, sb, sx, so)));                  \
    }                                                                                                           \
    return (int)hipGetLastError();                                                                              \
  }
  if (R.table_f && R.ucell_own) {  // mixed precision
    const dim3 g(nblk(R.nu_own, 64), nb);
    if (nrhs == 2) {
      PECS_DISPATCH_K(k, (k_schur_recover_f32<KK, 2><<<g, 64, 0, s>>>(R, b1, b2, xs1, xs2, mu1, mu2, out1, out2, sb,
                                                                      sx, so)));
    } else {
      PECS_DISPATCH_K(k, (k_schur_recover_f32<KK, 1><<<g, 64, 0, s>>>(R, b1, b1, xs1, xs1, mu1, mu1, out1, out1, sb,
                                                                      sx, so)));
    }
    return (int)hipGetLastError();
  }
  // a pass per right-hand side once a batch fills the chip (4 or more instances)
  if (R.ucell_own && nb >= 4) PECS_RECOVER(true, true, R.nu_own)
  if (R.ucell_own) PECS_RECOVER(true, false, R.nu_own)
  PECS_RECOVER(false, false, R.ncell)
#undef PECS_RECOVER
}

template <int NRHS, bool FWD>
static int sweep_dispatch(const DevFactor& F, const int4* tl, int nt, int w, int maxr, const double* v0,
                          const double* v1, double* t0, double* t1, hipStream_t s, bool pf) {
#define PECS_SWEEP_CASE(WW, MM)                                                          \
  if (w == WW && maxr == MM) {                                                           \
    if (pf) k_sweep<NRHS, WW, MM, FWD, true><<<nt, 256, 0, s>>>(F, tl, v0, v1, t0, t1);  \
    else k_sweep<NRHS, WW, MM, FWD><<<nt, 256, 0, s>>>(F, tl, v0, v1, t0, t1);           \
    return 0;                                                                            \
  }
  PECS_SWEEP_CASE(8, 1) PECS_SWEEP_CASE(8, 2) PECS_SWEEP_CASE(8, 4)
  PECS_SWEEP_CASE(16, 1) PECS_SWEEP_CASE(16, 2) PECS_SWEEP_CASE(16, 4)
  PECS_SWEEP_CASE(32, 1) PECS_SWEEP_CASE(32, 2) PECS_SWEEP_CASE(32, 4)
  PECS_SWEEP_CASE(64, 1) PECS_SWEEP_CASE(64, 2) PECS_SWEEP_CASE(64, 4)
  PECS_SWEEP_CASE(32, 8) PECS_SWEEP_CASE(64, 8)
#undef PECS_SWEEP_CASE
  return (int)hipErrorInvalidValue;
}

template <bool FWD>
static int sweep_mfma_dispatch(const DevFactor& F, const int4* tl, int nt, int mb, int nrp, bool split, double* Y,
                               int ks, const int4* rtl, int nrt, hipStream_t s) {
  if (ks > 0 && (!rtl || !F.Part)) return (int)hipErrorInvalidValue;
#define PECS_MFMA_CASE(NN, MM)                                                              \
  if (!split && nrp == NN && mb == MM) {                                                    \
    k_sweep_mfma<NN, MM, FWD><<<nt, 256, 0, s>>>(F, tl, Y, ks, F.Part);                     \
    if (ks > 0 && nrt > 0) k_sweep_reduce<NN, FWD><<<dim3(nrt, REDUCE_SPLIT), 256, 0, s>>>(F, rtl, 64 * MM, F.Part, Y); \
    return 0;                                                                               \
  }
#define PECS_MFMA_SPLIT(NN, MM, HH)                                                         \
  if (split && nrp == NN && mb == MM) {                                                     \
    k_sweep_mfma<NN, MM, FWD, HH><<<nt, 256, 0, s>>>(F, tl, Y, ks, F.Part);                 \
    if (ks > 0 && nrt > 0) k_sweep_reduce<NN, FWD><<<dim3(nrt, REDUCE_SPLIT), 256, 0, s>>>(F, rtl, 64 * MM, F.Part, Y); \
    return 0;                                                                               \
  }
  PECS_MFMA_CASE(16, 1) PECS_MFMA_CASE(16, 2) PECS_MFMA_CASE(16, 4)
  PECS_MFMA_CASE(32, 1) PECS_MFMA_CASE(32, 2) PECS_MFMA_CASE(32, 4)
  PECS_MFMA_CASE(64, 1) PECS_MFMA_CASE(64, 2) PECS_MFMA_CASE(64, 4)
  PECS_MFMA_CASE(128, 1) PECS_MFMA_CASE(128, 2)
  PECS_MFMA_SPLIT(32, 1, 1) PECS_MFMA_SPLIT(32, 2, 1) PECS_MFMA_SPLIT(32, 4, 1)
  PECS_MFMA_SPLIT(64, 1, 1) PECS_MFMA_SPLIT(64, 2, 1) PECS_MFMA_SPLIT(64, 4, 1)
  PECS_MFMA_SPLIT(128, 1, 2) PECS_MFMA_SPLIT(128, 2, 2) PECS_MFMA_SPLIT(128, 4, 2)
#undef PECS_MFMA_SPLIT
#undef PECS_MFMA_CASE
  return (int)hipErrorInvalidValue;
}

int mfma_rhs_pad(int nrhs) { return nrhs <= 16 ? 16 : nrhs <= 32 ? 32 : nrhs <= 64 ? 64 : 128; }

// 4 blocks per wave at 128 right-hand sides only with the groups split over
// waves (2 groups per wave): unsplit, their accumulators and operands would
// not fit the register file
bool mfma_shape_ok(int nrp, int mb, bool split) { return mb == 1 || mb == 2 || (mb == 4 && (nrp <= 64 || split)); }

int mfma_tile_rows(int nrp, int mb, bool split) {
  if (!split || nrp <= 16) return 64 * mb;
  const int hpw = nrp >= 128 ? 2 : 1, wpr = nrp / 16 / hpw;
  return 4 / wpr * 16 * mb;
}

int launch_sweep(const DevFactor& F, const int4* tiles, int nt, int w, int maxr, bool fwd, int nrhs,
                 double* t, hipStream_t s, int ks, const int4* rtiles, int nrt, const double* src, bool pf) {
  if (nt <= 0) return 0;
  if (w <= 0) {  // 0: MFMA sweep, -1: its right-hand-side groups split over waves, -2: LDS-staged vector
    if (nrhs > MAX_MFMA_RHS) return (int)hipErrorInvalidValue;
    const int nrp = mfma_rhs_pad(nrhs);
    // a factor tuned for many right-hand sides and solved with 16 or fewer
    // (single solves on a batched context): the LDS-staged shape runs as the
    // plain MFMA sweep, whose tiles of 64 maxr rows and partial-sum slots
    // are the same
    if (w == -2 && nrp < 32) w = 0;
    if (w == -2) {
      if ((maxr != 1 && maxr != 2) || (ks > 0 && (!rtiles || !F.Part))) return (int)hipErrorInvalidValue;
#define PECS_MFMA_LDS(NN, MM)                                                                    \
  if (nrp == NN && maxr == MM) {                                                                 \
    if (fwd) {                                                                                   \
      k_sweep_mfma_lds<NN, true, MM><<<nt, 256, 0, s>>>(F, tiles, t, ks, F.Part);                \
      if (ks > 0 && nrt > 0) k_sweep_reduce<NN, true><<<dim3(nrt, REDUCE_SPLIT), 256, 0, s>>>(F, rtiles, 64 * MM, F.Part, t); \
    } else {                                                                                     \
      k_sweep_mfma_lds<NN, false, MM><<<nt, 256, 0, s>>>(F, tiles, t, ks, F.Part);               \
      if (ks > 0 && nrt > 0) k_sweep_reduce<NN, false><<<dim3(nrt, REDUCE_SPLIT), 256, 0, s>>>(F, rtiles, 64 * MM, F.Part, t); \
    }                                                                                            \
    return (int)hipGetLastError();                                                               \
  }
      PECS_MFMA_LDS(32, 1) PECS_MFMA_LDS(64, 1) PECS_MFMA_LDS(128, 1)
      PECS_MFMA_LDS(32, 2) PECS_MFMA_LDS(64, 2) PECS_MFMA_LDS(128, 2)
#undef PECS_MFMA_LDS
      return (int)hipErrorInvalidValue;
    }
    const bool split = w < 0 && nrp > 16;
    return fwd ? sweep_mfma_dispatch<true>(F, tiles, nt, maxr, nrp, split, t, ks, rtiles, nrt, s)
               : sweep_mfma_dispatch<false>(F, tiles, nt, maxr, nrp, split, t, ks, rtiles, nrt, s);
  }
  if (nrhs > 2 || ks > 0) return (int)hipErrorInvalidValue;
  const double* v0 = fwd ? src : nullptr;
  double* t1 = nrhs == 2 ? t + F.n : t;
  if (fwd) {
    return nrhs == 2 ? sweep_dispatch<2, true>(F, tiles, nt, w, maxr, v0, v0, t, t1, s, pf)
                     : sweep_dispatch<1, true>(F, tiles, nt, w, maxr, v0, v0, t, t1, s, pf);
  }
  return nrhs == 2 ? sweep_dispatch<2, false>(F, tiles, nt, w, maxr, v0, v0, t, t1, s, pf)
                   : sweep_dispatch<1, false>(F, tiles, nt, w, maxr, v0, v0, t, t1, s, pf);
}

template <int NRP>
static int chol_solve_il(const DevFactor& F, const FactorLevels& L, double* v, double* t, int nrhs, hipStream_t s) {
  k_pack_il<NRP><<<nblk(F.n, IL_TP), 256, 0, s>>>(F, v, nrhs);
  // the leaf level's rows are owed nothing (pack_layout checks it)
  for (int l = 0; l < L.nlevels; ++l) {
    const int ng = L.goff[l + 1] - L.goff[l];
    if (l > 0 && ng > 0)
      k_fwd_gather_il<NRP><<<nblk((int64_t)ng * NRP / 2, 256), 256, 0, s>>>(F, F.lrows + L.goff[l], ng, nrhs);
    const int nt = L.foff[l + 1] - L.foff[l];
    if (L.fw[l] > 0) return (int)hipErrorInvalidValue;
    const int e = launch_sweep(F, F.ftiles + L.foff[l], nt, L.fw[l], L.fm[l], true, nrhs, t, s, L.fks[l],
                               F.frtiles + L.froff[l], L.froff[l + 1] - L.froff[l]);
    if (e) return e;
  }
  for (int l = L.nlevels - 1; l >= 0; --l) {
    const int nx = L.xoff[l + 1] - L.xoff[l];
    // (the LDS-staged sweep, 32 or more right-hand sides, reads the panel
    // columns through the row index)
    if (nx > 0 && !(L.bw[l] == -2 && NRP >= 32))
      k_bwd_gather_il<NRP><<<nblk((int64_t)nx * NRP / 2, 256), 256, 0, s>>>(F, F.xidx + L.xoff[l], nx);
    const int nt = L.boff[l + 1] - L.boff[l];
    if (L.bw[l] > 0) return (int)hipErrorInvalidValue;
    const int e = launch_sweep(F, F.btiles + L.boff[l], nt, L.bw[l], L.bm[l], false, nrhs, t, s, L.bks[l],
                               F.brtiles + L.broff[l], L.broff[l + 1] - L.broff[l]);
    if (e) return e;
  }
  k_unpack_il<NRP><<<nblk(F.n, IL_TP), 256, 0, s>>>(F, v, nrhs);
  return (int)hipGetLastError();
}

int launch_chol_solve(const DevFactor& F, const FactorLevels& L, double* v, double* t, int nrhs, hipStream_t s) {
  if (L.nlevels > 0 && L.fw[0] <= 0) {  // MFMA-shaped factor: interleaved vectors
    if (nrhs > MAX_MFMA_RHS) return (int)hipErrorInvalidValue;
    switch (mfma_rhs_pad(nrhs)) {
      case 16: return chol_solve_il<16>(F, L, v, t, nrhs, s);
      case 32: return chol_solve_il<32>(F, L, v, t, nrhs, s);
      case 64: return chol_solve_il<64>(F, L, v, t, nrhs, s);
      default: return chol_solve_il<128>(F, L, v, t, nrhs, s);
    }
  }
  if (nrhs > 2) return (int)hipErrorInvalidValue;
  for (int l = 0; l < L.nlevels; ++l) {
    // level 0 (leaves) owes nothing to its rows: its sweep reads the right-hand side itself
    const int ng = L.goff[l + 1] - L.goff[l];
    if (l > 0 && ng > 0) {
      const int* lr = F.lrows + L.goff[l];
      if (ng >= 16384) {
        if (nrhs == 2) k_fwd_gather<1, 2><<<nblk(ng, 256), 256, 0, s>>>(F, lr, ng, v);
        else k_fwd_gather<1, 1><<<nblk(ng, 256), 256, 0, s>>>(F, lr, ng, v);
      } else {
        if (nrhs == 2) k_fwd_gather<8, 2><<<nblk(8 * (int64_t)ng, 256), 256, 0, s>>>(F, lr, ng, v);
        else k_fwd_gather<8, 1><<<nblk(8 * (int64_t)ng, 256), 256, 0, s>>>(F, lr, ng, v);
      }
    }
    const int nt = L.foff[l + 1] - L.foff[l];
    const int e = launch_sweep(F, F.ftiles + L.foff[l], nt, L.fw[l], L.fm[l], true, nrhs, t, s, 0, nullptr, 0,
                               l == 0 ? v : nullptr, L.fpf && L.fpf[l]);
    if (e) return e;
  }
  for (int l = L.nlevels - 1; l >= 0; --l) {
    // the backward sweep reads the panel rows' x straight from X (no gather)
    const int nt = L.boff[l + 1] - L.boff[l];
    const int e = launch_sweep(F, F.btiles + L.boff[l], nt, L.bw[l], L.bm[l], false, nrhs, t, s, 0, nullptr, 0,
                               nullptr, L.bpf && L.bpf[l]);
    if (e) return e;
  }
  // a kernel, not a memcpy node, so that the solve can be captured in a graph
  const int64_t nx = (int64_t)nrhs * F.n;
  k_copy<<<(unsigned)std::min<int64_t>(nblk(nx, 256), 4096), 256, 0, s>>>(F.X, v, nx);
  return (int)hipGetLastError();
}

int launch_poisson_cell(int k, const DevPoisson& P, const DevTabs& T, const DevPhys* phs, const double* un,
                        const double* up, const double* ur, const double* uo, int ncs, int nce, double* y,
                        hipStream_t s, int nb, int64_t ss, int64_t se, int64_t sy) {
  if (P.table_f && P.ucell) {  // mixed precision
    PECS_DISPATCH_K(k, (k_poisson_cell_f32<KK><<<dim3(nblk(P.nucell, 64), nb), 64, 0, s>>>(
                           P, T, phs, un, up, ur, uo, ncs, nce, y, ss, se, sy)));
    return (int)hipGetLastError();
  }
  if (P.ucell) {
    PECS_DISPATCH_K(k, (k_poisson_cell<KK, true><<<dim3(nblk(P.nucell, 64), nb), 64, 0, s>>>(
                           P, T, phs, un, up, ur, uo, ncs, nce, y, ss, se, sy)));
  } else {
    PECS_DISPATCH_K(k, (k_poisson_cell<KK, false><<<dim3(nblk(P.ncell, 64), nb), 64, 0, s>>>(
                           P, T, phs, un, up, ur, uo, ncs, nce, y, ss, se, sy)));
  }
  return (int)hipGetLastError();
}

int launch_poisson_cell_from_b(const DevPoisson& P, const double* b, double* y, hipStream_t s) {
  const int k = (P.nphi == 1) ? 1 : (P.nphi == 4 ? 2 : 3);
  PECS_DISPATCH_K(k, (k_poisson_cell_from_b<KK><<<nblk(P.ncell, 64), 64, 0, s>>>(P, b, y)));
  return (int)hipGetLastError();
}

static int launch_poisson_face_main(const DevPoisson& P, const double* y, double* r, hipStream_t s, int nb, int64_t sy,
                        int64_t sr) {
  k_poisson_face<<<dim3(nblk(P.nmult, 256), nb), 256, 0, s>>>(P, y, r, sy, sr);
  return (int)hipGetLastError();
}

static int launch_poisson_recover_main(int k, const DevPoisson& P, const double* lam, const double* y, double* out,
                           hipStream_t s, int nb, int64_t sl, int64_t sy, int64_t so) {
  if (P.table_f && P.ucell) {  // mixed precision
    PECS_DISPATCH_K(k, (k_poisson_recover_f32<KK><<<dim3(nblk(P.nucell, 64), nb), 64, 0, s>>>(P, lam, y, out, sl, sy,
                                                                                              so)));
    return (int)hipGetLastError();
  }
  if (P.ucell) {
    PECS_DISPATCH_K(k, (k_poisson_recover<KK, true><<<dim3(nblk(P.nucell, 64), nb), 64, 0, s>>>(P, lam, y, out, sl,
                                                                                                 sy, so)));
  } else {
    PECS_DISPATCH_K(k, (k_poisson_recover<KK, false><<<dim3(nblk(P.ncell, 64), nb), 64, 0, s>>>(P, lam, y, out, sl,
                                                                                                  sy, so)));
  }
  return (int)hipGetLastError();
}

int launch_schur_rhs(int k, const DevRegion& R, const double* b1, const double* b2, double mu1, double mu2,
                     double* g1, double* g2, int nrhs, hipStream_t s, int nb, int64_t sb, int64_t sg) {
  const int rc = launch_schur_rhs_main(k, R, b1, b2, mu1, mu2, g1, g2, nrhs, s, nb, sb, sg);
  if (rc || R.nhx == 0) return rc;
  const dim3 g(nblk(R.nhx, 64), nb);
  if (nrhs == 2) {
    PECS_DISPATCH_K(k, (k_schur_rhs_hang<KK, 2><<<g, 64, 0, s>>>(R, b1, b2, mu1, mu2, g1, g2, sb, sg)));
  } else {
    PECS_DISPATCH_K(k, (k_schur_rhs_hang<KK, 1><<<g, 64, 0, s>>>(R, b1, b1, mu1, mu1, g1, g1, sb, sg)));
  }
  return (int)hipGetLastError();
}

int launch_schur_recover(int k, const DevRegion& R, const double* b1, const double* b2, const double* xs1,
                         const double* xs2, double mu1, double mu2, double* out1, double* out2, int nrhs,
                         hipStream_t s, int nb, int64_t sb, int64_t sx, int64_t so) {
  const int rc = launch_schur_recover_main(k, R, b1, b2, xs1, xs2, mu1, mu2, out1, out2, nrhs, s, nb, sb, sx, so);
  if (rc || R.nhm == 0) return rc;
  const dim3 g(nblk(R.nhm, 64), nb);
  if (nrhs == 2) {
    PECS_DISPATCH_K(k, (k_schur_recover_hang<KK, 2><<<g, 64, 0, s>>>(R, xs1, xs2, mu1, mu2, out1, out2, sx, so)));
  } else {
    PECS_DISPATCH_K(k, (k_schur_recover_hang<KK, 1><<<g, 64, 0, s>>>(R, xs1, xs1, mu1, mu1, out1, out1, sx, so)));
  }
  return (int)hipGetLastError();
}

int launch_poisson_face(const DevPoisson& P, const double* y, double* r, hipStream_t s, int nb, int64_t sy,
                        int64_t sr) {
  const int rc = launch_poisson_face_main(P, y, r, s, nb, sy, sr);
  if (rc || P.nhang == 0) return rc;
  const int k = (P.nphi == 1) ? 1 : (P.nphi == 4 ? 2 : 3);
  PECS_DISPATCH_K(k, (k_poisson_face_hang<KK><<<dim3(nblk(P.nhang * k, 64), nb), 64, 0, s>>>(P, y, r, sy, sr)));
  return (int)hipGetLastError();
}

int launch_poisson_recover(int k, const DevPoisson& P, const double* lam, const double* y, double* out,
                           hipStream_t s, int nb, int64_t sl, int64_t sy, int64_t so) {
  const int rc = launch_poisson_recover_main(k, P, lam, y, out, s, nb, sl, sy, so);
  if (rc || P.nhc == 0) return rc;
  PECS_DISPATCH_K(k, (k_poisson_recover_hang<KK><<<dim3(nblk(P.nhc, 64), nb), 64, 0, s>>>(P, lam, out, sl, so)));
  return (int)hipGetLastError();
}

int launch_poisson_rhs_assemble(int k, const DevPoisson& P, const DevTabs& T, const DevPhys* phs, const double* un,
                                const double* up, const double* ur, const double* uo, int ncs, int nce, double* b,
                                hipStream_t s) {
  PECS_DISPATCH_K(k, (k_poisson_rhs_assemble<KK><<<nblk(P.ncell, 64), 64, 0, s>>>(P, T, phs, un, up, ur, uo, ncs,
                                                                                   nce, b)));
  return (int)hipGetLastError();
}

int launch_diff_norms(const double* a, const double* b, int64_t n, double* part, hipStream_t s) {
  k_diff_norms<<<RES_BLOCKS, 256, 0, s>>>(a, b, n, part);
  return (int)hipGetLastError();
}

int launch_copy(const double* src, double* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  k_copy<<<(unsigned)std::min<int64_t>(nblk(n, 256), 4096), 256, 0, s>>>(src, dst, n);
  return (int)hipGetLastError();
}

int launch_grid_sample(int k, const double* u, int r, int G, double* out, hipStream_t s) {
  PECS_DISPATCH_K(k, (k_grid_sample<KK><<<nblk((int64_t)G * G, 256), 256, 0, s>>>(u, r, G, out)));
  return (int)hipGetLastError();
}

int launch_grid_project(int k, const double* grid, int G, int r, double* u, hipStream_t s) {
  const int64_t n = ((int64_t)1 << (2 * r)) * (k + 1) * (k + 1);
  PECS_DISPATCH_K(k, (k_grid_project<KK><<<nblk(n, 256), 256, 0, s>>>(grid, G, r, u)));
  return (int)hipGetLastError();
}

int launch_sum_partials(const double* part, int nvec, double* acc, int ninst, hipStream_t s) {
  k_sum_partials<<<ninst, 256, 0, s>>>(part, nvec * RES_BLOCKS, acc);
  return (int)hipGetLastError();
}

}  // namespace pecs
