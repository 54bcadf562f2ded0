// Learned warm start: a network trained on steady states produced by the
// time stepper.  A fully connected encoder (nin inputs -> nlayers hidden
// ReLU layers) feeds either nout linear outputs (a plain MLP) or a
// convolutional decoder: the encoder's last layer (ReLU) is read as a
// c0 x g0 x g0 feature map, and each decoder stage is an optional nearest
// x2 upsampling followed by a 3x3 convolution (zero padding), ReLU after all
// but the last stage.  Training is FP32 (forward, backward, Adam); inference
// is FP32 or BF16 storage with FP32 accumulation, all in the kernels of this
// file; the C-ABI (include/pecs_abi.h, pecs_mlp_*) moves host arrays in and
// out.
//
// Every matrix product -- the dense layers, their gradients, and the
// convolutions with their weight and input gradients as implicit GEMMs
// (im2col gathered while staging the LDS tiles) -- is one kernel, k_mm, on
// the matrix cores: v_mfma_f32_16x16x4f32 for FP32, v_mfma_f32_16x16x16bf16_1k
// for BF16.  MFMA operand layout (both): lane l holds A[l%16][k], B[k][l%16]
// with k = l/16 (FP32) or 4(l/16)..4(l/16)+3 (BF16); accumulator lane l,
// register r holds C[4(l/16)+r][l%16].
//
// Parameters are one FP32 array: per dense layer W (n_in x n_out, row-major)
// then b (n_out); per convolution W (cout x cin x 3 x 3) then b (cout).
// Feature maps are NCHW per sample.  Loss: mean squared error over the batch
// and outputs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pecs_abi.h"

namespace {

thread_local std::string g_mlp_error;

#define MCHECK(x)                                                                                 \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Round to nearest even (NaN not expected here).
__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <class T>
__device__ __forceinline__ float ldv(const T* p, int64_t i);
template <>
__device__ __forceinline__ float ldv<float>(const float* p, int64_t i) {
  return p[i];
}
template <>
__device__ __forceinline__ float ldv<uint16_t>(const uint16_t* p, int64_t i) {
  return bf16_to_f32(p[i]);
}

// An operand matrix of k_mm, read element by element while staging tiles.
enum OpKind {
  OP_ROW = 0,     // (r, c) = p[r ld + c]
  OP_TRANS = 1,   // (r, c) = p[c ld + r]
  OP_IM2COL = 2,  // (ci 9 + t, b R^2 + px): input pixel under tap t of output pixel px (conv forward B)
  OP_IM2COL_T = 3,  // its transpose (weight gradient B)
  OP_CHAN = 4,    // (ch, b R^2 + px) of a C x R x R map per sample (weight gradient A)
  OP_WT = 5,      // (ci, co 9 + t) of W[co][ci][t], C = cin (input gradient A)
  OP_DCOL = 6     // (co 9 + t, b R^2 + px): output gradient at px - offset(t) (input gradient B)
};

struct Opnd {
  const void* p;
  int kind;
  int ld;  // OP_ROW / OP_TRANS
  int C;   // channels of the map read (OP_WT: cin)
  int R;   // side of the convolution's output (and upsampled input)
  int up;  // OP_IM2COL*: the input map is R/2 x R/2, upsampled by 2 (nearest)
};

template <class T>
__device__ __forceinline__ float opnd(const Opnd& o, int r, int c) {
  const T* p = static_cast<const T*>(o.p);
  switch (o.kind) {
    case OP_ROW:
      return ldv(p, (int64_t)r * o.ld + c);
    case OP_TRANS:
      return ldv(p, (int64_t)c * o.ld + r);
    case OP_IM2COL:
    case OP_IM2COL_T: {
      const int k = o.kind == OP_IM2COL ? r : c, n = o.kind == OP_IM2COL ? c : r;
      const int R2 = o.R * o.R;
      const int ci = k / 9, t = k - 9 * ci;
      const int b = n / R2, px = n - b * R2;
      const int y = px / o.R + t / 3 - 1, x = px % o.R + t % 3 - 1;
      if (y < 0 || y >= o.R || x < 0 || x >= o.R) return 0.f;
      const int Ri = o.R >> o.up;
      return ldv(p, (((int64_t)b * o.C + ci) * Ri + (y >> o.up)) * Ri + (x >> o.up));
    }
    case OP_CHAN: {
      const int R2 = o.R * o.R;
      const int b = c / R2, px = c - b * R2;
      return ldv(p, ((int64_t)b * o.C + r) * R2 + px);
    }
    case OP_WT: {
      const int co = c / 9, t = c - 9 * co;
      return ldv(p, ((int64_t)co * o.C + r) * 9 + t);
    }
    default: {  // OP_DCOL
      const int R2 = o.R * o.R;
      const int co = r / 9, t = r - 9 * co;
      const int b = c / R2, px = c - b * R2;
      const int y = px / o.R - (t / 3 - 1), x = px % o.R - (t % 3 - 1);
      if (y < 0 || y >= o.R || x < 0 || x >= o.R) return 0.f;
      return ldv(p, (((int64_t)b * o.C + co) * o.R + y) * o.R + x);
    }
  }
}

// Where k_mm writes C[m][n]: row-major (ld), or as channel m of the maps
// (n = b R^2 + px).  Bias per column (dense) or per row (convolution).
struct Dst {
  void* p;
  int chan;  // 0: p[m ld + n]; 1: p[(b C + m) R^2 + px]
  int ld, C, R;
  const float* bias;
  int bias_row;
  int relu;
  int bf16;  // store BF16 (hidden activations of BF16 inference)
};

constexpr int TM = 64, TN = 64, TK = 16;  // tile: 64 x 64 outputs, 16-deep k steps, 4 waves of 32 x 32

// Epilogue of C[m][n] = v: bias, ReLU, store (row-major or as map channel m).
__device__ __forceinline__ void mm_store(const Dst& D, int m, int n, float v) {
  if (D.bias) v += D.bias_row ? D.bias[m] : D.bias[n];
  if (D.relu && v < 0.f) v = 0.f;
  int64_t o;
  if (D.chan) {
    const int R2 = D.R * D.R;
    const int b = n / R2, px = n - b * R2;
    o = ((int64_t)b * D.C + m) * R2 + px;
  } else {
    o = (int64_t)m * D.ld + n;
  }
  if (D.bf16) static_cast<uint16_t*>(D.p)[o] = f32_to_bf16(v);
  else static_cast<float*>(D.p)[o] = v;
}

// C[M x N] = A[M x K] B[K x N] on the matrix cores.  T: FP32 (MFMA 16x16x4
// f32) or BF16 (MFMA 16x16x16 bf16, operands as stored, FP32 accumulation).
// The k loop runs in a fixed order: results are reproducible.
// Split k: block z of gridDim.z covers k in [z kchunk, (z + 1) kchunk) and
// writes its raw sums to part[(z M + m) N + n]; k_mm_reduce adds the slices
// in order (deterministic) and applies the epilogue.  Used where M x N has
// few tiles and K is long (the weight gradients, K = batch x pixels).
template <class T>
__global__ __launch_bounds__(256) void k_mm(Opnd A, Opnd B, Dst D, int M, int N, int K, int kchunk,
                                            float* __restrict__ part) {
  constexpr bool BF = sizeof(T) == 2;
  // BF16 rows padded by 4 halves (8-byte aligned 4-wide reads), FP32 by 1
  __shared__ T As[TM][TK + (BF ? 4 : 1)];
  __shared__ T Bs[TK][TN + (BF ? 4 : 1)];
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const int li = l & 15, lq = l >> 4;
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    for (int e = threadIdx.x; e < TM * TK; e += 256) {
      const int r = e / TK, kk = e % TK;
      const int gm = m0 + r, gk = k0 + kk;
      const float v = (gm < M && gk < kend) ? opnd<T>(A, gm, gk) : 0.f;
      if constexpr (BF) As[r][kk] = f32_to_bf16(v);
      else As[r][kk] = v;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      const int kk = e / TN, c = e % TN;
      const int gn = n0 + c, gk = k0 + kk;
      const float v = (gn < N && gk < kend) ? opnd<T>(B, gk, gn) : 0.f;
      if constexpr (BF) Bs[kk][c] = f32_to_bf16(v);
      else Bs[kk][c] = v;
    }
    __syncthreads();
    if constexpr (BF) {
      v4s a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          a[i][e] = (short)As[wm + 16 * i + li][4 * lq + e];
          b[i][e] = (short)Bs[4 * lq + e][wn + 16 * i + li];
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    } else {
#pragma unroll
      for (int ks = 0; ks < TK; ks += 4) {
        float a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          a[i] = As[wm + 16 * i + li][ks + lq];
          b[i] = Bs[ks + lq][wn + 16 * i + li];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * lq + r, n = n0 + wn + 16 * j + li;
        if (m >= M || n >= N) continue;
        if (part) part[((int64_t)blockIdx.z * M + m) * N + n] = acc[i][j][r];
        else mm_store(D, m, n, acc[i][j][r]);
      }
}

__global__ __launch_bounds__(256) void k_mm_reduce(const float* __restrict__ part, int splits, Dst D, int M, int N) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)M * N) return;
  float v = 0.f;
  for (int z = 0; z < splits; ++z) v += part[(int64_t)z * M * N + e];
  mm_store(D, (int)(e / N), (int)(e % N), v);
}

// db[c] = sum over rows of dZ[r][c] (fixed order)
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, float* __restrict__ db, int m, int n) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  float s = 0.f;
  for (int r = 0; r < m; ++r) s += dZ[(int64_t)r * n + c];
  db[c] = s;
}

// db[ch] = sum over samples and pixels of dZ[b][ch][px] (one block per
// channel, fixed-order tree reduction)
__global__ __launch_bounds__(256) void k_chansum(const float* __restrict__ dZ, float* __restrict__ db, int batch,
                                                 int C, int R2) {
  __shared__ float s[256];
  const int ch = blockIdx.x;
  float acc = 0.f;
  const int64_t n = (int64_t)batch * R2;
  for (int64_t e = threadIdx.x; e < n; e += 256) {
    const int64_t b = e / R2, px = e - b * R2;
    acc += dZ[(b * C + ch) * R2 + px];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[ch] = s[0];
}

// gradient of a nearest x2 upsampling: each input pixel sums its 2 x 2 block
__global__ __launch_bounds__(256) void k_down2_sum(const float* __restrict__ dUp, float* __restrict__ dIn,
                                                   int64_t maps, int Ri) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= maps * Ri * Ri) return;
  const int64_t mp = e / ((int64_t)Ri * Ri);
  const int px = (int)(e - mp * Ri * Ri), Y = px / Ri, X = px % Ri;
  const int R = 2 * Ri;
  const float* u = dUp + mp * R * R + (int64_t)(2 * Y) * R + 2 * X;
  dIn[e] = (u[0] + u[1]) + (u[R] + u[R + 1]);
}

// dZ = dA where A > 0 (ReLU backward; A is the layer's output)
__global__ __launch_bounds__(256) void k_relu_back(float* __restrict__ dA, const float* __restrict__ A, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n && A[i] <= 0.f) dA[i] = 0.f;
}

// dY = 2 (Y - T) / count; per-block partial sums of (Y - T)^2 into part
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             float* __restrict__ dY, float* __restrict__ part, int64_t n) {
  __shared__ float s[256];
  float acc = 0.f;
  const float sc = 2.f / (float)n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float d = Y[i] - T[i];
    dY[i] = sc * d;
    acc += d * d;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                                              float c1, float c2, float eps) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = f32_to_bf16(a[i]);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// k slices for an M x N x K product: enough workgroups to fill the chip
// when the output has few tiles, slices of at least 1024 deep (a multiple
// of TK), at most MAX_SPLIT
constexpr int MAX_SPLIT = 256;
inline int mm_splits(int M, int N, int K) {
  const int64_t tiles = (int64_t)nblk(N, TN) * nblk(M, TM);
  if (tiles >= 256 || K < 2048) return 1;
  int s = (int)std::min<int64_t>(512 / tiles, K / 1024);
  return std::max(1, std::min(s, MAX_SPLIT));
}

template <class T>
void mm(const Opnd& A, const Opnd& B, const Dst& D, int M, int N, int K, hipStream_t s, float* part = nullptr,
        int64_t part_cap = 0) {
  int splits = part ? mm_splits(M, N, K) : 1;
  if ((int64_t)splits * M * N > part_cap) splits = 1;
  const int kchunk = splits > 1 ? (((K + splits - 1) / splits + TK - 1) / TK) * TK : K;
  if (splits > 1) splits = (K + kchunk - 1) / kchunk;
  dim3 grid(nblk(N, TN), nblk(M, TM), splits);
  k_mm<T><<<grid, 256, 0, s>>>(A, B, D, M, N, K, kchunk, splits > 1 ? part : nullptr);
  MCHECK(hipGetLastError());
  if (splits > 1) {
    k_mm_reduce<<<nblk((int64_t)M * N, 256), 256, 0, s>>>(part, splits, D, M, N);
    MCHECK(hipGetLastError());
  }
}

Opnd op_row(const void* p, int ld) { return Opnd{p, OP_ROW, ld, 0, 0, 0}; }
Opnd op_trans(const void* p, int ld) { return Opnd{p, OP_TRANS, ld, 0, 0, 0}; }
Dst dst_row(void* p, int ld, const float* bias, bool relu, bool bf16) {
  return Dst{p, 0, ld, 0, 0, bias, 0, relu ? 1 : 0, bf16 ? 1 : 0};
}
Dst dst_chan(void* p, int C, int R, const float* bias, bool relu, bool bf16) {
  return Dst{p, 1, 0, C, R, bias, 1, relu ? 1 : 0, bf16 ? 1 : 0};
}

struct Layer {
  bool conv = false;
  int nin = 0, nout = 0;             // dense: widths
  int cin = 0, cout = 0, R = 0, up = 0;  // convolution: channels, output side, upsampled input
  int64_t woff = 0, boff = 0;
  int64_t in_size() const { return conv ? (int64_t)cin * (R >> up) * (R >> up) : nin; }
  int64_t out_size() const { return conv ? (int64_t)cout * R * R : nout; }
};

}  // namespace

struct pecs_mlp {
  std::vector<Layer> layers;
  int nin = 0, nout = 0;
  int64_t np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // parameters, gradients, Adam moments
  uint16_t* pb = nullptr;                                        // BF16 copy of the parameters (inference)
  bool pb_valid = false;
  int64_t t = 0;  // Adam step count
  // work buffers, sized for cap rows
  int cap = 0;
  std::vector<float*> act;  // act[l]: cap x size of layer l's input (act[L] = output)
  std::vector<float*> dact;
  std::vector<uint16_t*> actb;
  float* dup = nullptr;  // input gradient of an upsampling convolution at the upsampled size
  float* kpart = nullptr;  // split-k partial sums of the weight gradients
  int64_t kpart_cap = 0;
  float *target = nullptr, *part = nullptr;
  hipStream_t s = nullptr;
  ~pecs_mlp() {
    if (s) (void)hipStreamSynchronize(s);
    auto f = [](void* q) {
      if (q) (void)hipFree(q);
    };
    f(p), f(g), f(m), f(v), f(pb), f(target), f(part), f(dup), f(kpart);
    for (float* a : act) f(a);
    for (float* a : dact) f(a);
    for (uint16_t* a : actb) f(a);
    if (s) (void)hipStreamDestroy(s);
  }
  int L() const { return (int)layers.size(); }
  int64_t size(int l) const { return l == 0 ? nin : layers[l - 1].out_size(); }
  void ensure(int rows) {
    if (rows <= cap) return;
    (void)hipStreamSynchronize(s);
    for (float* a : act) (void)hipFree(a);
    for (float* a : dact) (void)hipFree(a);
    for (uint16_t* a : actb) (void)hipFree(a);
    if (target) (void)hipFree(target);
    if (dup) (void)hipFree(dup);
    dup = nullptr;
    act.assign(L() + 1, nullptr);
    dact.assign(L() + 1, nullptr);
    actb.assign(L() + 1, nullptr);
    for (int l = 0; l <= L(); ++l) {
      MCHECK(hipMalloc(&act[l], (size_t)rows * size(l) * sizeof(float)));
      MCHECK(hipMalloc(&dact[l], (size_t)rows * size(l) * sizeof(float)));
      MCHECK(hipMalloc(&actb[l], (size_t)rows * size(l) * sizeof(uint16_t)));
    }
    int64_t dmax = 0;
    for (const Layer& y : layers)
      if (y.conv && y.up) dmax = std::max(dmax, (int64_t)y.cin * y.R * y.R);
    if (dmax) MCHECK(hipMalloc(&dup, (size_t)rows * dmax * sizeof(float)));
    MCHECK(hipMalloc(&target, (size_t)rows * nout * sizeof(float)));
    cap = rows;
  }
  // layer l (1-based) from act[l-1] (or actb) to act[l] (or actb)
  template <class T>
  void layer_fwd(int l, int rows, const T* in, void* out, const T* w, bool bf16out) {
    const Layer& y = layers[l - 1];
    const bool relu = l < L();
    const float* bias = p + y.boff;
    if (!y.conv) {
      mm<T>(op_row(in, y.nin), op_row(w + y.woff, y.nout), dst_row(out, y.nout, bias, relu, bf16out), rows, y.nout,
            y.nin, s);
    } else {
      const Opnd B{in, OP_IM2COL, 0, y.cin, y.R, y.up};
      mm<T>(op_row(w + y.woff, y.cin * 9), B, dst_chan(out, y.cout, y.R, bias, relu, bf16out), y.cout,
            rows * y.R * y.R, y.cin * 9, s);
    }
  }
  void forward(int rows) {
    for (int l = 1; l <= L(); ++l) layer_fwd<float>(l, rows, act[l - 1], act[l], p, false);
  }
  // loss gradient in dact[L]; returns the loss (synchronises)
  float loss_grad(int rows) {
    const int64_t n = (int64_t)rows * nout;
    k_mse<<<256, 256, 0, s>>>(act[L()], target, dact[L()], part, n);
    MCHECK(hipGetLastError());
    std::vector<float> h(256);
    MCHECK(hipMemcpyAsync(h.data(), part, 256 * sizeof(float), hipMemcpyDeviceToHost, s));
    MCHECK(hipStreamSynchronize(s));
    double acc = 0.0;
    for (float x : h) acc += x;
    return (float)(acc / (double)n);
  }
  void backward(int rows) {
    for (int l = L(); l >= 1; --l) {
      const Layer& y = layers[l - 1];
      if (!y.conv) {
        // dW = act[l-1]^T dact[l]; db = colsum(dact[l]); dact[l-1] = dact[l] W^T
        mm<float>(op_trans(act[l - 1], y.nin), op_row(dact[l], y.nout), dst_row(g + y.woff, y.nout, nullptr, false, false),
                  y.nin, y.nout, rows, s, kpart, kpart_cap);
        k_colsum<<<nblk(y.nout, 256), 256, 0, s>>>(dact[l], g + y.boff, rows, y.nout);
        MCHECK(hipGetLastError());
        if (l > 1)
          mm<float>(op_row(dact[l], y.nout), op_trans(p + y.woff, y.nout),
                    dst_row(dact[l - 1], y.nin, nullptr, false, false), rows, y.nin, y.nout, s);
      } else {
        const int R2 = y.R * y.R;
        // dW[co][ci t] = sum over (b, px) of dOut[b][co][px] im2col(in)[ci t][b px]
        mm<float>(Opnd{dact[l], OP_CHAN, 0, y.cout, y.R, 0}, Opnd{act[l - 1], OP_IM2COL_T, 0, y.cin, y.R, y.up},
                  dst_row(g + y.woff, y.cin * 9, nullptr, false, false), y.cout, y.cin * 9, rows * R2, s, kpart,
                  kpart_cap);
        k_chansum<<<y.cout, 256, 0, s>>>(dact[l], g + y.boff, rows, y.cout, R2);
        MCHECK(hipGetLastError());
        if (l > 1) {
          // d(upsampled input)[ci][b px] = sum over (co, t) of W[co][ci][t] dOut[b][co][px - offset(t)]
          float* dst = y.up ? dup : dact[l - 1];
          mm<float>(Opnd{p + y.woff, OP_WT, 0, y.cin, 0, 0}, Opnd{dact[l], OP_DCOL, 0, y.cout, y.R, 0},
                    dst_chan(dst, y.cin, y.R, nullptr, false, false), y.cin, rows * R2, y.cout * 9, s);
          if (y.up) {
            const int Ri = y.R / 2;
            const int64_t n = (int64_t)rows * y.cin * Ri * Ri;
            k_down2_sum<<<nblk(n, 256), 256, 0, s>>>(dup, dact[l - 1], (int64_t)rows * y.cin, Ri);
            MCHECK(hipGetLastError());
          }
        }
      }
      if (l > 1) {  // every layer but the last ends in a ReLU
        const int64_t na = (int64_t)rows * size(l - 1);
        k_relu_back<<<nblk(na, 256), 256, 0, s>>>(dact[l - 1], act[l - 1], na);
        MCHECK(hipGetLastError());
      }
    }
  }
};

namespace {

int mlp_fail(const std::exception& e) {
  g_mlp_error = e.what();
  return -1;
}

// parameter offsets, He-uniform weights for the ReLU layers, Glorot-uniform
// for the output layer, zero biases; device arrays and the stream
void mlp_init(pecs_mlp* net, unsigned seed) {
  int64_t off = 0;
  for (Layer& y : net->layers) {
    y.woff = off;
    off += y.conv ? (int64_t)y.cout * y.cin * 9 : (int64_t)y.nin * y.nout;
    y.boff = off;
    off += y.conv ? y.cout : y.nout;
  }
  net->np = off;
  int64_t wmax = 0;
  for (const Layer& y : net->layers) wmax = std::max(wmax, y.boff - y.woff);
  net->kpart_cap = std::min<int64_t>((int64_t)MAX_SPLIT * wmax, (int64_t)1 << 24);  // larger products do not split
  MCHECK(hipMalloc(&net->kpart, net->kpart_cap * sizeof(float)));
  std::vector<float> h(off, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 1; l <= net->L(); ++l) {
    const Layer& y = net->layers[l - 1];
    const double fi = y.conv ? 9.0 * y.cin : y.nin, fo = y.conv ? 9.0 * y.cout : y.nout;
    const double a = l < net->L() ? std::sqrt(6.0 / fi) : std::sqrt(6.0 / (fi + fo));
    std::uniform_real_distribution<double> U(-a, a);
    for (int64_t i = y.woff; i < y.boff; ++i) h[i] = (float)U(rng);
  }
  MCHECK(hipStreamCreateWithFlags(&net->s, hipStreamNonBlocking));
  MCHECK(hipMalloc(&net->p, off * sizeof(float)));
  MCHECK(hipMalloc(&net->g, off * sizeof(float)));
  MCHECK(hipMalloc(&net->m, off * sizeof(float)));
  MCHECK(hipMalloc(&net->v, off * sizeof(float)));
  MCHECK(hipMalloc(&net->pb, off * sizeof(uint16_t)));
  MCHECK(hipMalloc(&net->part, 256 * sizeof(float)));
  MCHECK(hipMemcpy(net->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice));
  MCHECK(hipMemset(net->m, 0, off * sizeof(float)));
  MCHECK(hipMemset(net->v, 0, off * sizeof(float)));
  MCHECK(hipMemset(net->g, 0, off * sizeof(float)));
}

}  // namespace

extern "C" {

const char* pecs_mlp_last_error(void) { return g_mlp_error.c_str(); }

int pecs_mlp_create(int nin, int nhidden, int nlayers, int nout, unsigned seed, pecs_mlp** out) {
  if (!out) return -1;
  *out = nullptr;
  try {
    if (nin < 1 || nhidden < 1 || nlayers < 1 || nlayers > 8 || nout < 1)
      throw std::runtime_error("pecs_mlp_create: sizes must be positive (1..8 hidden layers)");
    auto net = std::make_unique<pecs_mlp>();
    net->nin = nin;
    net->nout = nout;
    int prev = nin;
    for (int l = 0; l <= nlayers; ++l) {
      Layer y;
      y.nin = prev;
      y.nout = l < nlayers ? nhidden : nout;
      net->layers.push_back(y);
      prev = y.nout;
    }
    mlp_init(net.get(), seed);
    *out = net.release();
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

int pecs_mlp_create_decoder(int nin, int nhidden, int nlayers, int c0, int g0, int nconv, const int* conv_cout,
                            const int* conv_up, unsigned seed, pecs_mlp** out) {
  if (!out) return -1;
  *out = nullptr;
  try {
    if (nin < 1 || nhidden < 1 || nlayers < 1 || nlayers > 8 || c0 < 1 || g0 < 1 || nconv < 1 || nconv > 8 ||
        !conv_cout || !conv_up)
      throw std::runtime_error("pecs_mlp_create_decoder: sizes must be positive (1..8 hidden layers, 1..8 stages)");
    auto net = std::make_unique<pecs_mlp>();
    net->nin = nin;
    int prev = nin;
    for (int l = 0; l <= nlayers; ++l) {  // the last dense layer is the c0 x g0 x g0 feature map
      Layer y;
      y.nin = prev;
      y.nout = l < nlayers ? nhidden : c0 * g0 * g0;
      net->layers.push_back(y);
      prev = y.nout;
    }
    int cin = c0, side = g0;
    for (int j = 0; j < nconv; ++j) {
      if (conv_cout[j] < 1 || (conv_up[j] != 0 && conv_up[j] != 1))
        throw std::runtime_error("pecs_mlp_create_decoder: stage channels must be positive and up 0 or 1");
      Layer y;
      y.conv = true;
      y.cin = cin;
      y.cout = conv_cout[j];
      y.up = conv_up[j];
      y.R = side << y.up;
      if ((int64_t)y.R * y.R > (1 << 20)) throw std::runtime_error("pecs_mlp_create_decoder: maps above 1024 x 1024");
      net->layers.push_back(y);
      cin = y.cout;
      side = y.R;
    }
    net->nout = cin * side * side;
    mlp_init(net.get(), seed);
    *out = net.release();
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

int pecs_mlp_destroy(pecs_mlp* net) {
  if (!net) return -1;
  delete net;
  return 0;
}

int64_t pecs_mlp_num_params(pecs_mlp* net) { return net ? net->np : -1; }

int64_t pecs_mlp_num_outputs(pecs_mlp* net) { return net ? net->nout : -1; }

int pecs_mlp_get_params(pecs_mlp* net, float* params) {
  if (!net || !params) return -1;
  try {
    MCHECK(hipStreamSynchronize(net->s));
    MCHECK(hipMemcpy(params, net->p, net->np * sizeof(float), hipMemcpyDeviceToHost));
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

int pecs_mlp_set_params(pecs_mlp* net, const float* params) {
  if (!net || !params) return -1;
  try {
    MCHECK(hipStreamSynchronize(net->s));
    MCHECK(hipMemcpy(net->p, params, net->np * sizeof(float), hipMemcpyHostToDevice));
    net->pb_valid = false;
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

int pecs_mlp_gradients(pecs_mlp* net, const float* x, const float* y, int batch, float* grads, float* loss) {
  if (!net || !x || !y || batch < 1) return -1;
  try {
    net->ensure(batch);
    MCHECK(hipMemcpyAsync(net->act[0], x, (size_t)batch * net->nin * sizeof(float), hipMemcpyHostToDevice, net->s));
    MCHECK(hipMemcpyAsync(net->target, y, (size_t)batch * net->nout * sizeof(float), hipMemcpyHostToDevice, net->s));
    net->forward(batch);
    const float l = net->loss_grad(batch);
    net->backward(batch);
    if (grads) MCHECK(hipMemcpyAsync(grads, net->g, net->np * sizeof(float), hipMemcpyDeviceToHost, net->s));
    MCHECK(hipStreamSynchronize(net->s));
    if (loss) *loss = l;
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

int pecs_mlp_train_step(pecs_mlp* net, const float* x, const float* y, int batch, float lr, float* loss) {
  if (!net || !x || !y || batch < 1 || !(lr > 0)) return -1;
  try {
    float l = 0.f;
    if (pecs_mlp_gradients(net, x, y, batch, nullptr, &l) != 0) throw std::runtime_error(g_mlp_error);
    net->t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)net->t), c2 = 1.f - std::pow(b2, (float)net->t);
    k_adam<<<nblk(net->np, 256), 256, 0, net->s>>>(net->p, net->g, net->m, net->v, net->np, lr, b1, b2, c1, c2,
                                                   1e-8f);
    MCHECK(hipGetLastError());
    net->pb_valid = false;
    if (loss) *loss = l;
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

int pecs_mlp_predict(pecs_mlp* net, const float* x, int batch, float* y, int bf16) {
  if (!net || !x || !y || batch < 1) return -1;
  try {
    net->ensure(batch);
    MCHECK(hipMemcpyAsync(net->act[0], x, (size_t)batch * net->nin * sizeof(float), hipMemcpyHostToDevice, net->s));
    const int L = net->L();
    if (!bf16) {
      net->forward(batch);
    } else {
      if (!net->pb_valid) {
        k_to_bf16<<<nblk(net->np, 256), 256, 0, net->s>>>(net->p, net->pb, net->np);
        MCHECK(hipGetLastError());
        net->pb_valid = true;
      }
      const int64_t n0 = (int64_t)batch * net->nin;
      k_to_bf16<<<nblk(n0, 256), 256, 0, net->s>>>(net->act[0], net->actb[0], n0);
      MCHECK(hipGetLastError());
      // BF16 weights and activations, FP32 accumulation and biases; hidden
      // outputs rounded to BF16, the last layer's output kept in FP32
      for (int l = 1; l <= L; ++l) {
        if (l < L) net->layer_fwd<uint16_t>(l, batch, net->actb[l - 1], net->actb[l], net->pb, true);
        else net->layer_fwd<uint16_t>(l, batch, net->actb[l - 1], net->act[l], net->pb, false);
      }
    }
    MCHECK(hipMemcpyAsync(y, net->act[L], (size_t)batch * net->nout * sizeof(float), hipMemcpyDeviceToHost, net->s));
    MCHECK(hipStreamSynchronize(net->s));
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

}  // extern "C"
