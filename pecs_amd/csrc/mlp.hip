// Learned warm start: a small fully connected network trained on steady
// states produced by the time stepper.  Training is FP32 (forward, backward,
// Adam) and inference is BF16 storage with FP32 accumulation, all in the
// kernels of this file; the C-ABI (include/pecs_abi.h, pecs_mlp_*) moves host
// arrays in and out.
//
// Network: nin inputs -> nlayers hidden layers of nh units (ReLU) -> nout
// linear outputs.  Weights W_l (n_{l-1} x n_l) and biases b_l, row-major, all
// in one parameter array (W_1, b_1, W_2, b_2, ...).  Loss: mean squared error
// over the batch and outputs.
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

constexpr int TM = 64, TN = 64, TK = 16;  // GEMM tile: 64 x 64 outputs, 16-deep k steps, 256 threads (4 x 4 each)

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
__device__ __forceinline__ float ld(const T* p, int64_t i);
template <>
__device__ __forceinline__ float ld<float>(const float* p, int64_t i) {
  return p[i];
}
template <>
__device__ __forceinline__ float ld<uint16_t>(const uint16_t* p, int64_t i) {
  return bf16_to_f32(p[i]);
}

// C[m x n] = op(A) op(B) (+ bias[n]) (ReLU), op = transpose when TA / TB.
// A is m x k (TA: stored k x m), B is k x n (TB: stored n x k), row-major.
// Each thread accumulates a 4 x 4 block in FP32; the k loop runs in a fixed
// order (results are reproducible).  With BF16 = true, inputs A and B are
// BF16 and the output is rounded to BF16 (inference); otherwise all FP32.
template <bool TA, bool TB, class TIn, class TOut>
__global__ __launch_bounds__(256) void k_gemm(const TIn* __restrict__ A, const TIn* __restrict__ B,
                                              const float* __restrict__ bias, TOut* __restrict__ C, int m, int n,
                                              int k, int relu) {
  __shared__ float As[TK][TM + 4];
  __shared__ float Bs[TK][TN + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int row0 = blockIdx.y * TM, col0 = blockIdx.x * TN;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < k; k0 += TK) {
    for (int e = threadIdx.x; e < TK * TM; e += 256) {
      const int kk = e / TM, r = e % TM;
      const int gr = row0 + r, gk = k0 + kk;
      float v = 0.f;
      if (gr < m && gk < k) v = TA ? ld(A, (int64_t)gk * m + gr) : ld(A, (int64_t)gr * k + gk);
      As[kk][r] = v;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      const int kk = e / TN, c = e % TN;
      const int gc = col0 + c, gk = k0 + kk;
      float v = 0.f;
      if (gc < n && gk < k) v = TB ? ld(B, (int64_t)gc * k + gk) : ld(B, (int64_t)gk * n + gc);
      Bs[kk][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = row0 + ty + 16 * i;
    if (r >= m) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = col0 + tx + 16 * j;
      if (c >= n) continue;
      float v = acc[i][j] + (bias ? bias[c] : 0.f);
      if (relu && v < 0.f) v = 0.f;
      if constexpr (sizeof(TOut) == 2)
        C[(int64_t)r * n + c] = f32_to_bf16(v);
      else
        C[(int64_t)r * n + c] = v;
    }
  }
}

// db[c] = sum over rows of dZ[r][c] (fixed order)
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, float* __restrict__ db, int m, int n) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  float s = 0.f;
  for (int r = 0; r < m; ++r) s += dZ[(int64_t)r * n + c];
  db[c] = s;
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

template <bool TA, bool TB, class TIn, class TOut>
void gemm(const TIn* A, const TIn* B, const float* bias, TOut* C, int m, int n, int k, bool relu, hipStream_t s) {
  dim3 grid(nblk(n, TN), nblk(m, TM));
  k_gemm<TA, TB, TIn, TOut><<<grid, 256, 0, s>>>(A, B, bias, C, m, n, k, relu ? 1 : 0);
  MCHECK(hipGetLastError());
}

}  // namespace

struct pecs_mlp {
  int nin = 0, nh = 0, nl = 0, nout = 0;
  std::vector<int> dims;            // nin, nh x nl, nout
  std::vector<int64_t> woff, boff;  // per layer offsets into the parameter array
  int64_t np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // parameters, gradients, Adam moments
  uint16_t* pb = nullptr;                                        // BF16 copy of the parameters (inference)
  bool pb_valid = false;
  int64_t t = 0;  // Adam step count
  // work buffers, sized for cap rows
  int cap = 0;
  std::vector<float*> act;  // act[l]: cap x dims[l] (act[0] = input)
  std::vector<float*> dact;
  std::vector<uint16_t*> actb;
  float *target = nullptr, *part = nullptr;
  hipStream_t s = nullptr;
  ~pecs_mlp() {
    if (s) (void)hipStreamSynchronize(s);
    auto f = [](void* q) {
      if (q) (void)hipFree(q);
    };
    f(p), f(g), f(m), f(v), f(pb), f(target), f(part);
    for (float* a : act) f(a);
    for (float* a : dact) f(a);
    for (uint16_t* a : actb) f(a);
    if (s) (void)hipStreamDestroy(s);
  }
  int L() const { return nl + 1; }
  void ensure(int rows) {
    if (rows <= cap) return;
    (void)hipStreamSynchronize(s);
    for (float* a : act) (void)hipFree(a);
    for (float* a : dact) (void)hipFree(a);
    for (uint16_t* a : actb) (void)hipFree(a);
    if (target) (void)hipFree(target);
    act.assign(L() + 1, nullptr);
    dact.assign(L() + 1, nullptr);
    actb.assign(L() + 1, nullptr);
    for (int l = 0; l <= L(); ++l) {
      MCHECK(hipMalloc(&act[l], (size_t)rows * dims[l] * sizeof(float)));
      MCHECK(hipMalloc(&dact[l], (size_t)rows * dims[l] * sizeof(float)));
      MCHECK(hipMalloc(&actb[l], (size_t)rows * dims[l] * sizeof(uint16_t)));
    }
    MCHECK(hipMalloc(&target, (size_t)rows * nout * sizeof(float)));
    cap = rows;
  }
  void forward(int rows) {
    for (int l = 1; l <= L(); ++l)
      gemm<false, false, float, float>(act[l - 1], p + woff[l - 1], p + boff[l - 1], act[l], rows, dims[l],
                                       dims[l - 1], l < L(), s);
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
      // dW_l = act[l-1]^T dact[l]; db_l = colsum(dact[l]); dact[l-1] = dact[l] W_l^T
      gemm<true, false, float, float>(act[l - 1], dact[l], nullptr, g + woff[l - 1], dims[l - 1], dims[l], rows, false,
                                      s);
      k_colsum<<<nblk(dims[l], 256), 256, 0, s>>>(dact[l], g + boff[l - 1], rows, dims[l]);
      MCHECK(hipGetLastError());
      if (l > 1) {
        gemm<false, true, float, float>(dact[l], p + woff[l - 1], nullptr, dact[l - 1], rows, dims[l - 1], dims[l],
                                        false, s);
        const int64_t na = (int64_t)rows * dims[l - 1];
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
    net->nh = nhidden;
    net->nl = nlayers;
    net->nout = nout;
    net->dims.push_back(nin);
    for (int l = 0; l < nlayers; ++l) net->dims.push_back(nhidden);
    net->dims.push_back(nout);
    int64_t off = 0;
    for (int l = 1; l <= net->L(); ++l) {
      net->woff.push_back(off);
      off += (int64_t)net->dims[l - 1] * net->dims[l];
      net->boff.push_back(off);
      off += net->dims[l];
    }
    net->np = off;
    // He-uniform weights for the ReLU layers, Glorot-uniform for the output, zero biases
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 1; l <= net->L(); ++l) {
      const int fi = net->dims[l - 1], fo = net->dims[l];
      const double a = l < net->L() ? std::sqrt(6.0 / fi) : std::sqrt(6.0 / (fi + fo));
      std::uniform_real_distribution<double> U(-a, a);
      for (int64_t i = 0; i < (int64_t)fi * fo; ++i) h[net->woff[l - 1] + i] = (float)U(rng);
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
      MCHECK(hipMemcpyAsync(y, net->act[L], (size_t)batch * net->nout * sizeof(float), hipMemcpyDeviceToHost,
                            net->s));
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
        const float* bias = net->p + net->boff[l - 1];
        if (l < L)
          gemm<false, false, uint16_t, uint16_t>(net->actb[l - 1], net->pb + net->woff[l - 1], bias, net->actb[l],
                                                 batch, net->dims[l], net->dims[l - 1], true, net->s);
        else
          gemm<false, false, uint16_t, float>(net->actb[l - 1], net->pb + net->woff[l - 1], bias, net->act[l], batch,
                                              net->dims[l], net->dims[l - 1], false, net->s);
      }
      MCHECK(hipMemcpyAsync(y, net->act[L], (size_t)batch * net->nout * sizeof(float), hipMemcpyDeviceToHost,
                            net->s));
    }
    MCHECK(hipStreamSynchronize(net->s));
  } catch (const std::exception& e) {
    return mlp_fail(e);
  }
  return 0;
}

}  // extern "C"
