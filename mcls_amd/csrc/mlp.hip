// Importance network of the learned proposal (include/mcls_mlp.h): forward,
// backward and Adam kernels for gfx950.  Matrix products use bf16 MFMA with
// FP32 accumulation on 16-row tiles staged in LDS; one wave owns one tile at
// a time.  Products over the 64 hidden units (K = 64) use gfx950's
// v_mfma_f32_16x16x32_bf16 (two per 16 x 16 output tile); the input layer
// (K = 16 padded inputs) and the weight gradients (K = the tile's 16 rows)
// use v_mfma_f32_16x16x16_bf16.
//
// MFMA operand layouts (lane l of 64), K-chunk of 16 or 32:
//   A (16 x K): row l % 16, k = (K / 4) (l / 16) + v, v = 0 .. K/4 - 1
//   B (K x 16): col l % 16, k = (K / 4) (l / 16) + v
//   D (16 x 16): col l % 16, row = 4 (l / 16) + v, v = 0..3
// The GPU tests check forward outputs and gradients against the CPU oracle,
// which pins these layouts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mcls_gpu.h"
#include "mcls_mlp.h"
#include "mlp.hpp"

namespace {

constexpr int IN = MCLS_MLP_INPUTS, HID = MCLS_MLP_HIDDEN, KIN = 16;
constexpr int P_W1 = 0, P_B1 = P_W1 + IN * HID, P_W2 = P_B1 + HID, P_B2 = P_W2 + HID * HID, P_W3 = P_B2 + HID,
              P_B3 = P_W3 + HID * HID, P_W4 = P_B3 + HID, P_B4 = P_W4 + HID, NPARAM = P_B4 + 1;
static_assert(NPARAM == MCLS_MLP_PARAMS, "parameter layout");
constexpr int WAVES = 4, THREADS = 64 * WAVES, TILE = 16;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// FP32 -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Fragment of M[r][k] stored row-major (ld elements per row): rows r0 + l%16,
// k = k0 + 4 (l/16) + v.  Serves A (rows = i) and B from a k-contiguous
// store (rows = j).
__device__ __forceinline__ s4 frag_row(const unsigned short* T, int ld, int r0, int k0, int lane) {
  const unsigned short* p = T + (r0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4);
  s4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}

// Fragment of M[r][k] = S[k][r] where S is row-major: serves A = S^T and B
// from an r-contiguous store.
__device__ __forceinline__ s4 frag_col(const unsigned short* S, int ld, int r0, int k0, int lane) {
  const int r = r0 + (lane & 15), k = k0 + 4 * (lane >> 4);
  s4 v;
  v[0] = (short)S[(k + 0) * ld + r];
  v[1] = (short)S[(k + 1) * ld + r];
  v[2] = (short)S[(k + 2) * ld + r];
  v[3] = (short)S[(k + 3) * ld + r];
  return v;
}

// Fragment of eight consecutive k of M[r][k] stored row-major: rows
// r0 + l%16, k = k0 + 8 (l/16) + v (the 16 x 32 operands).
__device__ __forceinline__ bf8 frag_row8(const unsigned short* T, int ld, int r0, int k0, int lane) {
  const unsigned short* p = T + (r0 + (lane & 15)) * ld + k0 + 8 * (lane >> 4);
  bf8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = __builtin_bit_cast(__bf16, p[e]);
  return v;
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct Weights {
  unsigned short w1t[HID][KIN];             // W1^T, inputs padded to 16
  unsigned short w2t[HID][HID], w2[HID][HID];  // W2^T (forward B), W2 (backward B)
  unsigned short w3t[HID][HID], w3[HID][HID];
  float b1[HID], b2[HID], b3[HID], w4[HID], b4;
};

struct WaveBuf {
  unsigned short xs[TILE][KIN];
  unsigned short a1[TILE][HID], a2[TILE][HID], a3[TILE][HID], dz[TILE][HID];
  float o[TILE], dO[TILE];
};

struct Grads {
  float w1[KIN][HID], w2[HID][HID], w3[HID][HID];
  float b1[HID], b2[HID], b3[HID], w4[HID], b4, loss;
};

__device__ void load_weights(Weights& W, const float* __restrict__ P) {
  for (int t = threadIdx.x; t < HID * KIN; t += THREADS) {
    const int j = t / KIN, k = t % KIN;
    W.w1t[j][k] = k < IN ? f2bf(P[P_W1 + k * HID + j]) : (unsigned short)0;
  }
  for (int t = threadIdx.x; t < HID * HID; t += THREADS) {
    const int k = t / HID, j = t % HID;  // W[k][j], [in][out]
    const unsigned short v2 = f2bf(P[P_W2 + t]), v3 = f2bf(P[P_W3 + t]);
    W.w2[k][j] = v2;
    W.w2t[j][k] = v2;
    W.w3[k][j] = v3;
    W.w3t[j][k] = v3;
  }
  for (int t = threadIdx.x; t < HID; t += THREADS) {
    W.b1[t] = P[P_B1 + t];
    W.b2[t] = P[P_B2 + t];
    W.b3[t] = P[P_B3 + t];
    W.w4[t] = P[P_W4 + t];
  }
  if (threadIdx.x == 0) W.b4 = P[P_B4];
}

// out = bf16(relu(in (16 x K) W + b)), W given transposed (wt[64][K]).
template <int K>
__device__ void dense_relu(const unsigned short* in, const unsigned short* wt, const float* b, unsigned short* out,
                           int lane) {
  for (int nt = 0; nt < HID / 16; ++nt) {
    const float bj = b[nt * 16 + (lane & 15)];
    f4 acc = {bj, bj, bj, bj};
    if constexpr (K % 32 == 0) {
      for (int kt = 0; kt < K / 32; ++kt)
        acc = mfma32(frag_row8(in, K, 0, kt * 32, lane), frag_row8(wt, K, nt * 16, kt * 32, lane), acc);
    } else {
      for (int kt = 0; kt < K / 16; ++kt)
        acc = mfma(frag_row(in, K, 0, kt * 16, lane), frag_row(wt, K, nt * 16, kt * 16, lane), acc);
    }
    for (int v = 0; v < 4; ++v) out[(4 * (lane >> 4) + v) * HID + nt * 16 + (lane & 15)] = f2bf(fmaxf(acc[v], 0.f));
  }
}

// Forward pass of rows [row0, row0+16) (rows beyond `rows` are zero).  Every
// wave of the block calls it the same number of times: the barriers order
// the LDS traffic between lanes.
__device__ void forward_tile(const Weights& W, WaveBuf& B, const float* __restrict__ X, int64_t row0, int64_t rows,
                             int lane) {
  for (int t = lane; t < TILE * KIN; t += 64) {
    const int i = t / KIN, k = t % KIN;
    const int64_t r = row0 + i;
    B.xs[i][k] = f2bf((k < IN && r < rows) ? X[r * IN + k] : 0.f);
  }
  __syncthreads();
  dense_relu<KIN>(&B.xs[0][0], &W.w1t[0][0], W.b1, &B.a1[0][0], lane);
  __syncthreads();
  dense_relu<HID>(&B.a1[0][0], &W.w2t[0][0], W.b2, &B.a2[0][0], lane);
  __syncthreads();
  dense_relu<HID>(&B.a2[0][0], &W.w3t[0][0], W.b3, &B.a3[0][0], lane);
  __syncthreads();
  const int i = lane >> 2, part = lane & 3;
  float s = 0.f;
  for (int j = part * 16; j < part * 16 + 16; ++j) s += bf2f(B.a3[i][j]) * W.w4[j];
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  if (part == 0) B.o[i] = s + W.b4;
  __syncthreads();
}

// G (64 x 64, ld 64) += A^T D with A, D 16 x 64 bf16 tiles (K = 16 rows).
__device__ void weight_grad(const unsigned short* A, const unsigned short* D, float* G, int lane) {
  for (int mt = 0; mt < HID / 16; ++mt)
    for (int nt = 0; nt < HID / 16; ++nt) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(frag_col(A, HID, mt * 16, 0, lane), frag_col(D, HID, nt * 16, 0, lane), acc);
      for (int v = 0; v < 4; ++v)
        atomicAdd(&G[(mt * 16 + 4 * (lane >> 4) + v) * HID + nt * 16 + (lane & 15)], acc[v]);
    }
}

// dA (16 x 64, D layout per 16-column tile) = dZ W^T, W row-major [in][out].
__device__ void input_grad(const unsigned short* dZ, const unsigned short* Wrm, f4 da[4], int lane) {
  for (int nt = 0; nt < HID / 16; ++nt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < HID / 32; ++kt)
      acc = mfma32(frag_row8(dZ, HID, 0, kt * 32, lane), frag_row8(Wrm, HID, nt * 16, kt * 32, lane), acc);
    da[nt] = acc;
  }
}

// dz = bf16(dA * relu'(z)); gb += sum over rows of the unrounded dz.
__device__ void relu_back(const f4 da[4], const unsigned short* act, unsigned short* dz, float* gb, int lane) {
  for (int nt = 0; nt < HID / 16; ++nt) {
    const int j = nt * 16 + (lane & 15);
    float s = 0.f;
    for (int v = 0; v < 4; ++v) {
      const int i = 4 * (lane >> 4) + v;
      const float d = bf2f(act[i * HID + j]) > 0.f ? da[nt][v] : 0.f;
      s += d;
      dz[i * HID + j] = f2bf(d);
    }
    atomicAdd(&gb[j], s);
  }
}

__device__ void backward_tile(const Weights& W, WaveBuf& B, Grads& G, const float* __restrict__ T, int64_t row0,
                              int64_t rows, float inv_b, int lane) {
  if (lane < TILE) {
    const int64_t r = row0 + lane;
    float d = 0.f;
    if (r < rows) {
      const float e = B.o[lane] - T[r];
      d = 2.f * e * inv_b;
      atomicAdd(&G.loss, e * e);
    }
    B.dO[lane] = d;
  }
  __syncthreads();
  {
    const int j = lane;
    float gw = 0.f, gb = 0.f;
    for (int i = 0; i < TILE; ++i) {
      const float a = bf2f(B.a3[i][j]);
      gw += a * B.dO[i];
      const float dz = a > 0.f ? B.dO[i] * W.w4[j] : 0.f;
      gb += dz;
      B.dz[i][j] = f2bf(dz);
    }
    atomicAdd(&G.w4[j], gw);
    atomicAdd(&G.b3[j], gb);
    if (j == 0) {
      float s = 0.f;
      for (int i = 0; i < TILE; ++i) s += B.dO[i];
      atomicAdd(&G.b4, s);
    }
  }
  __syncthreads();
  f4 da[4];
  weight_grad(&B.a2[0][0], &B.dz[0][0], &G.w3[0][0], lane);
  input_grad(&B.dz[0][0], &W.w3[0][0], da, lane);
  __syncthreads();
  relu_back(da, &B.a2[0][0], &B.dz[0][0], G.b2, lane);
  __syncthreads();
  weight_grad(&B.a1[0][0], &B.dz[0][0], &G.w2[0][0], lane);
  input_grad(&B.dz[0][0], &W.w2[0][0], da, lane);
  __syncthreads();
  relu_back(da, &B.a1[0][0], &B.dz[0][0], G.b1, lane);
  __syncthreads();
  for (int nt = 0; nt < HID / 16; ++nt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(frag_col(&B.xs[0][0], KIN, 0, 0, lane), frag_col(&B.dz[0][0], HID, nt * 16, 0, lane), acc);
    for (int v = 0; v < 4; ++v) {
      const int m = 4 * (lane >> 4) + v;
      if (m < IN) atomicAdd(&G.w1[m][nt * 16 + (lane & 15)], acc[v]);
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(THREADS) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                         int64_t rows, float* __restrict__ out) {
  __shared__ Weights W;
  __shared__ WaveBuf B[WAVES];
  load_weights(W, P);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ntiles = (rows + TILE - 1) / TILE;
  for (int64_t base = (int64_t)blockIdx.x * WAVES; base < ntiles; base += (int64_t)gridDim.x * WAVES) {
    const int64_t tile = base + wave;
    const bool active = tile < ntiles;
    forward_tile(W, B[wave], X, tile * TILE, active ? rows : 0, lane);
    if (active && lane < TILE && tile * TILE + lane < rows) out[tile * TILE + lane] = B[wave].o[lane];
  }
}

__device__ void grad_body(const float* __restrict__ P, const float* __restrict__ X, const float* __restrict__ T,
                          int64_t rows, float inv_b, float* __restrict__ grad, double* loss);

__global__ __launch_bounds__(THREADS) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                      const float* __restrict__ T, int64_t rows, float inv_b,
                                                      float* __restrict__ grad, double* loss) {
  grad_body(P, X, T, rows, inv_b, grad, loss);
}

// The gradient of minibatch order[bi] of the training set (rows
// order[bi] * batch .. + batch), the minibatch found on the device so that
// an epoch's launches can be replayed from a graph with a new order.
__global__ __launch_bounds__(THREADS) void k_mlp_grad_batch(const float* __restrict__ P, const float* __restrict__ X,
                                                            const float* __restrict__ T, int64_t total,
                                                            int64_t batch, const int64_t* __restrict__ order, int64_t bi,
                                                            float* __restrict__ grad, double* loss) {
  const int64_t r0 = order[bi] * batch;
  const int64_t rows = total - r0 < batch ? total - r0 : batch;
  grad_body(P, X + IN * r0, T + r0, rows, 1.0f / (float)rows, grad, loss);
}

__device__ void grad_body(const float* __restrict__ P, const float* __restrict__ X, const float* __restrict__ T,
                          int64_t rows, float inv_b, float* __restrict__ grad, double* loss) {
  __shared__ Weights W;
  __shared__ WaveBuf B[WAVES];
  __shared__ Grads G;
  load_weights(W, P);
  {
    float* g = &G.w1[0][0];
    const int total = (int)(sizeof(Grads) / sizeof(float));
    for (int t = threadIdx.x; t < total; t += THREADS) g[t] = 0.f;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ntiles = (rows + TILE - 1) / TILE;
  for (int64_t base = (int64_t)blockIdx.x * WAVES; base < ntiles; base += (int64_t)gridDim.x * WAVES) {
    const int64_t tile = base + wave;
    const int64_t lim = tile < ntiles ? rows : 0;
    forward_tile(W, B[wave], X, tile * TILE, lim, lane);
    backward_tile(W, B[wave], G, T, tile * TILE, lim, inv_b, lane);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < IN * HID; t += THREADS) atomicAdd(&grad[P_W1 + t], G.w1[t / HID][t % HID]);
  for (int t = threadIdx.x; t < HID * HID; t += THREADS) {
    atomicAdd(&grad[P_W2 + t], G.w2[t / HID][t % HID]);
    atomicAdd(&grad[P_W3 + t], G.w3[t / HID][t % HID]);
  }
  for (int t = threadIdx.x; t < HID; t += THREADS) {
    atomicAdd(&grad[P_B1 + t], G.b1[t]);
    atomicAdd(&grad[P_B2 + t], G.b2[t]);
    atomicAdd(&grad[P_B3 + t], G.b3[t]);
    atomicAdd(&grad[P_W4 + t], G.w4[t]);
  }
  if (threadIdx.x == 0) {
    atomicAdd(&grad[P_B4], G.b4);
    atomicAdd(loss, (double)G.loss);
  }
}

// Adam (Kingma & Ba 2015) with bias correction, FP32 master parameters.
// One Adam step by a single workgroup, its step number kept on the device
// (bias corrections from it), the gradient zeroed once read so that the next
// minibatch accumulates into a clear buffer.
__global__ __launch_bounds__(1024) void k_adam_step(float* __restrict__ p, float* __restrict__ g,
                                                    float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                    int64_t* step) {
  __shared__ float c1, c2;
  if (threadIdx.x == 0) {
    const float st = (float)(*step + 1);
    c1 = 1.f - powf(0.9f, st);
    c2 = 1.f - powf(0.999f, st);
  }
  __syncthreads();
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float gi = g[i];
    g[i] = 0.f;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
  __syncthreads();
  if (threadIdx.x == 0) *step += 1;
}

template <typename T>
__global__ void k_zero(T* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

#define MLP_HIP(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return mcls::set_error(MCLS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct mcls_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float* params = nullptr;
  float* grad = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  double* loss = nullptr;
  int64_t step = 0;
  float* X = nullptr;  // training set rows x 8
  float* T = nullptr;  // rows
  int64_t rows = 0, cap = 0;
  float* scratch_X = nullptr;  // host-pointer staging
  float* scratch_T = nullptr;
  float* scratch_o = nullptr;
  int64_t scratch_rows = 0;
  // training: one epoch of (gradient, Adam step) launches captured in a
  // hipGraph and replayed with each epoch's minibatch order
  int64_t* step_dev = nullptr;  // Adam step count
  int64_t* order_dev = nullptr;
  int64_t order_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int64_t graph_rows = -1, graph_batch = -1;
  float graph_lr = 0.f;
};

namespace mcls {

int mlp_device(const mcls_mlp* m) { return m->device; }
const float* mlp_params_device(const mcls_mlp* m) { return m->params; }

hipError_t launch_mlp_forward(const float* params, const float* X, int64_t rows, float* out, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  const int64_t tiles = (rows + TILE - 1) / TILE;
  int64_t g = (tiles + WAVES - 1) / WAVES;
  if (g > 4096) g = 4096;
  k_mlp_forward<<<(unsigned)g, THREADS, 0, s>>>(params, X, rows, out);
  return hipGetLastError();
}

}  // namespace mcls

namespace {

int ensure_scratch(mcls_mlp* m, int64_t rows) {
  if (rows <= m->scratch_rows) return MCLS_OK;
  (void)hipFree(m->scratch_X);
  (void)hipFree(m->scratch_T);
  (void)hipFree(m->scratch_o);
  m->scratch_X = m->scratch_T = m->scratch_o = nullptr;
  m->scratch_rows = 0;
  MLP_HIP(hipMalloc((void**)&m->scratch_X, sizeof(float) * IN * (size_t)rows));
  MLP_HIP(hipMalloc((void**)&m->scratch_T, sizeof(float) * (size_t)rows));
  MLP_HIP(hipMalloc((void**)&m->scratch_o, sizeof(float) * (size_t)rows));
  m->scratch_rows = rows;
  return MCLS_OK;
}

int grad_batch(mcls_mlp* m, const float* X, const float* T, int64_t rows) {
  const int64_t tiles = (rows + TILE - 1) / TILE;
  int64_t g = (tiles + WAVES - 1) / WAVES;
  if (g > 128) g = 128;
  k_zero<float><<<64, 256, 0, m->stream>>>(m->grad, NPARAM);
  k_mlp_grad<<<(unsigned)g, THREADS, 0, m->stream>>>(m->params, X, T, rows, 1.0f / (float)rows, m->grad, m->loss);
  MLP_HIP(hipGetLastError());
  return MCLS_OK;
}

}  // namespace

namespace {

void drop_graph(mcls_mlp* m) {
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
  m->graph_rows = m->graph_batch = -1;
}

// Capture one epoch: nb minibatch gradients, each followed by its Adam step.
int capture_epoch(mcls_mlp* m, int64_t nb, int64_t batch_rows, float lr) {
  drop_graph(m);
  const int64_t tiles = (batch_rows + TILE - 1) / TILE;
  int64_t g = (tiles + WAVES - 1) / WAVES;
  if (g > 128) g = 128;
  MLP_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  for (int64_t bi = 0; bi < nb; ++bi) {
    k_mlp_grad_batch<<<(unsigned)g, THREADS, 0, m->stream>>>(m->params, m->X, m->T, m->rows, batch_rows, m->order_dev,
                                                             bi, m->grad, m->loss);
    k_adam_step<<<1, 1024, 0, m->stream>>>(m->params, m->grad, m->m1, m->m2, NPARAM, lr, m->step_dev);
  }
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(m->stream, &gr);
  MLP_HIP(e);
  m->graph = gr;
  MLP_HIP(hipGraphInstantiate(&m->graph_exec, m->graph, nullptr, nullptr, 0));
  m->graph_rows = m->rows;
  m->graph_batch = batch_rows;
  m->graph_lr = lr;
  return MCLS_OK;
}

}  // namespace

extern "C" {

int mcls_mlp_create(mcls_mlp** out, int device, uint64_t seed) {
  if (!out) return mcls::set_error(MCLS_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0 || device < 0 || device >= count)
    return mcls::set_error(MCLS_ERR_NO_DEVICE, "no HIP device available");
  MLP_HIP(hipSetDevice(device));
  auto* m = new mcls_mlp();
  m->device = device;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&m->params, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc((void**)&m->grad, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc((void**)&m->m1, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc((void**)&m->m2, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc((void**)&m->loss, sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&m->step_dev, sizeof(int64_t)) != hipSuccess) {
    mcls_mlp_destroy(m);
    return mcls::set_error(MCLS_ERR_HIP, "mcls_mlp_create: allocation failed");
  }
  // He-normal weights (std sqrt(2 / fan_in); output layer sqrt(1 / 64)),
  // zero biases; normals by Box-Muller over SplitMix64 draws.
  std::vector<float> p(NPARAM, 0.f);
  uint64_t st = seed;
  auto normal = [&]() {
    st += 0x9E3779B97F4A7C15ULL;
    const double u1 = ((double)(splitmix64(st) >> 11) + 0.5) * 0x1.0p-53;
    st += 0x9E3779B97F4A7C15ULL;
    const double u2 = (double)(splitmix64(st) >> 11) * 0x1.0p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  };
  for (int i = 0; i < IN * HID; ++i) p[P_W1 + i] = (float)(normal() * std::sqrt(2.0 / IN));
  for (int i = 0; i < HID * HID; ++i) p[P_W2 + i] = (float)(normal() * std::sqrt(2.0 / HID));
  for (int i = 0; i < HID * HID; ++i) p[P_W3 + i] = (float)(normal() * std::sqrt(2.0 / HID));
  for (int i = 0; i < HID; ++i) p[P_W4 + i] = (float)(normal() * std::sqrt(1.0 / HID));
  int rc = mcls_mlp_set_params(m, p.data());
  if (rc != MCLS_OK) {
    mcls_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return MCLS_OK;
}

int mcls_mlp_destroy(mcls_mlp* m) {
  if (!m) return MCLS_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  drop_graph(m);
  for (void* p : {(void*)m->params, (void*)m->grad, (void*)m->m1, (void*)m->m2, (void*)m->loss, (void*)m->X,
                  (void*)m->T, (void*)m->scratch_X, (void*)m->scratch_T, (void*)m->scratch_o, (void*)m->step_dev,
                  (void*)m->order_dev})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return MCLS_OK;
}

int mcls_mlp_get_params(mcls_mlp* m, float* params) {
  if (!m || !params) return mcls::set_error(MCLS_ERR_INVALID_ARG, "null argument");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(params, m->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return MCLS_OK;
}

int mcls_mlp_set_params(mcls_mlp* m, const float* params) {
  if (!m || !params) return mcls::set_error(MCLS_ERR_INVALID_ARG, "null argument");
  for (int i = 0; i < NPARAM; ++i)
    if (!std::isfinite(params[i])) return mcls::set_error(MCLS_ERR_NUMERIC, "non-finite network parameter");
  MLP_HIP(hipSetDevice(m->device));
  MLP_HIP(hipMemcpyAsync(m->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, m->stream));
  k_zero<float><<<64, 256, 0, m->stream>>>(m->m1, NPARAM);
  k_zero<float><<<64, 256, 0, m->stream>>>(m->m2, NPARAM);
  k_zero<int64_t><<<1, 64, 0, m->stream>>>(m->step_dev, 1);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(m->stream));
  m->step = 0;
  return MCLS_OK;
}

int mcls_mlp_forward(mcls_mlp* m, const float* X, int64_t rows, float* out, int ptr_kind) {
  if (!m || !X || !out || rows < 0) return mcls::set_error(MCLS_ERR_INVALID_ARG, "bad forward arguments");
  if (ptr_kind != MCLS_HOST && ptr_kind != MCLS_DEVICE) return mcls::set_error(MCLS_ERR_INVALID_ARG, "ptr_kind");
  MLP_HIP(hipSetDevice(m->device));
  const float* dX = X;
  float* dO = out;
  if (ptr_kind == MCLS_HOST) {
    int rc = ensure_scratch(m, rows);
    if (rc != MCLS_OK) return rc;
    MLP_HIP(hipMemcpyAsync(m->scratch_X, X, sizeof(float) * IN * (size_t)rows, hipMemcpyHostToDevice, m->stream));
    dX = m->scratch_X;
    dO = m->scratch_o;
  }
  MLP_HIP(mcls::launch_mlp_forward(m->params, dX, rows, dO, m->stream));
  if (ptr_kind == MCLS_HOST)
    MLP_HIP(hipMemcpyAsync(out, dO, sizeof(float) * (size_t)rows, hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return MCLS_OK;
}

int mcls_mlp_loss_grad(mcls_mlp* m, const float* X, const float* T, int64_t rows, double* loss, float* grad,
                       int ptr_kind) {
  if (!m || !X || !T || rows <= 0) return mcls::set_error(MCLS_ERR_INVALID_ARG, "bad loss_grad arguments");
  if (ptr_kind != MCLS_HOST && ptr_kind != MCLS_DEVICE) return mcls::set_error(MCLS_ERR_INVALID_ARG, "ptr_kind");
  MLP_HIP(hipSetDevice(m->device));
  const float *dX = X, *dT = T;
  if (ptr_kind == MCLS_HOST) {
    int rc = ensure_scratch(m, rows);
    if (rc != MCLS_OK) return rc;
    MLP_HIP(hipMemcpyAsync(m->scratch_X, X, sizeof(float) * IN * (size_t)rows, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(hipMemcpyAsync(m->scratch_T, T, sizeof(float) * (size_t)rows, hipMemcpyHostToDevice, m->stream));
    dX = m->scratch_X;
    dT = m->scratch_T;
  }
  k_zero<double><<<1, 1, 0, m->stream>>>(m->loss, 1);
  int rc = grad_batch(m, dX, dT, rows);
  if (rc != MCLS_OK) return rc;
  double l = 0.0;
  MLP_HIP(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (grad) MLP_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (loss) *loss = l / (double)rows;
  return MCLS_OK;
}

int mcls_mlp_dataset_clear(mcls_mlp* m) {
  if (!m) return mcls::set_error(MCLS_ERR_INVALID_ARG, "null network");
  m->rows = 0;
  drop_graph(m);
  return MCLS_OK;
}

int mcls_mlp_dataset_rows(mcls_mlp* m, int64_t* rows) {
  if (!m || !rows) return mcls::set_error(MCLS_ERR_INVALID_ARG, "null argument");
  *rows = m->rows;
  return MCLS_OK;
}

int mcls_mlp_dataset_append(mcls_mlp* m, const float* X, const float* T, int64_t rows, int ptr_kind) {
  if (!m || !X || !T || rows < 0) return mcls::set_error(MCLS_ERR_INVALID_ARG, "bad dataset arguments");
  if (ptr_kind != MCLS_HOST && ptr_kind != MCLS_DEVICE) return mcls::set_error(MCLS_ERR_INVALID_ARG, "ptr_kind");
  MLP_HIP(hipSetDevice(m->device));
  if (m->rows + rows > m->cap) {
    int64_t cap = std::max<int64_t>(m->rows + rows, 2 * m->cap);
    float *nX = nullptr, *nT = nullptr;
    MLP_HIP(hipMalloc((void**)&nX, sizeof(float) * IN * (size_t)cap));
    MLP_HIP(hipMalloc((void**)&nT, sizeof(float) * (size_t)cap));
    if (m->rows) {
      MLP_HIP(hipMemcpyAsync(nX, m->X, sizeof(float) * IN * (size_t)m->rows, hipMemcpyDeviceToDevice, m->stream));
      MLP_HIP(hipMemcpyAsync(nT, m->T, sizeof(float) * (size_t)m->rows, hipMemcpyDeviceToDevice, m->stream));
      MLP_HIP(hipStreamSynchronize(m->stream));
    }
    (void)hipFree(m->X);
    (void)hipFree(m->T);
    m->X = nX;
    m->T = nT;
    m->cap = cap;
    drop_graph(m);  // the captured epoch names the old buffers
  }
  const hipMemcpyKind k = ptr_kind == MCLS_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  MLP_HIP(hipMemcpyAsync(m->X + IN * m->rows, X, sizeof(float) * IN * (size_t)rows, k, m->stream));
  MLP_HIP(hipMemcpyAsync(m->T + m->rows, T, sizeof(float) * (size_t)rows, k, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  m->rows += rows;
  return MCLS_OK;
}


int mcls_mlp_train(mcls_mlp* m, int epochs, int64_t batch_rows, float lr, uint64_t seed, double* final_loss) {
  if (!m || epochs < 0 || batch_rows <= 0 || !(lr > 0.f)) return mcls::set_error(MCLS_ERR_INVALID_ARG, "bad train arguments");
  if (m->rows == 0) return mcls::set_error(MCLS_ERR_STATE, "empty training set");
  MLP_HIP(hipSetDevice(m->device));
  const int64_t nb = (m->rows + batch_rows - 1) / batch_rows;
  if (nb > m->order_cap) {
    (void)hipFree(m->order_dev);
    m->order_dev = nullptr;
    MLP_HIP(hipMalloc((void**)&m->order_dev, sizeof(int64_t) * (size_t)nb));
    m->order_cap = nb;
    drop_graph(m);  // the graph names the old order buffer
  }
  if (!m->graph_exec || m->graph_rows != m->rows || m->graph_batch != batch_rows || m->graph_lr != lr) {
    int rc = capture_epoch(m, nb, batch_rows, lr);
    if (rc != MCLS_OK) return rc;
  }
  // a clear gradient before the first minibatch (each Adam step clears it after)
  k_zero<float><<<64, 256, 0, m->stream>>>(m->grad, NPARAM);
  std::vector<int64_t> order((size_t)nb);
  uint64_t st = seed;
  double last = 0.0;
  for (int e = 0; e < epochs; ++e) {
    for (int64_t b = 0; b < nb; ++b) order[(size_t)b] = b;
    for (int64_t b = nb - 1; b > 0; --b) {  // Fisher-Yates
      st += 0x9E3779B97F4A7C15ULL;
      const int64_t j = (int64_t)(splitmix64(st) % (uint64_t)(b + 1));
      std::swap(order[(size_t)b], order[(size_t)j]);
    }
    MLP_HIP(hipMemcpyAsync(m->order_dev, order.data(), sizeof(int64_t) * (size_t)nb, hipMemcpyHostToDevice, m->stream));
    k_zero<double><<<1, 1, 0, m->stream>>>(m->loss, 1);
    MLP_HIP(hipGraphLaunch(m->graph_exec, m->stream));
    m->step += nb;
    double l = 0.0;
    MLP_HIP(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    last = l / (double)m->rows;
  }
  if (final_loss) *final_loss = last;
  return MCLS_OK;
}

}  // extern "C"
