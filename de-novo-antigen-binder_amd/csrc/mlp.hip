// Learned prescreen (no reference counterpart): a 32 -> 256 -> 256 -> 256 -> 1
// SiLU MLP over per-candidate features read from antigen maps (maps.hip),
// with bf16 weights and activations and FP32 accumulation on the MFMA units,
// trained on the device (Adam) on energies that the exact pair kernel
// computes, and the screening pipeline features -> MLP -> top-n -> exact
// re-score of the kept candidates -> top-k.
//
// GEMM (training): C[M x N] = op(A)[M x K] op(B)[K x N], bf16 operands staged
// through LDS with k contiguous, v_mfma_f32_16x16x32_bf16 (gfx950), a
// workgroup of 4 waves computes a 64 x 64 tile, each wave 32 x 32 as 2 x 2
// MFMA blocks. Inference runs the whole network per 64-row tile in one
// kernel (k_mlp_infer), activations in LDS, with the same MFMA sequence and
// epilogues, so its outputs equal the layer-by-layer GEMM path bit for bit.
// MFMA lane layout (16x16x32): lane l supplies A[row l%16][k 8(l/16)..+7]
// and B[k 8(l/16)..+7][col l%16]; it receives C[rows 4(l/16)..+3][col l%16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "dnab_internal.h"
#include "mlp_kernels.h"

namespace dnab {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;
constexpr int kKStep = 32;
constexpr int kLdsStride = kKStep + 8;  // bf16 elements per LDS row (pad: 80 B)

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kTile * kLdsStride];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile * kLdsStride];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4_t acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += kKStep) {
    __syncthreads();
    // ---- A tile (64 m x 32 k) -> As[m][k]
    if (!TA) {
      const int r = t >> 2, c = (t & 3) * 8;
      const int m = m0 + r, k = k0 + c;
      uint16_t v[8];
      if (g.vec_ok && m < g.M && k + 8 <= g.K) {
        *reinterpret_cast<uint4*>(v) = *reinterpret_cast<const uint4*>(&g.A[(size_t)m * g.lda + k]);
      } else {
        for (int e = 0; e < 8; ++e)
          v[e] = (m < g.M && k + e < g.K) ? g.A[(size_t)m * g.lda + k + e] : 0;
      }
      *reinterpret_cast<uint4*>(&As[r * kLdsStride + c]) = *reinterpret_cast<uint4*>(v);
    } else {
      const int kk = t >> 3, c = (t & 7) * 8;
      const int k = k0 + kk, m = m0 + c;
      uint16_t v[8];
      if (g.vec_ok && k < g.K && m + 8 <= g.M) {
        *reinterpret_cast<uint4*>(v) = *reinterpret_cast<const uint4*>(&g.A[(size_t)k * g.lda + m]);
      } else {
        for (int e = 0; e < 8; ++e)
          v[e] = (k < g.K && m + e < g.M) ? g.A[(size_t)k * g.lda + m + e] : 0;
      }
      for (int e = 0; e < 8; ++e) As[(c + e) * kLdsStride + kk] = v[e];
    }
    // ---- B tile (32 k x 64 n) -> Bs[n][k]
    if (!TB) {
      const int kk = t >> 3, c = (t & 7) * 8;
      const int k = k0 + kk, n = n0 + c;
      uint16_t v[8];
      if (g.vec_ok && k < g.K && n + 8 <= g.N) {
        *reinterpret_cast<uint4*>(v) = *reinterpret_cast<const uint4*>(&g.B[(size_t)k * g.ldb + n]);
      } else {
        for (int e = 0; e < 8; ++e)
          v[e] = (k < g.K && n + e < g.N) ? g.B[(size_t)k * g.ldb + n + e] : 0;
      }
      for (int e = 0; e < 8; ++e) Bs[(c + e) * kLdsStride + kk] = v[e];
    } else {
      const int r = t >> 2, c = (t & 3) * 8;
      const int n = n0 + r, k = k0 + c;
      uint16_t v[8];
      if (g.vec_ok && n < g.N && k + 8 <= g.K) {
        *reinterpret_cast<uint4*>(v) = *reinterpret_cast<const uint4*>(&g.B[(size_t)n * g.ldb + k]);
      } else {
        for (int e = 0; e < 8; ++e)
          v[e] = (n < g.N && k + e < g.K) ? g.B[(size_t)n * g.ldb + k + e] : 0;
      }
      *reinterpret_cast<uint4*>(&Bs[r * kLdsStride + c]) = *reinterpret_cast<uint4*>(v);
    }
    __syncthreads();
    const int kq = (lane >> 4) * 8;
    bf16x8_t af[2], bf[2];
    for (int i = 0; i < 2; ++i) {
      const int row = wm + 16 * i + (lane & 15);
      af[i] = *reinterpret_cast<const bf16x8_t*>(&As[row * kLdsStride + kq]);
    }
    for (int j = 0; j < 2; ++j) {
      const int col = wn + 16 * j + (lane & 15);
      bf[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[col * kLdsStride + kq]);
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
  }
  // ---- epilogue: lane holds C[4(l/16) + r][l % 16] of each 16 x 16 block
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + 16 * j + (lane & 15);
      if (n >= g.N) continue;
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + (lane >> 4) * 4 + r;
        if (m >= g.M) continue;
        const size_t o = (size_t)m * g.ldc + n;
        const float a = acc[i][j][r];
        if (EPI == kEpiStoreF32) {
          g.C32[o] = a;
        } else if (EPI == kEpiBias) {
          g.C32[o] = a + g.bias[n];
        } else if (EPI == kEpiBiasSilu) {
          const float z = a + g.bias[n];
          if (g.C32) g.C32[o] = z;
          g.Cbf[o] = f32_to_bf16(z * sigmoidf_(z));
        } else {  // kEpiSiluGrad
          const float z = g.C32[o];
          const float s = sigmoidf_(z);
          g.Cbf[o] = f32_to_bf16(a * (s * (1.0f + z * (1.0f - s))));
        }
      }
    }
}

// ---- features (maps.hip) ------------------------------------------------------

constexpr int kFeat = kPrescreenFeatures;  // 32

// ---- small kernels ------------------------------------------------------------------

__global__ void k_gather_rows(const uint16_t* __restrict__ X, const float* __restrict__ y,
                              const uint32_t* __restrict__ idx, uint32_t B,
                              uint16_t* __restrict__ Xb, float* __restrict__ yb) {
  const uint32_t b = blockIdx.x;
  if (b >= B) return;
  const uint32_t src = idx[b];
  const uint4* s = reinterpret_cast<const uint4*>(X + static_cast<size_t>(src) * kFeat);
  uint4* d = reinterpret_cast<uint4*>(Xb + static_cast<size_t>(b) * kFeat);
  for (int v = threadIdx.x; v < kFeat / 8; v += blockDim.x) d[v] = s[v];
  if (threadIdx.x == 0) yb[b] = y[src];
}

// y[m] = sum_n A[m][n] w[n] + b (one wave per row)
__global__ void k_last_forward(const uint16_t* __restrict__ A, const uint16_t* __restrict__ w,
                               const float* __restrict__ b, uint32_t M, uint32_t H,
                               float* __restrict__ y) {
  const uint32_t m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (m >= M) return;
  float s = 0.0f;
  for (uint32_t n = lane; n < H; n += kWave)
    s += bf16_to_f32(A[static_cast<size_t>(m) * H + n]) * bf16_to_f32(w[n]);
  s = wave_reduce_f32<OpAdd>(s);
  if (lane == 0) y[m] = s + b[0];
}

// Squared loss: dy = 2 (y - t) / B; per-row (y - t)^2 into sq.
__global__ void k_loss_grad(const float* __restrict__ y, const float* __restrict__ t, uint32_t B,
                            float* __restrict__ dy, float* __restrict__ sq) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= B) return;
  const float r = y[m] - t[m];
  dy[m] = 2.0f * r / static_cast<float>(B);
  sq[m] = r * r;
}

// Column sums over rows for 64 columns per workgroup: wave w of the 16 adds
// rows w, w + 16, ... of its lane's column, then the 16 partial sums are
// added in wave order (a fixed order).
constexpr int kColWaves = 16;
__device__ inline float colsum_finish(float v, float (*part)[64], uint32_t wave, uint32_t lane) {
  part[wave][lane] = v;
  __syncthreads();
  float s = 0.0f;
  if (wave == 0)
    for (int w = 0; w < kColWaves; ++w) s += part[w][lane];
  return s;
}

// Last-layer backward: gw[n] = sum_m A[m][n] dy[m], gb = sum_m dy[m];
// dZ[m][n] = bf16(dy[m] w[n] silu'(Z[m][n])). Workgroup = 64 columns.
__global__ void __launch_bounds__(kColWaves * 64)
k_last_backward(const uint16_t* __restrict__ A, const float* __restrict__ Z,
                const uint16_t* __restrict__ w, const float* __restrict__ dy, uint32_t M,
                uint32_t H, float* __restrict__ gw, float* __restrict__ gb,
                uint16_t* __restrict__ dZ) {
  __shared__ float part[kColWaves][64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t n = blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (n < H) {
    const float wn = bf16_to_f32(w[n]);
    for (uint32_t m = wave; m < M; m += kColWaves) {
      const size_t o = static_cast<size_t>(m) * H + n;
      const float d = dy[m];
      s += bf16_to_f32(A[o]) * d;
      const float z = Z[o];
      const float sg = sigmoidf_(z);
      dZ[o] = f32_to_bf16(d * wn * (sg * (1.0f + z * (1.0f - sg))));
    }
  }
  s = colsum_finish(s, part, wave, lane);
  if (wave == 0 && n < H) gw[n] = s;
  if (blockIdx.x == 0) {
    __syncthreads();
    float b = 0.0f;
    for (uint32_t m = threadIdx.x; m < M; m += kColWaves * 64) b += dy[m];
    b = colsum_finish(b, part, wave, lane);
    if (wave == 0) {
      b = wave_reduce_f32<OpAdd>(b);
      if (lane == 0) gb[0] = b;
    }
  }
}

// Column sums of a bf16 [M x N] matrix (bias gradients), fixed order.
// Workgroup = 64 columns.
__global__ void __launch_bounds__(kColWaves * 64)
k_colsum_bf16(const uint16_t* __restrict__ D, uint32_t M, uint32_t N, float* __restrict__ out) {
  __shared__ float part[kColWaves][64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t n = blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (n < N)
    for (uint32_t m = wave; m < M; m += kColWaves) s += bf16_to_f32(D[static_cast<size_t>(m) * N + n]);
  s = colsum_finish(s, part, wave, lane);
  if (wave == 0 && n < N) out[n] = s;
}

// Adam (Kingma & Ba 2015) on the fp32 master weights; refresh bf16 copies.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ pbf, size_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  pbf[i] = f32_to_bf16(pi);
}

__global__ void k_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = f32_to_bf16(a[i]);
}

// target transform: asinh(min(E, cap) / scale); NaN counts as above the cap.
__global__ void k_targets(const double* __restrict__ e, uint32_t n, float scale, float cap,
                          float* __restrict__ t) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = e[i];
  if (!(v <= static_cast<double>(cap))) v = cap;  // also NaN
  t[i] = static_cast<float>(asinh(fmax(v, -1e30) / scale));
}

// ---- fused inference -----------------------------------------------------------

constexpr int kInferRows = 64;
constexpr int kHid = 256;
constexpr int kXStride = kFeat + 8;  // bf16 per LDS row (16-byte aligned, banks spread)
constexpr int kHStride = kHid + 8;

// W[K][N] (row-major bf16) -> WT[N][K]
__global__ void k_transpose_bf16(const uint16_t* __restrict__ W, uint32_t K, uint32_t N,
                                 uint16_t* __restrict__ WT) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N) return;
  const uint32_t k = i / N, n = i - k * N;
  WT[static_cast<size_t>(n) * K + k] = W[i];
}

// One dense SiLU layer of the tile: out[64][256] = bf16(silu(in[64][K] W + b)),
// W given transposed (WT[256][K], k contiguous, read from global memory in
// the MFMA B layout). Wave w computes columns 64w..64w+63 as 4 x 4 blocks.
template <int K, int IN_STRIDE>
__device__ __attribute__((always_inline)) inline void infer_layer(const uint16_t* __restrict__ in, const uint16_t* __restrict__ WT,
                                   const float* __restrict__ bias, uint16_t* __restrict__ out,
                                   uint32_t lane, uint32_t wave) {
  const int n0 = static_cast<int>(wave) * 64;
  const int kq = static_cast<int>(lane >> 4) * 8;
  floatx4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += kKStep) {
    bf16x8_t af[4], bw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      bw[j] = *reinterpret_cast<const bf16x8_t*>(WT + static_cast<size_t>(n0 + 16 * j + (lane & 15)) * K + k0 + kq);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      af[i] = *reinterpret_cast<const bf16x8_t*>(in + (16 * i + (lane & 15)) * IN_STRIDE + k0 + kq);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bw[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + 16 * j + static_cast<int>(lane & 15);
      const float b = bias[n];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 16 * i + static_cast<int>(lane >> 4) * 4 + r;
        const float z = acc[i][j][r] + b;
        out[m * kHStride + n] = f32_to_bf16(z * sigmoidf_(z));
      }
    }
}

struct InferArgs {
  const uint16_t* X;    // [M][kFeat] bf16
  uint32_t M;
  const uint16_t* wt1;  // [256][kFeat]
  const uint16_t* wt2;  // [256][256]
  const uint16_t* wt3;  // [256][256]
  const uint16_t* w4;   // [256] bf16
  const float* b1;
  const float* b2;
  const float* b3;
  const float* b4;
  float* y;             // [M]
};

// The whole network on 64 rows per workgroup (4 waves); the last layer as
// k_last_forward forms it (lane-strided products, then the wave sum).
__global__ void __launch_bounds__(256, 2) k_mlp_infer(InferArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t sX[kInferRows * kXStride];
  __shared__ __attribute__((aligned(16))) uint16_t sA[kInferRows * kHStride];
  __shared__ __attribute__((aligned(16))) uint16_t sB[kInferRows * kHStride];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t r0 = blockIdx.x * kInferRows;
  constexpr int kVec = kFeat / 8;  // 16-byte pieces per input row
  for (int v = static_cast<int>(t); v < kInferRows * kVec; v += 256) {
    const int r = v / kVec, c = (v - r * kVec) * 8;
    uint4 w = make_uint4(0u, 0u, 0u, 0u);
    if (r0 + r < g.M) w = *reinterpret_cast<const uint4*>(g.X + static_cast<size_t>(r0 + r) * kFeat + c);
    *reinterpret_cast<uint4*>(&sX[r * kXStride + c]) = w;
  }
  __syncthreads();
  infer_layer<kFeat, kXStride>(sX, g.wt1, g.b1, sA, lane, wave);
  __syncthreads();
  infer_layer<kHid, kHStride>(sA, g.wt2, g.b2, sB, lane, wave);
  __syncthreads();
  infer_layer<kHid, kHStride>(sB, g.wt3, g.b3, sA, lane, wave);
  __syncthreads();
  // w4 (lane-strided) and b4 held in registers across the rows
  float w4[kHid / kWave];
#pragma unroll
  for (int q = 0; q < kHid / kWave; ++q) w4[q] = bf16_to_f32(g.w4[lane + q * kWave]);
  const float b4 = g.b4[0];
  for (uint32_t q = 0; q < kInferRows / 4; ++q) {
    const uint32_t m = wave * (kInferRows / 4) + q;
    float s = 0.0f;
#pragma unroll
    for (int e = 0; e < kHid / kWave; ++e) s += bf16_to_f32(sA[m * kHStride + lane + e * kWave]) * w4[e];
    s = wave_reduce_f32<OpAdd>(s);
    if (lane == 0 && r0 + m < g.M) g.y[r0 + m] = s + b4;
  }
}

// kept candidates' energies (from the population's per-candidate array)
__global__ void k_gather_energy(const double* __restrict__ e, const uint32_t* __restrict__ idx,
                                uint32_t n, double* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = e[idx[i]];
}

}  // namespace dnab

// ---- host side ------------------------------------------------------------------------

namespace {
using namespace dnab;

constexpr size_t kD = kFeat, kH = 256;
constexpr size_t oW1 = 0, ob1 = kD * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
                 ob3 = oW3 + kH * kH, ow4 = ob3 + kH, ob4 = ow4 + kH, kParams = ob4 + 1;

inline uint32_t nb(uint64_t n, uint32_t t) {
  uint64_t b = (n + t - 1) / t;
  return static_cast<uint32_t>(b ? b : 1);
}

template <bool TA, bool TB, int EPI>
void gemm(dnab_ctx* ctx, const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N,
          int K, float* C32, uint16_t* Cbf, int ldc, const float* bias) {
  GemmArgs g{A, B, M, N, K, lda, ldb, ldc, C32, Cbf, bias, (lda % 8 == 0 && ldb % 8 == 0) ? 1 : 0};
  dim3 grid(nb(N, kTile), nb(M, kTile));
  hipLaunchKernelGGL((k_gemm_bf16<TA, TB, EPI>), grid, dim3(256), 0, ctx->stream, g);
}

}  // namespace

struct dnab_mlp {
  dnab_ctx* ctx = nullptr;
  DevBuf<float> p, g, m, v;
  DevBuf<uint16_t> pbf;
  uint64_t step = 0;
  uint32_t rows = 0;  // workspace capacity
  DevBuf<float> Z1, Z2, Z3, y, dy, sq, tb;
  DevBuf<uint16_t> A1, A2, A3, dZ1, dZ2, dZ3, Xb;
  DevBuf<uint32_t> perm;
  DevBuf<float> loss_hist;
  DevBuf<uint16_t> wt;  // inference copies of W1..W3, transposed (k contiguous)
};

struct dnab_dataset {
  dnab_ctx* ctx = nullptr;
  size_t capacity = 0, size = 0;
  float energy_scale = 10.0f;
  float energy_cap = 0.0f;
  DevBuf<uint16_t> X;
  DevBuf<float> t;
  DevBuf<double> e;
};

namespace {

int ensure_rows(dnab_ctx* ctx, dnab_mlp* n, uint32_t rows) {
  if (rows <= n->rows) return DNAB_OK;
  const size_t r = rows;
  if (n->Z1.alloc(r * kH) || n->Z2.alloc(r * kH) || n->Z3.alloc(r * kH) || n->y.alloc(r) ||
      n->dy.alloc(r) || n->sq.alloc(r) || n->tb.alloc(r) || n->A1.alloc(r * kH) ||
      n->A2.alloc(r * kH) || n->A3.alloc(r * kH) || n->dZ1.alloc(r * kH) ||
      n->dZ2.alloc(r * kH) || n->dZ3.alloc(r * kH) || n->Xb.alloc(r * kD))
    return dnab_set_error(ctx, DNAB_ERR_ALLOC, "mlp: workspace allocation failed");
  n->rows = rows;
  return DNAB_OK;
}

// Forward pass on M rows of X; Z* stored when keep_z.
void forward(dnab_ctx* ctx, dnab_mlp* n, const uint16_t* X, uint32_t M, bool keep_z) {
  const uint16_t* w = n->pbf.p;
  const float* p = n->p.p;
  gemm<false, false, kEpiBiasSilu>(ctx, X, kD, w + oW1, kH, M, kH, kD, keep_z ? n->Z1.p : nullptr,
                                   n->A1.p, kH, p + ob1);
  gemm<false, false, kEpiBiasSilu>(ctx, n->A1.p, kH, w + oW2, kH, M, kH, kH,
                                   keep_z ? n->Z2.p : nullptr, n->A2.p, kH, p + ob2);
  gemm<false, false, kEpiBiasSilu>(ctx, n->A2.p, kH, w + oW3, kH, M, kH, kH,
                                   keep_z ? n->Z3.p : nullptr, n->A3.p, kH, p + ob3);
  hipLaunchKernelGGL(k_last_forward, dim3(nb(M, 4)), dim3(256), 0, ctx->stream, n->A3.p,
                     w + ow4, p + ob4, M, static_cast<uint32_t>(kH), n->y.p);
}

// Backward pass for the squared loss on M rows (targets in tb) into n->g.
void backward(dnab_ctx* ctx, dnab_mlp* n, const uint16_t* X, uint32_t M) {
  const uint16_t* w = n->pbf.p;
  float* g = n->g.p;
  hipLaunchKernelGGL(k_loss_grad, dim3(nb(M, 256)), dim3(256), 0, ctx->stream, n->y.p, n->tb.p,
                     M, n->dy.p, n->sq.p);
  hipLaunchKernelGGL(k_last_backward, dim3(nb(kH, 64)), dim3(kColWaves * 64), 0, ctx->stream, n->A3.p,
                     n->Z3.p, w + ow4, n->dy.p, M, static_cast<uint32_t>(kH), g + ow4, g + ob4,
                     n->dZ3.p);
  // layer 3
  gemm<true, false, kEpiStoreF32>(ctx, n->A2.p, kH, n->dZ3.p, kH, kH, kH, M, g + oW3, nullptr,
                                  kH, nullptr);
  hipLaunchKernelGGL(k_colsum_bf16, dim3(nb(kH, 64)), dim3(kColWaves * 64), 0, ctx->stream, n->dZ3.p, M,
                     static_cast<uint32_t>(kH), g + ob3);
  gemm<false, true, kEpiSiluGrad>(ctx, n->dZ3.p, kH, w + oW3, kH, M, kH, kH, n->Z2.p, n->dZ2.p,
                                  kH, nullptr);
  // layer 2
  gemm<true, false, kEpiStoreF32>(ctx, n->A1.p, kH, n->dZ2.p, kH, kH, kH, M, g + oW2, nullptr,
                                  kH, nullptr);
  hipLaunchKernelGGL(k_colsum_bf16, dim3(nb(kH, 64)), dim3(kColWaves * 64), 0, ctx->stream, n->dZ2.p, M,
                     static_cast<uint32_t>(kH), g + ob2);
  gemm<false, true, kEpiSiluGrad>(ctx, n->dZ2.p, kH, w + oW2, kH, M, kH, kH, n->Z1.p, n->dZ1.p,
                                  kH, nullptr);
  // layer 1 (no input gradient)
  gemm<true, false, kEpiStoreF32>(ctx, X, kD, n->dZ1.p, kH, kD, kH, M, g + oW1, nullptr, kH,
                                  nullptr);
  hipLaunchKernelGGL(k_colsum_bf16, dim3(nb(kH, 64)), dim3(kColWaves * 64), 0, ctx->stream, n->dZ1.p, M,
                     static_cast<uint32_t>(kH), g + ob1);
}

__global__ void k_sum_into(const float* __restrict__ a, uint32_t n, float scale,
                           float* __restrict__ out) {
  __shared__ float s[256];
  float v = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) v += a[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0] * scale;
}

int population_features(dnab_ctx* ctx, dnab_population* pop, uint16_t* dst) {
  return population_map_features(ctx, pop, dst, nullptr);
}

// The bf16 weights W1..W3 transposed for k_mlp_infer (after every update).
int refresh_inference_weights(dnab_ctx* ctx, dnab_mlp* n) {
  DNAB_ALLOC_CHECK(ctx, n->wt.ensure(kD * kH + 2 * kH * kH));
  hipLaunchKernelGGL(k_transpose_bf16, dim3(nb(kD * kH, 256)), dim3(256), 0, ctx->stream,
                     n->pbf.p + oW1, static_cast<uint32_t>(kD), static_cast<uint32_t>(kH), n->wt.p);
  hipLaunchKernelGGL(k_transpose_bf16, dim3(nb(kH * kH, 256)), dim3(256), 0, ctx->stream,
                     n->pbf.p + oW2, static_cast<uint32_t>(kH), static_cast<uint32_t>(kH),
                     n->wt.p + kD * kH);
  hipLaunchKernelGGL(k_transpose_bf16, dim3(nb(kH * kH, 256)), dim3(256), 0, ctx->stream,
                     n->pbf.p + oW3, static_cast<uint32_t>(kH), static_cast<uint32_t>(kH),
                     n->wt.p + kD * kH + kH * kH);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return DNAB_OK;
}

// Predictions for M rows of X (device) into y, one fused launch.
void infer(dnab_ctx* ctx, dnab_mlp* n, const uint16_t* X, uint32_t M, float* y) {
  InferArgs g;
  g.X = X;
  g.M = M;
  g.wt1 = n->wt.p;
  g.wt2 = n->wt.p + kD * kH;
  g.wt3 = n->wt.p + kD * kH + kH * kH;
  g.w4 = n->pbf.p + ow4;
  g.b1 = n->p.p + ob1;
  g.b2 = n->p.p + ob2;
  g.b3 = n->p.p + ob3;
  g.b4 = n->p.p + ob4;
  g.y = y;
  hipEvent_t ev_end;
  hipEvent_t ev = timer_begin(ctx, "mlp_infer", &ev_end);
  hipLaunchKernelGGL(k_mlp_infer, dim3(nb(M, kInferRows)), dim3(256), 0, ctx->stream, g);
  timer_end(ctx, "mlp_infer", ev, ev_end);
}

int predict_device(dnab_ctx* ctx, dnab_mlp* n, dnab_population* pop) {
  const size_t nc = pop->n_cand;
  DNAB_ALLOC_CHECK(ctx, pop->feat.ensure(nc * kD));
  DNAB_ALLOC_CHECK(ctx, pop->scores.ensure(nc));
  DNAB_RETURN_IF(population_features(ctx, pop, pop->feat.p));
  if (nc) infer(ctx, n, pop->feat.p, static_cast<uint32_t>(nc), pop->scores.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return DNAB_OK;
}

}  // namespace

extern "C" {

int dnab_mlp_param_count(size_t* n) {
  if (!n) return DNAB_ERR_INVALID;
  *n = kParams;
  return DNAB_OK;
}

int dnab_mlp_create(dnab_ctx* ctx, uint64_t seed, dnab_mlp** out) {
  if (!ctx || !out) return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  dnab_mlp* n = new (std::nothrow) dnab_mlp();
  if (!n) return DNAB_ERR_ALLOC;
  n->ctx = ctx_retain(ctx);
  if (n->p.alloc(kParams) || n->g.alloc(kParams) || n->m.alloc(kParams) || n->v.alloc(kParams) ||
      n->pbf.alloc(kParams)) {
    dnab_mlp_destroy(n);
    return dnab_set_error(ctx, DNAB_ERR_ALLOC, "mlp: allocation failed");
  }
  // Glorot-uniform weights, zero biases (host SplitMix64).
  std::vector<float> h(kParams, 0.0f);
  SplitMix64 rng{seed ^ 0x6d6c7000ull};
  auto fill = [&](size_t off, size_t fan_in, size_t fan_out) {
    const float a = std::sqrt(6.0f / static_cast<float>(fan_in + fan_out));
    for (size_t i = 0; i < fan_in * fan_out; ++i) h[off + i] = rng.uniform(-a, a);
  };
  fill(oW1, kD, kH);
  fill(oW2, kH, kH);
  fill(oW3, kH, kH);
  fill(ow4, kH, 1);
  std::vector<float> zero(kParams, 0.0f);
  int rc = upload(ctx, n->p.p, h.data(), kParams * 4);
  if (!rc) rc = upload(ctx, n->m.p, zero.data(), kParams * 4);
  if (!rc) rc = upload(ctx, n->v.p, zero.data(), kParams * 4);
  if (!rc) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nb(kParams, 256)), dim3(256), 0, ctx->stream, n->p.p,
                       n->pbf.p, kParams);
    rc = refresh_inference_weights(ctx, n);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = DNAB_ERR_HIP;
  }
  if (rc) {
    dnab_mlp_destroy(n);
    return rc;
  }
  *out = n;
  return DNAB_OK;
}

void dnab_mlp_destroy(dnab_mlp* n) {
  if (!n) return;
  dnab_ctx* ctx = n->ctx;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  delete n;
  ctx_release(ctx);
}

int dnab_mlp_get_params(dnab_ctx* ctx, const dnab_mlp* n, float* out, size_t count) {
  if (!ctx || !n || !out || count != kParams) return DNAB_ERR_INVALID;
  return download(ctx, out, n->p.p, kParams * 4);
}

int dnab_mlp_set_params(dnab_ctx* ctx, dnab_mlp* n, const float* in, size_t count) {
  if (!ctx || !n || !in || count != kParams) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(upload(ctx, n->p.p, in, kParams * 4));
  hipLaunchKernelGGL(k_to_bf16, dim3(nb(kParams, 256)), dim3(256), 0, ctx->stream, n->p.p,
                     n->pbf.p, kParams);
  DNAB_RETURN_IF(refresh_inference_weights(ctx, n));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_mlp_loss_grad(dnab_ctx* ctx, dnab_mlp* n, const uint16_t* x, const float* t,
                       uint32_t batch, float* grads, float* loss, float* pred) {
  if (!ctx || !n || !x || !t || batch == 0) return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  DNAB_RETURN_IF(ensure_rows(ctx, n, batch));
  DNAB_RETURN_IF(upload(ctx, n->Xb.p, x, static_cast<size_t>(batch) * kD * 2));
  DNAB_RETURN_IF(upload(ctx, n->tb.p, t, batch * 4));
  forward(ctx, n, n->Xb.p, batch, true);
  backward(ctx, n, n->Xb.p, batch);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  if (grads) DNAB_RETURN_IF(download(ctx, grads, n->g.p, kParams * 4));
  if (pred) DNAB_RETURN_IF(download(ctx, pred, n->y.p, batch * 4));
  if (loss) {
    std::vector<float> sq(batch);
    DNAB_RETURN_IF(download(ctx, sq.data(), n->sq.p, batch * 4));
    double s = 0;
    for (float v : sq) s += v;
    *loss = static_cast<float>(s / batch);
  }
  return DNAB_OK;
}

int dnab_dataset_create(dnab_ctx* ctx, size_t capacity, float energy_scale, float energy_cap,
                        dnab_dataset** out) {
  if (!ctx || !out || capacity == 0 || !(energy_scale > 0.0f) || !std::isfinite(energy_cap))
    return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  dnab_dataset* d = new (std::nothrow) dnab_dataset();
  if (!d) return DNAB_ERR_ALLOC;
  d->ctx = ctx_retain(ctx);
  d->capacity = capacity;
  d->energy_scale = energy_scale;
  d->energy_cap = energy_cap;
  if (d->X.alloc(capacity * kD) || d->t.alloc(capacity) || d->e.alloc(capacity)) {
    dnab_dataset_destroy(d);
    return dnab_set_error(ctx, DNAB_ERR_ALLOC, "dataset: allocation of %zu rows failed", capacity);
  }
  *out = d;
  return DNAB_OK;
}

void dnab_dataset_destroy(dnab_dataset* d) {
  if (!d) return;
  dnab_ctx* ctx = d->ctx;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  delete d;
  ctx_release(ctx);
}

int dnab_dataset_size(const dnab_dataset* d, size_t* n) {
  if (!d || !n) return DNAB_ERR_INVALID;
  *n = d->size;
  return DNAB_OK;
}

int dnab_dataset_add_population(dnab_ctx* ctx, dnab_dataset* d, dnab_population* pop) {
  if (!ctx || !d || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "dataset_add_population"));
  if (d->size + pop->n_cand > d->capacity)
    return dnab_set_error(ctx, DNAB_ERR_CAPACITY, "dataset full (%zu + %zu > %zu)", d->size,
                          pop->n_cand, d->capacity);
  (void)hipSetDevice(ctx->device);
  DNAB_RETURN_IF(population_features(ctx, pop, d->X.p + d->size * kD));
  pop->masked = false;
  DNAB_RETURN_IF(population_sort(ctx, pop));
  DNAB_RETURN_IF(population_evaluate(ctx, pop, 0));
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  DNAB_HIP_CHECK(ctx, hipMemcpyAsync(d->e.p + d->size, pop->energy.p, nc * sizeof(double),
                                     hipMemcpyDeviceToDevice, ctx->stream));
  hipLaunchKernelGGL(k_targets, dim3(nb(nc, 256)), dim3(256), 0, ctx->stream, pop->energy.p, nc,
                     d->energy_scale, d->energy_cap, d->t.p + d->size);
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  d->size += nc;
  return DNAB_OK;
}

int dnab_train_default_params(dnab_train_params* p) {
  if (!p) return DNAB_ERR_INVALID;
  p->n_steps = 2000;
  p->batch = 512;
  p->lr = 1e-3f;
  p->beta1 = 0.9f;
  p->beta2 = 0.999f;
  p->eps = 1e-8f;
  p->seed = 1;
  return DNAB_OK;
}

int dnab_mlp_train(dnab_ctx* ctx, dnab_mlp* n, dnab_dataset* d, const dnab_train_params* tp,
                   float* loss_history) {
  if (!ctx || !n || !d || !tp || tp->batch == 0 || tp->batch % 64 != 0)
    return DNAB_ERR_INVALID;
  if (d->size < tp->batch)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "dataset smaller than one batch");
  (void)hipSetDevice(ctx->device);
  const uint32_t B = tp->batch;
  DNAB_RETURN_IF(ensure_rows(ctx, n, B));
  DNAB_ALLOC_CHECK(ctx, n->perm.ensure(d->size));
  DNAB_ALLOC_CHECK(ctx, n->loss_hist.ensure(tp->n_steps ? tp->n_steps : 1));
  SplitMix64 rng{tp->seed ^ 0x747261696eull};
  std::vector<uint32_t> perm(d->size);
  const size_t per_epoch = d->size / B;
  size_t pos = per_epoch;  // force a shuffle first
  for (uint32_t s = 0; s < tp->n_steps; ++s) {
    if (pos >= per_epoch) {
      for (size_t i = 0; i < perm.size(); ++i) perm[i] = static_cast<uint32_t>(i);
      for (size_t i = perm.size(); i > 1; --i) std::swap(perm[i - 1], perm[rng.below(static_cast<uint32_t>(i))]);
      DNAB_RETURN_IF(upload(ctx, n->perm.p, perm.data(), perm.size() * 4));
      pos = 0;
    }
    hipLaunchKernelGGL(k_gather_rows, dim3(B), dim3(256), 0, ctx->stream, d->X.p, d->t.p,
                       n->perm.p + pos * B, B, n->Xb.p, n->tb.p);
    ++pos;
    forward(ctx, n, n->Xb.p, B, true);
    backward(ctx, n, n->Xb.p, B);
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(256), 0, ctx->stream, n->sq.p, B,
                       1.0f / static_cast<float>(B), n->loss_hist.p + s);
    n->step += 1;
    const float c1 = 1.0f - std::pow(tp->beta1, static_cast<float>(n->step));
    const float c2 = 1.0f - std::pow(tp->beta2, static_cast<float>(n->step));
    hipLaunchKernelGGL(k_adam, dim3(nb(kParams, 256)), dim3(256), 0, ctx->stream, n->p.p, n->g.p,
                       n->m.p, n->v.p, n->pbf.p, kParams, tp->lr, tp->beta1, tp->beta2, tp->eps,
                       c1, c2);
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(refresh_inference_weights(ctx, n));
  if (loss_history && tp->n_steps)
    DNAB_RETURN_IF(download(ctx, loss_history, n->loss_hist.p, tp->n_steps * 4));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_population_features(dnab_ctx* ctx, dnab_population* pop, uint16_t* host) {
  if (!ctx || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "population_features"));
  (void)hipSetDevice(ctx->device);
  DNAB_ALLOC_CHECK(ctx, pop->feat.ensure(pop->n_cand * kD));
  DNAB_RETURN_IF(population_features(ctx, pop, pop->feat.p));
  if (host) DNAB_RETURN_IF(download(ctx, host, pop->feat.p, pop->n_cand * kD * 2));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_mlp_predict(dnab_ctx* ctx, dnab_mlp* n, dnab_population* pop, float* scores) {
  if (!ctx || !n || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "mlp_predict"));
  (void)hipSetDevice(ctx->device);
  DNAB_RETURN_IF(predict_device(ctx, n, pop));
  if (scores) DNAB_RETURN_IF(download(ctx, scores, pop->scores.p, pop->n_cand * 4));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_prescreen(dnab_ctx* ctx, dnab_mlp* n, dnab_population* pop, uint32_t n_keep,
                   uint32_t top_k, uint32_t* idx, double* energy, uint32_t* n_out) {
  if (!ctx || !n || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "prescreen"));
  if (n_keep == 0 || top_k == 0 || n_keep > 2048 || top_k > n_keep)
    return dnab_set_error(ctx, DNAB_ERR_INVALID,
                          "prescreen: need 1 <= top_k <= n_keep <= 2048 (got %u, %u)", top_k, n_keep);
  (void)hipSetDevice(ctx->device);
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  const uint32_t keep = std::min(n_keep, nc);
  const uint32_t k = std::min(top_k, keep);
  if (n_out) *n_out = k;
  if (nc == 0) return DNAB_OK;
  // features and predictions for every candidate
  DNAB_RETURN_IF(predict_device(ctx, n, pop));
  DNAB_ALLOC_CHECK(ctx, pop->scores64.ensure(nc));
  DNAB_ALLOC_CHECK(ctx, pop->keep_idx.ensure(keep));
  DNAB_ALLOC_CHECK(ctx, pop->keep_e.ensure(keep));
  DNAB_ALLOC_CHECK(ctx, pop->topk_idx.ensure(k));
  DNAB_ALLOC_CHECK(ctx, pop->topk_e.ensure(k));
  // the n_keep best predictions (in index order)
  DNAB_RETURN_IF(select_smallest(ctx, pop->scores.p, nc, keep, pop->keep_idx.p, pop->sel_scratch));
  TopkPlan pe;
  DNAB_RETURN_IF(topk_plan(ctx, pe, keep, k));
  // their exact energies (the fused kernel's final-state evaluation, one
  // workgroup per kept candidate: the same sums as a full evaluation), then
  // the top_k of them, ties by candidate index
  DNAB_RETURN_IF(trajectory_eval_subset(ctx, pop, pop->keep_idx.p, keep, pop->scores64.p));
  hipLaunchKernelGGL(k_gather_energy, dim3(nb(keep, 256)), dim3(256), 0, ctx->stream,
                     pop->energy.p, pop->keep_idx.p, keep, pop->keep_e.p);
  topk_launch(ctx, pe, pop->keep_e.p, pop->topk_e.p, pop->topk_idx.p, nullptr, 1, 0,
              pop->keep_idx.p);
  // the other candidates' outputs were not refreshed
  pop->evaluated = false;
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(download2(ctx, idx, pop->topk_idx.p, k * 4, energy, pop->topk_e.p, k * 8));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_exact_topk(dnab_ctx* ctx, dnab_population* pop, uint32_t top_k, uint32_t* idx,
                    double* energy, uint32_t* n_out) {
  if (!ctx || !pop || pop->ctx != ctx || top_k == 0 || top_k > 4096) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "exact_topk"));
  (void)hipSetDevice(ctx->device);
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  const uint32_t k = std::min(top_k, nc);
  if (n_out) *n_out = k;
  if (nc == 0) return DNAB_OK;
  DNAB_ALLOC_CHECK(ctx, pop->topk_idx.ensure(k));
  DNAB_ALLOC_CHECK(ctx, pop->topk_e.ensure(k));
  TopkPlan pe;
  DNAB_RETURN_IF(topk_plan(ctx, pe, nc, k));
  pop->masked = false;
  DNAB_RETURN_IF(population_sort(ctx, pop));
  DNAB_RETURN_IF(population_evaluate(ctx, pop, 0));
  topk_launch(ctx, pe, pop->energy.p, pop->topk_e.p, pop->topk_idx.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(download2(ctx, idx, pop->topk_idx.p, k * 4, energy, pop->topk_e.p, k * 8));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_select_smallest(dnab_ctx* ctx, const float* scores, uint32_t n, uint32_t m,
                         uint32_t* idx, uint32_t* n_out) {
  if (!ctx || (n && !scores) || !idx || m == 0 || m > 2048) return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  const uint32_t k = std::min(n, m);
  if (n_out) *n_out = k;
  if (n == 0) return DNAB_OK;
  DevBuf<float> d_s;
  DevBuf<uint32_t> d_i, scratch;
  DNAB_ALLOC_CHECK(ctx, d_s.alloc(n));
  DNAB_ALLOC_CHECK(ctx, d_i.alloc(k));
  DNAB_RETURN_IF(upload(ctx, d_s.p, scores, static_cast<size_t>(n) * sizeof(float)));
  DNAB_RETURN_IF(select_smallest(ctx, d_s.p, n, m, d_i.p, scratch));
  DNAB_RETURN_IF(download(ctx, idx, d_i.p, k * sizeof(uint32_t)));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

}  // extern "C"
