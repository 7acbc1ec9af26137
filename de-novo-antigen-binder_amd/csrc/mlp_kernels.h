// Learned prescreen building blocks (no reference counterpart: the reference
// has no learned components, Cargo.toml:4): bf16 storage, a bf16 MFMA GEMM
// with fused epilogues for the MLP's forward and backward passes, and small
// elementwise / reduction kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace dnab {

// bf16 <-> f32, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  const uint32_t r = 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>((u + r) >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  uint32_t u = static_cast<uint32_t>(h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// logistic function; v_rcp_f32 (1 ulp) rather than a correctly rounded division
__device__ inline float sigmoidf_(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

enum GemmEpilogue : int {
  kEpiStoreF32 = 0,  // C32 = acc
  kEpiBiasSilu = 1,  // Z32 = acc + bias; Cbf = bf16(silu(Z32))
  kEpiSiluGrad = 2,  // Cbf = bf16(acc * silu'(Z32)) (Z32 read)
  kEpiBias = 3,      // C32 = acc + bias
};

struct GemmArgs {
  const uint16_t* A;  // bf16
  const uint16_t* B;  // bf16
  int M, N, K;
  int lda, ldb, ldc;
  float* C32;         // fp32 output (or Z32 for the SiLU epilogues)
  uint16_t* Cbf;      // bf16 output
  const float* bias;  // N
  int vec_ok;         // 16-byte global loads allowed (lda/ldb multiples of 8)
};

}  // namespace dnab
