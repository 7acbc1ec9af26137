// Device helpers shared by the gfx950 kernels: SpatialHashGrid cell keys,
// Morton codes, wave64 DPP reductions, TwoSum, Philox4x32-10.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dnab {

constexpr int kWave = 64;

// Rust `f32 as i32` (saturating, NaN -> 0), the cast applied to the floored
// key in SpatialHashGrid::cell_of (src/spatial.rs:61-67).
__host__ __device__ inline int32_t f32_to_i32_sat(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return static_cast<int32_t>(f);
}

// cell_of one coordinate: floor(x * inv_cell) with inv_cell = fl(1/cell)
// computed once in f32 (src/spatial.rs:26, 63-65).
__host__ __device__ inline int32_t cell_key(float x, float inv_cell) {
  float p = x * inv_cell;
  return f32_to_i32_sat(floorf(p));
}

// Rust i32 `+` in a release build wraps (overflow checks are off).
__host__ __device__ inline int32_t wrap_add(int32_t a, int32_t b) {
  return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b));
}

// Spread the low 10 bits of v to every third bit.
__host__ __device__ inline uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__host__ __device__ inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

// DPP controls (GCN/CDNA encoding).
constexpr int kDppQuadXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141;  // lane i <-> 7-i within 8
constexpr int kDppRowMirror = 0x140;      // lane i <-> 15-i within 16

template <int CTRL>
__device__ inline float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ inline double dpp_f64(double v) {
  int2 p = *reinterpret_cast<int2*>(&v);
  int2 q;
  q.x = __builtin_amdgcn_mov_dpp(p.x, CTRL, 0xF, 0xF, false);
  q.y = __builtin_amdgcn_mov_dpp(p.y, CTRL, 0xF, 0xF, false);
  return *reinterpret_cast<double*>(&q);
}

__device__ inline float readlane_f32(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ inline double readlane_f64(double v, int lane) {
  int2 p = *reinterpret_cast<int2*>(&v);
  int2 q;
  q.x = __builtin_amdgcn_readlane(p.x, lane);
  q.y = __builtin_amdgcn_readlane(p.y, lane);
  return *reinterpret_cast<double*>(&q);
}

// Wave64 reductions: DPP butterflies inside each row of 16 lanes, then the
// four row results in a fixed order. Must be called with all 64 lanes
// active; the result is uniform across the wave.
struct OpAdd {
  template <class T> __device__ static T f(T a, T b) { return a + b; }
};
struct OpMin {
  __device__ static float f(float a, float b) { return fminf(a, b); }
};
struct OpMax {
  __device__ static float f(float a, float b) { return fmaxf(a, b); }
};

template <class Op>
__device__ inline float wave_reduce_f32(float v) {
  v = Op::f(v, dpp_f32<kDppQuadXor1>(v));
  v = Op::f(v, dpp_f32<kDppQuadXor2>(v));
  v = Op::f(v, dpp_f32<kDppRowHalfMirror>(v));
  v = Op::f(v, dpp_f32<kDppRowMirror>(v));
  float r0 = readlane_f32(v, 0), r1 = readlane_f32(v, 16);
  float r2 = readlane_f32(v, 32), r3 = readlane_f32(v, 48);
  return Op::f(Op::f(r0, r1), Op::f(r2, r3));
}

__device__ inline double wave_sum_f64(double v) {
  v = v + dpp_f64<kDppQuadXor1>(v);
  v = v + dpp_f64<kDppQuadXor2>(v);
  v = v + dpp_f64<kDppRowHalfMirror>(v);
  v = v + dpp_f64<kDppRowMirror>(v);
  double r0 = readlane_f64(v, 0), r1 = readlane_f64(v, 16);
  double r2 = readlane_f64(v, 32), r3 = readlane_f64(v, 48);
  return (r0 + r1) + (r2 + r3);
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
  v += static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), kDppQuadXor1, 0xF, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), kDppQuadXor2, 0xF, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), kDppRowHalfMirror, 0xF, 0xF, false));
  v += static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), kDppRowMirror, 0xF, 0xF, false));
  uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
  uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
  return (r0 + r1) + (r2 + r3);
}

// (the compare mask itself, with no round trip through a 0/1 register)
__device__ inline uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// This lane's bit of a wave mask (the inverse of ballot64).
__device__ inline bool lane_in(uint64_t mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

// TwoSum (Knuth, TAOCP vol. 2, 4.2.2): s + e = a + b exactly, no branch and
// no magnitude precondition. Requires IEEE evaluation without contraction or
// reassociation (these sources are compiled with -ffp-contract=off).
__host__ __device__ inline void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bp = s - a;
  float ap = s - bp;
  e = (a - ap) + (b - bp);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based RNG.
struct Philox4 {
  uint32_t v[4];
};
__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
}
__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                 uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint32_t h0 = mulhi32(M0, c0), l0 = M0 * c0;
    uint32_t h1 = mulhi32(M1, c2), l1 = M1 * c2;
    uint32_t n0 = h1 ^ c1 ^ k0, n1 = l1, n2 = h0 ^ c3 ^ k1, n3 = l0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  Philox4 o;
  o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
  return o;
}

// Uniform float in [0, 1) from the top 24 bits.
__host__ __device__ inline float u01_from_u32(uint32_t u) {
  return static_cast<float>(u >> 8) * (1.0f / 16777216.0f);
}

}  // namespace dnab
