// Pair terms and exact chunk sums shared by the pair kernels (pair.hip) and
// the fused trajectory kernel (traj.hip).
//
// Reference semantics: interaction_energy_atoms_breakdown
// (src/energy.rs:261-314) and the WGSL kernel (shaders/energy.wgsl:46-118).
// A lane holds one candidate atom; the antigen atoms of one chunk (up to 64
// consecutive atoms of one grid cell) are staged in LDS and broadcast. Each
// (atom, chunk) sum is formed in FP32 in chunk order and rounded once to a
// 64-bit integer in units of 2^-20, so sums over chunks are exact and do not
// depend on how chunks are assigned to waves or kernels.
#pragma once
#include "device_common.h"

namespace dnab {

constexpr float kMinRSq = 0.25f;   // src/energy.rs:17
constexpr float kHydroSq = 36.0f;  // src/energy.rs:20
constexpr int kLoadAhead = 4;   // chunk loads in flight per round trip
constexpr double kFixScale = 1048576.0;          // 2^20
constexpr double kFixUnit = 1.0 / 1048576.0;
constexpr long long kFixLim = 1LL << 59;         // accumulator bound; sums of 4 stay below 2^61

__device__ inline long long fix_round(double v) {
  const double lim = static_cast<double>(kFixLim);
  v = fmin(fmax(v * kFixScale, -lim), lim);   // NaN -> -lim (fmax drops NaN)
  return __double2ll_rn(v);
}
__device__ inline long long fix_add(long long a, long long b) {
  const long long s = a + b;  // |a|, |b| <= 2^59 + 2^25: no overflow
  return s > kFixLim ? kFixLim : (s < -kFixLim ? -kFixLim : s);
}

// Per-lane integer accumulators of one tile slot.
struct FixAcc {
  long long e = 0, a = 0, x = 0, y = 0, z = 0;
};

// The pair predicate in f32, exactly the reference's (src/energy.rs:279-284).
__device__ inline void pair_pred(float4 B0, float xb, float yb, float zb, float rc2_lane,
                                 float& dx, float& dy, float& dz, float& r2, bool& in) {
  dx = xb - B0.x;
  dy = yb - B0.y;
  dz = zb - B0.z;
  r2 = (dx * dx + dy * dy) + dz * dz;
  in = r2 <= rc2_lane && r2 >= kMinRSq;
}

// One pair term (the reference's formulas, src/energy.rs:261-314) after the
// predicate: the f32 energy ep and, with FORCES, the scalar f with
// F = f (dx, dy, dz); both exactly zero when the pair is out of range.
template <bool FORCES>
__device__ inline void pair_value(float4 B0, float4 B1, float qb, float rmhb, float sqeb,
                                  float sqe12b, bool hydb, float r2, bool in, float& ep, float& f,
                                  uint32_t& h) {
  float rinv = __builtin_amdgcn_rsqf(r2);
  rinv = in ? rinv : 0.0f;
  const float rinv2 = rinv * rinv;
  const float R = B1.x + rmhb;
  const float s2 = (R * R) * rinv2;
  const float s6 = (s2 * s2) * s2;
  const float s12 = s6 * s6;
  const float elj = (B1.y * sqeb) * __builtin_fmaf(-2.0f, s6, s12);
  const float ecl = (B0.w * qb) * rinv;
  f = 0.0f;
  if (FORCES) {
    const float flj = ((B1.y * sqe12b) * (s12 - s6)) * rinv2;
    f = __builtin_fmaf(ecl, rinv2, flj);
  }
  h = 0;
  if (B1.z != 0.0f) {  // uniform: the antigen atom is the same in every lane
    h = (in && hydb && r2 < kHydroSq) ? 1u : 0u;
    if (FORCES) f = h ? __builtin_fmaf(-0.1f, rinv, f) : f;
  }
  ep = elj + ecl;
}

// Pairs of the lane's atom with the `n_s` antigen atoms staged in `st` (in
// chunk order); folds the chunk's sums into `acc`. The predicate is formed
// for two atoms at a time; an atom in range of no lane of the wave is
// skipped (each of its terms would add an exact zero), the others are
// added in chunk order, which keeps the sums those of a plain loop. A chunk
// with no pair in range in the whole wave leaves `acc` as it is (its sums
// are exact zeros), without the integer conversions.
template <bool FORCES, bool COUNT>
__device__ inline void chunk_pairs(const float4 (*st)[2], uint32_t n_s, float xb, float yb,
                                   float zb, float qb, float rmhb, float sqeb, float sqe12b,
                                   bool hydb, float rc2_lane, bool valid, FixAcc& acc,
                                   uint32_t& n_in, uint32_t& n_tested) {
  float fx = 0.f, fy = 0.f, fz = 0.f, es = 0.f, ec = 0.f, eabs = 0.f;
  uint32_t nh = 0;
  uint64_t any = 0;  // wave-uniform
  auto add = [&](const float4& B0, const float4& B1, float dx, float dy, float dz, float r2,
                 bool in) {
    float ep, f;
    uint32_t h;
    pair_value<FORCES>(B0, B1, qb, rmhb, sqeb, sqe12b, hydb, r2, in, ep, f, h);
    nh += h;
    if (FORCES) {
      fx = __builtin_fmaf(f, dx, fx);
      fy = __builtin_fmaf(f, dy, fy);
      fz = __builtin_fmaf(f, dz, fz);
    }
    float sum, err;
    two_sum(es, ep, sum, err);
    es = sum;
    ec += err;
    eabs += fabsf(ep);
  };
  uint32_t q = 0;
  for (; q + 2 <= n_s; q += 2) {
    const float4 A0 = st[q][0], B0 = st[q + 1][0];
    float dxa, dya, dza, r2a, dxb, dyb, dzb, r2b;
    bool ina, inb;
    pair_pred(A0, xb, yb, zb, rc2_lane, dxa, dya, dza, r2a, ina);
    pair_pred(B0, xb, yb, zb, rc2_lane, dxb, dyb, dzb, r2b, inb);
    if (COUNT) {
      n_tested += valid ? 2u : 0u;
      n_in += (ina ? 1u : 0u) + (inb ? 1u : 0u);
    }
    const uint64_t ba = ballot64(ina), bb = ballot64(inb);
    any |= ba | bb;
    if (ba) add(A0, st[q][1], dxa, dya, dza, r2a, ina);
    if (bb) add(B0, st[q + 1][1], dxb, dyb, dzb, r2b, inb);
  }
  if (q < n_s) {
    const float4 A0 = st[q][0];
    float dx, dy, dz, r2;
    bool in;
    pair_pred(A0, xb, yb, zb, rc2_lane, dx, dy, dz, r2, in);
    if (COUNT) {
      n_tested += valid ? 1u : 0u;
      n_in += in ? 1u : 0u;
    }
    const uint64_t b1 = ballot64(in);
    any |= b1;
    if (b1) add(A0, st[q][1], dx, dy, dz, r2, in);
  }
  if (!any) return;
  // hydrophobic: -0.5 per pair, exactly -2^19 in fixed point
  const long long eh = -static_cast<long long>(nh) << 19;
  acc.e = fix_add(acc.e, fix_round(static_cast<double>(es) + static_cast<double>(ec)) + eh);
  acc.a = fix_add(acc.a, fix_round(static_cast<double>(eabs)));
  if (FORCES) {
    acc.x = fix_add(acc.x, fix_round(static_cast<double>(fx)));
    acc.y = fix_add(acc.y, fix_round(static_cast<double>(fy)));
    acc.z = fix_add(acc.z, fix_round(static_cast<double>(fz)));
  }
}

// Stage the chunk atoms selected by `mask` (in chunk order) into `st`.
__device__ inline void stage_chunk(float4 (*st)[2], uint64_t mask, uint32_t lane, float4 A0,
                                   float4 A1) {
  if ((mask >> lane) & 1ull) {
    const uint32_t slot = __builtin_amdgcn_mbcnt_hi(
        static_cast<uint32_t>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
    st[slot][0] = A0;
    st[slot][1] = A1;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace dnab
