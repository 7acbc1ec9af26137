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
//
// Clash pairs. A pair deep inside the repulsive wall, (R/r)^2 > kHotS2, is
// evaluated in FP64 from the exact coordinate differences and the raw
// epsilons: its LJ terms reach 1e10 kcal/mol and 1e11 kcal/mol/A, where the
// FP32 terms' own rounding (~1e-6 of each) would exceed the tolerance of a
// force component in which two such pairs nearly cancel. Every other pair has
// |F| below ~70 kcal/mol/A and stays in FP32. Per (atom, chunk) the FP32 sum
// and the FP64 sum of the clash pairs are each rounded once to the integer
// unit; the clash pairs are found in the pair loop at the cost of two
// instructions per pair and evaluated after it (hot_chunk). Evaluations
// whose results are reported take this path (CLASH); the loop's steps do not
// need it and keep every pair in FP32.
//
// Saturation. An integer sum is bounded by +-2^59 units (+-5.5e11); a sum
// that reaches the bound is clamped and its component is marked in the
// accumulator's sticky status, which the writers turn into +-inf, so a
// saturated result is never reported as a finite number.
#pragma once
#include "device_common.h"

namespace dnab {

constexpr float kMinRSq = 0.25f;   // src/energy.rs:17
constexpr float kHydroSq = 36.0f;  // src/energy.rs:20
constexpr float kHotS2 = 2.0f;     // (R/r)^2 above which a pair is evaluated in FP64
constexpr int kLoadAhead = 4;   // chunk loads in flight per round trip
constexpr double kFixScale = 1048576.0;          // 2^20
constexpr double kFixUnit = 1.0 / 1048576.0;
constexpr long long kFixLim = 1LL << 59;         // accumulator bound; sums of 4 stay below 2^61
// error estimate per energy rounding, in units of sum |e| (times 32 ulp(1) = 2^-19
// gives 2^-21, half a fixed-point unit)
constexpr float kRoundAbs = 0.25f;

// Sticky status bits of an integer accumulator: the component reached the
// bound (energy, fx, fy, fz), or a non-finite value was rounded.
constexpr uint32_t kStE = 1u, kStX = 2u, kStY = 4u, kStZ = 8u, kStNaN = 16u;

// The antigen atom record's last float: AtomType (0..15, or 16 when not
// given) plus kTagHyd when the atom is hydrophobic.
constexpr float kTagHyd = 32.0f;
constexpr float kTagUntyped = 16.0f;
__host__ __device__ inline float ag_tag(bool hyd, int type) {
  return (type >= 0 && type < 16 ? static_cast<float>(type) : kTagUntyped) + (hyd ? kTagHyd : 0.0f);
}
__host__ __device__ inline bool tag_hyd(float tag) { return tag >= kTagHyd; }
// AtomType, or -1 when not given
__host__ __device__ inline int tag_type(float tag) {
  const int t = static_cast<int>(tag >= kTagHyd ? tag - kTagHyd : tag);
  return t < 16 ? t : -1;
}

__device__ inline long long fix_round(double v, uint32_t& st, uint32_t bit) {
  const double lim = static_cast<double>(kFixLim);
  const double s = v * kFixScale;
  if (!(fabs(s) < lim)) st |= (s != s) ? (bit | kStNaN) : bit;
  return __double2ll_rn(fmin(fmax(s, -lim), lim));   // NaN -> -lim (fmax drops NaN), flagged
}
__device__ inline long long fix_add(long long a, long long b, uint32_t& st, uint32_t bit) {
  const long long s = a + b;  // |a|, |b| <= 2^59 + 2^25: no overflow
  if (s >= kFixLim || s <= -kFixLim) st |= bit;
  return s > kFixLim ? kFixLim : (s < -kFixLim ? -kFixLim : s);
}
// A clamped integer sum as a reported value: +-inf when its status bit is set
// (NaN with kStNaN), else v units.
__device__ inline double fix_value(long long v, uint32_t st, uint32_t bit) {
  if (st & kStNaN) return __builtin_nan("");
  if (st & bit) return v < 0 ? -__builtin_huge_val() : __builtin_huge_val();
  return static_cast<double>(v) * kFixUnit;
}

// Per-lane integer accumulators of one tile slot.
struct FixAcc {
  long long e = 0, a = 0, x = 0, y = 0, z = 0;
  uint32_t st = 0;
};

// The pair predicate in f32: r^2 = (dx*dx + dy*dy) + dz*dz and the
// reference's skip test `r_sq > CUTOFF_SQ || r_sq < MIN_R_SQ`
// (src/energy.rs:279-284), negated, so a NaN distance is kept as the CPU
// reference keeps it (the WGSL kernel, shaders/energy.wgsl:74, drops it).
// An empty slot passes rc2_lane = -1 and has no pair.
__device__ inline void pair_pred(float4 B0, float xb, float yb, float zb, float rc2_lane,
                                 float& dx, float& dy, float& dz, float& r2, bool& in) {
  dx = xb - B0.x;
  dy = yb - B0.y;
  dz = zb - B0.z;
  r2 = (dx * dx + dy * dy) + dz * dz;
  in = !(r2 > rc2_lane || r2 < kMinRSq);
}

// The lanes whose pair passes the cutoff half of the predicate of pair_pred
// (a NaN distance passes), as a wave mask: the comparison's own result. The
// pair loop skips a staged atom outside it for every lane; the other half,
// r^2 >= kMinRSq, is applied where the pair is evaluated (pair_in).
__device__ inline uint64_t pred_mask(float r2, float rc2_lane) { return ballot64(!(r2 > rc2_lane)); }
// The lane's pair passes the whole predicate, given its bit of pred_mask.
__device__ inline bool pair_in(uint64_t mask, float r2) { return lane_in(mask) && !(r2 < kMinRSq); }

// One pair term (the reference's formulas, src/energy.rs:261-314) after the
// predicate: the f32 energy ep and, with FORCES, the scalar f with
// F = f (dx, dy, dz); both exactly zero when the pair is out of range or
// `hot` (evaluated in FP64 instead). h: the pair takes the hydrophobic step.
// B0 = (x, y, z, 332 q), B1 = (Rmin/2, sqrt(eps), eps, tag) of the antigen atom.
template <bool FORCES>
__device__ inline void pair_value(float4 B0, float4 B1, float qb, float rmhb, float sqeb,
                                  float sqe12b, bool hydb, float r2, bool in, bool hot, float& ep,
                                  float& f, uint32_t& h) {
  float rinv = __builtin_amdgcn_rsqf(r2);
  rinv = (in && !hot) ? rinv : 0.0f;
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
  if (tag_hyd(B1.w)) {  // uniform: the antigen atom is the same in every lane
    h = (in && hydb && r2 < kHydroSq) ? 1u : 0u;
    if (FORCES) f = h ? __builtin_fmaf(-0.1f, rinv, f) : f;
  }
  ep = elj + ecl;
}

// 1/sqrt(x) in FP64 for x > 0 (normal): the hardware estimate refined by
// two Newton steps y <- y (3 - x y^2) / 2, each of which squares the relative
// error, to below 1e-16 relative from the estimate's ~2^-26.
__device__ inline double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * __builtin_fma(-0.5 * x, y * y, 1.5);
  y = y * __builtin_fma(-0.5 * x, y * y, 1.5);
  return y;
}

// The FP64 evaluation of a clash pair, from the exact differences of the
// f32 coordinates and eps_ij = sqrt(eps_i eps_j) from the raw epsilons, as
// the all-FP64 kernel forms them (1/r and the square root through
// rsqrt_f64): adds its energy to he and, with FORCES, its force to
// (hx, hy, hz); returns its share of the error estimate's sum |e| (the FP32
// rounding of 332 q and a margin on the LJ term).
template <bool FORCES>
__device__ inline double hot_pair(float4 B0, float4 B1, float xb, float yb, float zb, float qb,
                                  float rmhb, float epsb, uint32_t h, double& he, double& hx,
                                  double& hy, double& hz) {
  const double ddx = static_cast<double>(xb) - static_cast<double>(B0.x);
  const double ddy = static_cast<double>(yb) - static_cast<double>(B0.y);
  const double ddz = static_cast<double>(zb) - static_cast<double>(B0.z);
  const double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;  // >= 0.25 (the predicate)
  const double rinv = rsqrt_f64(r2d);
  const double rinv2 = rinv * rinv;
  const double R = static_cast<double>(B1.x) + static_cast<double>(rmhb);
  const double s2 = (R * R) * rinv2;
  const double s6 = (s2 * s2) * s2;
  const double s12 = s6 * s6;
  const double pe = static_cast<double>(B1.z) * static_cast<double>(epsb);  // exact
  const double eps = pe > 0.0 ? pe * rsqrt_f64(pe) : 0.0;
  const double elj = eps * (s12 - 2.0 * s6);
  const double ecl = (static_cast<double>(B0.w) * static_cast<double>(qb)) * rinv;
  he += elj + ecl;
  if (FORCES) {
    double f = (12.0 * eps * (s12 - s6) + ecl) * rinv2;
    if (h) f -= 0.1 * rinv;
    hx += f * ddx;
    hy += f * ddy;
    hz += f * ddz;
  }
  return fabs(ecl) * (1.0 / 32.0) + fabs(elj) * 0x1p-26;
}

// FP64 sums of one lane's clash pairs with one chunk, in chunk order.
struct HotSum {
  double e = 0.0, x = 0.0, y = 0.0, z = 0.0, a = 0.0;
  bool mine = false;  // the lane has a clash pair in the chunk
};

// The pair of the lane's atom with antigen atom (B0, B1) when it is a clash
// pair: the same f32 predicate and test as the pair loop, then hot_pair.
template <bool FORCES>
__device__ inline void hot_add(float4 B0, float4 B1, float xb, float yb, float zb, float qb,
                               float rmhb, float epsb, bool hydb, float rc2_lane, HotSum& hs) {
  float dx, dy, dz, r2;
  bool in;
  pair_pred(B0, xb, yb, zb, rc2_lane, dx, dy, dz, r2, in);
  const float R = B1.x + rmhb;
  const bool hot = in && (R * R) > kHotS2 * r2;
  if (ballot64(hot)) {
    if (hot) {
      const uint32_t h = (tag_hyd(B1.w) && hydb && r2 < kHydroSq) ? 1u : 0u;
      hs.a += hot_pair<FORCES>(B0, B1, xb, yb, zb, qb, rmhb, epsb, h, hs.e, hs.x, hs.y, hs.z);
      hs.mine = true;
    }
  }
}

// Rounds a chunk's clash sums once into `acc` (separately from the chunk's
// FP32 sum) and clears them.
template <bool FORCES>
__device__ inline void hot_fold(HotSum& hs, FixAcc& acc) {
  if (hs.mine) {
    acc.e = fix_add(acc.e, fix_round(hs.e, acc.st, kStE), acc.st, kStE);
    uint32_t ignore = 0;
    acc.a = fix_add(acc.a, fix_round(hs.a + kRoundAbs, ignore, 0u), ignore, 0u);
    if (FORCES) {
      acc.x = fix_add(acc.x, fix_round(hs.x, acc.st, kStX), acc.st, kStX);
      acc.y = fix_add(acc.y, fix_round(hs.y, acc.st, kStY), acc.st, kStY);
      acc.z = fix_add(acc.z, fix_round(hs.z, acc.st, kStZ), acc.st, kStZ);
    }
  }
  hs = HotSum();
}

// The clash pairs of the lane's atom with the `n_s` antigen atoms staged in
// `st`: summed in FP64 in chunk order and rounded once per (atom, chunk) into
// `acc`, separately from the chunk's FP32 sum. A clash pair is always among
// the staged atoms, whichever tile box staged them, so every path that sums
// a chunk's clash pairs in chunk order forms the same sum.
template <bool FORCES>
__device__ inline void hot_chunk(const float4 (*st)[2], uint32_t n_s, float xb, float yb, float zb,
                                 float qb, float rmhb, float epsb, bool hydb, float rc2_lane,
                                 FixAcc& acc) {
  HotSum hs;
  for (uint32_t k = 0; k < n_s; ++k)
    hot_add<FORCES>(st[k][0], st[k][1], xb, yb, zb, qb, rmhb, epsb, hydb, rc2_lane, hs);
  hot_fold<FORCES>(hs, acc);
}

// Pairs of the lane's atom with the `n_s` antigen atoms staged in `st` (in
// chunk order); folds the chunk's sums into `acc`. The predicate is formed
// for two atoms at a time; an atom within the cutoff of no lane of the wave
// is skipped (each of its terms would add an exact zero), the others are
// added in chunk order, which keeps the sums those of a plain loop. With
// CLASH a clash pair adds exact zeros to these FP32 sums and the chunk's
// clash pairs are then summed in FP64 by hot_chunk; without it every pair is
// an FP32 term (the loop's steps: a clash force moves its residue by the
// clamped step whatever its rounding). A chunk with no atom within the
// cutoff of any lane leaves `acc` as it is (its sums are exact zeros),
// without the integer conversions. Without ENERGY (forces only: the loop's
// steps whose energies nothing reads) the energy, sum |e| and hydrophobic
// count are not formed and acc.e, acc.a stay as they are; the forces are
// formed by the same operations either way.
template <bool FORCES, bool COUNT, bool CLASH, bool ENERGY = true>
__device__ inline void chunk_pairs(const float4 (*st)[2], uint32_t n_s, float xb, float yb,
                                   float zb, float qb, float rmhb, float sqeb, float sqe12b,
                                   float epsb, bool hydb, float rc2_lane, bool valid, FixAcc& acc,
                                   uint32_t& n_in, uint32_t& n_tested) {
  static_assert(ENERGY || !CLASH, "clash pairs are evaluated only where energies are reported");
  float fx = 0.f, fy = 0.f, fz = 0.f, es = 0.f, ec = 0.f, eabs = 0.f;
  uint32_t nh = 0;
  uint64_t any = 0;     // wave-uniform
  bool mine = false;    // the lane's atom has a pair in this chunk
  bool anyhot = false;  // ... a clash pair (CLASH)
  auto add = [&](const float4& B0, const float4& B1, float dx, float dy, float dz, float r2,
                 bool in) {
    bool hot = false;
    if (CLASH) {
      const float R = B1.x + rmhb;
      hot = in && (R * R) > kHotS2 * r2;
      anyhot = anyhot || hot;
    }
    mine = mine || in;
    float ep, f;
    uint32_t h;
    pair_value<FORCES>(B0, B1, qb, rmhb, sqeb, sqe12b, hydb, r2, in, hot, ep, f, h);
    if (FORCES) {
      fx = __builtin_fmaf(f, dx, fx);
      fy = __builtin_fmaf(f, dy, fy);
      fz = __builtin_fmaf(f, dz, fz);
    }
    if (ENERGY) {
      nh += h;
      float sum, err;
      two_sum(es, ep, sum, err);
      es = sum;
      ec += err;
      eabs += fabsf(ep);
    }
  };
  uint32_t q = 0;
  for (; q + 2 <= n_s; q += 2) {
    const float4 A0 = st[q][0], B0 = st[q + 1][0];
    float dxa, dya, dza, r2a, dxb, dyb, dzb, r2b;
    bool ina, inb;
    pair_pred(A0, xb, yb, zb, rc2_lane, dxa, dya, dza, r2a, ina);
    pair_pred(B0, xb, yb, zb, rc2_lane, dxb, dyb, dzb, r2b, inb);
    const uint64_t ba = pred_mask(r2a, rc2_lane), bb = pred_mask(r2b, rc2_lane);
    if (COUNT) {
      n_tested += valid ? 2u : 0u;
      n_in += (pair_in(ba, r2a) ? 1u : 0u) + (pair_in(bb, r2b) ? 1u : 0u);
    }
    any |= ba | bb;
    // (the lane's flag taken back from the mask: no per-atom copy of the
    // flag is kept in a vector register while the mask decides the branch)
    if (ba) add(A0, st[q][1], dxa, dya, dza, r2a, pair_in(ba, r2a));
    if (bb) add(B0, st[q + 1][1], dxb, dyb, dzb, r2b, pair_in(bb, r2b));
  }
  if (q < n_s) {
    const float4 A0 = st[q][0];
    float dx, dy, dz, r2;
    bool in;
    pair_pred(A0, xb, yb, zb, rc2_lane, dx, dy, dz, r2, in);
    const uint64_t b1 = pred_mask(r2, rc2_lane);
    if (COUNT) {
      n_tested += valid ? 1u : 0u;
      n_in += pair_in(b1, r2) ? 1u : 0u;
    }
    any |= b1;
    if (b1) add(A0, st[q][1], dx, dy, dz, r2, pair_in(b1, r2));
  }
  if (!any) return;
  if (ENERGY) {
    // hydrophobic: -0.5 per pair, exactly -2^19 in fixed point
    const long long eh = -static_cast<long long>(nh) << 19;
    const double e_chunk = static_cast<double>(es) + static_cast<double>(ec);
    acc.e = fix_add(acc.e, fix_round(e_chunk, acc.st, kStE) + eh, acc.st, kStE);
    // sum |e| feeds the error estimate only; an atom's energy rounding can err
    // only in a chunk where it has a pair (a chunk is fixed by the antigen, so
    // this count does not depend on the tiles)
    uint32_t ignore = 0;
    acc.a = fix_add(acc.a, fix_round(static_cast<double>(mine ? eabs + kRoundAbs : eabs), ignore, 0u),
                    ignore, 0u);
  }
  if (FORCES) {
    acc.x = fix_add(acc.x, fix_round(static_cast<double>(fx), acc.st, kStX), acc.st, kStX);
    acc.y = fix_add(acc.y, fix_round(static_cast<double>(fy), acc.st, kStY), acc.st, kStY);
    acc.z = fix_add(acc.z, fix_round(static_cast<double>(fz), acc.st, kStZ), acc.st, kStZ);
  }
  if (CLASH && ballot64(anyhot))
    hot_chunk<FORCES>(st, n_s, xb, yb, zb, qb, rmhb, epsb, hydb, rc2_lane, acc);
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
