// All-atom refinement with Monte-Carlo moves on the device: every iteration
// of allatom_diffusion_step (src/diffusion.rs:290-416) for a batch of
// candidates in one launch, one workgroup of 1, 4 or 8 waves per candidate:
//   1. residue-level forces on each Calpha from the antigen Calpha grid
//      (compute_forces_with_grid, src/energy.rs:322-398), the radial
//      restraint, clamp, optional noise, set_ca_pos (the gradient step);
//   2. the interface mask: Calpha within 8 A of an antigen Calpha;
//   3. per residue in order, a Metropolis rotamer move or amino-acid
//      mutation scored by residue_contribution_allatom (src/diffusion.rs:
//      422-456) plus disulfide_energy (src/allatom.rs:155-180);
//   4. per residue in order, a Metropolis phi/psi torsion move of the chain
//      downstream (perturb_phi / perturb_psi, src/allatom.rs:392-456).
// Each residue holds up to 14 atom slots in LDS in topology order, so a
// mutation rewrites the slots of one residue instead of splicing arrays
// (mutate_residue, src/allatom.rs:299-384). Residues are rebuilt with the
// host builder's NERF (geometry.h) and sines/cosines the host tabulates, so
// they match host-built ones bit for bit. Random draws are Philox4x32-10 by
// (iteration, residue, kind); the energies of the Metropolis tests are FP64
// sums of the reference's f32 pair terms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "device_common.h"
#include "dnab_internal.h"
#include "ff_tables.h"
#include "ff_terms.h"
#include "geometry.h"
#include "grid_device.h"
#include "host_parallel.h"
#include "step_device.h"

namespace dnab {
namespace {

constexpr int kSlots = 14;        // atom slots per residue (the largest topology, Trp)
constexpr int kMaxDefs = 12;      // side-chain placements per residue type
constexpr uint32_t kMaxRefineRes = 512;
constexpr uint8_t kBuildNone = 0, kBuildSide = 1, kBuildIdeal = 2;  // deferred side-chain builds
constexpr uint32_t kTagMc = 0x6d636d76u;  // draws of the side-chain moves
constexpr uint32_t kTagBb = 0x62626d76u;  // draws of the torsion moves

struct RefDef {
  uint8_t a, b, c, chi;  // reference atoms; chi index or 255 (fixed dihedral)
  float bond, sa, ca;    // bond length, sine and cosine of the bond angle
  float sd, cd;          // sine and cosine of the fixed dihedral
};

// Per residue type: side-chain placements, chi sines/cosines per rotamer, the
// Calpha atom's AMBER parameters and the residue-level (CG) parameters.
struct RefTables {
  RefDef defs[kNumAminoAcids][kMaxDefs];
  uint8_t ndef[kNumAminoAcids], natoms[kNumAminoAcids], nrot[kNumAminoAcids];
  float chi_sc[kNumAminoAcids][5][4][2];  // (sin, cos) of chi k of rotamer j
  float ca_rmh[kNumAminoAcids], ca_eps[kNumAminoAcids], ca_q[kNumAminoAcids];
  uint8_t ca_hyd[kNumAminoAcids], ca_type[kNumAminoAcids];
  float cg_q[kNumAminoAcids], cg_eps[kNumAminoAcids], cg_sig[kNumAminoAcids];
  uint8_t cg_hyd[kNumAminoAcids];
};

static_assert(sizeof(RefTables) % 4 == 0, "tables are copied to LDS in 4-byte words");

struct RefineArgs {
  const RefTables* tab;
  // antigen Calphas: positions, CA-atom parameters (rmh, eps, q, hyd + 256 type)
  const float *ag_x, *ag_y, *ag_z;
  const float4* ag_par;
  uint32_t n_ag, n_ag_lds;  // n_ag_lds: n_ag when the Calphas are copied to LDS, else 0
  uint32_t ncells_lds;      // entries of the grid's cell_start copied with them
  GridView ag_grid;
  CgCloud ag_cg;
  float cx, cy, cz, r0;
  // candidates: residue ranges, per residue type / rotamer, 14 slots per residue
  const uint32_t* cand_res_off;
  uint8_t* res_aa;
  uint8_t* res_rot;
  float4* slots;
  uint32_t n_cand, max_res;
  uint32_t pair_cap;  // force-pass list entries (a multiple of 64)
  // schedule, per step
  const float* temp;
  const float* sigma;
  uint32_t n_steps, iter0;
  StepParams sp;
  float mut_p, rot_p, bb_p, max_torsion, iface_sq;
  uint32_t moves;
  uint32_t finite;  // every input coordinate finite: energies stay finite
};

// Orders this wave's LDS accesses around it (a wave's LDS operations complete
// in issue order): the synchronisation of work that one wave owns alone.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP64 sums over the workgroup's NW waves: each wave's DPP sum, then the
// waves' sums in wave order (red: NW doubles of LDS), the same on every lane.
template <int NW>
__device__ inline double block_sum(double v, double* red, uint32_t lane, uint32_t wv) {
  v = wave_sum_f64(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) t += red[w];
    __syncthreads();
    return t;
  }
}
template <int NW>
__device__ inline void block_sum2(double& a, double& b, double* red, uint32_t lane, uint32_t wv) {
  a = wave_sum_f64(a);
  b = wave_sum_f64(b);
  if constexpr (NW > 1) {
    if (lane == 0) {
      red[wv] = a;
      red[NW + wv] = b;
    }
    __syncthreads();
    a = red[0];
    b = red[NW];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      a += red[w];
      b += red[NW + w];
    }
    __syncthreads();
  }
}

// A value every lane holds alike, moved to a scalar register so that the
// control flow depending on it is scalar.
__device__ inline uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uniform_f32(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ inline float4 u4(uint32_t it, uint32_t r, uint32_t tag, const StepParams& sp) {
  const Philox4 u = philox4x32_10(it, r, tag, 0u, sp.seed_lo, sp.seed_hi);
  return make_float4(u01_from_u32(u.v[0]), u01_from_u32(u.v[1]), u01_from_u32(u.v[2]),
                     u01_from_u32(u.v[3]));
}

// The antigen Calphas the residue energy runs over (in LDS when they fit).
struct CaSet {
  const float *x, *y, *z;
  const float4* par;  // r_min_half, epsilon, charge, hydrophobic + 256 * type
  uint32_t n;
};

// One antigen Calpha's term of residue_contribution_allatom (src/diffusion.rs:
// 422-456) on a Calpha of type `aa` at squared distance r_sq, for a pair the
// reference does not skip (0.25 <= r_sq <= 100); q holds the antigen Calpha's
// r_min_half, epsilon, charge and hydrophobic flag + 256 * atom type.
__device__ inline double pair_term(const RefTables& T, uint32_t aa, const float4 q, float r_sq) {
  const float brm = T.ca_rmh[aa], be = T.ca_eps[aa], bq = T.ca_q[aa];
  const uint8_t bh = T.ca_hyd[aa];
  const uint32_t bt = T.ca_type[aa];
  const float r_ij = brm + q.x;
  const float eps = sqrtf(be * q.y);
  const float r2 = (r_ij * r_ij) / r_sq;
  const float r6 = r2 * r2 * r2;
  double e = static_cast<double>(eps * (r6 * r6 - 2.0f * r6));
  if (bq != 0.0f && q.z != 0.0f) e += static_cast<double>(332.0f * bq * q.z / sqrtf(r_sq));
  const uint32_t packed = static_cast<uint32_t>(q.w);
  const uint32_t ah = packed & 1u, at = packed >> 8;
  if (bh == 1 && ah == 1 && r_sq < 36.0f) e -= 0.5;
  const bool hb = (((kHbondDonorMask >> bt) & (kHbondAcceptorMask >> at)) |
                   ((kHbondAcceptorMask >> bt) & (kHbondDonorMask >> at))) & 1u;
  if (hb && r_sq <= kHbCutoffSq) {
    const float z = (sqrtf(r_sq) - kHbR0) / kHbSigma;
    e += static_cast<double>(-kHbDepth * expf(-0.5f * z * z));
  }
  return e;
}

// residue_contribution_allatom of residues [r0, r1) at their Calphas, this
// lane's FP64 part over the units u_first, u_first + u_step, ... of (antigen
// chunk of 64, group of 4 residues) pairs (a workgroup's waves split them): a
// cheap squared-distance pass (the reference's skip test) compacts the pairs
// it keeps into the wave's buf, then their terms are evaluated lane-parallel.
// With kTwo, one residue (r0) is scored as type ta into ea and as type tb
// into eb in the same pass; else each residue is scored as its current type
// into ea.
template <bool kTwo>
__device__ inline void contrib(const RefTables& T, const CaSet& g, const uint8_t* aa,
                               const float* X, const float* Y, const float* Z, uint32_t r0,
                               uint32_t r1, uint32_t ta, uint32_t tb, uint32_t lane,
                               uint32_t u_first, uint32_t u_step, uint2* buf, uint32_t bcap,
                               double& ea, double& eb) {
  ea = eb = 0.0;
  uint32_t cnt = 0;
  auto flush = [&]() {
    wave_sync();
    for (uint32_t k = lane; k < cnt; k += kWave) {
      const uint2 p = buf[k];
      const uint32_t t = p.y * kSlots + 1;
      const float dx = X[t] - g.x[p.x], dy = Y[t] - g.y[p.x], dz = Z[t] - g.z[p.x];
      const float r_sq = dx * dx + dy * dy + dz * dz;
      const float4 q = g.par[p.x];
      if (kTwo) {
        ea += pair_term(T, ta, q, r_sq);
        eb += pair_term(T, tb, q, r_sq);
      } else {
        ea += pair_term(T, aa[p.y], q, r_sq);
      }
    }
    wave_sync();
    cnt = 0;
  };
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t cb = r0; cb < r1; cb += kWave) {
    // the chunk's Calphas, residue cb + j in lane j, broadcast by readlane
    const uint32_t nrc = min(r1 - cb, static_cast<uint32_t>(kWave));
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (lane < nrc) {
      const uint32_t t = (cb + lane) * kSlots + 1;
      px = X[t];
      py = Y[t];
      pz = Z[t];
    }
    const uint32_t ngr = (nrc + 3) / 4, nunits = ngr * ((g.n + kWave - 1) / kWave);
    for (uint32_t u = u_first; u < nunits; u += u_step) {
      const uint32_t i = (u / ngr) * kWave + lane, j0 = (u % ngr) * 4;
      const bool live = i < g.n;
      const float ax = live ? g.x[i] : 0.0f, ay = live ? g.y[i] : 0.0f,
                  az = live ? g.z[i] : 0.0f;
      // four residues per unit (independent chains); room for their pairs first
      if (cnt + 4 * kWave > bcap) flush();
#pragma unroll
      for (uint32_t v = 0; v < 4; ++v) {
        const uint32_t j = j0 + v, jl = min(j, nrc - 1);
        const float dx = readlane_f32(px, jl) - ax, dy = readlane_f32(py, jl) - ay,
                    dz = readlane_f32(pz, jl) - az;
        const float r_sq = dx * dx + dy * dy + dz * dz;
        const bool keep = live & (j < nrc) & !(r_sq > 100.0f) & !(r_sq < 0.25f);  // NaN kept
        const uint64_t b = ballot64(keep);
        if (keep) buf[cnt + static_cast<uint32_t>(__popcll(b & below))] = make_uint2(i, cb + j);
        cnt += static_cast<uint32_t>(__popcll(b));
      }
    }
  }
  flush();
}

// residue_contribution_allatom of a Calpha at p as type ta and as type tb in
// one pass, FP64 totals on every thread.
template <int NW>
__device__ inline void contribution2(const RefTables& T, const CaSet& g, uint32_t ta, uint32_t tb,
                                     float px, float py, float pz, uint32_t lane, uint32_t wv,
                                     double* red, double& ea, double& eb) {
  ea = eb = 0.0;
  for (uint32_t i = wv * kWave + lane; i < g.n; i += NW * kWave) {
    const float dx = px - g.x[i], dy = py - g.y[i], dz = pz - g.z[i];
    const float r_sq = dx * dx + dy * dy + dz * dz;
    if (r_sq > 100.0f || r_sq < 0.25f) continue;
    const float4 q = g.par[i];
    ea += pair_term(T, ta, q, r_sq);
    eb += pair_term(T, tb, q, r_sq);
  }
  block_sum2<NW>(ea, eb, red, lane, wv);
}

// ---- residue builds on a quad of lanes ------------------------------------
// Lane q = 0, 1, 2 of a quad holds component q (x, y, z) of every vector and
// lane 3 follows along without writing. A lane takes the components it needs
// from the others by DPP quad permutes and applies the host builder's f32
// operations in the host builder's order, so a residue equals geo::place's
// bit for bit, with one division per lane where geo::normalize does three.
constexpr int kQuadFrom0 = 0x00, kQuadFrom1 = 0x55, kQuadFrom2 = 0xAA;
constexpr int kQuadNext = 0xC9;   // lane q reads lane (q + 1) mod 3; lane 3 itself
constexpr int kQuadNext2 = 0xD2;  // lane q reads lane (q + 2) mod 3; lane 3 itself

__device__ inline float q_normalize(float v) {
  const float s = v * v;
  float l = sqrtf((dpp_f32<kQuadFrom0>(s) + dpp_f32<kQuadFrom1>(s)) + dpp_f32<kQuadFrom2>(s));
  if (l < 1e-8f) l = 1e-8f;
  return v / l;
}
__device__ inline float q_cross(float a, float b) {
  return dpp_f32<kQuadNext>(a) * dpp_f32<kQuadNext2>(b) -
         dpp_f32<kQuadNext2>(a) * dpp_f32<kQuadNext>(b);
}
// geo::place, component q of D
__device__ inline float q_place(float A, float B, float C, float bond_len, float sa, float ca,
                                float sd, float cd) {
  const float bc = q_normalize(B - C);
  const float n = q_normalize(q_cross(B - A, bc));
  const float m = q_cross(bc, n);
  return C + bond_len * (-ca * bc + sa * cd * m + sa * sd * n);
}
// rebuild_side_chain (src/allatom.rs:244-271) on a quad: the side chain of
// type aa, rotamer rot from the backbone atoms 0..nb-1; P = this lane's
// component array of the residue
__device__ inline void q_side_chain(const RefTables& T, uint32_t aa, uint32_t rot, int nb,
                                    float* P, uint32_t q) {
  int n = nb;
  for (int k = 0; k < T.ndef[aa] && n < T.natoms[aa]; ++k) {
    const RefDef& d = T.defs[aa][k];
    float sd = d.sd, cd = d.cd;
    if (d.chi != 255) {
      if (T.nrot[aa]) {
        sd = T.chi_sc[aa][rot][d.chi][0];
        cd = T.chi_sc[aa][rot][d.chi][1];
      } else {
        sd = 0.0f;
        cd = 1.0f;
      }
    }
    const float v = q_place(P[d.a], P[d.b], P[d.c], d.bond, d.sa, d.ca, sd, cd);
    if (q < 3) P[n] = v;
    ++n;
  }
}
// The other-type branch of mutate_residue (src/allatom.rs:316-384) on a
// quad: the ideal backbone around the Calpha in P[1] (geo::ideal_backbone on
// every lane), then the side chain
__device__ inline void q_build_ideal(const RefTables& T, uint32_t new_aa, uint32_t new_rot,
                                     float* P, uint32_t q) {
  const float c = P[1];
  const float cx = dpp_f32<kQuadFrom0>(c), cy = dpp_f32<kQuadFrom1>(c), cz = dpp_f32<kQuadFrom2>(c);
  geo::V3 bb[5];
  int nb = geo::ideal_backbone(cx, cy, cz, new_aa == kGly, bb);
  if (nb > T.natoms[new_aa]) nb = T.natoms[new_aa];
  for (int k = 0; k < nb; ++k)
    if (q < 3) P[k] = q == 0 ? bb[k].x : q == 1 ? bb[k].y : bb[k].z;
  q_side_chain(T, new_aa, new_rot, nb, P, q);
}

// disulfide_energy (src/allatom.rs:155-180) over the Cys SG slots (slot 5):
// the pair set by the f32 distance and each well in FP64, as there. The
// wave lists its Cys residues in order (cys: one entry per residue of LDS,
// written alike by every wave), then lane 8a + b takes the pair (8 ta + a,
// 8 tb + b) of each 8 x 8 tile; the wells are summed by DPP.
__device__ inline double disulfide(const uint8_t* aa, const float* X, const float* Y,
                                   const float* Z, uint32_t nr, uint32_t lane, uint16_t* cys) {
  uint32_t k = 0;
  for (uint32_t cb = 0; cb < nr; cb += kWave) {
    const bool is = cb + lane < nr && aa[cb + lane] == kCys;
    const uint64_t m = ballot64(is);
    if (is) cys[k + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)))] =
        static_cast<uint16_t>(cb + lane);
    k += static_cast<uint32_t>(__popcll(m));
  }
  wave_sync();
  double e = 0.0;
  const uint32_t nt = (k + 7) / 8, a = lane >> 3, b = lane & 7;
  for (uint32_t ta = 0; ta < nt; ++ta)
    for (uint32_t tb = ta; tb < nt; ++tb) {
      const uint32_t pi = ta * 8 + a, pj = tb * 8 + b;
      if (!(pi < pj && pj < k)) continue;
      const uint32_t si = cys[pi] * kSlots + kCysSgSlot, sj = cys[pj] * kSlots + kCysSgSlot;
      const float dx = X[si] - X[sj], dy = Y[si] - Y[sj], dz = Z[si] - Z[sj];
      if (sqrtf(dx * dx + dy * dy + dz * dz) > kSsCutoff) continue;
      const double ex = static_cast<double>(X[si]) - X[sj], ey = static_cast<double>(Y[si]) - Y[sj],
                   ez = static_cast<double>(Z[si]) - Z[sj];
      const double t = (sqrt(ex * ex + ey * ey + ez * ez) - static_cast<double>(kSsD0)) /
                       static_cast<double>(kSsSigma);
      e -= static_cast<double>(kSsDepth) * exp(-0.5 * t * t);
    }
  wave_sync();
  return wave_sum_f64(e);
}

// Sum of residue_contribution_allatom over residues r..nr-1 plus the
// disulfide term: the energy a torsion move at r is judged by. ncys: the
// candidate's Cys count (no pair, no disulfide term below two).
template <int NW>
__device__ inline double chain_energy(const RefTables& T, const CaSet& g, const uint8_t* aa,
                                      const float* X, const float* Y, const float* Z, uint32_t r,
                                      uint32_t nr, uint32_t ncys, uint32_t lane, uint32_t wv,
                                      uint2* buf, uint32_t bcap, double* red, uint16_t* cysl) {
  double part, unused;
  contrib<false>(T, g, aa, X, Y, Z, r, nr, 0, 0, lane, wv, NW, buf, bcap, part, unused);
  return block_sum<NW>(part, red, lane, wv) + (ncys >= 2 ? disulfide(aa, X, Y, Z, nr, lane, cysl) : 0.0);
}

// perturb_phi / perturb_psi (src/allatom.rs:392-456) by rad: phi turns the
// atoms of residue r after its Calpha and every later residue about the N-CA
// axis through CA; psi turns r's O and every later residue about the CA-C
// axis through C. Split over the threads by slot; the axis atoms do not move.
__device__ inline void torsion(float* X, float* Y, float* Z, uint32_t r, uint32_t nr, float rad,
                               bool phi, uint32_t tid, uint32_t nt) {
  const uint32_t b = r * kSlots;
  const geo::V3 p0 = phi ? geo::V3{X[b], Y[b], Z[b]} : geo::V3{X[b + 1], Y[b + 1], Z[b + 1]};
  const geo::V3 p1 = phi ? geo::V3{X[b + 1], Y[b + 1], Z[b + 1]} : geo::V3{X[b + 2], Y[b + 2], Z[b + 2]};
  const geo::V3 axis = geo::normalize(geo::sub(p1, p0));
  const float cs = static_cast<float>(cos(static_cast<double>(rad)));
  const float sn = static_cast<float>(sin(static_cast<double>(rad)));
  __syncthreads();
  for (uint32_t q = b + tid; q < nr * kSlots; q += nt) {
    const uint32_t rr = q / kSlots, k = q - rr * kSlots;
    if (rr == r && (phi ? k < 2 : k != 3)) continue;  // slots past the atoms turn too, unread
    const geo::V3 o = geo::rodrigues({X[q], Y[q], Z[q]}, p1, axis, cs, sn);
    X[q] = o.x;
    Y[q] = o.y;
    Z[q] = o.z;
  }
  __syncthreads();
}

// Scratch of one wave's residue-level force pass, in LDS: a flattened pair
// list of `cap` entries and the per-residue rows of one chunk of 64 residues.
struct CgScratch {
  float4* CT;    // cap: in-predicate force terms, compacted, in list order
  float4* RQ;    // 64: Calpha x, y, z and the residue-level charge
  float4* RP;    // 64: epsilon, sigma, hydrophobic flag
  uint32_t* H;   // cap: at each cell's first pair, its packed entry (below)
  uint32_t* HM;  // cap / 32: which list positions start a cell
  uint32_t* CS;  // 64: each residue's first compacted term
  uint32_t cap;  // a multiple of 64, with n + cap < 2^25 for the grid's n points
};

// A cell's entry in the list: bits 0-24 hold (first sorted index) - (first
// list position) + cap, bits 25-30 the residue's lane, bit 31 marks the
// residue's first non-empty cell.
constexpr uint32_t kEntryOffMask = 0x1FFFFFFu;
constexpr uint32_t kEntryFirst = 0x80000000u;

__device__ inline uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
  for (uint32_t d = 1; d < kWave; d <<= 1) {
    const uint32_t t = __shfl_up(v, d);
    if (lane >= d) v += t;
  }
  return v;
}

// compute_forces_with_grid (src/energy.rs:322-398) for the Calphas of the
// residues r_first, r_first + r_step, ... < nr, by one wave with its own
// scratch (a workgroup's waves take interleaved residues; each residue's force
// is one wave's work). The neighbour pairs of a chunk of 64 of them (one per
// lane) form a
// flat list in the reference's order: residues in order, the 27 cells in the
// reference's dx, dy, dz order, each cell's points in sorted order. Each lane
// evaluates list positions lane, lane + 64, ...: the cell a position belongs
// to is the last cell start at or before it (a bit scan over the head bits),
// so no lane walks a residue's list. In-predicate terms are compacted in list
// order; then each residue adds its terms one by one, so every force is the
// reference's f32 sum in the reference's order. A chunk whose list exceeds
// `cap` is split into batches of whole residues; a residue whose list alone
// exceeds it is summed by one lane in place.
__device__ inline void cg_forces(const RefTables& T, const GridView& grid, const CgCloud& cg,
                                 const uint8_t* aa, const float* X, const float* Y,
                                 const float* Z, uint32_t nr, uint32_t r_first, uint32_t r_step,
                                 uint32_t lane, float* F, const CgScratch& s) {
  const uint32_t nmine = nr > r_first ? (nr - r_first + r_step - 1) / r_step : 0;
  auto res = [&](uint32_t m) { return r_first + m * r_step; };
  for (uint32_t cb = 0; cb < nmine; cb += kWave) {
    const uint32_t nrc = min(nmine - cb, static_cast<uint32_t>(kWave));
    const bool act = lane < nrc;
    // this lane's residue: its 27 cell ranges, kept in registers
    uint32_t cbeg[27], ccnt[27];
    uint32_t n = 0;
#pragma unroll
    for (int c = 0; c < 27; ++c) ccnt[c] = cbeg[c] = 0;
    if (act) {
      const uint32_t r = res(cb + lane), t = r * kSlots + 1, ty = aa[r];
      const float px = X[t], py = Y[t], pz = Z[t];
      s.RQ[lane] = make_float4(px, py, pz, T.cg_q[ty]);
      s.RP[lane] = make_float4(T.cg_eps[ty], T.cg_sig[ty], static_cast<float>(T.cg_hyd[ty]), 0.0f);
      if (grid.n) {
        // cell_range of the 27 neighbours (dx, dy, dz order), with the key
        // range test and Morton spread done once per axis offset
        const int32_t k[3] = {cell_key(px, grid.inv), cell_key(py, grid.inv),
                              cell_key(pz, grid.inv)};
        uint32_t sp[3][3];
        bool ok[3][3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int32_t kk = wrap_add(k[ax], d - 1);
            ok[ax][d] = kk >= grid.kmin[ax] && kk <= grid.kmax[ax];
            sp[ax][d] = spread3(static_cast<uint32_t>(kk) - static_cast<uint32_t>(grid.kmin[ax])) << ax;
          }
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
          for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dz = 0; dz < 3; ++dz) {
              const int c = dx * 9 + dy * 3 + dz;
              if (ok[0][dx] & ok[1][dy] & ok[2][dz]) {
                const uint32_t code = sp[0][dx] | sp[1][dy] | sp[2][dz];
                const uint32_t b = grid.cell_start[code], e = grid.cell_start[code + 1];
                if (e > b) {
                  cbeg[c] = b;
                  ccnt[c] = e - b;
                }
              }
              n += ccnt[c];
            }
      }
    }
    const uint32_t incl = wave_incl_scan(n, lane), excl = incl - n;
    wave_sync();  // RQ, RP
    for (uint32_t ls = 0; ls < nrc;) {
      // the batch [ls, le): the residues from ls whose lists fit together
      const uint32_t base = __shfl(excl, static_cast<int>(ls));
      const uint32_t le = static_cast<uint32_t>(
          __popcll(ballot64(lane < ls || (act && incl - base <= s.cap))));
      if (le == ls) {  // one residue over the buffer: its own lane, in place
        if (lane == 0) {
          const uint32_t r = res(cb + ls), t = r * kSlots + 1, k = aa[r];
          cg_force_point(grid, cg, X[t], Y[t], Z[t], T.cg_q[k], T.cg_eps[k], T.cg_sig[k],
                         T.cg_hyd[k], F[3 * r], F[3 * r + 1], F[3 * r + 2]);
        }
        wave_sync();
        ++ls;
        continue;
      }
      const uint32_t tot = __shfl(incl, static_cast<int>(le - 1)) - base;
      for (uint32_t w = lane; w < (tot + 31) / 32; w += kWave) s.HM[w] = 0u;
      wave_sync();
      const bool mine = lane >= ls && lane < le;
      if (mine && n) {
        uint32_t k0 = excl - base;
        uint32_t first = kEntryFirst;
#pragma unroll
        for (int c = 0; c < 27; ++c)
          if (ccnt[c]) {
            s.H[k0] = (cbeg[c] + s.cap - k0) | (lane << 25) | first;
            atomicOr(&s.HM[k0 >> 5], 1u << (k0 & 31u));
            first = 0u;
            k0 += ccnt[c];
          }
      }
      wave_sync();
      uint32_t cnt = 0, carry = 0;
      for (uint32_t kb = 0; kb < tot; kb += kWave) {
        const uint32_t k = kb + lane;
        const bool live = k < tot;
        const uint64_t heads =
            (static_cast<uint64_t>(s.HM[kb >> 5]) | (static_cast<uint64_t>(s.HM[(kb >> 5) + 1]) << 32)) &
            ballot64(live);
        const uint64_t upto = heads & ((2ull << lane) - 1ull);  // heads at lanes <= this one
        const uint32_t hk = s.H[live ? k : kb];
        const uint32_t hv0 = __shfl(hk, upto ? 63 - __builtin_clzll(upto) : 0);
        const uint32_t hv = upto ? hv0 : carry;
        carry = __shfl(hv, 63);
        float tx = 0.0f, ty = 0.0f, tz = 0.0f;
        bool in = false;
        const uint32_t l = (hv >> 25) & 63u;
        if (live) {
          const uint32_t q = (hv & kEntryOffMask) + k - s.cap;
          if (q < grid.n) {
            const float4 rq = s.RQ[l], rp = s.RP[l];
            in = cg_term(cg, grid.sorted_idx[q], rq.x, rq.y, rq.z, rq.w, rp.x, rp.y,
                         static_cast<uint8_t>(rp.z), tx, ty, tz);
          }
        }
        const uint64_t inb = ballot64(in);
        const uint32_t pos = cnt + static_cast<uint32_t>(__popcll(inb & ((1ull << lane) - 1ull)));
        if (in) s.CT[pos] = make_float4(tx, ty, tz, 0.0f);
        if (live && ((heads >> lane) & 1ull) && (hk & kEntryFirst)) s.CS[l] = pos;
        cnt += static_cast<uint32_t>(__popcll(inb));
      }
      wave_sync();
      // each residue's terms [CS[l], next residue's CS) added in list order
      const bool has = mine && n;
      const uint32_t st = has ? s.CS[lane] : 0u;
      const uint64_t later = ballot64(has) & ~((2ull << lane) - 1ull);
      const uint32_t stn = __shfl(st, later ? __builtin_ctzll(later) : 0);
      const uint32_t en = has ? (later ? stn : cnt) : st;
      if (mine) {
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll 4
        for (uint32_t k = st; k < en; ++k) {
          const float4 v = s.CT[k];
          ax += v.x;
          ay += v.y;
          az += v.z;
        }
        const uint32_t r = res(cb + lane);
        F[3 * r] = ax;
        F[3 * r + 1] = ay;
        F[3 * r + 2] = az;
      }
      wave_sync();
      ls = le;
    }
  }
}

// kLds: the antigen Calpha set and grid live in LDS for the whole launch
// (their addresses then resolve to LDS at compile time; a pointer that may
// be either would turn every access into a flat one). NW: waves per
// candidate. Loops over residues, slots and antigen Calphas are split over
// the workgroup's threads; the force pass and the interface mask give each
// wave interleaved residues; the move decisions are scalar code every wave
// runs alike on the same LDS state, and one thread writes their outcome.
template <bool kLds, int NW>
__global__ void __launch_bounds__(NW * 64) k_refine(RefineArgs a) {
  extern __shared__ float lds_all[];
  constexpr uint32_t kThreads = NW * kWave;
  const uint32_t c = blockIdx.x;
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  if (c >= a.n_cand) return;
  // the residue tables first (read throughout: NERF, energies), then the rest
  constexpr uint32_t kTabWords = (sizeof(RefTables) + 15) / 16 * 4;
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.tab);
    uint32_t* dst = reinterpret_cast<uint32_t*>(lds_all);
    for (uint32_t k = tid; k < sizeof(RefTables) / 4; k += kThreads) dst[k] = src[k];
  }
  const RefTables& T = *reinterpret_cast<const RefTables*>(lds_all);
  float* lds = lds_all + kTabWords;
  const uint32_t g0 = uniform_u32(a.cand_res_off[c]),
                 nr = uniform_u32(a.cand_res_off[c + 1]) - g0;
  const uint32_t ms = a.max_res * kSlots;
  // LDS: the antigen Calpha set, its residue-level parameters and grid (when
  // they fit, n_ag_lds = n_ag), then the candidate's slots, forces, the
  // step's random draws and per-residue bytes.
  const uint32_t na_l = kLds ? a.n_ag : 0, nc_l = kLds ? a.ncells_lds : 0;
  float4* gpar = reinterpret_cast<float4*>(lds);  // 16-byte entries first
  float4* U = gpar + na_l;                         // 2 * max_res draws
  float4* CT = U + 2 * a.max_res;                  // pair_cap terms, a slice per wave
  float4* RQ = CT + a.pair_cap;                    // 64 rows per wave
  float4* RP = RQ + NW * kWave;
  double* red = reinterpret_cast<double*>(RP + NW * kWave);  // 16: the block sums
  const uint32_t wcap = a.pair_cap / NW;           // each wave's list
  CgScratch cs;
  cs.cap = wcap;
  cs.CT = CT + wv * wcap;
  cs.RQ = RQ + wv * kWave;
  cs.RP = RP + wv * kWave;
  // outside the force pass: the undo build's slots (in RQ) and the Cys list
  // (in RP); the undo is built by another wave's thread when there is one
  float* SX = reinterpret_cast<float*>(RQ);
  float* SY = SX + kSlots;
  float* SZ = SY + kSlots;
  uint16_t* cysl = reinterpret_cast<uint16_t*>(RP);
  constexpr uint32_t kUndoThread = NW > 1 ? kWave : 4;  // first lane of the undo's quad
  double* EC = red + 16;                           // 2 * max_res move energies
  float* gx = reinterpret_cast<float*>(EC + 2 * a.max_res);
  float* gy = gx + na_l;
  float* gz = gy + na_l;
  float* gq = gz + na_l;
  float* ge = gq + na_l;
  float* gs = ge + na_l;
  uint32_t* gcs = reinterpret_cast<uint32_t*>(gs + na_l);
  uint32_t* gsi = gcs + nc_l;
  float* X = reinterpret_cast<float*>(gsi + na_l);
  float* Y = X + ms;
  float* Z = Y + ms;
  float* F = Z + ms;                 // 3 * max_res forces
  uint32_t* H = reinterpret_cast<uint32_t*>(F + 3 * a.max_res);
  uint32_t* HM = H + a.pair_cap;
  uint32_t* CS = HM + a.pair_cap / 32;
  cs.H = H + wv * wcap;
  cs.HM = HM + wv * (wcap / 32);
  cs.CS = CS + wv * kWave;
  uint8_t* gh = reinterpret_cast<uint8_t*>(CS + NW * kWave);
  uint8_t* aa = gh + na_l;
  uint8_t* rot = aa + a.max_res;
  uint8_t* iface = rot + a.max_res;
  uint8_t* pend = iface + a.max_res;  // deferred build of each residue
  if (kLds) {
    for (uint32_t i = tid; i < a.n_ag; i += kThreads) {
      gpar[i] = a.ag_par[i];
      gx[i] = a.ag_x[i];
      gy[i] = a.ag_y[i];
      gz[i] = a.ag_z[i];
      gq[i] = a.ag_cg.q[i];
      ge[i] = a.ag_cg.eps[i];
      gs[i] = a.ag_cg.sig[i];
      gh[i] = a.ag_cg.hyd[i];
      gsi[i] = a.ag_grid.sorted_idx[i];
    }
    for (uint32_t k = tid; k < nc_l; k += kThreads) gcs[k] = a.ag_grid.cell_start[k];
  }
  const CaSet g = kLds ? CaSet{gx, gy, gz, gpar, a.n_ag} : CaSet{a.ag_x, a.ag_y, a.ag_z, a.ag_par, a.n_ag};
  const CgCloud cg = kLds ? CgCloud{gx, gy, gz, gq, ge, gs, gh} : a.ag_cg;
  GridView grid = a.ag_grid;
  if (kLds) {
    grid.cell_start = gcs;
    grid.sorted_idx = gsi;
  }
  for (uint32_t t = tid; t < nr * kSlots; t += kThreads) {
    const float4 v = a.slots[static_cast<size_t>(g0) * kSlots + t];
    X[t] = v.x;
    Y[t] = v.y;
    Z[t] = v.z;
  }
  for (uint32_t r = tid; r < nr; r += kThreads) {
    aa[r] = a.res_aa[g0 + r];
    rot[r] = a.res_rot[g0 + r];
    pend[r] = kBuildNone;
  }
  __syncthreads();

  // each wave's pair list of the filtered contributions reuses its slice of
  // the force pass's term list (free outside it); the candidate's Cys count
  // (uniform)
  uint2* lbuf = reinterpret_cast<uint2*>(cs.CT);
  const uint32_t lcap = 2 * wcap;
  uint32_t ncys = 0;
  for (uint32_t cb = 0; cb < nr; cb += kWave)
    ncys += static_cast<uint32_t>(__popcll(ballot64(cb + lane < nr && aa[cb + lane] == kCys)));

  for (uint32_t s = 0; s < a.n_steps; ++s) {
    const uint32_t it = a.iter0 + s;
    const float temp = a.temp[s];
    // 1. the gradient step (src/diffusion.rs:305-339): residue-level forces
    //    on the Calphas (compute_forces_with_grid)
    cg_forces(T, grid, cg, aa, X, Y, Z, nr, wv, NW, lane, F, cs);
    __syncthreads();
    for (uint32_t r = tid; r < nr; r += kThreads) {
      const uint32_t t = r * kSlots + 1;
      const float4 d = residue_delta(a.cx, a.cy, a.cz, a.sp, a.r0,
                                     make_float4(X[t], Y[t], Z[t], 0.0f), F[3 * r], F[3 * r + 1],
                                     F[3 * r + 2], a.sigma[s], it, g0 + r);
      F[3 * r] = d.x;
      F[3 * r + 1] = d.y;
      F[3 * r + 2] = d.z;
    }
    __syncthreads();
    // set_ca_pos: every slot moves by its residue's displacement (slots past
    // the residue's atoms are never read before a rebuild writes them)
    for (uint32_t q = tid; q < nr * kSlots; q += kThreads) {
      const uint32_t r = q / kSlots;
      X[q] += F[3 * r];
      Y[q] += F[3 * r + 1];
      Z[q] += F[3 * r + 2];
    }
    __syncthreads();
    // 2. interface mask (src/diffusion.rs:341-351): each wave takes the
    //    residues wv, wv + NW, ...; per chunk of 64 of them (Calphas in
    //    lanes, broadcast by readlane) each lane tests its antigen Calphas
    //    against every residue; the wave's compare masks set the residue bits
    const uint32_t nmine = nr > wv ? (nr - wv + NW - 1) / NW : 0;
    for (uint32_t cb = 0; cb < nmine; cb += kWave) {
      const uint32_t nrc = min(nmine - cb, static_cast<uint32_t>(kWave));
      float px = 0.0f, py = 0.0f, pz = 0.0f;
      if (lane < nrc) {
        const uint32_t t = (wv + (cb + lane) * NW) * kSlots + 1;
        px = X[t];
        py = Y[t];
        pz = Z[t];
      }
      uint64_t m = 0;
      for (uint32_t i0 = 0; i0 < g.n; i0 += kWave) {
        const uint32_t i = i0 + lane;
        const bool live = i < g.n;
        const float ax = live ? g.x[i] : 0.0f, ay = live ? g.y[i] : 0.0f,
                    az = live ? g.z[i] : 0.0f;
        for (uint32_t j0 = 0; j0 < nrc; j0 += 4) {
#pragma unroll
          for (uint32_t v = 0; v < 4; ++v) {
            const uint32_t j = j0 + v, jl = min(j, nrc - 1);
            const float dx = readlane_f32(px, jl) - ax, dy = readlane_f32(py, jl) - ay,
                        dz = readlane_f32(pz, jl) - az;
            if (ballot64(live & (dx * dx + dy * dy + dz * dz < a.iface_sq))) m |= 1ull << jl;
          }
        }
      }
      if (lane < nrc) iface[wv + (cb + lane) * NW] = static_cast<uint8_t>((m >> lane) & 1ull);
    }
    __syncthreads();
    // this step's draws of both move kinds, one residue per lane
    for (uint32_t r = tid; r < nr; r += kThreads) {
      U[2 * r] = u4(it, g0 + r, kTagMc, a.sp);
      U[2 * r + 1] = u4(it, g0 + r, kTagBb, a.sp);
    }
    __syncthreads();
    // 3. side-chain moves (src/diffusion.rs:353-398)
    //    (the residues whose first draw asks for a move, from a ballot per 64,
    //    visited in order)
    if (a.moves & DNAB_MC_SIDECHAIN) {
      // The energies of this step's mutations that involve no Cys, first and
      // all at once: each depends only on its residue's Calpha and two types,
      // which no decision of this phase changes (a residue moves once per
      // step). Proposal k is scored by wave k mod NW, its FP64 sums inside
      // that wave; EC holds (old type, new type) per residue.
      if (a.finite) {
        uint32_t k = 0;
        for (uint32_t cb = 0; cb < nr; cb += kWave) {
          bool mut = false;
          if (cb + lane < nr) {
            const uint32_t rl = cb + lane;
            const float4 ul = U[2 * rl];
            const bool in = iface[rl] != 0;
            const uint32_t na = min(static_cast<uint32_t>(ul.z * static_cast<float>(kNumAminoAcids)),
                                    static_cast<uint32_t>(kNumAminoAcids - 1));
            mut = ul.x < (in ? a.mut_p * 3.0f : a.mut_p * 0.4f) &&
                  !(ul.y < (in ? a.rot_p * 2.0f : a.rot_p)) && aa[rl] != kCys && na != kCys;
          }
          for (uint64_t mv = ballot64(mut); mv; mv &= mv - 1, ++k) {
            if (k % NW != wv) continue;
            const uint32_t r = cb + static_cast<uint32_t>(__builtin_ctzll(mv));
            const uint32_t oa = uniform_u32(aa[r]);
            const uint32_t na =
                min(static_cast<uint32_t>(uniform_f32(U[2 * r].z) * static_cast<float>(kNumAminoAcids)),
                    static_cast<uint32_t>(kNumAminoAcids - 1));
            double eo, en;
            contrib<true>(T, g, aa, X, Y, Z, r, r + 1, oa, na, lane, 0, 1, lbuf, lcap, eo, en);
            eo = wave_sum_f64(eo);
            en = wave_sum_f64(en);
            if (lane == 0) {
              EC[2 * r] = eo;
              EC[2 * r + 1] = en;
            }
          }
        }
        __syncthreads();
      }
      double ds_fast = 0.0;  // the disulfide term while no Cys move intervenes
      bool ds_known = false;
      for (uint32_t cb = 0; cb < nr; cb += kWave) {
        bool want = false;
        if (cb + lane < nr) {
          const bool in = iface[cb + lane] != 0;
          want = U[2 * (cb + lane)].x < (in ? a.mut_p * 3.0f : a.mut_p * 0.4f);
        }
        for (uint64_t mv = ballot64(want); mv; mv &= mv - 1) {
          const uint32_t r = cb + static_cast<uint32_t>(__builtin_ctzll(mv));
          const float4 uu = U[2 * r];
          const float4 u = make_float4(uniform_f32(uu.x), uniform_f32(uu.y), uniform_f32(uu.z),
                                       uniform_f32(uu.w));
          const bool in = uniform_u32(iface[r]) != 0;
          const float eff_rot = in ? a.rot_p * 2.0f : a.rot_p;
          const uint32_t t = r * kSlots + 1;
          const uint32_t old_aa = uniform_u32(aa[r]), old_rot = uniform_u32(rot[r]);
          const bool rot_move = u.y < eff_rot;
          uint32_t new_aa = old_aa, new_rot;
          if (rot_move) {  // rotamer move
            const uint32_t nrot = uniform_u32(T.nrot[old_aa]);
            if (nrot == 0) continue;
            new_rot = min(static_cast<uint32_t>(u.z * static_cast<float>(nrot)), nrot - 1);
          } else {  // mutation: the new type's first rotamer
            new_aa = min(static_cast<uint32_t>(u.z * static_cast<float>(kNumAminoAcids)),
                         static_cast<uint32_t>(kNumAminoAcids - 1));
            new_rot = 0;
          }
          if (a.finite && old_aa != kCys && new_aa != kCys) {
            // Neither the Calpha nor any Cys SG moves, so the reference's
            // energies before and after differ only in residue r's type: a
            // rotamer move scores the same finite energy twice (delta 0,
            // accepted) and a mutation is decided on both types' energies
            // before anything is rebuilt; then the outcome the reference ends
            // with is built once. (From finite coordinates every step stays
            // finite; a non-finite input takes the literal path below.)
            bool accept = true;
            if (!rot_move) {
              const double e_old = EC[2 * r], e_new = EC[2 * r + 1];
              if (!ds_known) {
                ds_fast = ncys >= 2 ? disulfide(aa, X, Y, Z, nr, lane, cysl) : 0.0;
                ds_known = true;
              }
              const double delta = (e_new + ds_fast) - (e_old + ds_fast);
              accept = delta <= 0.0 || static_cast<double>(u.w) < exp(-delta / temp);
            }
            // The build itself waits for the end of the phase: until then
            // only this residue's Calpha (unchanged by any build) and type are
            // read, and no other move touches it this step. No barrier per
            // move: what thread 0 writes here (pend, aa, rot) is published to
            // the other waves by the barrier that closes the phase, before the
            // build loop reads it; no write here changes which residues are
            // Cys.
            if (tid == 0) {
              if (rot_move) {  // apply_rotamer: the new chi on the current backbone
                rot[r] = static_cast<uint8_t>(new_rot);
                pend[r] = kBuildSide;
              } else if (accept) {  // mutate_residue to (new_aa, rotamer 0)
                pend[r] = new_aa == old_aa ? kBuildSide : kBuildIdeal;
                aa[r] = static_cast<uint8_t>(new_aa);
                rot[r] = 0;
              } else {  // undone: the old type and rotamer, rebuilt as the reference does
                pend[r] = new_aa == old_aa ? kBuildSide : kBuildIdeal;
              }
            }
            continue;
          }
          // A Cys gains, loses or moves its SG: the disulfide term needs the
          // built residue. The Calpha does not move, so both types' energies
          // come from one pass first; thread 0 builds the move while another
          // wave's thread builds the reference's undo into scratch (SX/SY/SZ),
          // from a copy of the residue taken before either starts.
          double c_old, c_new;
          contribution2<NW>(T, g, old_aa, new_aa, X[t], Y[t], Z[t], lane, wv, red, c_old, c_new);
          const double old_e = c_old + (ncys >= 2 ? disulfide(aa, X, Y, Z, nr, lane, cysl) : 0.0);
          for (uint32_t q = tid; q < kSlots; q += kThreads) {
            SX[q] = X[r * kSlots + q];
            SY[q] = Y[r * kSlots + q];
            SZ[q] = Z[r * kSlots + q];
          }
          __syncthreads();
          if (tid < 4) {  // the move, on quad 0
            const uint32_t q = tid;
            float* P = (q == 1 ? Y : q == 2 ? Z : X) + r * kSlots;
            if (rot_move || new_aa == old_aa)  // apply_rotamer, or mutate_residue to the same type
              q_side_chain(T, old_aa, new_rot, old_aa == kGly ? 4 : 5, P, q);
            else  // mutate_residue to another type
              q_build_ideal(T, new_aa, new_rot, P, q);
            if (q == 0) {
              aa[r] = static_cast<uint8_t>(new_aa);
              rot[r] = static_cast<uint8_t>(new_rot);
            }
          }
          if (tid >= kUndoThread && tid < kUndoThread + 4) {  // the undo, on another quad
            const uint32_t q = tid - kUndoThread;
            float* P = q == 1 ? SY : q == 2 ? SZ : SX;
            if (rot_move || new_aa == old_aa)  // the old chi on the same backbone
              q_side_chain(T, old_aa, old_rot, old_aa == kGly ? 4 : 5, P, q);
            else  // mutate_residue back: the old type on an ideal backbone
              q_build_ideal(T, old_aa, old_rot, P, q);
          }
          __syncthreads();
          const uint32_t ncys_new = ncys + (uniform_u32(aa[r]) == kCys) - (old_aa == kCys);
          const double new_e =
              c_new + (ncys_new >= 2 ? disulfide(aa, X, Y, Z, nr, lane, cysl) : 0.0);
          const double delta = new_e - old_e;
          const bool accept = delta <= 0.0 || static_cast<double>(u.w) < exp(-delta / temp);
          __syncthreads();
          ds_known = false;
          if (accept) {
            ncys = ncys_new;
          } else {
            for (uint32_t q = tid; q < kSlots; q += kThreads) {
              X[r * kSlots + q] = SX[q];
              Y[r * kSlots + q] = SY[q];
              Z[r * kSlots + q] = SZ[q];
            }
            if (tid == 0) {
              aa[r] = static_cast<uint8_t>(old_aa);
              rot[r] = static_cast<uint8_t>(old_rot);
            }
          }
          __syncthreads();
        }
      }
    }
    // Close the phase: thread 0's pend/aa/rot writes of the fast moves become
    // visible to every wave before the build loop reads them.
    __syncthreads();
    // the side chains the moves above decided, built in parallel, a quad of
    // lanes per residue
    for (uint32_t r = tid / 4; r < nr; r += kThreads / 4) {
      const uint32_t q = tid & 3u;
      float* P = (q == 1 ? Y : q == 2 ? Z : X) + r * kSlots;
      const uint8_t b = pend[r];
      const uint32_t ra = aa[r], rr = rot[r];
      if (b == kBuildSide)
        q_side_chain(T, ra, rr, ra == kGly ? 4 : 5, P, q);
      else if (b == kBuildIdeal)
        q_build_ideal(T, ra, rr, P, q);
    }
    __syncthreads();
    for (uint32_t r = tid; r < nr; r += kThreads) pend[r] = kBuildNone;
    // 4. torsion moves (src/diffusion.rs:400-415)
    if (a.moves & DNAB_MC_BACKBONE) {
      for (uint32_t cb = 0; cb < nr; cb += kWave) {
        const bool want = cb + lane < nr && U[2 * (cb + lane) + 1].x < a.bb_p;
        for (uint64_t mv = ballot64(want); mv; mv &= mv - 1) {
          const uint32_t r = cb + static_cast<uint32_t>(__builtin_ctzll(mv));
          const float4 uu = U[2 * r + 1];
          const float4 u = make_float4(uu.x, uniform_f32(uu.y), uniform_f32(uu.z), uniform_f32(uu.w));
          const float deg = -a.max_torsion + (a.max_torsion + a.max_torsion) * u.y;
          const float rad = geo::deg2rad(deg);
          const bool phi = u.z < 0.5f;
          const double old_e =
              chain_energy<NW>(T, g, aa, X, Y, Z, r, nr, ncys, lane, wv, lbuf, lcap, red, cysl);
          torsion(X, Y, Z, r, nr, rad, phi, tid, kThreads);
          const double new_e =
              chain_energy<NW>(T, g, aa, X, Y, Z, r, nr, ncys, lane, wv, lbuf, lcap, red, cysl);
          const double delta = new_e - old_e;
          const bool accept = delta <= 0.0 || static_cast<double>(u.w) < exp(-delta / temp);
          if (!accept) torsion(X, Y, Z, r, nr, -rad, phi, tid, kThreads);  // rotate back
        }
      }
    }
  }
  for (uint32_t t = tid; t < nr * kSlots; t += kThreads)
    a.slots[static_cast<size_t>(g0) * kSlots + t] = make_float4(X[t], Y[t], Z[t], 0.0f);
  for (uint32_t r = tid; r < nr; r += kThreads) {
    a.res_aa[g0 + r] = aa[r];
    a.res_rot[g0 + r] = rot[r];
  }
}

void fill_tables(RefTables& T) {
  std::memset(&T, 0, sizeof(T));
  for (int aa = 0; aa < kNumAminoAcids; ++aa) {
    const ResTopology& topo = kResidueTopology[aa];
    T.natoms[aa] = topo.n_atoms;
    T.ndef[aa] = kSidechainCount[aa];
    T.nrot[aa] = kRotamerCount[aa];
    for (int k = 0; k < kSidechainCount[aa]; ++k) {
      const BondDef& d = kSidechainDefs[aa][k];
      RefDef& o = T.defs[aa][k];
      o.a = d.ref_a;
      o.b = d.ref_b;
      o.c = d.ref_c;
      o.chi = d.chi_idx;
      o.bond = d.bond_len;
      // the host builder's sines and cosines (builder.cpp place_atom)
      const float ang = geo::deg2rad(d.angle_deg), dih = geo::deg2rad(d.dihedral_fixed);
      o.sa = std::sin(ang);
      o.ca = std::cos(ang);
      o.sd = std::sin(dih);
      o.cd = std::cos(dih);
    }
    for (int j = 0; j < kRotamerCount[aa] && j < 5; ++j)
      for (int k = 0; k < 4; ++k) {
        const float dih = geo::deg2rad(kRotamerLib[aa][j].chi[k]);
        T.chi_sc[aa][j][k][0] = std::sin(dih);
        T.chi_sc[aa][j][k][1] = std::cos(dih);
      }
    const ResAtomDef& cad = topo.atoms[topo.ca_idx];
    T.ca_rmh[aa] = kAmberNb[cad.atom_type].r_min_half;
    T.ca_eps[aa] = kAmberNb[cad.atom_type].epsilon;
    T.ca_q[aa] = cad.charge;
    T.ca_hyd[aa] = cad.hydrophobic;
    T.ca_type[aa] = cad.atom_type;
    uint8_t h = 0;
    (void)dnab_residue_cg_params(static_cast<uint8_t>(aa), &T.cg_q[aa], &T.cg_eps[aa],
                                 &T.cg_sig[aa], &h);
    T.cg_hyd[aa] = h;
  }
}

struct PopGuard {
  dnab_population* p = nullptr;
  ~PopGuard() { dnab_population_destroy(p); }
};

}  // namespace
}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_mc_default_params(dnab_mc_params* p) {
  if (!p) return DNAB_ERR_INVALID;
  p->mutation_p = 0.08f;       // src/diffusion.rs:41
  p->rotamer_move_p = 0.12f;   // src/diffusion.rs:43
  p->backbone_move_p = 0.04f;  // src/diffusion.rs:45
  p->max_torsion_deg = 15.0f;  // src/diffusion.rs:47
  p->iface_cutoff_sq = 64.0f;  // src/diffusion.rs:51
  p->moves = DNAB_MC_SIDECHAIN | DNAB_MC_BACKBONE;
  p->waves = 0;
  return DNAB_OK;
}

int dnab_refine(dnab_ctx* ctx, dnab_antigen* ag, const dnab_population_desc* ab,
                const uint8_t* res_rotamer, const dnab_langevin_params* lp,
                const dnab_mc_params* mc, dnab_protein_out* out, uint8_t* out_rotamer,
                double* out_energy, size_t* n_atoms_out) {
  if (!ctx || !ag || !ab || !lp || !mc || ag->ctx != ctx) return DNAB_ERR_INVALID;
  const size_t nc = ab->n_cand, nr = ab->n_res, na = ab->atoms.n;
  if (nc && !ab->cand_res_off) return DNAB_ERR_INVALID;
  if (nr && (!ab->res_atom_off || !ab->res_ca_atom || !ab->res_amino_acid))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: residues need offsets, Calpha and types");
  if (!ag->has_ca_aa || ag->n_ca == 0)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: the antigen needs its Calpha cloud and types");
  if (lp->schedule_total == 0 || !(lp->t_hot > 0.0f) || !(lp->t_cold > 0.0f))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: bad temperature schedule");
  if (nc >= 0xFFFFFF00ull || nr >= 0xFFFFFF00ull / kSlots)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "refine: too large");
  if (nc && (ab->cand_res_off[0] != 0 || ab->cand_res_off[nc] != nr))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: cand_res_off must span 0..n_res");
  if (nr && (ab->res_atom_off[0] != 0 || ab->res_atom_off[nr] != na))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: res_atom_off must span 0..n_atoms");
  if (na && (!ab->atoms.x || !ab->atoms.y || !ab->atoms.z)) return DNAB_ERR_INVALID;
  uint32_t max_res = 1;
  for (size_t c = 0; c < nc; ++c) {
    if (ab->cand_res_off[c + 1] < ab->cand_res_off[c])
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: cand_res_off not monotone");
    max_res = std::max(max_res, ab->cand_res_off[c + 1] - ab->cand_res_off[c]);
  }
  for (int t = 0; t < kNumAminoAcids; ++t)
    if (kSidechainCount[t] > kMaxDefs || kRotamerCount[t] > 5 || kResidueTopology[t].n_atoms > kSlots)
      return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "refine: residue tables exceed the slots");
  if (max_res > kMaxRefineRes)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "refine: more than %u residues in a candidate",
                          kMaxRefineRes);
  // every residue carries its topology's atoms in topology order (the slots)
  auto t_pack = std::chrono::steady_clock::now();
  std::vector<float4> slots(std::max<size_t>(nr * kSlots, 1), make_float4(0, 0, 0, 0));
  std::vector<uint8_t> raa(std::max<size_t>(nr, 1)), rrot(std::max<size_t>(nr, 1), 0);
  bool finite = true;
  for (size_t r = 0; r < nr; ++r) {
    const uint8_t t = ab->res_amino_acid[r];
    if (t >= kNumAminoAcids)
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: residue %zu type out of range", r);
    const uint32_t b = ab->res_atom_off[r], e = ab->res_atom_off[r + 1];
    if (e < b || e - b != kResidueTopology[t].n_atoms || ab->res_ca_atom[r] != b + 1)
      return dnab_set_error(ctx, DNAB_ERR_INVALID,
                            "refine: residue %zu does not hold its topology's atoms", r);
    raa[r] = t;
    if (res_rotamer) {
      if (kRotamerCount[t] ? res_rotamer[r] >= kRotamerCount[t] : res_rotamer[r] != 0)
        return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: residue %zu rotamer out of range", r);
      rrot[r] = res_rotamer[r];
    }
  }
  {
    std::atomic<bool> all_finite{true};
    parallel_for(nr, [&](size_t rb, size_t re) {
      bool ok = true;
      for (size_t r = rb; r < re; ++r) {
        const uint32_t b = ab->res_atom_off[r], e = ab->res_atom_off[r + 1];
        for (uint32_t k = 0; k < e - b; ++k) {
          const float x = ab->atoms.x[b + k], y = ab->atoms.y[b + k], z = ab->atoms.z[b + k];
          ok = ok && std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
          slots[r * kSlots + k] = make_float4(x, y, z, 0.0f);
        }
      }
      if (!ok) all_finite.store(false, std::memory_order_relaxed);
    });
    finite = all_finite.load();
  }
  timer_host(ctx, "refine_host_pack", ms_since(t_pack));
  if (nc == 0) {
    if (n_atoms_out) *n_atoms_out = 0;
    if (out && out->res_start) out->res_start[0] = 0;
    return DNAB_OK;
  }
  (void)hipSetDevice(ctx->device);
  // tables, antigen Calpha parameters, schedule
  RefTables tab;
  fill_tables(tab);
  const size_t n_ag = ag->n_ca;
  std::vector<uint8_t> ag_aa(n_ag);
  DNAB_RETURN_IF(download(ctx, ag_aa.data(), ag->ca_aa.p, n_ag));
  {
    std::vector<float> c(3 * n_ag);
    DNAB_RETURN_IF(download(ctx, c.data(), ag->ca_x.p, n_ag * sizeof(float)));
    DNAB_RETURN_IF(download(ctx, c.data() + n_ag, ag->ca_y.p, n_ag * sizeof(float)));
    DNAB_RETURN_IF(download(ctx, c.data() + 2 * n_ag, ag->ca_z.p, n_ag * sizeof(float)));
    for (float v : c) finite = finite && std::isfinite(v);
  }
  std::vector<float4> ag_par(n_ag);
  for (size_t i = 0; i < n_ag; ++i) {
    const uint8_t t = ag_aa[i] < kNumAminoAcids ? ag_aa[i] : 0;
    ag_par[i] = make_float4(tab.ca_rmh[t], tab.ca_eps[t], tab.ca_q[t],
                            static_cast<float>(tab.ca_hyd[t] | (tab.ca_type[t] << 8)));
  }
  const uint32_t n_steps = lp->n_steps;
  std::vector<float> temp(std::max<uint32_t>(n_steps, 1)), sigma(std::max<uint32_t>(n_steps, 1));
  for (uint32_t s = 0; s < n_steps; ++s) {
    // temperature (src/diffusion.rs:102-104) and the noise scale (src/diffusion.rs:315)
    const float it = static_cast<float>(s + lp->schedule_offset);
    temp[s] = lp->t_hot * std::pow(lp->t_cold / lp->t_hot, it / static_cast<float>(lp->schedule_total));
    sigma[s] = lp->with_noise ? lp->noise_base * std::sqrt(temp[s]) : 0.0f;
  }
  // the call's buffers in one allocation (one free at the end)
  DevBuf<RefTables> d_tab;
  DevBuf<float4> d_slots, d_agpar;
  DevBuf<uint8_t> d_aa, d_rot;
  DevBuf<uint32_t> d_off;
  DevBuf<float> d_temp, d_sigma;
  DevSlab slab;
  slab.carve(d_tab, 1);
  slab.carve(d_slots, slots.size());
  slab.carve(d_agpar, n_ag);
  slab.carve(d_aa, raa.size());
  slab.carve(d_rot, rrot.size());
  slab.carve(d_off, nc + 1);
  slab.carve(d_temp, temp.size());
  slab.carve(d_sigma, sigma.size());
  DNAB_ALLOC_CHECK(ctx, slab.alloc());
  DNAB_RETURN_IF(upload(ctx, d_tab.p, &tab, sizeof(tab)));
  DNAB_RETURN_IF(upload(ctx, d_slots.p, slots.data(), slots.size() * sizeof(float4)));
  DNAB_RETURN_IF(upload(ctx, d_agpar.p, ag_par.data(), n_ag * sizeof(float4)));
  DNAB_RETURN_IF(upload(ctx, d_aa.p, raa.data(), raa.size()));
  DNAB_RETURN_IF(upload(ctx, d_rot.p, rrot.data(), rrot.size()));
  DNAB_RETURN_IF(upload(ctx, d_off.p, ab->cand_res_off, (nc + 1) * sizeof(uint32_t)));
  DNAB_RETURN_IF(upload(ctx, d_temp.p, temp.data(), temp.size() * sizeof(float)));
  DNAB_RETURN_IF(upload(ctx, d_sigma.p, sigma.data(), sigma.size() * sizeof(float)));

  RefineArgs a{};
  a.tab = d_tab.p;
  a.ag_x = ag->ca_x.p;
  a.ag_y = ag->ca_y.p;
  a.ag_z = ag->ca_z.p;
  a.ag_par = d_agpar.p;
  a.n_ag = static_cast<uint32_t>(n_ag);
  a.ag_grid = make_view(ag->ca_grid);
  a.ag_cg = CgCloud{ag->ca_x.p, ag->ca_y.p, ag->ca_z.p, ag->ca_q.p, ag->ca_eps.p, ag->ca_sig.p,
                    ag->ca_hyd.p};
  a.cx = ag->center[0];
  a.cy = ag->center[1];
  a.cz = ag->center[2];
  a.r0 = ag->restraint_radius > 0.0f ? ag->restraint_radius : lp->restraint_radius;
  a.cand_res_off = d_off.p;
  a.res_aa = d_aa.p;
  a.res_rot = d_rot.p;
  a.slots = d_slots.p;
  a.n_cand = static_cast<uint32_t>(nc);
  a.max_res = max_res;
  a.temp = d_temp.p;
  a.sigma = d_sigma.p;
  a.n_steps = n_steps;
  a.iter0 = lp->schedule_offset;
  a.sp.step = lp->step;
  a.sp.max_disp = lp->max_disp;
  a.sp.restraint_k = lp->restraint_k;
  a.sp.restraint_radius = lp->restraint_radius;
  a.sp.with_noise = lp->with_noise;
  a.sp.force_model = DNAB_FORCE_CG;
  a.sp.seed_lo = static_cast<uint32_t>(lp->seed);
  a.sp.seed_hi = static_cast<uint32_t>(lp->seed >> 32);
  a.mut_p = mc->mutation_p;
  a.rot_p = mc->rotamer_move_p;
  a.bb_p = mc->backbone_move_p;
  a.max_torsion = mc->max_torsion_deg;
  a.iface_sq = mc->iface_cutoff_sq;
  a.moves = mc->moves;
  a.finite = finite ? 1u : 0u;
  // LDS: the tables, the candidate's rows, the force pass's pair list, and
  // the antigen Calpha set with its grid when that leaves room for a list of
  // 1024 pairs. Blocks resident per CU: as many as the batch needs to fill
  // the device once (up to 4), fewer when the list would drop below 1024.
  int lds_max = 0;
  DNAB_HIP_CHECK(ctx, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                            ctx->device));
  const size_t lds_cu = static_cast<size_t>(lds_max);
  if (mc->waves != 0 && mc->waves != 1 && mc->waves != 2 && mc->waves != 4 && mc->waves != 8)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "refine: waves must be 0, 1, 2, 4 or 8");
  const size_t cu = ctx->cu_count > 0 ? static_cast<size_t>(ctx->cu_count) : 256;
  // a lone wave per SIMD is latency-bound: small batches split each candidate
  // over eight or four waves; larger ones over two (two waves per SIMD, which
  // the kernel's ~210 VGPRs allow)
  const uint32_t nw = mc->waves ? mc->waves : (nc <= cu ? 8u : nc <= 2 * cu ? 4u : 2u);
  const size_t fixed = (sizeof(RefTables) + 15) / 16 * 16 +
                       (2 * kWave * nw + 8) * sizeof(float4) + kWave * nw * sizeof(uint32_t) +
                       static_cast<size_t>(max_res) * (2 * 16 + 2 * 8 + (3 * kSlots + 3) * sizeof(float) + 4);
  const size_t ncells = ag->ca_grid.n ? ag->ca_grid.ncodes + 1 : 0;
  const size_t ag_bytes = n_ag * (16 + 6 * 4 + 4 + 1) + ncells * 4;
  auto list_bytes = [](size_t cap) { return cap * (sizeof(float4) + sizeof(uint32_t)) + cap / 8; };
  if (n_ag >= (size_t(1) << 24))
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "refine: more than 2^24 antigen Calphas");
  size_t cap = 0;
  bool ag_in = false;
  for (size_t bpc = std::min<size_t>(4, (nc + cu - 1) / cu); bpc >= 1; --bpc) {
    const size_t budget = lds_cu / bpc;
    ag_in = ncells && fixed + ag_bytes + list_bytes(1024) <= budget;
    const size_t used = fixed + (ag_in ? ag_bytes : 0);
    const size_t unit = 64 * nw;  // each wave's list a multiple of 64
    cap = used < budget ? (budget - used) * 8 / 161 / unit * unit : 0;
    cap = std::min<size_t>(cap, 8192);
    if (cap >= 1024) break;
  }
  if (cap < 128 * nw)  // a wave's filtered contributions stage up to 4 x 64 pairs in 2 x its list
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "refine: candidates too long for LDS");
  a.pair_cap = static_cast<uint32_t>(cap);
  size_t lds = fixed + list_bytes(cap);
  if (ag_in) {
    a.n_ag_lds = static_cast<uint32_t>(n_ag);
    a.ncells_lds = static_cast<uint32_t>(ncells);
    lds += ag_bytes;
  }
  const void* kfn = nullptr;
  if (a.n_ag_lds)
    kfn = nw == 8 ? reinterpret_cast<const void*>(&k_refine<true, 8>)
        : nw == 4 ? reinterpret_cast<const void*>(&k_refine<true, 4>)
        : nw == 2 ? reinterpret_cast<const void*>(&k_refine<true, 2>)
                  : reinterpret_cast<const void*>(&k_refine<true, 1>);
  else
    kfn = nw == 8 ? reinterpret_cast<const void*>(&k_refine<false, 8>)
        : nw == 4 ? reinterpret_cast<const void*>(&k_refine<false, 4>)
        : nw == 2 ? reinterpret_cast<const void*>(&k_refine<false, 2>)
                  : reinterpret_cast<const void*>(&k_refine<false, 1>);
  if (lds > 65536)
    DNAB_HIP_CHECK(ctx, hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(lds)));
  hipEvent_t ev_end = nullptr;
  hipEvent_t ev = timer_begin(ctx, "refine", &ev_end);
  const dim3 grid_dim(static_cast<uint32_t>(nc)), block_dim(kWave * nw);
  if (a.n_ag_lds && nw == 8)
    hipLaunchKernelGGL((k_refine<true, 8>), grid_dim, block_dim, lds, ctx->stream, a);
  else if (a.n_ag_lds && nw == 4)
    hipLaunchKernelGGL((k_refine<true, 4>), grid_dim, block_dim, lds, ctx->stream, a);
  else if (a.n_ag_lds && nw == 2)
    hipLaunchKernelGGL((k_refine<true, 2>), grid_dim, block_dim, lds, ctx->stream, a);
  else if (a.n_ag_lds)
    hipLaunchKernelGGL((k_refine<true, 1>), grid_dim, block_dim, lds, ctx->stream, a);
  else if (nw == 8)
    hipLaunchKernelGGL((k_refine<false, 8>), grid_dim, block_dim, lds, ctx->stream, a);
  else if (nw == 4)
    hipLaunchKernelGGL((k_refine<false, 4>), grid_dim, block_dim, lds, ctx->stream, a);
  else if (nw == 2)
    hipLaunchKernelGGL((k_refine<false, 2>), grid_dim, block_dim, lds, ctx->stream, a);
  else
    hipLaunchKernelGGL((k_refine<false, 1>), grid_dim, block_dim, lds, ctx->stream, a);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  timer_end(ctx, "refine", ev, ev_end);
  DNAB_RETURN_IF(download(ctx, slots.data(), d_slots.p, slots.size() * sizeof(float4)));
  DNAB_RETURN_IF(download(ctx, raa.data(), d_aa.p, raa.size()));
  DNAB_RETURN_IF(download(ctx, rrot.data(), d_rot.p, rrot.size()));

  // the final proteins, residue by residue in topology order
  size_t n_out = 0;
  std::vector<uint32_t> res_off(nr + 1, 0);
  for (size_t r = 0; r < nr; ++r) {
    n_out += kResidueTopology[raa[r]].n_atoms;
    res_off[r + 1] = static_cast<uint32_t>(n_out);
  }
  if (n_atoms_out) *n_atoms_out = n_out;
  if (out_rotamer)
    for (size_t r = 0; r < nr; ++r) out_rotamer[r] = rrot[r];
  auto t_asm = std::chrono::steady_clock::now();
  const size_t no = std::max<size_t>(n_out, 1);
  std::unique_ptr<float[]> fx(new float[no]), fy(new float[no]), fz(new float[no]),
      fq(new float[no]), frm(new float[no]), fe(new float[no]);
  std::unique_ptr<uint8_t[]> fh(new uint8_t[no]), ft(new uint8_t[no]);
  std::unique_ptr<uint32_t[]> fca(new uint32_t[std::max<size_t>(nr, 1)]), fri(new uint32_t[no]);
  parallel_for(nc, [&](size_t cb, size_t ce) {
    for (size_t c = cb; c < ce; ++c)
      for (uint32_t r = ab->cand_res_off[c]; r < ab->cand_res_off[c + 1]; ++r) {
        const ResTopology& topo = kResidueTopology[raa[r]];
        fca[r] = res_off[r] + topo.ca_idx;
        for (uint32_t k = 0; k < topo.n_atoms; ++k) {
          const size_t o = res_off[r] + k;
          const float4 v = slots[r * kSlots + k];
          const ResAtomDef& d = topo.atoms[k];
          fx[o] = v.x;
          fy[o] = v.y;
          fz[o] = v.z;
          fq[o] = d.charge;
          frm[o] = kAmberNb[d.atom_type].r_min_half;
          fe[o] = kAmberNb[d.atom_type].epsilon;
          fh[o] = d.hydrophobic;
          ft[o] = d.atom_type;
          fri[o] = r - ab->cand_res_off[c];
        }
      }
  });
  timer_host(ctx, "refine_host_assemble", ms_since(t_asm));
  if (out_energy) {
    // allatom_energy of each refined candidate (src/diffusion.rs:1092)
    std::vector<uint32_t> cand_ag(nc, 0);
    dnab_population_desc pd{};
    pd.n_cand = nc;
    pd.n_res = nr;
    pd.cand_antigen = cand_ag.data();
    pd.cand_res_off = ab->cand_res_off;
    pd.res_atom_off = res_off.data();
    pd.res_ca_atom = fca.get();
    pd.res_amino_acid = raa.data();
    pd.atoms.n = n_out;
    pd.atoms.x = fx.get();
    pd.atoms.y = fy.get();
    pd.atoms.z = fz.get();
    pd.atoms.charge = fq.get();
    pd.atoms.r_min_half = frm.get();
    pd.atoms.epsilon = fe.get();
    pd.atoms.hydrophobic = fh.get();
    pd.atoms.atom_type = ft.get();
    PopGuard pop;
    dnab_antigen* ags[1] = {ag};
    DNAB_RETURN_IF(dnab_population_create(ctx, ags, 1, &pd, &pop.p));
    DNAB_RETURN_IF(dnab_population_evaluate(ctx, pop.p, DNAB_EVAL_BREAKDOWN));
    std::vector<double> terms(6 * nc);
    DNAB_RETURN_IF(dnab_population_read_terms(ctx, pop.p, terms.data()));
    for (size_t c = 0; c < nc; ++c) {
      const double* t = &terms[6 * c];
      out_energy[c] = ((((t[0] + t[1]) + t[2]) + t[3]) + t[4]) + t[5];
    }
  }
  if (!out) return DNAB_OK;
  if (!out->x || !out->y || !out->z) return DNAB_ERR_INVALID;
  if (n_out > out->atom_capacity) return DNAB_ERR_CAPACITY;
  std::memcpy(out->x, fx.get(), n_out * sizeof(float));
  std::memcpy(out->y, fy.get(), n_out * sizeof(float));
  std::memcpy(out->z, fz.get(), n_out * sizeof(float));
  if (out->charge) std::memcpy(out->charge, fq.get(), n_out * sizeof(float));
  if (out->r_min_half) std::memcpy(out->r_min_half, frm.get(), n_out * sizeof(float));
  if (out->epsilon) std::memcpy(out->epsilon, fe.get(), n_out * sizeof(float));
  if (out->hydrophobic) std::memcpy(out->hydrophobic, fh.get(), n_out);
  if (out->atom_type) std::memcpy(out->atom_type, ft.get(), n_out);
  if (out->residue_idx) std::memcpy(out->residue_idx, fri.get(), n_out * sizeof(uint32_t));
  if (out->res_start) std::memcpy(out->res_start, res_off.data(), (nr + 1) * sizeof(uint32_t));
  if (out->ca_atom_idx && nr) std::memcpy(out->ca_atom_idx, fca.get(), nr * sizeof(uint32_t));
  if (out->amino_acid && nr) std::memcpy(out->amino_acid, raa.data(), nr);
  return DNAB_OK;
}

}  // extern "C"
