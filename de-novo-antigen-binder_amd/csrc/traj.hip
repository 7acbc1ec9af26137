// Fused gradient-step trajectory: one workgroup per candidate runs every
// step of the loop (src/diffusion.rs:1018-1057) in a single launch.
//
// Candidates do not interact (each sees only the fixed antigen), so a
// candidate's trajectory needs no synchronisation with any other: its atoms
// stay in LDS for the whole run and a step costs workgroup barriers instead
// of six dependent kernel launches and device-wide passes.
//
// Per step, inside the workgroup:
//  1. tile boxes: the candidate's atoms in tiles of 64 (one atom per lane),
//     each tile's bounding box over the atoms that can have a pair (those
//     within the cutoff of the antigen's atom box) and the grid cells within
//     the cutoff of it;
//  2. pairs: work items of up to kTrajCells cells of one tile, taken from an
//     LDS counter by the waves; a wave walks the cells' chunks exactly as the
//     population pair kernels do (chunk_pairs, pair_device.h) and adds its
//     integer sums into the tile's LDS accumulators. Because every (atom,
//     chunk) sum is rounded once to 2^-20 and chunk sums add exactly, the
//     per-atom results are bit-identical to the population kernels' on the
//     loop's steps (whose pair terms are all FP32: a clash force moves its
//     residue by the clamped step whatever its rounding). A step whose
//     energies nothing reads (not a scoring step, not the evaluated final
//     state) forms forces only, by the same operations;
//  3. per-atom energies and forces from the integer sums;
//  4. the residue update (residue_delta, step_device.h), one thread per residue;
//  5. the rigid residue moves, one thread per atom.
// After the last step the positions are written back; a launch with no
// steps, whose clash pairs are FP64 (CLASH), evaluates the final state and
// writes the reported energies and forces, the candidate energy reduced in
// the order of the population reduction kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <type_traits>
#include <vector>

#include "device_common.h"
#include "dnab_internal.h"
#include "pair_device.h"
#include "step_device.h"
#include "host_parallel.h"

namespace dnab {

constexpr uint32_t kTrajMaxTiles = 64;    // tiles per workgroup (one ballot finds a tile)
constexpr uint32_t kTrajMaxCpw = 16;      // candidates per workgroup
constexpr uint16_t kTrajPad = 0xFFFFu;    // residue slot of a padding atom
constexpr uint32_t kTileChunks = 192;     // chunk list entries per tile
constexpr uint32_t kTileOverflow = 0xFFFFFFFFu;
// The fused loop lists a tile's chunks for its box widened by this much (A),
// and lists again when the box leaves the widened one.
constexpr float kListMargin = 1.0f;
constexpr uint32_t kTrajMaxAgAtoms = 1u << 26;  // a chunk entry holds (first atom << 6 | len - 1)
constexpr size_t kTrajLdsLimit = 160 * 1024;
constexpr float kTrajErrPerAbs = 32.0f / 16777216.0f;  // 32 ulp of f32 per unit of sum |e|

struct TrajTile {
  float lo[3], hi[3];
  float llo[3], lhi[3];  // the box the chunk list was made for
  int32_t k0[3];
  int32_t nx, ny, nz, nxy, ncell;
  uint32_t nchunk;  // listed chunks, or kTileOverflow
  uint32_t cand;    // candidate slot in the workgroup
};

struct TrajCand {
  float center[3];     // antigen centre
  float r_restr;       // restraint radius
  uint32_t id;         // candidate index
  uint32_t a0, n;      // global atom range
  uint32_t r0, nres;   // global residue range
  uint32_t a_off;      // first index of its atoms in group order (a multiple of 64
                       // unless packed; the slot of index i is s_perm[i])
  uint32_t r_off;      // first LDS residue slot
  uint32_t ag;         // antigen view
};

// The antigen fields a step reads, copied into LDS once per candidate slot.
struct TrajAg {
  const AgAtom* atoms;
  const uint32_t* cell_start;
  const float4* cell_box;
  float excl_lo[3], excl_hi[3];
  float rc, rc2, inv;
  int32_t kmin[3], kmax[3];
};

struct TrajArgs {
  const AgView* views;
  const uint32_t* cand_ag;
  const uint32_t* cand_atom_off;
  const uint32_t* cand_res_off;
  const uint32_t* res_atom_off;
  const uint32_t* res_ca;
  const uint32_t* atom_res;
  float4* pos;
  const float4* par;
  float4* force;
  double* e_atom;
  double* energy;
  float* err;
  const float* sigma_table;
  unsigned long long* stats;
  double* score_e;  // [scorings][n_cand]: candidate energy at each scoring, the final last
  // workgroup g runs candidates group_cand[group_off[g] .. group_off[g + 1]),
  // or, without a table, candidates g * cpw .. g * cpw + cpw - 1
  const uint32_t* group_off;
  const uint32_t* group_cand;
  uint32_t* work;   // optional: per candidate, antigen atoms taken into pair loops (all steps)
  uint32_t n_cand, n_steps, score_every;
  uint32_t s0;              // loop index of this launch's first step
  uint32_t n_total;         // steps of the whole loop (the final scoring's row)
  int32_t final_eval;       // evaluate the state after the last step: the zero-step launches
                            // that measure each candidate's work and that evaluate the
                            // loop's final state
  uint32_t n_groups;        // with a table
  int32_t cells;            // grid cells per work item (tiles whose chunk list overflowed)
  uint32_t chunks;          // listed chunks per work item (1..kLoadAhead)
  uint32_t cpw;             // candidates per workgroup (the most, with a table)
  uint32_t a_cap, r_cap;    // LDS capacity: atom slots (multiple of 64), residues
  int32_t packed;           // with a table: the group's atoms are packed (no padding per
                            // candidate) and bisected into compact tiles at the start
  StepParams sp;
};

// Bytes of dynamic LDS for a capacity of a_cap atom slots and r_cap residues:
// 5 integer sums, energy, position, parameters, force, status per atom slot; the
// displacement and the (atom range, Calpha, index, candidate) record per
// residue; the residue slot and the slot of each group-order index per atom
// slot; the chunk list of each tile (a_cap / 64 tiles).
__host__ __device__ inline size_t traj_lds_bytes(uint32_t a_cap, uint32_t r_cap) {
  const size_t res = (static_cast<size_t>(a_cap) * 4 + 3) & ~static_cast<size_t>(3);
  return static_cast<size_t>(a_cap) * (5 * 8 + 8 + 3 * 16 + 4) + static_cast<size_t>(r_cap) * 32 + res +
         static_cast<size_t>(a_cap / kWave) * kTileChunks * 4;
}
// Static LDS of a workgroup of `waves` waves (staged chunks, tiles,
// candidates, antigen fields, counters).
__host__ __device__ inline size_t traj_static_lds(int waves) {
  return static_cast<size_t>(waves) * kWave * 32 + kTrajMaxTiles * sizeof(TrajTile) +
         kTrajMaxCpw * (sizeof(TrajCand) + sizeof(TrajAg) + 4) + 256;
}

// Float keys that sort as unsigned integers, and back.
__device__ inline uint32_t float_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float key_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Compact tiles for n points (recursive bisection): every segment, a run of
// whole tiles of 64 slots, is cut after half of its tiles (rounded down)
// across the longest axis of its points' box, until each segment is one
// tile. A cut needs no sort: a 64-bin histogram along the axis gives the bin
// holding the cut; the points below it go left, those above right, and the
// bin's own points (within 1/64 of the box of the cut) fill the left side in
// slot order. Deterministic (counts and scans, no ordering by atomics).
// Returns ord or ord2, whichever holds the result: [p] (p < n) is the point
// in slot p. Scratch: the other of ord and ord2 (n each), pbin (n), hist (64
// per tile), and per tile seg (the first tile of its segment), bb (6: the
// segment's box) and cnt (9). Called by every thread.
template <int NT>
__device__ __attribute__((noinline)) uint16_t* traj_bisect(
    const float4* pts, uint32_t n, uint16_t* ord, uint16_t* ord2, uint8_t* pbin, uint32_t* hist,
    uint32_t* seg, uint32_t* bb, uint32_t* cnt, uint32_t* more, uint32_t tid) {
  constexpr uint32_t NB = 64, WAVES = NT / kWave;
  const uint32_t lane = tid & 63u, wave = tid / kWave;
  const uint32_t T = (n + kWave - 1) / kWave;
  // per segment start s: cut bin, points of it going left, left size, points
  // in the cut bin, points below it; per tile: points left of the cut bin,
  // in it, and the counts of both in earlier tiles of the segment
  uint32_t *c_bin = cnt, *c_r = cnt + T, *c_k = cnt + 2 * T, *c_nb = cnt + 3 * T,
           *c_below = cnt + 4 * T, *t_l = cnt + 5 * T, *t_b = cnt + 6 * T, *o_l = cnt + 7 * T,
           *o_b = cnt + 8 * T;
  for (uint32_t p = tid; p < n; p += NT) ord[p] = static_cast<uint16_t>(p);
  for (uint32_t t = tid; t < T; t += NT) seg[t] = 0;
  if (tid == 0) *more = T > 1 ? 1u : 0u;
  __syncthreads();
  // the histogram bin of the point in slot p along segment s's longest axis
  auto bin_of = [&](uint32_t p, uint32_t sg) -> uint32_t {
    const uint32_t* b = bb + 6 * sg;
    const float lx = key_float(b[0]), ly = key_float(b[1]), lz = key_float(b[2]);
    const float ex = key_float(b[3]) - lx, ey = key_float(b[4]) - ly, ez = key_float(b[5]) - lz;
    const float4 q = pts[ord[p]];
    float c, lo, e;
    if (ex >= ey && ex >= ez) {
      c = q.x; lo = lx; e = ex;
    } else if (ey >= ez) {
      c = q.y; lo = ly; e = ey;
    } else {
      c = q.z; lo = lz; e = ez;
    }
    const float f = e > 0.0f ? (c - lo) * (static_cast<float>(NB) / e) : 0.0f;
    return static_cast<uint32_t>(fminf(fmaxf(f, 0.0f), static_cast<float>(NB - 1)));
  };
  while (*more) {
    for (uint32_t t = tid; t < T; t += NT) {
      bb[6 * t + 0] = bb[6 * t + 1] = bb[6 * t + 2] = 0xFFFFFFFFu;
      bb[6 * t + 3] = bb[6 * t + 4] = bb[6 * t + 5] = 0u;
    }
    for (uint32_t i = tid; i < T * NB; i += NT) hist[i] = 0;
    __syncthreads();
    // boxes: a tile's (one segment's points) by wave reductions, then per tile
    for (uint32_t t = wave; t < T; t += WAVES) {
      const uint32_t p = t * kWave + lane;
      const bool live = p < n;
      const float4 q = pts[ord[live ? p : t * kWave]];
      const float inf = __builtin_inff();
      const float x0 = wave_reduce_f32<OpMin>(live ? q.x : inf);
      const float y0 = wave_reduce_f32<OpMin>(live ? q.y : inf);
      const float z0 = wave_reduce_f32<OpMin>(live ? q.z : inf);
      const float x1 = wave_reduce_f32<OpMax>(live ? q.x : -inf);
      const float y1 = wave_reduce_f32<OpMax>(live ? q.y : -inf);
      const float z1 = wave_reduce_f32<OpMax>(live ? q.z : -inf);
      if (lane == 0) {
        uint32_t* b = bb + 6 * seg[t];
        atomicMin(&b[0], float_key(x0));
        atomicMin(&b[1], float_key(y0));
        atomicMin(&b[2], float_key(z0));
        atomicMax(&b[3], float_key(x1));
        atomicMax(&b[4], float_key(y1));
        atomicMax(&b[5], float_key(z1));
      }
    }
    __syncthreads();
    // histograms: a tile's points (clustered in few bins) counted per
    // distinct bin, one atomic each
    for (uint32_t t = wave; t < T; t += WAVES) {
      const uint32_t p = t * kWave + lane, sg = seg[t];
      const bool live = p < n;
      const uint32_t bn = live ? bin_of(p, sg) : 0u;
      if (live) pbin[p] = static_cast<uint8_t>(bn);
      uint64_t rem = ballot64(live);
      while (rem) {
        const int l = __builtin_ctzll(rem);
        const uint32_t lb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bn), l));
        const uint64_t grp = ballot64(live && bn == lb);
        if (lane == static_cast<uint32_t>(l))
          atomicAdd(&hist[sg * NB + lb], static_cast<uint32_t>(__builtin_popcountll(grp)));
        rem &= ~grp;
      }
    }
    __syncthreads();
    // the cut of each segment of two or more tiles (one wave per segment)
    for (uint32_t sg = wave; sg < T; sg += WAVES) {
      if (seg[sg] != sg) continue;  // not a segment's first tile (wave-uniform)
      const uint64_t past = ballot64(lane < T && lane > sg && seg[lane] != sg);
      const uint32_t e = past ? static_cast<uint32_t>(__builtin_ctzll(past)) : T;
      if (e - sg < 2) {
        if (lane == 0) c_bin[sg] = NB;  // one tile: stays
        continue;
      }
      const uint32_t k = (e - sg) / 2 * kWave;
      const uint32_t h = hist[sg * NB + lane];
      uint32_t incl = h;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if (lane >= static_cast<uint32_t>(o)) incl += u;
      }
      const int b = __builtin_ctzll(ballot64(incl >= k));  // exists: k < the segment's points
      const uint32_t hb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(h), b));
      const uint32_t below = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), b)) - hb;
      if (lane == 0) {
        c_bin[sg] = static_cast<uint32_t>(b);
        c_r[sg] = k - below;
        c_k[sg] = k;
        c_nb[sg] = hb;
        c_below[sg] = below;
      }
    }
    __syncthreads();
    // per tile: its points below the cut bin and in it
    for (uint32_t t = wave; t < T; t += WAVES) {
      const uint32_t p = t * kWave + lane, sg = seg[t], cb = c_bin[sg];
      const bool live = p < n && cb < NB;
      const uint32_t bn = live ? pbin[p] : 0u;
      const uint64_t ml = ballot64(live && bn < cb), mb = ballot64(live && bn == cb);
      if (lane == 0) {
        t_l[t] = static_cast<uint32_t>(__builtin_popcountll(ml));
        t_b[t] = static_cast<uint32_t>(__builtin_popcountll(mb));
      }
    }
    __syncthreads();
    if (wave == 0) {
      // exclusive scans of the tile counts within each segment
      const uint32_t vl = lane < T ? t_l[lane] : 0u, vb = lane < T ? t_b[lane] : 0u;
      uint32_t il = vl, ib = vb;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t ul = __shfl_up(il, o, 64), ub = __shfl_up(ib, o, 64);
        if (lane >= static_cast<uint32_t>(o)) {
          il += ul;
          ib += ub;
        }
      }
      const uint32_t el = il - vl, eb = ib - vb;
      const int first = lane < T ? static_cast<int>(seg[lane]) : 0;
      const uint32_t sl = static_cast<uint32_t>(__shfl(static_cast<int>(el), first, 64));
      const uint32_t sb = static_cast<uint32_t>(__shfl(static_cast<int>(eb), first, 64));
      if (lane < T) {
        o_l[lane] = el - sl;
        o_b[lane] = eb - sb;
      }
    }
    __syncthreads();
    // new slots: left of the cut bin, the cut bin's first r, then its rest,
    // then the points right of it
    for (uint32_t t = wave; t < T; t += WAVES) {
      const uint32_t p = t * kWave + lane, sg = seg[t], cb = c_bin[sg];
      const bool live = p < n;
      uint32_t dest = p;
      const bool split = cb < NB;
      const uint32_t bn = live && split ? pbin[p] : 0u;
      const bool isl = live && split && bn < cb, isb = live && split && bn == cb;
      const uint64_t ml = ballot64(isl), mb = ballot64(isb);
      const uint32_t lr = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(ml >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ml), 0u));
      const uint32_t br = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mb >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mb), 0u));
      if (live && split) {
        const uint32_t base = sg * kWave, rl = o_l[t] + lr, rb = o_b[t] + br;
        const uint32_t k = c_k[sg], r = c_r[sg];
        if (isl) {
          dest = base + rl;
        } else if (isb) {
          dest = rb < r ? base + c_below[sg] + rb : base + k + (rb - r);
        } else {
          dest = base + k + (c_nb[sg] - r) + ((p - base) - rl - rb);
        }
      }
      if (live) ord2[dest] = ord[p];
    }
    __syncthreads();
    uint16_t* const t_ord = ord;  // (every thread swaps alike)
    ord = ord2;
    ord2 = t_ord;
    if (tid == 0) {
      uint32_t m = 0;
      for (uint32_t s0 = 0; s0 < T;) {
        uint32_t e = s0 + 1;
        while (e < T && seg[e] == s0) ++e;
        const uint32_t left = (e - s0) / 2;
        if (left) {
          for (uint32_t q = s0 + left; q < e; ++q) seg[q] = s0 + left;
          if (left > 1 || e - s0 - left > 1) m = 1;
        }
        s0 = e;
      }
      *more = m;
    }
    __syncthreads();
  }
  return ord;
}

__device__ inline bool traj_can_pair(const TrajAg& V, float4 p) {
  const float bx = fmaxf(fmaxf(V.excl_lo[0] - p.x, p.x - V.excl_hi[0]), 0.0f);
  const float by = fmaxf(fmaxf(V.excl_lo[1] - p.y, p.y - V.excl_hi[1]), 0.0f);
  const float bz = fmaxf(fmaxf(V.excl_lo[2] - p.z, p.z - V.excl_hi[2]), 0.0f);
  return ((bx * bx + by * by) + bz * bz) <= V.rc2;
}

// One workgroup runs the trajectories of g.cpw consecutive candidates: their
// atoms share the LDS slot space (each candidate padded to whole tiles) and
// all WAVES waves share the pair work of all of them, so a candidate in deep
// contact is worked on by every wave of its workgroup.
// CLASH: clash pairs in FP64, for the launch that evaluates the final state
// (no steps); the loop's launches keep every pair term in FP32.
template <bool COUNT, int WAVES, bool CLASH>
__global__ void __launch_bounds__(WAVES * 64)
k_trajectory(TrajArgs g) {
  extern __shared__ __align__(16) unsigned char s_dyn[];
  __shared__ float4 s_stage[WAVES][kWave][2];
  __shared__ TrajTile s_tile[kTrajMaxTiles];
  __shared__ TrajCand s_cand[kTrajMaxCpw];
  __shared__ TrajAg s_ag[kTrajMaxCpw];
  __shared__ uint32_t s_next, s_ncand, s_nslot, s_nres, s_norder, s_more;
  __shared__ uint32_t s_work[kTrajMaxCpw];

  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t AC = g.a_cap;
  long long* s_acc = reinterpret_cast<long long*>(s_dyn);    // [5][a_cap]
  double* s_e = reinterpret_cast<double*>(s_acc + 5 * AC);  // [a_cap]
  float4* s_pos = reinterpret_cast<float4*>(s_e + AC);      // [a_cap]
  float4* s_par = s_pos + AC;                               // [a_cap]
  float4* s_force = s_par + AC;                             // [a_cap]
  uint32_t* s_st = reinterpret_cast<uint32_t*>(s_force + AC);  // [a_cap] sticky status
  float4* s_delta = reinterpret_cast<float4*>(s_st + AC);      // [r_cap] (a_cap: a multiple of 64)
  uint4* s_rinfo = reinterpret_cast<uint4*>(s_delta + g.r_cap);       // [r_cap]
  uint16_t* s_res = reinterpret_cast<uint16_t*>(s_rinfo + g.r_cap);  // [a_cap]
  uint16_t* s_perm = s_res + AC;                                      // [a_cap] index -> slot
  uint32_t* s_desc = reinterpret_cast<uint32_t*>(s_dyn + ((reinterpret_cast<uintptr_t>(s_perm + AC) -
                                                             reinterpret_cast<uintptr_t>(s_dyn) + 3) & ~uintptr_t(3)));
  const bool packed = g.packed && g.group_off;

  if (tid == 0) {
    uint32_t first, nc;
    if (g.group_off) {
      first = g.group_off[blockIdx.x];
      nc = g.group_off[blockIdx.x + 1] - first;
    } else {
      first = blockIdx.x * g.cpw;
      nc = min(g.cpw, g.n_cand - first);
    }
    uint32_t a_off = 0, r_off = 0;
    for (uint32_t j = 0; j < nc; ++j) {
      const uint32_t c = g.group_off ? g.group_cand[first + j] : first + j;
      TrajCand C;
      C.id = c;
      C.a0 = g.cand_atom_off[c];
      C.n = g.cand_atom_off[c + 1] - C.a0;
      C.r0 = g.cand_res_off[c];
      C.nres = g.cand_res_off[c + 1] - C.r0;
      C.a_off = a_off;
      C.r_off = r_off;
      C.ag = g.cand_ag[c];
      const AgView& V = g.views[C.ag];
      C.center[0] = V.center[0];
      C.center[1] = V.center[1];
      C.center[2] = V.center[2];
      C.r_restr = V.restraint_radius > 0.0f ? V.restraint_radius : g.sp.restraint_radius;
      s_cand[j] = C;
      TrajAg& G = s_ag[j];
      G.atoms = V.atoms;
      G.cell_start = V.cell_start;
      G.cell_box = V.cell_box;
      for (int d = 0; d < 3; ++d) {
        G.excl_lo[d] = V.excl_lo[d];
        G.excl_hi[d] = V.excl_hi[d];
        G.kmin[d] = V.kmin[d];
        G.kmax[d] = V.kmax[d];
      }
      G.rc = V.rc;
      G.rc2 = V.rc2;
      G.inv = V.inv;
      s_work[j] = 0;
      const uint32_t slots = (C.n + kWave - 1) & ~(kWave - 1);
      for (uint32_t t = a_off / kWave; !packed && t < (a_off + slots) / kWave; ++t) s_tile[t].cand = j;
      a_off += packed ? C.n : slots;
      r_off += C.nres;
    }
    const uint32_t slots = (a_off + kWave - 1) & ~(kWave - 1);
    for (uint32_t t = 0; t < slots / kWave; ++t) {
      // no chunk list yet: no box lies inside an empty one
      for (int d = 0; d < 3; ++d) {
        s_tile[t].llo[d] = __builtin_inff();
        s_tile[t].lhi[d] = -__builtin_inff();
      }
      s_tile[t].nchunk = 0;
    }
    s_ncand = nc;
    s_norder = a_off;
    s_nslot = slots;
    s_nres = r_off;
  }
  __syncthreads();
  const uint32_t NC = s_ncand, A = s_nslot, R = s_nres, T = A / kWave;
  // the candidate holding group-order index o
  auto cand_of = [&](uint32_t o) {
    uint32_t j = 0;
    while (j + 1 < NC && s_cand[j + 1].a_off <= o) ++j;
    return j;
  };

  if (!packed) {
    // each candidate from a whole tile, atoms in order: slot = index
    for (uint32_t k = tid; k < A; k += WAVES * kWave) {
      const TrajCand C = s_cand[s_tile[k / kWave].cand];
      const uint32_t i = k - C.a_off;
      s_perm[k] = static_cast<uint16_t>(k);
      if (i < C.n) {
        s_pos[k] = g.pos[C.a0 + i];
        s_par[k] = g.par[C.a0 + i];
        s_res[k] = static_cast<uint16_t>(C.r_off + (g.atom_res[C.a0 + i] - C.r0));
      } else {
        s_pos[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        s_par[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        s_res[k] = kTrajPad;
      }
    }
  } else {
    // packed: positions in group order (in the force array, unused until
    // the first step), compact tiles by bisection, then the slots
    const uint32_t N = s_norder;
    for (uint32_t o = tid; o < N; o += WAVES * kWave) {
      const TrajCand& C = s_cand[cand_of(o)];
      s_force[o] = g.pos[C.a0 + (o - C.a_off)];
    }
    __syncthreads();
    // scratch in the sums (free until the first step: 4 bytes per slot for
    // ord and ord2, 4 for the histograms (256 per tile), 1 for the bins,
    // within their 40 per slot) and in the chunk lists (16 words per tile,
    // within their 192)
    const uint32_t TN = (N + kWave - 1) / kWave;
    uint16_t* const ord0 = reinterpret_cast<uint16_t*>(s_acc);
    uint32_t* const hist = reinterpret_cast<uint32_t*>(s_acc) + AC;
    uint8_t* const pbin = reinterpret_cast<uint8_t*>(s_acc) + 8 * static_cast<size_t>(AC);
    const uint16_t* ord = traj_bisect<WAVES * kWave>(s_force, N, ord0, ord0 + AC, pbin, hist, s_desc,
                                                     s_desc + TN, s_desc + 7 * TN, &s_more, tid);
    for (uint32_t k = tid; k < A; k += WAVES * kWave) {
      if (k < N) {
        const uint32_t o = ord[k], j = cand_of(o);
        const TrajCand& C = s_cand[j];
        const uint32_t i = o - C.a_off;
        s_pos[k] = s_force[o];
        s_par[k] = g.par[C.a0 + i];
        s_res[k] = static_cast<uint16_t>(C.r_off + (g.atom_res[C.a0 + i] - C.r0));
        s_perm[o] = static_cast<uint16_t>(k);
        if (k % kWave == 0) s_tile[k / kWave].cand = j;  // the tile's antigen (one per group)
      } else {
        s_pos[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        s_par[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        s_res[k] = kTrajPad;
      }
    }
  }
  __syncthreads();
  // per residue: its atoms' group-order range, Calpha slot, global index, candidate slot
  for (uint32_t r = tid; r < R; r += WAVES * kWave) {
    uint32_t j = 0;
    while (j + 1 < NC && s_cand[j + 1].r_off <= r) ++j;
    const TrajCand C = s_cand[j];
    const uint32_t rg = C.r0 + (r - C.r_off);
    const uint32_t b0 = C.a_off + (g.res_atom_off[rg] - C.a0);
    const uint32_t b1 = C.a_off + (g.res_atom_off[rg + 1] - C.a0);
    s_rinfo[r] = make_uint4(b0 | (b1 << 16), s_perm[C.a_off + (g.res_ca[rg] - C.a0)], rg, j);
  }
  __syncthreads();

  // candidate j's energy and sum |e| (one wave), in the order of the
  // population reduction kernel: lane-strided FP64 sums, then a DPP wave sum
  auto candidate_sums = [&](uint32_t j, double& se, double& sa) {
    const TrajCand C = s_cand[j];
    se = 0.0;
    sa = 0.0;
    for (uint32_t i = lane; i < C.n; i += kWave) {
      const uint32_t k = s_perm[C.a_off + i];
      se += s_e[k];
      sa += static_cast<double>(s_force[k].w);
    }
    se = wave_sum_f64(se);
    sa = wave_sum_f64(sa);
  };

  uint32_t n_in = 0, n_tested = 0;
  for (uint32_t s = 0;; ++s) {
    if (s == g.n_steps && !g.final_eval) break;  // a later launch continues the loop
    const uint32_t it = g.s0 + s;                // the loop's step index
    // this step's noise amplitude, loaded early (used in phase 4)
    const float sigma = (g.sp.with_noise && s < g.n_steps) ? g.sigma_table[it] : 0.0f;
    // this step's energies are read: a scoring step, or the evaluated final state
    const bool need_e = s == g.n_steps || (g.score_every && it % g.score_every == 0);
    // 1. tile boxes, the chunks of the grid cells within the cutoff of
    //    each box (listed in LDS), and cleared accumulators
    for (uint32_t t = wave; t < T; t += WAVES) {
      const uint32_t k = t * kWave + lane;
      const TrajAg& V = s_ag[s_tile[t].cand];
      const float4 P = s_pos[k];
      const bool valid = s_res[k] != kTrajPad && traj_can_pair(V, P);
      const float inf = __builtin_inff();
      const float lox = wave_reduce_f32<OpMin>(valid ? P.x : inf);
      const float loy = wave_reduce_f32<OpMin>(valid ? P.y : inf);
      const float loz = wave_reduce_f32<OpMin>(valid ? P.z : inf);
      const float hix = wave_reduce_f32<OpMax>(valid ? P.x : -inf);
      const float hiy = wave_reduce_f32<OpMax>(valid ? P.y : -inf);
      const float hiz = wave_reduce_f32<OpMax>(valid ? P.z : -inf);
      const bool any = ballot64(valid) != 0;
      // The chunk list is made for the tile box widened by kListMargin, and
      // holds the chunks of the cells whose atom box lies within rc of that
      // box. It stands while the tile box stays inside: an antigen atom
      // within rc of the tile box is then within rc of the listing box, and
      // so is its cell's box (the same rounded differences and sums, each
      // monotone in its operands), so the list holds every chunk with an atom
      // the pair phase stages. (The antigen does not move.)
      const TrajTile& tc = s_tile[t];
      const bool same = any && lox >= tc.llo[0] && loy >= tc.llo[1] && loz >= tc.llo[2] &&
                        hix <= tc.lhi[0] && hiy <= tc.lhi[1] && hiz <= tc.lhi[2];
      uint32_t* desc = s_desc + t * kTileChunks;
      uint32_t nt = 0;
      bool over = false;
      if (same) {
        nt = tc.nchunk;
        over = nt == kTileOverflow;
      } else {
        const float inf = __builtin_inff();
        const float llx = any ? lox - kListMargin : inf, lly = any ? loy - kListMargin : inf,
                    llz = any ? loz - kListMargin : inf;
        const float lhx = any ? hix + kListMargin : -inf, lhy = any ? hiy + kListMargin : -inf,
                    lhz = any ? hiz + kListMargin : -inf;
        // cells that can hold an antigen atom within rc of the listing box
        // (cell_key is monotone, so a slightly widened range is conservative)
        const float margin = V.rc * 1.001f + 1e-3f;
        const int32_t kx0 = max(cell_key(llx - margin, V.inv), V.kmin[0]);
        const int32_t ky0 = max(cell_key(lly - margin, V.inv), V.kmin[1]);
        const int32_t kz0 = max(cell_key(llz - margin, V.inv), V.kmin[2]);
        const int32_t kx1 = min(cell_key(lhx + margin, V.inv), V.kmax[0]);
        const int32_t ky1 = min(cell_key(lhy + margin, V.inv), V.kmax[1]);
        const int32_t kz1 = min(cell_key(lhz + margin, V.inv), V.kmax[2]);
        // (a tile without a valid atom has an empty box and no cells)
        const int32_t nx = any ? kx1 - kx0 + 1 : 0;
        const int32_t ny = any ? ky1 - ky0 + 1 : 0;
        const int32_t nz = any ? kz1 - kz0 + 1 : 0;
        const int32_t nxy = nx * ny;
        const int32_t ncell = (nx > 0 && ny > 0 && nz > 0) ? nxy * nz : 0;
        for (int32_t cb0 = 0; cb0 < ncell; cb0 += kWave) {
          // lane = a cell, in (kz, ky, kx) order
          const int32_t cl = cb0 + static_cast<int32_t>(lane);
          uint32_t cb = 0, ce = 0;
          if (cl < ncell) {
            const int32_t iz = cl / nxy, rem = cl - iz * nxy;
            const int32_t iy = rem / nx, ix = rem - iy * nx;
            const uint32_t code = morton3(static_cast<uint32_t>(kx0 + ix - V.kmin[0]),
                                          static_cast<uint32_t>(ky0 + iy - V.kmin[1]),
                                          static_cast<uint32_t>(kz0 + iz - V.kmin[2]));
            cb = V.cell_start[code];
            ce = V.cell_start[code + 1];
            const float4 clo = V.cell_box[2 * code], chi = V.cell_box[2 * code + 1];
            const float dx = fmaxf(fmaxf(llx - chi.x, clo.x - lhx), 0.0f);
            const float dy = fmaxf(fmaxf(lly - chi.y, clo.y - lhy), 0.0f);
            const float dz = fmaxf(fmaxf(llz - chi.z, clo.z - lhz), 0.0f);
            if (!(((dx * dx + dy * dy) + dz * dz) <= V.rc2)) ce = cb;  // (an empty cell fails too)
          }
          const uint32_t nch = (ce - cb + kWave - 1) / kWave;
          uint32_t incl = nch;
          for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o, 64);
            if (lane >= static_cast<uint32_t>(o)) incl += u;
          }
          const uint32_t tot = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
          if (!over && nt + tot <= kTileChunks) {
            const uint32_t off = nt + incl - nch;
            for (uint32_t m = 0; m < nch; ++m) {
              const uint32_t f = cb + kWave * m;
              desc[off + m] = (f << 6) | (min(static_cast<uint32_t>(kWave), ce - f) - 1u);
            }
          } else {
            over = true;  // the tile's chunks go through cell items
          }
          nt += tot;
        }
        if (lane == 0) {
          TrajTile& tb = s_tile[t];
          tb.llo[0] = llx; tb.llo[1] = lly; tb.llo[2] = llz;
          tb.lhi[0] = lhx; tb.lhi[1] = lhy; tb.lhi[2] = lhz;
          tb.k0[0] = kx0; tb.k0[1] = ky0; tb.k0[2] = kz0;
          tb.nx = nx;
          tb.ny = ny;
          tb.nz = nz;
          tb.nxy = nxy;
          tb.ncell = ncell;
          tb.nchunk = over ? kTileOverflow : nt;
        }
      }
      if (lane == 0) {
        TrajTile& tb = s_tile[t];
        tb.lo[0] = lox; tb.lo[1] = loy; tb.lo[2] = loz;
        tb.hi[0] = hix; tb.hi[1] = hiy; tb.hi[2] = hiz;
      }
      for (int q = 0; q < 5; ++q) s_acc[q * AC + k] = 0;
      s_st[k] = 0;
    }
    if (tid == 0) s_next = 0;
    __syncthreads();

    // 2. pairs: work items, tile-major: runs of g.chunks listed chunks of a
    //    tile, or runs of g.cells cells of a tile whose list overflowed
    {
      uint32_t nb = 0;
      if (lane < T) {
        const TrajTile& tl = s_tile[lane];
        nb = tl.nchunk != kTileOverflow ? (tl.nchunk + g.chunks - 1) / g.chunks
                                       : static_cast<uint32_t>((tl.ncell + g.cells - 1) / g.cells);
      }
      uint32_t incl = nb;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if (lane >= static_cast<uint32_t>(o)) incl += u;
      }
      const uint32_t first = incl - nb;  // first item of tile `lane`
      const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
      auto grab = [&]() -> uint32_t {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(&s_next, 1u);
        return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(i)));
      };

      // a step whose energies nothing reads (no scoring, not the evaluated
      // final state) forms forces only: the same forces, fewer operations
      auto run_items = [&](auto energy) {
        FixAcc acc;
        int32_t cur = -1;
        float xb = 0.f, yb = 0.f, zb = 0.f, qb = 0.f, rmhb = 0.f, sqeb = 0.f, sqe12b = 0.f, epsb = 0.f;
        bool hydb = false, valid = false;
        float rc2_lane = -1.0f;
        auto flush = [&]() {
          const uint32_t k = static_cast<uint32_t>(cur) * kWave + lane;
          // one flush per (wave, tile) and step: at most WAVES sums of <= 2^59
          if (acc.e | acc.a | acc.x | acc.y | acc.z) {
            if (decltype(energy)::value) {  // (forces only: acc.e and acc.a stay zero)
              atomicAdd(reinterpret_cast<unsigned long long*>(&s_acc[0 * AC + k]), static_cast<unsigned long long>(acc.e));
              atomicAdd(reinterpret_cast<unsigned long long*>(&s_acc[1 * AC + k]), static_cast<unsigned long long>(acc.a));
            }
            atomicAdd(reinterpret_cast<unsigned long long*>(&s_acc[2 * AC + k]), static_cast<unsigned long long>(acc.x));
            atomicAdd(reinterpret_cast<unsigned long long*>(&s_acc[3 * AC + k]), static_cast<unsigned long long>(acc.y));
            atomicAdd(reinterpret_cast<unsigned long long*>(&s_acc[4 * AC + k]), static_cast<unsigned long long>(acc.z));
          }
          if (acc.st) atomicOr(&s_st[k], acc.st);
          acc = FixAcc();
        };

        for (uint32_t i = grab(); i < total; i = grab()) {
          // the item's tile: the last tile whose first item is <= i
          const uint64_t m = ballot64(lane < T && first <= i);
          const int32_t t = 63 - static_cast<int32_t>(__builtin_clzll(m));
          const uint32_t b = i - static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(first), t));
          const TrajTile& tb = s_tile[t];
          const TrajAg& V = s_ag[tb.cand];
          if (t != cur) {
            if (cur >= 0) flush();
            cur = t;
            const uint32_t k = static_cast<uint32_t>(t) * kWave + lane;
            const float4 P = s_pos[k];
            const float4 Q = s_par[k];
            valid = s_res[k] != kTrajPad && traj_can_pair(V, P);
            xb = P.x; yb = P.y; zb = P.z; qb = P.w;
            rmhb = Q.x;
            sqeb = sqrtf(Q.y);
            sqe12b = 12.0f * sqeb;
            epsb = Q.y;
            hydb = Q.z != 0.0f;
            rc2_lane = valid ? V.rc2 : -1.0f;  // an atom without pairs takes none
          }
          const AgAtom* __restrict__ atoms = V.atoms;
          const float rc2 = V.rc2;
          const float lox = tb.lo[0], loy = tb.lo[1], loz = tb.lo[2];
          const float hix = tb.hi[0], hiy = tb.hi[1], hiz = tb.hi[2];
          // up to kLoadAhead chunks [cbv, cev): loads in flight together, then
          // the atoms near the tile box are staged and paired in chunk order
          auto run_chunks = [&](const uint32_t (&cbv)[kLoadAhead], const uint32_t (&cev)[kLoadAhead]) {
            float4 L0[kLoadAhead], L1[kLoadAhead];
            #pragma unroll
            for (int u = 0; u < kLoadAhead; ++u) {
              L0[u] = make_float4(0.f, 0.f, 0.f, 0.f);
              L1[u] = L0[u];
              if (cbv[u] + lane < cev[u]) {
                L0[u] = *reinterpret_cast<const float4*>(&atoms[cbv[u] + lane].x);
                L1[u] = *reinterpret_cast<const float4*>(&atoms[cbv[u] + lane].rmh);
              }
            }
            #pragma unroll
            for (int u = 0; u < kLoadAhead; ++u) {
              if (cev[u] <= cbv[u]) continue;
              const float bx = fmaxf(fmaxf(lox - L0[u].x, L0[u].x - hix), 0.0f);
              const float by = fmaxf(fmaxf(loy - L0[u].y, L0[u].y - hiy), 0.0f);
              const float bz = fmaxf(fmaxf(loz - L0[u].z, L0[u].z - hiz), 0.0f);
              const bool near = cbv[u] + lane < cev[u] && ((bx * bx + by * by) + bz * bz) <= rc2;
              const uint64_t mask = ballot64(near);
              const uint32_t n_s = static_cast<uint32_t>(__builtin_popcountll(mask));
              if (n_s == 0) continue;
              if (lane == 0 && g.work) atomicAdd(&s_work[tb.cand], n_s);
              stage_chunk(s_stage[wave], mask, lane, L0[u], L1[u]);
              chunk_pairs<true, COUNT, CLASH, decltype(energy)::value>(s_stage[wave], n_s, xb, yb, zb, qb, rmhb, sqeb, sqe12b, epsb,
                                       hydb, rc2_lane, valid, acc, n_in, n_tested);
            }
          };
          uint32_t cbv[kLoadAhead], cev[kLoadAhead];
          if (tb.nchunk != kTileOverflow) {
            // listed chunks b * g.chunks ...
            const uint32_t base = b * g.chunks;
            const uint32_t* desc = s_desc + t * kTileChunks;
            #pragma unroll
            for (int u = 0; u < kLoadAhead; ++u) {
              cbv[u] = 0;
              cev[u] = 0;
              if (static_cast<uint32_t>(u) < g.chunks && base + u < tb.nchunk) {
                const uint32_t d = desc[base + u];
                cbv[u] = d >> 6;
                cev[u] = cbv[u] + (d & 63u) + 1u;
              }
            }
            run_chunks(cbv, cev);
          } else {
            // cells b * g.cells ...: their chunks, kLoadAhead at a time
            const uint32_t* __restrict__ cstart = V.cell_start;
            const int32_t nx = tb.nx, nxy = tb.nxy;
            const int32_t cbeg = static_cast<int32_t>(b) * g.cells;
            const int32_t cend = min(cbeg + g.cells, tb.ncell);
            const int32_t cl = cbeg + static_cast<int32_t>(lane);
            uint32_t cb = 0, ce = 0;
            if (cl < cend) {
              const int32_t iz = cl / nxy, rem = cl - iz * nxy;
              const int32_t iy = rem / nx, ix = rem - iy * nx;
              const uint32_t code = morton3(static_cast<uint32_t>(tb.k0[0] + ix - V.kmin[0]),
                                            static_cast<uint32_t>(tb.k0[1] + iy - V.kmin[1]),
                                            static_cast<uint32_t>(tb.k0[2] + iz - V.kmin[2]));
              cb = cstart[code];
              ce = cstart[code + 1];
            }
            const uint32_t nch = (ce - cb + kWave - 1) / kWave;
            uint32_t cincl = nch;
            for (int o = 1; o < 64; o <<= 1) {
              const uint32_t u = __shfl_up(cincl, o, 64);
              if (lane >= static_cast<uint32_t>(o)) cincl += u;
            }
            const uint32_t off = cincl - nch;
            const uint32_t nchunk = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cincl), 63));
            for (uint32_t kc = 0; kc < nchunk; kc += kLoadAhead) {
              #pragma unroll
              for (int u = 0; u < kLoadAhead; ++u) {
                const uint32_t kk = kc + u;
                cbv[u] = 0;
                cev[u] = 0;
                if (kk < nchunk) {
                  // the cell holding chunk kk, and the chunk's atom range
                  const uint64_t mc = ballot64(nch > 0 && off <= kk);
                  const int ln = 63 - static_cast<int>(__builtin_clzll(mc));
                  cbv[u] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cb), ln)) +
                           kWave * (kk - static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off), ln)));
                  cev[u] = min(cbv[u] + kWave, static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ce), ln)));
                }
              }
              run_chunks(cbv, cev);
            }
          }
        }
        if (cur >= 0) flush();
      };
      if constexpr (CLASH) {  // (a launch evaluating the final state: no steps)
        run_items(std::true_type{});
      } else {
        if (need_e)
          run_items(std::true_type{});
        else
          run_items(std::false_type{});
      }
    }
    __syncthreads();

    // 3. per-atom energy and force from the integer sums
    //    (a saturated component as +-inf, as the population kernels write it)
    for (uint32_t k = tid; k < A; k += WAVES * kWave) {
      uint32_t st = s_st[k], ignore = 0;
      const long long ve = fix_add(s_acc[0 * AC + k], 0, st, kStE);
      const long long va = fix_add(s_acc[1 * AC + k], 0, ignore, 0u);
      const long long vx = fix_add(s_acc[2 * AC + k], 0, st, kStX);
      const long long vy = fix_add(s_acc[3 * AC + k], 0, st, kStY);
      const long long vz = fix_add(s_acc[4 * AC + k], 0, st, kStZ);
      s_e[k] = fix_value(ve, st, kStE);
      s_force[k] = make_float4(static_cast<float>(fix_value(vx, st, kStX)),
                               static_cast<float>(fix_value(vy, st, kStY)),
                               static_cast<float>(fix_value(vz, st, kStZ)),
                               static_cast<float>(static_cast<double>(va) * kFixUnit));
    }
    __syncthreads();

    if (s == g.n_steps) break;
    if (g.score_every && it % g.score_every == 0) {
      for (uint32_t j = wave; j < NC; j += WAVES) {
        double se, sa;
        candidate_sums(j, se, sa);
        if (lane == 0) g.score_e[static_cast<size_t>(it / g.score_every) * g.n_cand + s_cand[j].id] = se;
      }
    }

    // 4. residue update (src/diffusion.rs:300-339)
    for (uint32_t r = tid; r < R; r += WAVES * kWave) {
      const uint4 ri = s_rinfo[r];
      const TrajCand& C = s_cand[ri.w];
      const uint32_t b0 = ri.x & 0xFFFFu, b1 = ri.x >> 16;
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
      for (uint32_t o = b0; o < b1; ++o) {  // in atom order
        const float4 f = s_force[s_perm[o]];
        fx += dyn_force(f.x);
        fy += dyn_force(f.y);
        fz += dyn_force(f.z);
      }
      s_delta[r] = residue_delta(C.center[0], C.center[1], C.center[2], g.sp, C.r_restr,
                                 s_pos[ri.y], fx, fy, fz, sigma, it, ri.z);
    }
    __syncthreads();

    // 5. rigid residue moves
    for (uint32_t k = tid; k < A; k += WAVES * kWave) {
      const uint32_t r = s_res[k];
      if (r == kTrajPad) continue;
      const float4 d = s_delta[r];
      float4 q = s_pos[k];
      q.x += d.x;
      q.y += d.y;
      q.z += d.z;
      s_pos[k] = q;
    }
    __syncthreads();
  }

  // the final state's candidate energies (when this launch ends the loop)
  const uint32_t n_mid = g.score_every ? (g.n_total + g.score_every - 1) / g.score_every : 0;
  for (uint32_t j = wave; j < NC; j += WAVES) {
    double se = 0.0, sa = 0.0;
    if (g.final_eval) candidate_sums(j, se, sa);
    if (lane == 0) {
      const uint32_t c = s_cand[j].id;
      if (g.final_eval) {
        g.energy[c] = se;
        g.err[c] = static_cast<float>(sa) * kTrajErrPerAbs;
        g.score_e[static_cast<size_t>(n_mid) * g.n_cand + c] = se;
      }
      if (g.work) g.work[c] = s_work[j];
    }
  }
  for (uint32_t j = wave; j < NC; j += WAVES) {
    const TrajCand& C = s_cand[j];
    for (uint32_t i = lane; i < C.n; i += kWave) {
      const uint32_t k = s_perm[C.a_off + i];
      g.pos[C.a0 + i] = s_pos[k];
      if (g.final_eval) {
        g.force[C.a0 + i] = s_force[k];
        g.e_atom[C.a0 + i] = s_e[k];
      }
    }
  }
  if (COUNT) {
    const uint32_t ai = wave_sum_u32(n_in), at = wave_sum_u32(n_tested);
    if (lane == 0 && (ai | at)) {
      atomicAdd(&g.stats[0], static_cast<unsigned long long>(ai));
      atomicAdd(&g.stats[1], static_cast<unsigned long long>(at));
    }
  }
}

template <bool COUNT, int WAVES, bool CLASH>
static hipError_t launch_traj(const TrajArgs& a, size_t lds, hipStream_t stream) {
  auto kern = k_trajectory<COUNT, WAVES, CLASH>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  if (e != hipSuccess) return e;
  const uint32_t groups = a.group_off ? a.n_groups : (a.n_cand + a.cpw - 1) / a.cpw;
  hipLaunchKernelGGL(kern, dim3(groups), dim3(WAVES * kWave), lds, stream, a);
  return hipGetLastError();
}

struct TrajPlan {
  uint32_t waves = 0, cpw = 0, a_cap = 0, r_cap = 0, n_groups = 0;
  size_t lds = 0;
};

// The largest number of candidates per workgroup, up to `cpw`, whose groups
// (consecutive candidates, each padded to whole tiles) fit the workgroup's
// tiles, residue slots and LDS.
static bool traj_plan(const dnab_population* pop, uint32_t waves, uint32_t cpw, TrajPlan& pl) {
  const std::vector<uint32_t>& ao = pop->h_cand_atom_off;
  const std::vector<uint32_t>& ro = pop->h_cand_res_off;
  const size_t nc = pop->n_cand;
  if (ao.size() != nc + 1 || ro.size() != nc + 1) return false;
  for (uint32_t m = std::min(cpw, kTrajMaxCpw); m >= 1; --m) {
    uint32_t a_cap = kWave, r_cap = 1;
    for (size_t c0 = 0; c0 < nc; c0 += m) {
      uint32_t a = 0, r = 0;
      for (size_t c = c0; c < std::min(nc, c0 + m); ++c) {
        a += (ao[c + 1] - ao[c] + kWave - 1) & ~(kWave - 1);
        r += ro[c + 1] - ro[c];
      }
      a_cap = std::max(a_cap, a);
      r_cap = std::max(r_cap, r);
    }
    const size_t lds = traj_lds_bytes(a_cap, r_cap);
    if (a_cap / kWave > kTrajMaxTiles || r_cap > 0xFFFFu ||
        lds + traj_static_lds(static_cast<int>(waves)) > kTrajLdsLimit)
      continue;
    pl.waves = waves;
    pl.cpw = m;
    pl.a_cap = a_cap;
    pl.r_cap = r_cap;
    pl.lds = lds;
    return true;
  }
  return false;
}

static bool traj_choose(const dnab_population* pop, TrajPlan& pl) {
  return traj_plan(pop, pop->traj_waves, pop->traj_cpw, pl) || traj_plan(pop, 4, 1, pl);
}

bool trajectory_supported(const dnab_population* pop, const dnab_langevin_params* p) {
  if (p->force_model != DNAB_FORCE_ALLATOM || pop->n_cand == 0) return false;
  for (const AgView& v : pop->views_host)
    if (v.n_atoms >= kTrajMaxAgAtoms) return false;
  TrajPlan pl;
  return traj_choose(pop, pl);
}

// Candidate groups balanced by measured work (longest processing time
// first): candidates in descending work each join the open group with the
// least work that still has room (candidates, tiles, residue slots, LDS);
// groups are launched heaviest first. Fails (and the caller keeps consecutive
// groups) when the largest groups together would not fit the LDS limit.
static bool traj_balance(const dnab_population* pop, uint32_t waves, uint32_t cpw, uint32_t n_bins,
                         const std::vector<uint32_t>& work, std::vector<uint32_t>& off,
                         std::vector<uint32_t>& cand, TrajPlan& pl) {
  const size_t nc = pop->n_cand;
  const std::vector<uint32_t>& ao = pop->h_cand_atom_off;
  const std::vector<uint32_t>& ro = pop->h_cand_res_off;
  const size_t stat = traj_static_lds(static_cast<int>(waves));
  std::vector<uint32_t> order(nc);
  for (size_t c = 0; c < nc; ++c) order[c] = static_cast<uint32_t>(c);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return work[x] > work[y]; });
  struct Bin {
    uint64_t w = 0;
    uint32_t a = 0, r = 0, n = 0;
  };
  std::vector<uint32_t> assign(nc);
  std::vector<Bin> bins(std::max<uint32_t>(n_bins, 1));
  // min-heap of (work, bin) over groups with room; a group that cannot take
  // the current candidate is set aside and returned to the heap afterwards
  using Key = std::pair<uint64_t, uint32_t>;
  std::vector<Key> heap, aside;
  heap.reserve(bins.size() + 8);
  for (uint32_t b = 0; b < bins.size(); ++b) heap.push_back({0, b});
  std::make_heap(heap.begin(), heap.end(), std::greater<Key>());
  for (uint32_t c : order) {
    const uint32_t slots = (ao[c + 1] - ao[c] + kWave - 1) & ~(kWave - 1);
    const uint32_t nres = ro[c + 1] - ro[c];
    auto fits = [&](const Bin& B) {
      const uint32_t a = B.a + slots, r = B.r + nres;
      return a / kWave <= kTrajMaxTiles && r <= 0xFFFFu &&
             traj_lds_bytes(a, r) + stat <= kTrajLdsLimit;
    };
    uint32_t pick = UINT32_MAX;
    while (!heap.empty()) {
      std::pop_heap(heap.begin(), heap.end(), std::greater<Key>());
      const Key top = heap.back();
      heap.pop_back();
      if (fits(bins[top.second])) {
        pick = top.second;
        break;
      }
      aside.push_back(top);
    }
    for (const Key& k : aside) {
      heap.push_back(k);
      std::push_heap(heap.begin(), heap.end(), std::greater<Key>());
    }
    aside.clear();
    if (pick == UINT32_MAX) {  // no open group has room: a new one
      pick = static_cast<uint32_t>(bins.size());
      bins.emplace_back();
    }
    Bin& B = bins[pick];
    B.w += work[c];
    B.a += slots;
    B.r += nres;
    ++B.n;
    assign[c] = pick;
    if (B.n < cpw) {
      heap.push_back({B.w, pick});
      std::push_heap(heap.begin(), heap.end(), std::greater<Key>());
    }
  }
  std::vector<uint32_t> ids;
  for (uint32_t b = 0; b < bins.size(); ++b)
    if (bins[b].n) ids.push_back(b);
  std::stable_sort(ids.begin(), ids.end(), [&](uint32_t x, uint32_t y) { return bins[x].w > bins[y].w; });
  // group g = bin ids[g]: offsets, then the candidates in ascending index
  std::vector<uint32_t> slot(bins.size());
  off.assign(ids.size() + 1, 0);
  uint32_t a_cap = kWave, r_cap = 1, most = 1;
  for (size_t gi = 0; gi < ids.size(); ++gi) {
    const Bin& B = bins[ids[gi]];
    slot[ids[gi]] = off[gi];
    off[gi + 1] = off[gi] + B.n;
    a_cap = std::max(a_cap, B.a);
    r_cap = std::max(r_cap, B.r);
    most = std::max(most, B.n);
  }
  cand.resize(nc);
  for (size_t c = 0; c < nc; ++c) cand[slot[assign[c]]++] = static_cast<uint32_t>(c);
  const size_t lds = traj_lds_bytes(a_cap, r_cap);
  if (lds + stat > kTrajLdsLimit) return false;
  pl.waves = waves;
  pl.cpw = most;
  pl.a_cap = a_cap;
  pl.r_cap = r_cap;
  pl.lds = lds;
  return true;
}

// Per candidate, for grouping candidates by position: its antigen in the
// high word, and the Morton code of the mean of its atom positions on a grid
// of 2 A cells from 64 A below the antigen's atom box (clamped to 1024
// cells a side) in the low word. One wave per candidate.
__global__ void __launch_bounds__(256)
k_cand_key(const AgView* __restrict__ views, const uint32_t* __restrict__ cand_ag,
           const float4* __restrict__ pos, const uint32_t* __restrict__ cand_atom_off,
           uint32_t n_cand, unsigned long long* __restrict__ out) {
  const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (c >= n_cand) return;  // whole wave
  const uint32_t a0 = cand_atom_off[c], a1 = cand_atom_off[c + 1];
  float x = 0.0f, y = 0.0f, z = 0.0f;
  for (uint32_t k = a0 + lane; k < a1; k += kWave) {
    const float4 p = pos[k];
    x += p.x;
    y += p.y;
    z += p.z;
  }
  x = wave_reduce_f32<OpAdd>(x);
  y = wave_reduce_f32<OpAdd>(y);
  z = wave_reduce_f32<OpAdd>(z);
  if (lane == 0) {
    const float inv = 1.0f / static_cast<float>(max(a1 - a0, 1u));
    const AgView& V = views[cand_ag[c]];
    const float m[3] = {x * inv, y * inv, z * inv};
    uint32_t q[3];
    for (int d = 0; d < 3; ++d)  // (NaN to cell 0)
      q[d] = static_cast<uint32_t>(fminf(fmaxf((m[d] - (V.excl_lo[d] - 64.0f)) * 0.5f, 0.0f), 1023.0f));
    out[c] = (static_cast<unsigned long long>(cand_ag[c]) << 32) | morton3(q[0], q[1], q[2]);
  }
}

// Candidate indices in ascending key order, ties by index (LSD radix sort,
// 8-bit digits; digits equal across all keys are skipped).
static void sort_by_key(const std::vector<uint64_t>& key, std::vector<uint32_t>& order) {
  const size_t n = key.size();
  std::vector<uint32_t> tmp(n);
  order.resize(n);
  for (size_t i = 0; i < n; ++i) order[i] = static_cast<uint32_t>(i);
  for (int sh = 0; sh < 64; sh += 8) {
    size_t cnt[257] = {0};
    for (size_t i = 0; i < n; ++i) ++cnt[((key[i] >> sh) & 0xFFu) + 1];
    if (n == 0 || cnt[((key[0] >> sh) & 0xFFu) + 1] == n) continue;
    for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
    for (size_t i = 0; i < n; ++i) {
      const uint32_t c = order[i];
      tmp[cnt[(key[c] >> sh) & 0xFFu]++] = c;
    }
    order.swap(tmp);
  }
}

// Groups by position: candidates of one antigen, in the order of their keys
// (k_cand_key: antigen, then the Morton code of the centroid), up to cpw per
// group while the group's atoms, packed without padding per candidate, fit
// the tiles, residue slots and LDS. Groups are launched in that order.
static bool traj_spatial(const dnab_population* pop, uint32_t waves, uint32_t cpw,
                         const std::vector<uint64_t>& keys, std::vector<uint32_t>& off,
                         std::vector<uint32_t>& cand, TrajPlan& pl) {
  const size_t nc = pop->n_cand;
  const std::vector<uint32_t>& ao = pop->h_cand_atom_off;
  const std::vector<uint32_t>& ro = pop->h_cand_res_off;
  const std::vector<uint32_t>& ag = pop->h_cand_ag;
  if (ag.size() != nc || keys.size() != nc) return false;
  const size_t stat = traj_static_lds(static_cast<int>(waves));
  std::vector<uint32_t> order;
  sort_by_key(keys, order);
  off.assign(1, 0);
  cand.clear();
  cand.reserve(nc);
  uint32_t a_cap = kWave, r_cap = 1, most = 1, a = 0, r = 0, n = 0;
  for (size_t i = 0; i < nc; ++i) {
    const uint32_t c = order[i];
    const uint32_t na = ao[c + 1] - ao[c], nr = ro[c + 1] - ro[c];
    const uint32_t a2 = a + na, r2 = r + nr, slots = (a2 + kWave - 1) & ~(kWave - 1);
    const bool same = n == 0 || ag[cand.back()] == ag[c];
    if (n && (!same || n >= cpw || slots / kWave > kTrajMaxTiles || r2 > 0xFFFFu ||
              traj_lds_bytes(slots, r2) + stat > kTrajLdsLimit)) {
      off.push_back(static_cast<uint32_t>(cand.size()));
      a = r = n = 0;
    }
    a += na;
    r += nr;
    ++n;
    cand.push_back(c);
    a_cap = std::max(a_cap, (a + kWave - 1) & ~(kWave - 1));
    r_cap = std::max(r_cap, r);
    most = std::max(most, n);
  }
  off.push_back(static_cast<uint32_t>(cand.size()));
  if (a_cap / kWave > kTrajMaxTiles || r_cap > 0xFFFFu ||
      traj_lds_bytes(a_cap, r_cap) + stat > kTrajLdsLimit)
    return false;
  pl.waves = waves;
  pl.cpw = most;
  pl.a_cap = a_cap;
  pl.r_cap = r_cap;
  pl.lds = traj_lds_bytes(a_cap, r_cap);
  return true;
}

template <bool CLASH>
static hipError_t launch_plan_c(const TrajArgs& a, const TrajPlan& pl, bool count, hipStream_t st) {
  switch (pl.waves) {
    case 4:
      return count ? launch_traj<true, 4, CLASH>(a, pl.lds, st) : launch_traj<false, 4, CLASH>(a, pl.lds, st);
    case 8:
      return count ? launch_traj<true, 8, CLASH>(a, pl.lds, st) : launch_traj<false, 8, CLASH>(a, pl.lds, st);
    case 16:
      return count ? launch_traj<true, 16, CLASH>(a, pl.lds, st) : launch_traj<false, 16, CLASH>(a, pl.lds, st);
    default:
      return hipErrorInvalidValue;
  }
}
static hipError_t launch_plan(const TrajArgs& a, const TrajPlan& pl, bool count, bool clash,
                              hipStream_t st) {
  return clash ? launch_plan_c<true>(a, pl, count, st) : launch_plan_c<false>(a, pl, count, st);
}

int trajectory_run(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                   const StepParams& sp) {
  TrajPlan pl;
  if (!traj_choose(pop, pl))
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "trajectory: candidates too large for one workgroup");
  TrajArgs a;
  a.views = pop->views.p;
  a.cand_ag = pop->cand_ag.p;
  a.cand_atom_off = pop->cand_atom_off.p;
  a.cand_res_off = pop->cand_res_off.p;
  a.res_atom_off = pop->res_atom_off.p;
  a.res_ca = pop->res_ca.p;
  a.atom_res = pop->atom_res.p;
  a.pos = pop->pos.p;
  a.par = pop->par.p;
  a.force = pop->force.p;
  a.e_atom = pop->e_atom.p;
  a.energy = pop->energy.p;
  a.err = pop->err_bound.p;
  a.sigma_table = pop->sigma_table.p;
  a.stats = pop->pair_stats.p;
  a.score_e = pop->score_e.p;
  a.group_off = nullptr;
  a.group_cand = nullptr;
  a.work = nullptr;
  a.n_groups = 0;
  a.cells = static_cast<int32_t>(pop->traj_cells);
  a.chunks = pop->traj_chunks;
  a.n_cand = static_cast<uint32_t>(pop->n_cand);
  a.n_steps = p->n_steps;
  a.score_every = p->score_every;
  a.cpw = pl.cpw;
  a.a_cap = pl.a_cap;
  a.r_cap = pl.r_cap;
  a.packed = 0;
  a.sp = sp;
  a.s0 = 0;
  a.n_total = p->n_steps;
  a.final_eval = 1;
  const bool count = p->count_pairs != 0;
  const size_t nc = pop->n_cand;
  // grouping: 0 consecutive candidates, 1 by measured work, 2 by position
  const uint32_t mode = nc > 1 && p->n_steps > 0 ? pop->traj_balance : 0u;
  const bool balance = mode == 1;
  std::vector<uint32_t> work;
  auto set_groups = [&](const std::vector<uint32_t>& off, const std::vector<uint32_t>& cand,
                        const TrajPlan& np) -> int {
    DNAB_ALLOC_CHECK(ctx, pop->traj_group_off.ensure(off.size()));
    DNAB_ALLOC_CHECK(ctx, pop->traj_group_cand.ensure(cand.size()));
    DNAB_RETURN_IF(upload(ctx, pop->traj_group_off.p, off.data(), off.size() * sizeof(uint32_t)));
    DNAB_RETURN_IF(upload(ctx, pop->traj_group_cand.p, cand.data(), cand.size() * sizeof(uint32_t)));
    a.group_off = pop->traj_group_off.p;
    a.group_cand = pop->traj_group_cand.p;
    a.n_groups = static_cast<uint32_t>(off.size() - 1);
    a.cpw = np.cpw;
    a.a_cap = np.a_cap;
    a.r_cap = np.r_cap;
    pl = np;
    return DNAB_OK;
  };
  // group candidates by their centroids now; on failure the launch keeps
  // the groups it has
  std::vector<uint64_t> keys;
  auto regroup_space = [&]() -> int {
    std::vector<uint32_t>& off = pop->h_group_off;
    std::vector<uint32_t>& cand = pop->h_group_cand;
    DNAB_ALLOC_CHECK(ctx, pop->traj_key.ensure(nc));
    hipLaunchKernelGGL(k_cand_key, dim3(static_cast<uint32_t>((nc + 3) / 4)), dim3(256), 0,
                       ctx->stream, pop->views.p, pop->cand_ag.p, pop->pos.p, pop->cand_atom_off.p,
                       static_cast<uint32_t>(nc), pop->traj_key.p);
    DNAB_HIP_CHECK(ctx, hipGetLastError());
    keys.resize(nc);
    DNAB_RETURN_IF(download(ctx, keys.data(), pop->traj_key.p, nc * sizeof(uint64_t)));
    TrajPlan sp2;
    if (!traj_spatial(pop, pop->traj_waves, pop->traj_cpw, keys, off, cand, sp2)) return DNAB_OK;
    DNAB_RETURN_IF(set_groups(off, cand, sp2));
    a.packed = 1;
    return DNAB_OK;
  };
  // group candidates by the work in `work` (per candidate, staged antigen
  // atoms); on failure the launch keeps consecutive groups
  auto regroup = [&]() -> int {
    std::vector<uint32_t>& off = pop->h_group_off;   // kept: the copies read them
    std::vector<uint32_t>& cand = pop->h_group_cand; // after this returns
    TrajPlan bal;
    const uint32_t bins = std::max<uint32_t>(static_cast<uint32_t>(ctx->cu_count),
                                             static_cast<uint32_t>((nc + pop->traj_cpw - 1) / pop->traj_cpw));
    if (!traj_balance(pop, pop->traj_waves, pop->traj_cpw, bins, work, off, cand, bal)) return DNAB_OK;
    return set_groups(off, cand, bal);
  };
  if (balance) {
    // measure each candidate's work on the start state (one evaluation, whose
    // outputs the run overwrites), then group candidates by it
    DNAB_ALLOC_CHECK(ctx, pop->traj_work.ensure(nc));
    work.resize(nc);
    TrajArgs e = a;
    e.n_steps = 0;
    e.n_total = 0;
    e.score_every = 0;
    e.work = pop->traj_work.p;
    DNAB_HIP_CHECK(ctx, launch_plan(e, pl, false, false, ctx->stream));
    auto t0 = std::chrono::steady_clock::now();
    DNAB_RETURN_IF(download(ctx, work.data(), pop->traj_work.p, nc * sizeof(uint32_t)));
    timer_host(ctx, "trajectory_host_work", ms_since(t0));
    t0 = std::chrono::steady_clock::now();
    DNAB_RETURN_IF(regroup());
    timer_host(ctx, "trajectory_host_group", ms_since(t0));
  } else if (mode == 2) {
    DNAB_RETURN_IF(regroup_space());
  }
  // the loop in segments of pop->traj_rebalance steps (all at once when 0);
  // between segments the groups follow the work measured in the last one,
  // or the candidates' new positions
  const uint32_t seg = (mode && pop->traj_rebalance) ? pop->traj_rebalance : std::max<uint32_t>(p->n_steps, 1);
  uint32_t done = 0;
  do {
    const uint32_t len = std::min(seg, p->n_steps - done);
    a.s0 = done;
    a.n_steps = len;
    a.final_eval = 0;
    a.work = balance && done + len < p->n_steps ? pop->traj_work.p : nullptr;
    hipEvent_t ev_end;
    hipEvent_t ev = timer_begin(ctx, "trajectory", &ev_end);
    hipError_t e = launch_plan(a, pl, count, false, ctx->stream);
    timer_end(ctx, "trajectory", ev, ev_end);
    if (e != hipSuccess)
      return dnab_set_error(ctx, DNAB_ERR_HIP, "trajectory launch failed: %s", hipGetErrorString(e));
    done += len;
    if (a.work) {
      DNAB_RETURN_IF(download(ctx, work.data(), pop->traj_work.p, nc * sizeof(uint32_t)));
      DNAB_RETURN_IF(regroup());
    } else if (mode == 2 && done < p->n_steps) {
      DNAB_RETURN_IF(regroup_space());
    }
  } while (done < p->n_steps);
  // the final state: energies, forces and the last scoring row, with the
  // clash pairs in FP64 as every reported evaluation has them
  if (a.packed) {
    // a population large enough for packed groups evaluates it with the
    // population kernels (tiles sorted across candidates; the same results)
    hipEvent_t ev_end;
    hipEvent_t ev = timer_begin(ctx, "trajectory_final", &ev_end);
    DNAB_RETURN_IF(population_sort(ctx, pop));
    DNAB_RETURN_IF(population_pair(ctx, pop, count ? kEvalCount : DNAB_EVAL_FORCES));
    DNAB_RETURN_IF(population_reduce_energies(ctx, pop, DNAB_EVAL_FORCES));
    const uint32_t n_mid = p->score_every ? (p->n_steps + p->score_every - 1) / p->score_every : 0;
    DNAB_HIP_CHECK(ctx, hipMemcpyAsync(pop->score_e.p + static_cast<size_t>(n_mid) * nc, pop->energy.p,
                                       nc * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    timer_end(ctx, "trajectory_final", ev, ev_end);
    return DNAB_OK;
  }
  a.s0 = p->n_steps;
  a.n_steps = 0;
  a.final_eval = 1;
  a.work = nullptr;
  hipEvent_t ev_end;
  hipEvent_t ev = timer_begin(ctx, "trajectory_final", &ev_end);
  hipError_t e = launch_plan(a, pl, count, true, ctx->stream);
  timer_end(ctx, "trajectory_final", ev, ev_end);
  if (e != hipSuccess)
    return dnab_set_error(ctx, DNAB_ERR_HIP, "trajectory launch failed: %s", hipGetErrorString(e));
  return DNAB_OK;
}

// Group table of one candidate per group from a device index list.
__global__ void k_subset_groups(const uint32_t* __restrict__ idx, uint32_t n,
                                uint32_t* __restrict__ off, uint32_t* __restrict__ cand) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) off[i] = i;
  if (i < n) cand[i] = idx[i];
}

int trajectory_eval_subset(dnab_ctx* ctx, dnab_population* pop, const uint32_t* d_idx, uint32_t n,
                           double* score_scratch) {
  if (n == 0) return DNAB_OK;
  for (const AgView& v : pop->views_host)
    if (v.n_atoms >= kTrajMaxAgAtoms)
      return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "subset evaluation: antigen too large");
  TrajPlan pl;
  pl.waves = 4;
  pl.cpw = 1;
  pl.a_cap = std::max<uint32_t>(kWave, (pop->max_cand_atoms + kWave - 1) & ~(kWave - 1));
  pl.r_cap = std::max<uint32_t>(1, pop->max_cand_res);
  pl.lds = traj_lds_bytes(pl.a_cap, pl.r_cap);
  if (pl.a_cap / kWave > kTrajMaxTiles || pl.r_cap > 0xFFFFu ||
      pl.lds + traj_static_lds(4) > kTrajLdsLimit)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "subset evaluation: candidate too large");
  DNAB_ALLOC_CHECK(ctx, pop->traj_group_off.ensure(n + 1));
  DNAB_ALLOC_CHECK(ctx, pop->traj_group_cand.ensure(n));
  hipLaunchKernelGGL(k_subset_groups, dim3((n + 256) / 256), dim3(256), 0, ctx->stream, d_idx, n,
                     pop->traj_group_off.p, pop->traj_group_cand.p);
  TrajArgs a{};
  a.views = pop->views.p;
  a.cand_ag = pop->cand_ag.p;
  a.cand_atom_off = pop->cand_atom_off.p;
  a.cand_res_off = pop->cand_res_off.p;
  a.res_atom_off = pop->res_atom_off.p;
  a.res_ca = pop->res_ca.p;
  a.atom_res = pop->atom_res.p;
  a.pos = pop->pos.p;
  a.par = pop->par.p;
  a.force = pop->force.p;
  a.e_atom = pop->e_atom.p;
  a.energy = pop->energy.p;
  a.err = pop->err_bound.p;
  a.sigma_table = nullptr;  // no steps
  a.stats = nullptr;
  a.score_e = score_scratch;  // row 0 (score_every 0)
  a.group_off = pop->traj_group_off.p;
  a.group_cand = pop->traj_group_cand.p;
  a.work = nullptr;
  a.n_groups = n;
  a.cells = static_cast<int32_t>(pop->traj_cells);
  a.chunks = pop->traj_chunks;
  a.n_cand = static_cast<uint32_t>(pop->n_cand);
  a.n_steps = 0;
  a.score_every = 0;
  a.s0 = 0;
  a.n_total = 0;
  a.final_eval = 1;
  a.cpw = 1;
  a.a_cap = pl.a_cap;
  a.r_cap = pl.r_cap;
  a.packed = 0;
  hipEvent_t ev_end;
  hipEvent_t ev = timer_begin(ctx, "trajectory_subset", &ev_end);
  hipError_t e = launch_plan(a, pl, false, true, ctx->stream);
  timer_end(ctx, "trajectory_subset", ev, ev_end);
  if (e != hipSuccess)
    return dnab_set_error(ctx, DNAB_ERR_HIP, "subset evaluation launch failed: %s", hipGetErrorString(e));
  return DNAB_OK;
}

}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_population_get_trajectory_tuning(const dnab_population* pop, dnab_traj_tuning* t) {
  if (!pop || !t) return DNAB_ERR_INVALID;
  t->waves = pop->traj_waves;
  t->cand_per_group = pop->traj_cpw;
  t->balance = pop->traj_balance;
  t->cells_per_item = pop->traj_cells;
  t->chunks_per_item = pop->traj_chunks;
  t->rebalance_steps = pop->traj_rebalance;
  t->engine = pop->traj_engine;
  return DNAB_OK;
}

int dnab_population_set_trajectory_tuning(dnab_ctx* ctx, dnab_population* pop,
                                          const dnab_traj_tuning* t) {
  if (!ctx || !pop || !t || pop->ctx != ctx) return DNAB_ERR_INVALID;
  if (t->waves != 4 && t->waves != 8 && t->waves != 16)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: waves must be 4, 8 or 16");
  if (t->cand_per_group < 1 || t->cand_per_group > kTrajMaxCpw)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: cand_per_group must be 1..%u",
                          kTrajMaxCpw);
  if (t->balance > 2)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: balance must be 0, 1 or 2");
  if (t->cells_per_item < 1 || t->cells_per_item > 64)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: cells_per_item must be 1..64");
  if (t->chunks_per_item < 1 || t->chunks_per_item > static_cast<uint32_t>(kLoadAhead))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: chunks_per_item must be 1..%d",
                          kLoadAhead);
  if (t->engine > 1)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: engine must be 0 or 1");
  pop->traj_chunks = t->chunks_per_item;
  pop->traj_rebalance = t->rebalance_steps;
  pop->traj_waves = t->waves;
  pop->traj_cells = t->cells_per_item;
  pop->traj_cpw = t->cand_per_group;
  pop->traj_balance = t->balance;
  pop->traj_engine = t->engine;
  return DNAB_OK;
}

}  // extern "C"
