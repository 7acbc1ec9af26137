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
//     residue by the clamped step whatever its rounding);
//  3. per-atom energies and forces from the integer sums;
//  4. the residue update (residue_delta, step_device.h), one thread per residue;
//  5. the rigid residue moves, one thread per atom.
// After the last step the positions are written back; a launch with no
// steps, whose clash pairs are FP64 (CLASH), evaluates the final state and
// writes the reported energies and forces, the candidate energy reduced in
// the order of the population reduction kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <vector>

#include "device_common.h"
#include "dnab_internal.h"
#include "pair_device.h"
#include "step_device.h"

namespace dnab {

constexpr uint32_t kTrajMaxTiles = 64;    // tiles per workgroup (one ballot finds a tile)
constexpr uint32_t kTrajMaxCpw = 16;      // candidates per workgroup
constexpr uint16_t kTrajPad = 0xFFFFu;    // residue slot of a padding atom
constexpr uint32_t kTileChunks = 192;     // chunk list entries per tile
constexpr uint32_t kTileOverflow = 0xFFFFFFFFu;
constexpr uint32_t kTrajMaxAgAtoms = 1u << 26;  // a chunk entry holds (first atom << 6 | len - 1)
constexpr size_t kTrajLdsLimit = 160 * 1024;
constexpr float kTrajErrPerAbs = 32.0f / 16777216.0f;  // 32 ulp of f32 per unit of sum |e|

struct TrajTile {
  float lo[3], hi[3];
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
  uint32_t a_off;      // first LDS atom slot (a multiple of 64)
  uint32_t r_off;      // first LDS residue slot
  uint32_t ag;         // antigen view
};

// The antigen fields a step reads, copied into LDS once per candidate slot.
struct TrajAg {
  const AgAtom* atoms;
  const uint32_t* cell_start;
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
  StepParams sp;
};

// Bytes of dynamic LDS for a capacity of a_cap atom slots and r_cap residues:
// 5 integer sums, energy, position, parameters, force, status per atom slot; the
// displacement and the (atom range, Calpha, index, candidate) record per
// residue; the residue slot per atom slot; the chunk list of each tile
// (a_cap / 64 tiles).
__host__ __device__ inline size_t traj_lds_bytes(uint32_t a_cap, uint32_t r_cap) {
  const size_t res = (static_cast<size_t>(a_cap) * 2 + 3) & ~static_cast<size_t>(3);
  return static_cast<size_t>(a_cap) * (5 * 8 + 8 + 3 * 16 + 4) + static_cast<size_t>(r_cap) * 32 + res +
         static_cast<size_t>(a_cap / kWave) * kTileChunks * 4;
}
// Static LDS of a workgroup of `waves` waves (staged chunks, tiles,
// candidates, antigen fields, counters).
__host__ __device__ inline size_t traj_static_lds(int waves) {
  return static_cast<size_t>(waves) * kWave * 32 + kTrajMaxTiles * sizeof(TrajTile) +
         kTrajMaxCpw * (sizeof(TrajCand) + sizeof(TrajAg) + 4) + 256;
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
  __shared__ uint32_t s_next, s_ncand, s_nslot, s_nres;
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
  uint32_t* s_desc = reinterpret_cast<uint32_t*>(s_dyn + ((reinterpret_cast<uintptr_t>(s_res + AC) -
                                                             reinterpret_cast<uintptr_t>(s_dyn) + 3) & ~uintptr_t(3)));

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
      for (uint32_t t = a_off / kWave; t < (a_off + slots) / kWave; ++t) {
        s_tile[t].cand = j;
        s_tile[t].nx = INT32_MIN;  // no chunk list yet (no cell range has this width)
        s_tile[t].nchunk = 0;
      }
      a_off += slots;
      r_off += C.nres;
    }
    s_ncand = nc;
    s_nslot = a_off;
    s_nres = r_off;
  }
  __syncthreads();
  const uint32_t NC = s_ncand, A = s_nslot, R = s_nres, T = A / kWave;

  for (uint32_t k = tid; k < A; k += WAVES * kWave) {
    const TrajCand C = s_cand[s_tile[k / kWave].cand];
    const uint32_t i = k - C.a_off;
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
  // per residue: LDS atom range and Calpha slot, global index, candidate slot
  for (uint32_t r = tid; r < R; r += WAVES * kWave) {
    uint32_t j = 0;
    while (j + 1 < NC && s_cand[j + 1].r_off <= r) ++j;
    const TrajCand C = s_cand[j];
    const uint32_t rg = C.r0 + (r - C.r_off);
    const uint32_t b0 = C.a_off + (g.res_atom_off[rg] - C.a0);
    const uint32_t b1 = C.a_off + (g.res_atom_off[rg + 1] - C.a0);
    s_rinfo[r] = make_uint4(b0 | (b1 << 16), C.a_off + (g.res_ca[rg] - C.a0), rg, j);
  }
  __syncthreads();

  // candidate j's energy and sum |e| (one wave), in the order of the
  // population reduction kernel: lane-strided FP64 sums, then a DPP wave sum
  auto candidate_sums = [&](uint32_t j, double& se, double& sa) {
    const TrajCand C = s_cand[j];
    se = 0.0;
    sa = 0.0;
    for (uint32_t i = lane; i < C.n; i += kWave) {
      se += s_e[C.a_off + i];
      sa += static_cast<double>(s_force[C.a_off + i].w);
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
      // cells that can hold an antigen atom within rc of the box (cell_key
      // is monotone, so a slightly widened range is conservative)
      const float margin = V.rc * 1.001f + 1e-3f;
      const int32_t kx0 = max(cell_key(lox - margin, V.inv), V.kmin[0]);
      const int32_t ky0 = max(cell_key(loy - margin, V.inv), V.kmin[1]);
      const int32_t kz0 = max(cell_key(loz - margin, V.inv), V.kmin[2]);
      const int32_t kx1 = min(cell_key(hix + margin, V.inv), V.kmax[0]);
      const int32_t ky1 = min(cell_key(hiy + margin, V.inv), V.kmax[1]);
      const int32_t kz1 = min(cell_key(hiz + margin, V.inv), V.kmax[2]);
      // (a tile without a valid atom has an empty box and no cells)
      const int32_t nx = any ? kx1 - kx0 + 1 : 0;
      const int32_t ny = any ? ky1 - ky0 + 1 : 0;
      const int32_t nz = any ? kz1 - kz0 + 1 : 0;
      const int32_t nxy = nx * ny;
      const int32_t ncell = (nx > 0 && ny > 0 && nz > 0) ? nxy * nz : 0;
      uint32_t* desc = s_desc + t * kTileChunks;
      uint32_t nt = 0;
      bool over = false;
      // the antigen does not move: the list stands while the cell range does
      const TrajTile& tc = s_tile[t];
      const bool same = tc.nx == nx && tc.ny == ny && tc.nz == nz && tc.k0[0] == kx0 &&
                        tc.k0[1] == ky0 && tc.k0[2] == kz0;
      if (same) {
        nt = tc.nchunk;
        over = nt == kTileOverflow;
      }
      for (int32_t cb0 = 0; !same && cb0 < ncell; cb0 += kWave) {
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
        tb.lo[0] = lox; tb.lo[1] = loy; tb.lo[2] = loz;
        tb.hi[0] = hix; tb.hi[1] = hiy; tb.hi[2] = hiz;
        tb.k0[0] = kx0; tb.k0[1] = ky0; tb.k0[2] = kz0;
        tb.nx = nx;
        tb.ny = ny;
        tb.nz = nz;
        tb.nxy = nxy;
        tb.ncell = ncell;
        tb.nchunk = over ? kTileOverflow : nt;
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

      FixAcc acc;
      int32_t cur = -1;
      float xb = 0.f, yb = 0.f, zb = 0.f, qb = 0.f, rmhb = 0.f, sqeb = 0.f, sqe12b = 0.f, epsb = 0.f;
      bool hydb = false, valid = false;
      float rc2_lane = -1.0f;
      auto flush = [&]() {
        const uint32_t k = static_cast<uint32_t>(cur) * kWave + lane;
        // one flush per (wave, tile) and step: at most WAVES sums of <= 2^59
        if (acc.e | acc.a | acc.x | acc.y | acc.z) {
          atomicAdd(reinterpret_cast<unsigned long long*>(&s_acc[0 * AC + k]), static_cast<unsigned long long>(acc.e));
          atomicAdd(reinterpret_cast<unsigned long long*>(&s_acc[1 * AC + k]), static_cast<unsigned long long>(acc.a));
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
            chunk_pairs<true, COUNT, CLASH>(s_stage[wave], n_s, xb, yb, zb, qb, rmhb, sqeb, sqe12b, epsb,
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
      for (uint32_t k = b0; k < b1; ++k) {  // in atom order
        const float4 f = s_force[k];
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
  for (uint32_t k = tid; k < A; k += WAVES * kWave) {
    const TrajCand C = s_cand[s_tile[k / kWave].cand];
    const uint32_t i = k - C.a_off;
    if (i < C.n) {
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
  a.sp = sp;
  a.s0 = 0;
  a.n_total = p->n_steps;
  a.final_eval = 1;
  const bool count = p->count_pairs != 0;
  const size_t nc = pop->n_cand;
  const bool balance = pop->traj_balance && nc > 1 && p->n_steps > 0;
  std::vector<uint32_t> work;
  // group candidates by the work in `work` (per candidate, staged antigen
  // atoms); on failure the launch keeps consecutive groups
  auto regroup = [&]() -> int {
    std::vector<uint32_t>& off = pop->h_group_off;   // kept: the copies read them
    std::vector<uint32_t>& cand = pop->h_group_cand; // after this returns
    TrajPlan bal;
    const uint32_t bins = std::max<uint32_t>(static_cast<uint32_t>(ctx->cu_count),
                                             static_cast<uint32_t>((nc + pop->traj_cpw - 1) / pop->traj_cpw));
    if (!traj_balance(pop, pop->traj_waves, pop->traj_cpw, bins, work, off, cand, bal)) return DNAB_OK;
    DNAB_ALLOC_CHECK(ctx, pop->traj_group_off.ensure(off.size()));
    DNAB_ALLOC_CHECK(ctx, pop->traj_group_cand.ensure(cand.size()));
    DNAB_RETURN_IF(upload(ctx, pop->traj_group_off.p, off.data(), off.size() * sizeof(uint32_t)));
    DNAB_RETURN_IF(upload(ctx, pop->traj_group_cand.p, cand.data(), cand.size() * sizeof(uint32_t)));
    a.group_off = pop->traj_group_off.p;
    a.group_cand = pop->traj_group_cand.p;
    a.n_groups = static_cast<uint32_t>(off.size() - 1);
    a.cpw = bal.cpw;
    a.a_cap = bal.a_cap;
    a.r_cap = bal.r_cap;
    pl = bal;
    return DNAB_OK;
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
    DNAB_RETURN_IF(download(ctx, work.data(), pop->traj_work.p, nc * sizeof(uint32_t)));
    DNAB_RETURN_IF(regroup());
  }
  // the loop in segments of pop->traj_rebalance steps (all at once when 0);
  // between segments the groups follow the work measured in the last one
  const uint32_t seg = (balance && pop->traj_rebalance) ? pop->traj_rebalance : std::max<uint32_t>(p->n_steps, 1);
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
    }
  } while (done < p->n_steps);
  // the final state: energies, forces and the last scoring row, with the
  // clash pairs in FP64 as every reported evaluation has them
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

}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_population_get_trajectory_tuning(const dnab_population* pop, dnab_traj_tuning* t) {
  if (!pop || !t) return DNAB_ERR_INVALID;
  t->waves = pop->traj_waves;
  t->cand_per_group = pop->traj_cpw;
  t->balance = pop->traj_balance ? 1u : 0u;
  t->cells_per_item = pop->traj_cells;
  t->chunks_per_item = pop->traj_chunks;
  t->rebalance_steps = pop->traj_rebalance;
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
  if (t->cells_per_item < 1 || t->cells_per_item > 64)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: cells_per_item must be 1..64");
  if (t->chunks_per_item < 1 || t->chunks_per_item > static_cast<uint32_t>(kLoadAhead))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "trajectory tuning: chunks_per_item must be 1..%d",
                          kLoadAhead);
  pop->traj_chunks = t->chunks_per_item;
  pop->traj_rebalance = t->rebalance_steps;
  pop->traj_waves = t->waves;
  pop->traj_cells = t->cells_per_item;
  pop->traj_cpw = t->cand_per_group;
  pop->traj_balance = t->balance != 0;
  return DNAB_OK;
}

}  // extern "C"
