// All-atom AMBER LJ + Coulomb + hydrophobic energies and forces between
// candidate binder atoms and a fixed antigen.
//
// Reference semantics: interaction_energy_atoms_breakdown
// (src/energy.rs:261-314) and the WGSL kernel (shaders/energy.wgsl:46-118):
// for every (candidate atom j, antigen atom i) with 0.25 <= r^2 <= rc^2,
// r^2 = (dx*dx + dy*dy) + dz*dz in f32 with dx = x_j - x_i,
//   LJ   eps_ij ((R/r)^12 - 2 (R/r)^6), R = rmh_i + rmh_j, eps_ij = sqrt(eps_i eps_j)
//   Coul 332 q_i q_j / r
//   hydrophobic -0.5 when both atoms are flagged and r^2 < 36.
//
// Layout and mapping (gfx950, wave64):
//  * antigen atoms are sorted cell-major over its SpatialHashGrid (Morton
//    order of cells), one 32-byte AgAtom each; a wave loads a chunk (up to
//    64 atoms of one cell) one atom per lane, keeps the atoms near its tile
//    (ballot) and stages them in LDS, from where the pair loop broadcasts
//    them to all lanes;
//  * candidate atoms of the whole population are counting-sorted every
//    evaluation into 4 A Morton cells inside 16 A bricks around their
//    antigen; a tile is 64 consecutive sorted slots, one atom per lane, and
//    every brick starts on a tile boundary, so a tile's atoms are spatially
//    compact whichever candidates they belong to;
//  * k_pair_chunks runs one wave per tile and hands the chunks beyond its
//    share to k_pair_rest, where the waves of a workgroup share them;
//  * every (atom, chunk) sum is rounded once to a 64-bit integer in units of
//    2^-20 (pair_device.h), so each atom's result does not depend on its
//    tile or on which wave took which chunk, and repeated runs are
//    bit-identical;
//  * the per-atom energy is stored in FP64 and summed per candidate in FP64
//    with DPP wave reductions.
#include <hip/hip_runtime.h>

#include <chrono>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "dnab_internal.h"
#include "ff_tables.h"
#include "ff_terms.h"
#include "grid_device.h"
#include "pair_device.h"
#include "sort_device.h"

namespace dnab {

constexpr float kSortCell = 4.0f;  // candidate-sort cell edge
constexpr uint32_t kInvalid = 0xFFFFFFFFu;  // empty tile slot
constexpr int kTileWaves = 4;

__global__ void k_fill_u32b(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Packs the per-atom SoA (x, y, z, charge, r_min_half, epsilon as float
// planes of stride nf, then hydrophobic and atom type as byte planes) into
// pos = (x, y, z, q) and par = (r_min_half, epsilon, hydrophobic, type or -1),
// and finds each atom's residue by binary search in res_atom_off, with its
// candidate and antigen.
__global__ void __launch_bounds__(256)
k_pack_atoms(const float* __restrict__ soa, size_t nf, int typed, uint32_t na, uint32_t nr,
             const uint32_t* __restrict__ res_atom_off, const uint32_t* __restrict__ res_cand,
             const uint32_t* __restrict__ cand_ag, float4* __restrict__ pos,
             float4* __restrict__ par, uint32_t* __restrict__ atom_ag,
             uint32_t* __restrict__ atom_cand, uint32_t* __restrict__ atom_res) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  const uint8_t* bytes = reinterpret_cast<const uint8_t*>(soa + 6 * nf);
  pos[i] = make_float4(soa[i], soa[nf + i], soa[2 * nf + i], soa[3 * nf + i]);
  par[i] = make_float4(soa[4 * nf + i], soa[5 * nf + i], bytes[i] ? 1.0f : 0.0f,
                       typed ? static_cast<float>(bytes[nf + i]) : -1.0f);
  // last residue r with res_atom_off[r] <= i (residues are non-empty)
  uint32_t lo = 0, hi = nr;  // res_atom_off[lo] <= i < res_atom_off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (res_atom_off[mid] <= i) lo = mid;
    else hi = mid;
  }
  const uint32_t c = res_cand[lo];
  atom_res[i] = lo;
  atom_cand[i] = c;
  atom_ag[i] = cand_ag[c];
}

static inline uint32_t nblocks(uint64_t n, uint32_t t) {
  uint64_t b = (n + t - 1) / t;
  return static_cast<uint32_t>(b ? b : 1);
}

// ---- antigen records -------------------------------------------------------------

__global__ void k_ag_gather(const uint32_t* __restrict__ sorted, uint32_t n,
                            const float* __restrict__ x, const float* __restrict__ y,
                            const float* __restrict__ z, const float* __restrict__ q,
                            const float* __restrict__ rmh, const float* __restrict__ eps,
                            const uint8_t* __restrict__ hyd, const uint8_t* __restrict__ type,
                            AgAtom* __restrict__ out, float2* __restrict__ raw) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t i = sorted[s];
  AgAtom a;
  a.x = x[i];
  a.y = y[i];
  a.z = z[i];
  a.q332 = 332.0f * q[i];
  a.rmh = rmh[i];
  a.sqe = sqrtf(eps[i]);
  a.eps = eps[i];
  a.tag = ag_tag(hyd[i] != 0, type ? static_cast<int>(type[i]) : -1);  // AtomType, when given
  out[s] = a;
  raw[s] = make_float2(q[i], eps[i]);
}

// The box of each grid cell's atoms (thread per Morton code): lo, hi, from
// the grid-sorted records; +inf, -inf for an empty cell.
__global__ void k_cell_box(const AgAtom* __restrict__ atoms, const uint32_t* __restrict__ cell_start,
                           uint64_t ncodes, float4* __restrict__ box) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncodes) return;
  const float inf = __builtin_inff();
  float4 lo = make_float4(inf, inf, inf, 0.0f), hi = make_float4(-inf, -inf, -inf, 0.0f);
  for (uint32_t i = cell_start[c]; i < cell_start[c + 1]; ++i) {
    const AgAtom& a = atoms[i];
    lo.x = fminf(lo.x, a.x);
    lo.y = fminf(lo.y, a.y);
    lo.z = fminf(lo.z, a.z);
    hi.x = fmaxf(hi.x, a.x);
    hi.y = fmaxf(hi.y, a.y);
    hi.z = fmaxf(hi.z, a.z);
  }
  box[2 * c] = lo;
  box[2 * c + 1] = hi;
}

// ---- candidate-atom spatial sort ---------------------------------------------------
// count (one wave per candidate, or the step kernel for the next step) ->
// scan (one workgroup: exclusive starts per antigen segment, each segment
// starting on a 64-slot tile; zeroes the counts; fills the padding slots) ->
// scatter.

__global__ void __launch_bounds__(256)
k_sort_count(const AgView* __restrict__ views, const uint32_t* __restrict__ cand_ag,
             const uint32_t* __restrict__ cand_atom_off, uint32_t n_cand,
             const uint8_t* __restrict__ cand_mask, const float4* __restrict__ pos,
             uint32_t* __restrict__ count, uint32_t* __restrict__ code_out,
             uint32_t* __restrict__ rank_out, float4* __restrict__ force,
             double* __restrict__ e_atom, double* __restrict__ bd_atom,
             double* __restrict__ f64_atom) {
  const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (c >= n_cand) return;
  const AgView& V = views[cand_ag[c]];
  const bool out = cand_mask && !cand_mask[c];
  for (uint32_t j = cand_atom_off[c] + lane; j < cand_atom_off[c + 1]; j += kWave) {
    uint32_t code = kExcluded;
    if (!out) {
      const float4 p = pos[j];
      code = sort_code_of(V, p.x, p.y, p.z);
    }
    code_out[j] = code;
    if (code >= kNonFinite) fill_atom_outputs(j, code, V, force, e_atom, bd_atom, f64_atom);
    else rank_out[j] = atomicAdd(&count[code], 1u);
  }
}

// Exclusive starts of every sort code. Each 16 A brick (64 codes) starts on a
// 64-slot tile boundary, so a tile never mixes atoms of two bricks (its box is
// at most one brick wide); the slots a brick leaves empty hold kInvalid.
// Thread t owns codes 4t..4t+3 of a 4096-code chunk: 16 threads per brick.
__global__ void __launch_bounds__(1024)
k_sort_scan(const AgView* __restrict__ views, uint32_t n_ag, uint32_t* __restrict__ count,
            uint32_t* __restrict__ start, uint32_t* __restrict__ perm,
            uint32_t* __restrict__ ag_tile_start, uint32_t* __restrict__ counter,
            uint32_t* __restrict__ n_list) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t woff[16];
  __shared__ uint32_t block_total;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6, g = lane & 15u;
  uint32_t carry = 0;  // a multiple of 64 between bricks
  for (uint32_t a = 0; a < n_ag; ++a) {
    const uint32_t base = views[a].sort_code_base, ncode = views[a].sort_ncode;
    if (t == 0) ag_tile_start[a] = carry / kWave;
    for (uint32_t c0 = 0; c0 < ncode; c0 += 4096) {
      const uint32_t idx = c0 + 4 * t;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (idx < ncode) v = *reinterpret_cast<const uint4*>(&count[base + idx]);
      const uint32_t s0 = v.x, s1 = s0 + v.y, s2 = s1 + v.z, tsum = s2 + v.w;
      uint32_t incl = tsum;  // inclusive scan within the brick's 16 threads
      for (int o = 1; o < 16; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 16);
        if (g >= static_cast<uint32_t>(o)) incl += u;
      }
      const uint32_t brick_total = __shfl(incl, static_cast<int>((lane & ~15u) | 15u), 64);
      const uint32_t padded = (brick_total + kWave - 1) & ~(kWave - 1);
      const uint32_t p0 = __shfl(padded, 0, 64), p1 = __shfl(padded, 16, 64);
      const uint32_t p2 = __shfl(padded, 32, 64), p3 = __shfl(padded, 48, 64);
      const uint32_t gi = lane >> 4;
      const uint32_t brick_off = (gi > 0 ? p0 : 0u) + (gi > 1 ? p1 : 0u) + (gi > 2 ? p2 : 0u);
      if (lane == 0) wsum[w] = (p0 + p1) + (p2 + p3);
      __syncthreads();
      if (t == 0) {
        uint32_t run = 0;
        for (int k = 0; k < 16; ++k) {
          woff[k] = run;
          run += wsum[k];
        }
        block_total = run;
      }
      __syncthreads();
      const uint32_t bstart = carry + woff[w] + brick_off;
      if (idx < ncode) {
        const uint32_t off = bstart + (incl - tsum);
        *reinterpret_cast<uint4*>(&start[base + idx]) = make_uint4(off, off + s0, off + s1, off + s2);
        *reinterpret_cast<uint4*>(&count[base + idx]) = make_uint4(0, 0, 0, 0);
      }
      // the brick's empty slots (< 64), 4 per thread
      const uint32_t pad_end = bstart + padded;
      for (uint32_t q = bstart + brick_total + 4 * g, e = 0; e < 4 && q < pad_end; ++e, ++q)
        perm[q] = kInvalid;
      carry += block_total;
      __syncthreads();
    }
  }
  if (t == 0) {
    ag_tile_start[n_ag] = carry / kWave;
    const uint32_t next = counter[0];
    counter[1] = next;
    counter[0] = next + 1u;
    *n_list = 0;  // the pair kernels' list of heavy tiles
  }
}

// perm[slot] = atom, plus sorted copies of the atom's position and parameters
// so that the pair kernel reads its tile without going through perm.
__global__ void k_sort_scatter(const uint32_t* __restrict__ code_in,
                               const uint32_t* __restrict__ rank, uint32_t n,
                               const uint32_t* __restrict__ start, uint32_t* __restrict__ perm,
                               const float4* __restrict__ pos, const float4* __restrict__ par,
                               float4* __restrict__ spos, float4* __restrict__ spar) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t code = code_in[j];
  if (code < kNonFinite) {
    const uint32_t slot = start[code] + rank[j];
    perm[slot] = j;
    spos[slot] = pos[j];
    spar[slot] = par[j];
  }
}

// ---- the pair tile kernel ---------------------------------------------------------

#define DNAB_EEF1_DG(dg, vol, lam) dg,
#define DNAB_EEF1_VOL(dg, vol, lam) vol,
#define DNAB_EEF1_LAM(dg, vol, lam) lam,
__constant__ float c_eef1_dg[kNumAtomTypesDev] = {DNAB_EEF1_ROWS(DNAB_EEF1_DG)};
__constant__ float c_eef1_vol[kNumAtomTypesDev] = {DNAB_EEF1_ROWS(DNAB_EEF1_VOL)};
__constant__ float c_eef1_lam[kNumAtomTypesDev] = {DNAB_EEF1_ROWS(DNAB_EEF1_LAM)};
#undef DNAB_EEF1_DG
#undef DNAB_EEF1_VOL
#undef DNAB_EEF1_LAM

// H-bond term of one pair (src/energy.rs:32-39), f32 in the reference's order.
__device__ inline float hbond_term(float r2) {
  const float r = sqrtf(r2);
  const float z = (r - kHbR0) / kHbSigma;
  return -kHbDepth * expf(-0.5f * z * z);
}

// EEF1 cross term of one pair (src/solvation.rs:34-44): i the antigen atom,
// j the binder atom; powi(2) and powi(3) as repeated products.
__device__ inline float solvation_term(uint32_t ti, uint32_t tj, float r2) {
  const float r = sqrtf(r2);
  const float li = c_eef1_lam[ti], lj = c_eef1_lam[tj];
  const float ui = r / li, uj = r / lj;
  const float fi = expf(-(ui * ui)) / (kPi32 * (li * li * li));
  const float fj = expf(-(uj * uj)) / (kPi32 * (lj * lj * lj));
  return -(c_eef1_dg[ti] * c_eef1_vol[tj] * fi + c_eef1_dg[tj] * c_eef1_vol[ti] * fj);
}

struct TileOut {
  float4* force;      // per atom (fx, fy, fz, sum |e|)
  double* e_atom;     // per atom energy
  double* bd_atom;    // per atom (lj, coulomb, hydrophobic), BREAKDOWN only
  unsigned long long* stats;  // (in cutoff, tested), COUNT only
};

__device__ inline uint32_t tile_antigen(const uint32_t* __restrict__ ag_tile_start, uint32_t n_ag,
                                        uint32_t t) {
  uint32_t lo = 0, hi = n_ag;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (ag_tile_start[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <bool FORCES, bool BREAKDOWN, bool COUNT>
__global__ void __launch_bounds__(256)
k_pair_tile(const AgView* __restrict__ views, uint32_t n_ag,
            const uint32_t* __restrict__ ag_tile_start, const uint32_t* __restrict__ perm,
            const float4* __restrict__ pos, const float4* __restrict__ par, TileOut out) {
  const uint32_t t = blockIdx.x;
  const uint32_t n_tiles = ag_tile_start[n_ag];
  if (t >= n_tiles) return;
  const uint32_t a_id = tile_antigen(ag_tile_start, n_ag, t);
  const AgAtom* __restrict__ atoms = views[a_id].atoms;
  const uint32_t* __restrict__ cstart = views[a_id].cell_start;
  const float inv = views[a_id].inv;
  const float rc = views[a_id].rc, rc2 = views[a_id].rc2;
  const int bits = views[a_id].bits;
  const int32_t kmin0 = views[a_id].kmin[0], kmin1 = views[a_id].kmin[1], kmin2 = views[a_id].kmin[2];
  const int32_t kmax0 = views[a_id].kmax[0], kmax1 = views[a_id].kmax[1], kmax2 = views[a_id].kmax[2];

  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t j = perm[t * kWave + lane];
  const bool valid = j != kInvalid;
  float4 P = make_float4(0.f, 0.f, 0.f, 0.f), Q = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) {
    P = pos[j];
    Q = par[j];
  }
  const float xb = P.x, yb = P.y, zb = P.z, qb = P.w;
  const float rmhb = Q.x;
  const float sqeb = sqrtf(Q.y);
  const float sqe12b = 12.0f * sqeb;
  const bool hydb = Q.z != 0.0f;
  const bool typed_b = Q.w >= 0.0f;  // the binder atom's AtomType is known
  const uint32_t tb = typed_b ? static_cast<uint32_t>(Q.w) : 0u;

  const float inf = __builtin_inff();
  const float lox = wave_reduce_f32<OpMin>(valid ? xb : inf);
  const float loy = wave_reduce_f32<OpMin>(valid ? yb : inf);
  const float loz = wave_reduce_f32<OpMin>(valid ? zb : inf);
  const float hix = wave_reduce_f32<OpMax>(valid ? xb : -inf);
  const float hiy = wave_reduce_f32<OpMax>(valid ? yb : -inf);
  const float hiz = wave_reduce_f32<OpMax>(valid ? zb : -inf);

  // Cells that can hold an antigen atom within rc of the tile box. cell_key
  // is monotone, so a slightly widened range is conservative.
  const float margin = rc * 1.001f + 1e-3f;
  const int32_t kx0 = max(cell_key(lox - margin, inv), kmin0);
  const int32_t ky0 = max(cell_key(loy - margin, inv), kmin1);
  const int32_t kz0 = max(cell_key(loz - margin, inv), kmin2);
  const int32_t kx1 = min(cell_key(hix + margin, inv), kmax0);
  const int32_t ky1 = min(cell_key(hiy + margin, inv), kmax1);
  const int32_t kz1 = min(cell_key(hiz + margin, inv), kmax2);

  float fx = 0.f, fy = 0.f, fz = 0.f;
  float es = 0.f, ec = 0.f, eabs = 0.f;        // TwoSum (sum, compensation), sum |e|
  float ljs = 0.f, ljc = 0.f, cos_ = 0.f, coc = 0.f;  // BREAKDOWN
  float hbs = 0.f, hbc = 0.f, svs = 0.f, svc = 0.f;   // BREAKDOWN: H-bond, solvation
  uint32_t nh = 0;
  uint32_t n_in = 0, n_tested = 0;
  (void)bits;

  // Cells of the box go to waves by (kx + ky + kz) & 3, a function of the
  // cell alone; each wave visits its cells in (kz, ky, kx) order. Lane l
  // loads the range of the box's l-th cell of the current batch of 64.
  const int32_t nx = kx1 - kx0 + 1, ny = ky1 - ky0 + 1, nz = kz1 - kz0 + 1;
  const int32_t ncell = (nx > 0 && ny > 0 && nz > 0) ? nx * ny * nz : 0;
  for (int32_t cbase = 0; cbase < ncell; cbase += kWave) {
    const int32_t c = cbase + static_cast<int32_t>(lane);
    uint32_t my_b = 0, my_e = 0;
    if (c < ncell) {
      const int32_t ix = c % nx, iy = (c / nx) % ny, iz = c / (nx * ny);
      const int32_t kx = kx0 + ix, ky = ky0 + iy, kz = kz0 + iz;
      if (((kx + ky + kz) & 3) == static_cast<int32_t>(wave)) {
        const uint32_t code = morton3(static_cast<uint32_t>(kx - kmin0),
                                      static_cast<uint32_t>(ky - kmin1),
                                      static_cast<uint32_t>(kz - kmin2));
        my_b = cstart[code];
        my_e = cstart[code + 1];
      }
    }
    uint64_t cmask = ballot64(my_e > my_b);
    while (cmask) {
      const int cl = static_cast<int>(__builtin_ctzll(cmask));
      cmask &= cmask - 1;
      const uint32_t cb = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(my_b), cl));
      const uint32_t ce = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(my_e), cl));
      {
        for (uint32_t base = cb; base < ce; base += kWave) {
          // One coalesced load of the chunk's atoms (one per lane); the pair
          // loop below broadcasts atom k from lane k with v_readlane.
          const uint32_t i = base + lane;
          float4 A0 = make_float4(0.f, 0.f, 0.f, 0.f), A1 = make_float4(0.f, 0.f, 0.f, 0.f);
          bool near = false;
          if (i < ce) {
            A0 = *reinterpret_cast<const float4*>(&atoms[i].x);
            A1 = *reinterpret_cast<const float4*>(&atoms[i].rmh);
            const float bx = fmaxf(fmaxf(lox - A0.x, A0.x - hix), 0.0f);
            const float by = fmaxf(fmaxf(loy - A0.y, A0.y - hiy), 0.0f);
            const float bz = fmaxf(fmaxf(loz - A0.z, A0.z - hiz), 0.0f);
            const float d2 = (bx * bx + by * by) + bz * bz;
            near = d2 <= rc2;
          }
          uint64_t mask = ballot64(near);
          while (mask) {
            const int k = static_cast<int>(__builtin_ctzll(mask));
            mask &= mask - 1;
            AgAtom A;
            A.x = readlane_f32(A0.x, k);
            A.y = readlane_f32(A0.y, k);
            A.z = readlane_f32(A0.z, k);
            A.q332 = readlane_f32(A0.w, k);
            A.rmh = readlane_f32(A1.x, k);
            A.sqe = readlane_f32(A1.y, k);
            A.tag = readlane_f32(A1.w, k);
            // Pair predicate in f32: the reference's skip test, negated
            // (src/energy.rs:279-284).
            const float dx = xb - A.x;
            const float dy = yb - A.y;
            const float dz = zb - A.z;
            const float r2 = (dx * dx + dy * dy) + dz * dz;
            const bool in = valid && !(r2 > rc2 || r2 < kMinRSq);
            if (COUNT) {
              n_tested += valid ? 1u : 0u;
              n_in += in ? 1u : 0u;
            }
            float rinv = __builtin_amdgcn_rsqf(r2);
            rinv = in ? rinv : 0.0f;
            const float rinv2 = rinv * rinv;
            const float R = A.rmh + rmhb;
            const float s2 = (R * R) * rinv2;
            const float s6 = (s2 * s2) * s2;
            const float s12 = s6 * s6;
            const float eps = A.sqe * sqeb;
            const float elj = eps * __builtin_fmaf(-2.0f, s6, s12);
            const float ecl = (A.q332 * qb) * rinv;
            float f = 0.0f;
            if (FORCES) {
              const float flj = ((A.sqe * sqe12b) * (s12 - s6)) * rinv2;
              f = __builtin_fmaf(ecl, rinv2, flj);
            }
            if (tag_hyd(A.tag)) {
              const bool h = in && hydb && r2 < kHydroSq;
              nh += h ? 1u : 0u;
              if (FORCES) f = h ? __builtin_fmaf(-0.1f, rinv, f) : f;
            }
            if (FORCES) {
              fx = __builtin_fmaf(f, dx, fx);
              fy = __builtin_fmaf(f, dy, fy);
              fz = __builtin_fmaf(f, dz, fz);
            }
            if (BREAKDOWN) {
              float s, e;
              two_sum(ljs, elj, s, e);
              ljs = s;
              ljc += e;
              two_sum(cos_, ecl, s, e);
              cos_ = s;
              coc += e;
              const int ta_i = tag_type(A.tag);
              if (ta_i >= 0 && typed_b) {  // both atom types known
                const uint32_t ta = static_cast<uint32_t>(ta_i);
                // H-bond inside the pair predicate (src/energy.rs:303-307)
                const bool hb_types = (((kHbondDonorMask >> tb) & (kHbondAcceptorMask >> ta)) |
                                       ((kHbondAcceptorMask >> tb) & (kHbondDonorMask >> ta))) & 1u;
                const float ehb = (in && hb_types && r2 <= kHbCutoffSq) ? hbond_term(r2) : 0.0f;
                two_sum(hbs, ehb, s, e);
                hbs = s;
                hbc += e;
                // EEF1 with its own skip test, negated (src/solvation.rs:71-78)
                const bool sv = valid && !(r2 >= kSolvCutoffSq || r2 < kMinRSq);
                const float esv = sv ? solvation_term(ta, tb, r2) : 0.0f;
                two_sum(svs, esv, s, e);
                svs = s;
                svc += e;
              }
            } else {
              const float ep = elj + ecl;
              float s, e;
              two_sum(es, ep, s, e);
              es = s;
              ec += e;
              eabs += fabsf(ep);
            }
          }
        }
      }
    }
  }

  // Combine the four waves' partial sums in wave order.
  __shared__ float sh_f[kTileWaves][3][kWave];
  __shared__ float sh_e[kTileWaves][8][kWave];
  __shared__ uint32_t sh_nh[kTileWaves][kWave];
  if (FORCES) {
    sh_f[wave][0][lane] = fx;
    sh_f[wave][1][lane] = fy;
    sh_f[wave][2][lane] = fz;
  }
  if (BREAKDOWN) {
    sh_e[wave][0][lane] = ljs;
    sh_e[wave][1][lane] = ljc;
    sh_e[wave][2][lane] = cos_;
    sh_e[wave][3][lane] = coc;
    sh_e[wave][4][lane] = hbs;
    sh_e[wave][5][lane] = hbc;
    sh_e[wave][6][lane] = svs;
    sh_e[wave][7][lane] = svc;
  } else {
    sh_e[wave][0][lane] = es;
    sh_e[wave][1][lane] = ec;
    sh_e[wave][2][lane] = eabs;
  }
  sh_nh[wave][lane] = nh;
  if (COUNT) {
    uint32_t a = wave_sum_u32(n_in), b = wave_sum_u32(n_tested);
    if (lane == 0) {
      atomicAdd(&out.stats[0], static_cast<unsigned long long>(a));
      atomicAdd(&out.stats[1], static_cast<unsigned long long>(b));
    }
  }
  __syncthreads();
  if (wave != 0 || !valid) return;
  uint32_t nh_t = 0;
  for (int w = 0; w < kTileWaves; ++w) nh_t += sh_nh[w][lane];
  const double ehyd = -0.5 * static_cast<double>(nh_t);
  float4 F = make_float4(0.f, 0.f, 0.f, 0.f);
  if (FORCES) {
    for (int w = 0; w < kTileWaves; ++w) {
      F.x += sh_f[w][0][lane];
      F.y += sh_f[w][1][lane];
      F.z += sh_f[w][2][lane];
    }
  }
  if (BREAKDOWN) {
    double lj = 0.0, co = 0.0, hb = 0.0, sv = 0.0;
    for (int w = 0; w < kTileWaves; ++w) {
      lj += static_cast<double>(sh_e[w][0][lane]) + static_cast<double>(sh_e[w][1][lane]);
      co += static_cast<double>(sh_e[w][2][lane]) + static_cast<double>(sh_e[w][3][lane]);
      hb += static_cast<double>(sh_e[w][4][lane]) +
            static_cast<double>(sh_e[w][5][lane]);
      sv += static_cast<double>(sh_e[w][6][lane]) +
            static_cast<double>(sh_e[w][7][lane]);
    }
    double* bd = out.bd_atom + static_cast<size_t>(kBdTerms) * j;
    bd[0] = lj;
    bd[1] = co;
    bd[2] = ehyd;
    bd[3] = hb;
    bd[4] = sv;
    // the reference's total(): lj + coulomb + hydrophobic + hbond + solvation
    out.e_atom[j] = (((lj + co) + ehyd) + hb) + sv;
  } else {
    float S = sh_e[0][0][lane], C = sh_e[0][1][lane], Aabs = sh_e[0][2][lane];
    for (int w = 1; w < kTileWaves; ++w) {
      float s, e;
      two_sum(S, sh_e[w][0][lane], s, e);
      S = s;
      C = (C + e) + sh_e[w][1][lane];
      Aabs += sh_e[w][2][lane];
    }
    out.e_atom[j] = (static_cast<double>(S) + static_cast<double>(C)) + ehyd;
    F.w = Aabs;
  }
  out.force[j] = F;
}

// ---- the hot pair kernels -----------------------------------------------------------
//
// Work unit. A tile is 64 sorted candidate atoms, one per lane. The antigen
// atoms near it are cut into chunks: a chunk is up to 64 consecutive atoms of
// one grid cell, a unit fixed by the antigen alone. For each chunk a lane
// forms FP32 partial sums over the chunk's atoms in order (TwoSum-compensated
// energy, forces, sum |e|, hydrophobic count), skipping atoms outside the tile
// box (their terms would be exactly zero), and rounds them once to integers
// in units of 2^-20 (kcal/mol, kcal/mol/A). Chunk results are added as 64-bit
// integers, which is exact and order-free: an atom's result does not depend
// on its tile, on which wave took which chunk, or on whether a chunk was
// handled in place or through the queue, so repeated and differently batched
// runs agree bit for bit.
//
// Balance. k_pair_chunks runs one wave per tile (four tiles per workgroup,
// so every tile of a step is resident at once); it loads chunks four at a
// time and sums them until it has taken its share (PairTuning: a number of
// chunks or of surviving antigen atoms). A tile deep inside the antigen has ~10x the
// median work: its wave hands the remaining chunks and its partial sums to
// k_pair_rest, where a workgroup of WAVES waves shares them. The two
// kernels meet through plain stores and loads in device memory (the kernel
// boundary orders them), with no device-scope atomics or fences, which on
// this GPU would reach across the XCDs' separate L2 caches.
// An atom's outputs from its integer sums: a component whose sum saturated
// is written as +-inf (fix_value), so the candidate's energy and status show it.
__device__ inline void write_fixed(const FixAcc& acc, uint32_t j, TileOut out) {
  out.e_atom[j] = fix_value(acc.e, acc.st, kStE);
  out.force[j] = make_float4(static_cast<float>(fix_value(acc.x, acc.st, kStX)),
                             static_cast<float>(fix_value(acc.y, acc.st, kStY)),
                             static_cast<float>(fix_value(acc.z, acc.st, kStZ)),
                             static_cast<float>(static_cast<double>(acc.a) * kFixUnit));
}

template <bool FORCES, bool COUNT, bool CLASH>
__global__ void __launch_bounds__(256)
k_pair_chunks(const AgView* __restrict__ views, uint32_t n_ag,
              const uint32_t* __restrict__ ag_tile_start, const uint32_t* __restrict__ perm,
              const float4* __restrict__ spos, const float4* __restrict__ spar, TileOut out,
              PairHandover ho, PairTuning tune) {
  __shared__ float4 s_atoms[kTileWaves][kWave][2];  // a wave's staged chunk atoms

  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t t = blockIdx.x * kTileWaves + wave;  // one tile per wave
  const uint32_t n_tiles = ag_tile_start[n_ag];
  if (t >= n_tiles) return;  // whole wave; no workgroup barriers below
  const uint32_t a_id = tile_antigen(ag_tile_start, n_ag, t);
  const AgView& V = views[a_id];
  const AgAtom* __restrict__ atoms = V.atoms;
  const uint32_t* __restrict__ cstart = V.cell_start;
  const float inv = V.inv, rc = V.rc, rc2 = V.rc2;
  const int32_t kmin0 = V.kmin[0], kmin1 = V.kmin[1], kmin2 = V.kmin[2];
  const int32_t kmax0 = V.kmax[0], kmax1 = V.kmax[1], kmax2 = V.kmax[2];

  const uint32_t slot_id = t * kWave + lane;
  const uint32_t j = perm[slot_id];
  const float4 Ps = spos[slot_id], Qs = spar[slot_id];
  const bool valid = j != kInvalid;
  const float4 P = valid ? Ps : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 Q = valid ? Qs : make_float4(0.f, 0.f, 0.f, 0.f);
  const float xb = P.x, yb = P.y, zb = P.z, qb = P.w;
  const float rmhb = Q.x;
  const float sqeb = sqrtf(Q.y);
  const float sqe12b = 12.0f * sqeb;
  const float epsb = Q.y;
  const bool hydb = Q.z != 0.0f;
  const float rc2_lane = valid ? rc2 : -1.0f;  // an empty slot has no pair

  const float inf = __builtin_inff();
  const float lox = wave_reduce_f32<OpMin>(valid ? xb : inf);
  const float loy = wave_reduce_f32<OpMin>(valid ? yb : inf);
  const float loz = wave_reduce_f32<OpMin>(valid ? zb : inf);
  const float hix = wave_reduce_f32<OpMax>(valid ? xb : -inf);
  const float hiy = wave_reduce_f32<OpMax>(valid ? yb : -inf);
  const float hiz = wave_reduce_f32<OpMax>(valid ? zb : -inf);
  // cells that can hold an antigen atom within rc of the tile box (cell_key
  // is monotone, so a slightly widened range is conservative)
  const float margin = rc * 1.001f + 1e-3f;
  const int32_t kx0 = max(cell_key(lox - margin, inv), kmin0);
  const int32_t ky0 = max(cell_key(loy - margin, inv), kmin1);
  const int32_t kz0 = max(cell_key(loz - margin, inv), kmin2);
  const int32_t kx1 = min(cell_key(hix + margin, inv), kmax0);
  const int32_t ky1 = min(cell_key(hiy + margin, inv), kmax1);
  const int32_t kz1 = min(cell_key(hiz + margin, inv), kmax2);
  const int32_t nx = kx1 - kx0 + 1, ny = ky1 - ky0 + 1, nz = kz1 - kz0 + 1;
  const int32_t nxy = nx * ny;
  const int32_t ncell = (nx > 0 && ny > 0 && nz > 0) ? nxy * nz : 0;

  FixAcc acc;
  uint32_t n_in = 0, n_tested = 0, taken_atoms = 0, taken_chunks = 0, handed = 0;
  bool handing = false;

  for (int32_t cb0 = 0; cb0 < ncell; cb0 += kWave) {
    // lane = a cell of this batch, in (kz, ky, kx) order
    const int32_t c = cb0 + static_cast<int32_t>(lane);
    uint32_t b = 0, e = 0;
    if (c < ncell) {
      const int32_t iz = c / nxy, rem = c - iz * nxy;
      const int32_t iy = rem / nx, ix = rem - iy * nx;
      const uint32_t code = morton3(static_cast<uint32_t>(kx0 + ix - kmin0),
                                    static_cast<uint32_t>(ky0 + iy - kmin1),
                                    static_cast<uint32_t>(kz0 + iz - kmin2));
      b = cstart[code];
      e = cstart[code + 1];
    }
    const uint32_t nch = (e - b + kWave - 1) / kWave;
    uint32_t incl = nch;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(incl, o, 64);
      if (lane >= static_cast<uint32_t>(o)) incl += u;
    }
    const uint32_t off = incl - nch;
    const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
    // chunks [k, total) of this batch go to the rest kernel when the tile
    // has had its share and the handover table has room
    uint32_t k = 0, k_hand = total;  // chunks [k_hand, total) are handed over
    for (; k < total && k_hand == total; k += kLoadAhead) {
      if (taken_atoms >= tune.direct_survivors || taken_chunks >= tune.direct_chunks) handing = true;
      if (handing && handed + (total - k) <= kHandoverCap) {
        k_hand = k;
        break;
      }
      uint32_t cbv[kLoadAhead], cev[kLoadAhead];
      float4 L0[kLoadAhead], L1[kLoadAhead];
      #pragma unroll
      for (int u = 0; u < kLoadAhead; ++u) {
        const uint32_t kk = k + u;
        cbv[u] = 0;
        cev[u] = 0;
        L0[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        L1[u] = L0[u];
        if (kk < total) {
          const uint64_t m = ballot64(nch > 0 && off <= kk);
          const int cl = 63 - static_cast<int>(__builtin_clzll(m));
          cbv[u] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), cl)) +
                   kWave * (kk - static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(off), cl)));
          cev[u] = min(cbv[u] + kWave, static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(e), cl)));
          if (cbv[u] + lane < cev[u]) {
            L0[u] = *reinterpret_cast<const float4*>(&atoms[cbv[u] + lane].x);
            L1[u] = *reinterpret_cast<const float4*>(&atoms[cbv[u] + lane].rmh);
          }
        }
      }
      #pragma unroll
      for (int u = 0; u < kLoadAhead; ++u) {
        if (cev[u] <= cbv[u]) continue;
        // stop at this chunk once the tile has had its share (and the rest fits)
        if ((taken_atoms >= tune.direct_survivors || taken_chunks >= tune.direct_chunks) &&
            handed + (total - (k + u)) <= kHandoverCap) {
          handing = true;
          k_hand = k + u;
          break;
        }
        ++taken_chunks;
        const float bx = fmaxf(fmaxf(lox - L0[u].x, L0[u].x - hix), 0.0f);
        const float by = fmaxf(fmaxf(loy - L0[u].y, L0[u].y - hiy), 0.0f);
        const float bz = fmaxf(fmaxf(loz - L0[u].z, L0[u].z - hiz), 0.0f);
        const bool near = cbv[u] + lane < cev[u] && ((bx * bx + by * by) + bz * bz) <= rc2;
        const uint64_t mask = ballot64(near);
        const uint32_t n_s = static_cast<uint32_t>(__builtin_popcountll(mask));
        if (n_s == 0) continue;
        taken_atoms += n_s;
        stage_chunk(s_atoms[wave], mask, lane, L0[u], L1[u]);
        chunk_pairs<FORCES, COUNT, CLASH>(s_atoms[wave], n_s, xb, yb, zb, qb, rmhb, sqeb, sqe12b, epsb,
                                   hydb, rc2_lane, valid, acc, n_in, n_tested);
      }
    }
    if (k_hand < total) {
      // lane = cell writes its chunks with ordinal >= k_hand, in order
      for (uint32_t m = 0; m < nch; ++m) {
        const uint32_t o = off + m;
        if (o >= k_hand) {
          const uint32_t first = b + kWave * m;
          ho.chunks[t * kHandoverCap + handed + (o - k_hand)] =
              make_uint2(first, min(static_cast<uint32_t>(kWave), e - first));
        }
      }
      handed += total - k_hand;
    }
  }

  if (COUNT) {
    const uint32_t ai = wave_sum_u32(n_in), at = wave_sum_u32(n_tested);
    if (lane == 0 && (ai | at)) {
      atomicAdd(&out.stats[0], static_cast<unsigned long long>(ai));
      atomicAdd(&out.stats[1], static_cast<unsigned long long>(at));
    }
  }
  if (lane == 0) {
    ho.rest[t] = handed;
    if (handed) {  // once per tile and launch; the bound guards a missed reset
      const uint32_t idx = atomicAdd(ho.n_list, 1u);
      if (idx < ho.n_slots / kWave) ho.list[idx] = t;
    }
  }
  if (handed == 0) {
    if (valid) write_fixed(acc, j, out);
    return;
  }
  const size_t ns = ho.n_slots;
  ho.part[0 * ns + slot_id] = acc.e;
  ho.part[1 * ns + slot_id] = acc.a;
  ho.part[2 * ns + slot_id] = acc.x;
  ho.part[3 * ns + slot_id] = acc.y;
  ho.part[4 * ns + slot_id] = acc.z;
  ho.status[slot_id] = acc.st;
}

// The chunks a tile's wave handed over: WAVES waves take them from an
// LDS counter, sum per chunk into fixed-point integers, combine in LDS, add
// the wave's partial sums and write the tile's outputs.
template <bool FORCES, bool COUNT, bool CLASH, int WAVES>
__global__ void __launch_bounds__(WAVES * 64)
k_pair_rest(const AgView* __restrict__ views, uint32_t n_ag,
            const uint32_t* __restrict__ ag_tile_start, const uint32_t* __restrict__ perm,
            const float4* __restrict__ spos, const float4* __restrict__ spar, TileOut out,
            PairHandover ho) {
  __shared__ float4 s_atoms[WAVES][kWave][2];
  __shared__ long long s_sum[5][kWave];
  __shared__ uint32_t s_st[kWave];
  __shared__ uint32_t s_next;
  const uint32_t n_heavy = min(*ho.n_list, ho.n_slots / kWave);
  for (uint32_t h = blockIdx.x; h < n_heavy; h += gridDim.x) {  // workgroup-uniform
  const uint32_t t = ho.list[h];
  const uint32_t n = ho.rest[t];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const AgView& V = views[tile_antigen(ag_tile_start, n_ag, t)];
  const AgAtom* __restrict__ atoms = V.atoms;
  const float rc2 = V.rc2;
  const uint32_t slot_id = t * kWave + lane;
  const uint32_t j = perm[slot_id];
  const float4 Ps = spos[slot_id], Qs = spar[slot_id];
  const bool valid = j != kInvalid;
  const float4 P = valid ? Ps : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 Q = valid ? Qs : make_float4(0.f, 0.f, 0.f, 0.f);
  const float xb = P.x, yb = P.y, zb = P.z, qb = P.w;
  const float rmhb = Q.x;
  const float sqeb = sqrtf(Q.y);
  const float sqe12b = 12.0f * sqeb;
  const float epsb = Q.y;
  const bool hydb = Q.z != 0.0f;
  const float rc2_lane = valid ? rc2 : -1.0f;
  const float inf = __builtin_inff();
  const float lox = wave_reduce_f32<OpMin>(valid ? xb : inf);
  const float loy = wave_reduce_f32<OpMin>(valid ? yb : inf);
  const float loz = wave_reduce_f32<OpMin>(valid ? zb : inf);
  const float hix = wave_reduce_f32<OpMax>(valid ? xb : -inf);
  const float hiy = wave_reduce_f32<OpMax>(valid ? yb : -inf);
  const float hiz = wave_reduce_f32<OpMax>(valid ? zb : -inf);
  __syncthreads();  // the previous tile's LDS is no longer read
  for (uint32_t i = tid; i < 5 * kWave; i += WAVES * kWave) s_sum[i >> 6][i & 63u] = 0;
  if (tid < kWave) s_st[tid] = 0;
  if (tid == 0) s_next = 0;
  __syncthreads();

  FixAcc acc;
  uint32_t n_in = 0, n_tested = 0;
  auto grab = [&]() -> uint32_t {
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(&s_next, 1u);
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(k)));
  };
  const uint2* table = ho.chunks + static_cast<size_t>(t) * kHandoverCap;
  uint32_t k = grab();
  uint2 ch = k < n ? table[k] : make_uint2(0, 0);
  while (k < n) {
    float4 A0 = make_float4(0.f, 0.f, 0.f, 0.f), A1 = A0;
    if (lane < ch.y) {
      A0 = *reinterpret_cast<const float4*>(&atoms[ch.x + lane].x);
      A1 = *reinterpret_cast<const float4*>(&atoms[ch.x + lane].rmh);
    }
    const uint32_t k2 = grab();                     // next chunk's table entry,
    const uint2 ch2 = k2 < n ? table[k2] : make_uint2(0, 0);  // in flight meanwhile
    const float bx = fmaxf(fmaxf(lox - A0.x, A0.x - hix), 0.0f);
    const float by = fmaxf(fmaxf(loy - A0.y, A0.y - hiy), 0.0f);
    const float bz = fmaxf(fmaxf(loz - A0.z, A0.z - hiz), 0.0f);
    const bool near = lane < ch.y && ((bx * bx + by * by) + bz * bz) <= rc2;
    const uint64_t mask = ballot64(near);
    const uint32_t n_s = static_cast<uint32_t>(__builtin_popcountll(mask));
    if (n_s) {
      stage_chunk(s_atoms[wave], mask, lane, A0, A1);
      chunk_pairs<FORCES, COUNT, CLASH>(s_atoms[wave], n_s, xb, yb, zb, qb, rmhb, sqeb, sqe12b, epsb,
                                 hydb, rc2_lane, valid, acc, n_in, n_tested);
    }
    k = k2;
    ch = ch2;
  }
  // each wave's sums are at most 2^59 in size: eight of them and the partial stay below 2^63
  atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum[0][lane]), static_cast<unsigned long long>(acc.e));
  atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum[1][lane]), static_cast<unsigned long long>(acc.a));
  if (FORCES) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum[2][lane]), static_cast<unsigned long long>(acc.x));
    atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum[3][lane]), static_cast<unsigned long long>(acc.y));
    atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum[4][lane]), static_cast<unsigned long long>(acc.z));
  }
  if (acc.st) atomicOr(&s_st[lane], acc.st);
  if (COUNT) {
    const uint32_t ai = wave_sum_u32(n_in), at = wave_sum_u32(n_tested);
    if (lane == 0 && (ai | at)) {
      atomicAdd(&out.stats[0], static_cast<unsigned long long>(ai));
      atomicAdd(&out.stats[1], static_cast<unsigned long long>(at));
    }
  }
  __syncthreads();
  if (wave == 0 && valid) {
    const size_t ns = ho.n_slots;
    FixAcc tot;
    tot.e = s_sum[0][lane] + ho.part[0 * ns + slot_id];
    tot.a = s_sum[1][lane] + ho.part[1 * ns + slot_id];
    tot.x = s_sum[2][lane] + ho.part[2 * ns + slot_id];
    tot.y = s_sum[3][lane] + ho.part[3 * ns + slot_id];
    tot.z = s_sum[4][lane] + ho.part[4 * ns + slot_id];
    tot.st = s_st[lane] | ho.status[slot_id];
    uint32_t ignore = 0;
    tot.e = fix_add(tot.e, 0, tot.st, kStE);
    tot.a = fix_add(tot.a, 0, ignore, 0u);
    tot.x = fix_add(tot.x, 0, tot.st, kStX);
    tot.y = fix_add(tot.y, 0, tot.st, kStY);
    tot.z = fix_add(tot.z, 0, tot.st, kStZ);
    write_fixed(tot, j, out);
  }
  }  // next heavy tile
}

__global__ void k_pair_list_reset(uint32_t* n_list) {
  if (threadIdx.x == 0) *n_list = 0;
}

// All-FP64 oracle kernel: the same pair set (decided by the f32 predicate)
// with every term, mixing rule and sum in double from the raw per-atom
// parameters. Writes per atom (fx, fy, fz) and (lj, coulomb, hydrophobic).
__global__ void __launch_bounds__(256)
k_pair_tile_f64(const AgView* __restrict__ views, uint32_t n_ag,
                const uint32_t* __restrict__ ag_tile_start, const uint32_t* __restrict__ perm,
                const float4* __restrict__ pos, const float4* __restrict__ par,
                double* __restrict__ f64_atom, double* __restrict__ bd_atom,
                double* __restrict__ e_atom) {
  const uint32_t t = blockIdx.x;
  const uint32_t n_tiles = ag_tile_start[n_ag];
  if (t >= n_tiles) return;
  const uint32_t a_id = tile_antigen(ag_tile_start, n_ag, t);
  const AgView& V = views[a_id];
  const AgAtom* __restrict__ atoms = V.atoms;
  const float2* __restrict__ raw = V.raw;
  const uint32_t* __restrict__ cstart = V.cell_start;
  const float inv = V.inv, rc = V.rc, rc2 = V.rc2;
  const int32_t kmin0 = V.kmin[0], kmin1 = V.kmin[1], kmin2 = V.kmin[2];
  const int32_t kmax0 = V.kmax[0], kmax1 = V.kmax[1], kmax2 = V.kmax[2];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t j = perm[t * kWave + lane];
  const bool valid = j != kInvalid;
  float4 P = make_float4(0.f, 0.f, 0.f, 0.f), Q = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) {
    P = pos[j];
    Q = par[j];
  }
  const float xb = P.x, yb = P.y, zb = P.z;
  const double qb = P.w, rmhb = Q.x, epsb = Q.y;
  const bool hydb = Q.z != 0.0f;
  const float inf = __builtin_inff();
  const float lox = wave_reduce_f32<OpMin>(valid ? xb : inf);
  const float loy = wave_reduce_f32<OpMin>(valid ? yb : inf);
  const float loz = wave_reduce_f32<OpMin>(valid ? zb : inf);
  const float hix = wave_reduce_f32<OpMax>(valid ? xb : -inf);
  const float hiy = wave_reduce_f32<OpMax>(valid ? yb : -inf);
  const float hiz = wave_reduce_f32<OpMax>(valid ? zb : -inf);
  const float margin = rc * 1.001f + 1e-3f;
  const int32_t kx0 = max(cell_key(lox - margin, inv), kmin0);
  const int32_t ky0 = max(cell_key(loy - margin, inv), kmin1);
  const int32_t kz0 = max(cell_key(loz - margin, inv), kmin2);
  const int32_t kx1 = min(cell_key(hix + margin, inv), kmax0);
  const int32_t ky1 = min(cell_key(hiy + margin, inv), kmax1);
  const int32_t kz1 = min(cell_key(hiz + margin, inv), kmax2);
  double fx = 0, fy = 0, fz = 0, lj = 0, co = 0;
  uint32_t nh = 0;
  for (int32_t kz = kz0; kz <= kz1; ++kz)
    for (int32_t ky = ky0; ky <= ky1; ++ky)
      for (int32_t kx = kx0; kx <= kx1; ++kx) {
        const uint32_t code = morton3(static_cast<uint32_t>(kx - kmin0), static_cast<uint32_t>(ky - kmin1),
                                      static_cast<uint32_t>(kz - kmin2));
        const uint32_t cb = cstart[code], ce = cstart[code + 1];
        for (uint32_t base = cb; base < ce; base += kWave) {
          const uint32_t i = base + lane;
          bool near = false;
          if (i < ce) {
            const float4 A = *reinterpret_cast<const float4*>(&atoms[i]);
            const float bx = fmaxf(fmaxf(lox - A.x, A.x - hix), 0.0f);
            const float by = fmaxf(fmaxf(loy - A.y, A.y - hiy), 0.0f);
            const float bz = fmaxf(fmaxf(loz - A.z, A.z - hiz), 0.0f);
            near = (bx * bx + by * by) + bz * bz <= rc2 && ((i & 3u) == wave);
          }
          uint64_t mask = ballot64(near);
          while (mask) {
            const uint32_t k = static_cast<uint32_t>(__builtin_ctzll(mask));
            mask &= mask - 1;
            const AgAtom A = atoms[base + k];
            const float dx = xb - A.x, dy = yb - A.y, dz = zb - A.z;
            const float r2 = (dx * dx + dy * dy) + dz * dz;
            if (!valid || r2 > rc2 || r2 < kMinRSq) continue;
            const float2 rw = raw[base + k];
            const double ddx = static_cast<double>(xb) - static_cast<double>(A.x);
            const double ddy = static_cast<double>(yb) - static_cast<double>(A.y);
            const double ddz = static_cast<double>(zb) - static_cast<double>(A.z);
            const double r2d = ddx * ddx + ddy * ddy + ddz * ddz;
            const double r = sqrt(r2d);
            const double R = static_cast<double>(A.rmh) + rmhb;
            const double eps = sqrt(static_cast<double>(rw.y) * epsb);
            const double s2 = (R * R) / r2d;
            const double s6 = s2 * s2 * s2;
            const double s12 = s6 * s6;
            lj += eps * (s12 - 2.0 * s6);
            double f = 12.0 * eps * (s12 - s6) / r2d;
            const double ecl = 332.0 * qb * static_cast<double>(rw.x) / r;
            co += ecl;
            f += ecl / r2d;
            if (tag_hyd(A.tag) && hydb && r2 < kHydroSq) {
              ++nh;
              f += -0.1 / r;
            }
            fx += f * ddx;
            fy += f * ddy;
            fz += f * ddz;
          }
        }
      }
  __shared__ double sh[kTileWaves][5][kWave];
  __shared__ uint32_t sh_nh[kTileWaves][kWave];
  sh[wave][0][lane] = fx;
  sh[wave][1][lane] = fy;
  sh[wave][2][lane] = fz;
  sh[wave][3][lane] = lj;
  sh[wave][4][lane] = co;
  sh_nh[wave][lane] = nh;
  __syncthreads();
  if (wave != 0 || !valid) return;
  double acc[5] = {0, 0, 0, 0, 0};
  uint32_t nh_t = 0;
  for (int w = 0; w < kTileWaves; ++w) {
    for (int c = 0; c < 5; ++c) acc[c] += sh[w][c][lane];
    nh_t += sh_nh[w][lane];
  }
  const size_t jj = j;
  f64_atom[3 * jj] = acc[0];
  f64_atom[3 * jj + 1] = acc[1];
  f64_atom[3 * jj + 2] = acc[2];
  double* bd = bd_atom + static_cast<size_t>(kBdTerms) * jj;
  bd[0] = acc[3];
  bd[1] = acc[4];
  bd[2] = -0.5 * static_cast<double>(nh_t);
  bd[3] = 0.0;  // the FP64 kernel checks the three terms of the E+F path
  bd[4] = 0.0;
  e_atom[jj] = (acc[3] + acc[4]) + (-0.5 * static_cast<double>(nh_t));
}

// ---- per-candidate reductions (FP64, DPP) ----------------------------------------

// a-priori error estimate of an FP32 candidate energy: 32 ulp(1) per unit of
// sum |e_pair| (rsqrt, products and the compensated sum each a few ulp).
constexpr float kErrPerAbs = 32.0f / 16777216.0f;

__global__ void __launch_bounds__(256)
k_reduce_energy(const uint32_t* __restrict__ cand_atom_off, uint32_t n_cand,
                const double* __restrict__ e_atom, const float4* __restrict__ force,
                int with_abs, const uint8_t* __restrict__ cand_mask,
                double* __restrict__ energy, float* __restrict__ err) {
  const uint32_t c = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  if (c >= n_cand) return;  // whole wave exits together
  if (cand_mask && !cand_mask[c]) {  // not scored: ranks last
    if (lane == 0) {
      energy[c] = __builtin_huge_val();
      err[c] = 0.0f;
    }
    return;
  }
  const uint32_t b = cand_atom_off[c], e = cand_atom_off[c + 1];
  double s = 0.0, a = 0.0;
  for (uint32_t k = b + lane; k < e; k += kWave) {
    s += e_atom[k];
    if (with_abs) a += static_cast<double>(force[k].w);
  }
  s = wave_sum_f64(s);
  a = wave_sum_f64(a);
  if (lane == 0) {
    energy[c] = s;
    err[c] = with_abs ? static_cast<float>(a) * kErrPerAbs : 0.0f;
  }
}

__global__ void __launch_bounds__(256)
k_reduce_breakdown(const uint32_t* __restrict__ cand_atom_off, uint32_t n_cand,
                   const double* __restrict__ bd_atom, double* __restrict__ bd) {
  const uint32_t c = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  if (c >= n_cand) return;
  const uint32_t b = cand_atom_off[c], e = cand_atom_off[c + 1];
  double s[kBdTerms] = {0, 0, 0, 0, 0};
  for (uint32_t k = b + lane; k < e; k += kWave)
    for (int t = 0; t < kBdTerms; ++t) s[t] += bd_atom[static_cast<size_t>(kBdTerms) * k + t];
  for (int t = 0; t < kBdTerms; ++t) s[t] = wave_sum_f64(s[t]);
  if (lane == 0)
    for (int t = 0; t < kBdTerms; ++t) bd[static_cast<size_t>(kBdTerms) * c + t] = s[t];
}

// disulfide_energy (src/allatom.rs:155-180) of each candidate: Gaussian wells
// over the pairs of SG atoms (topology slot 5) of its Cys residues, pairs in
// residue order, f32 in the reference's order. One thread per candidate: a
// binder has few Cys. Without residue types the term is 0.
__global__ void __launch_bounds__(256)
k_disulfide(const uint32_t* __restrict__ cand_res_off, const uint32_t* __restrict__ res_atom_off,
            const uint8_t* __restrict__ res_aa, const float4* __restrict__ pos, uint32_t n_cand,
            double* __restrict__ out) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  float e = 0.0f;
  if (res_aa) {
    const uint32_t r0 = cand_res_off[c], r1 = cand_res_off[c + 1];
    auto has_sg = [&](uint32_t r) {
      return res_aa[r] == kCys && res_atom_off[r + 1] - res_atom_off[r] > kCysSgSlot;
    };
    for (uint32_t ri = r0; ri < r1; ++ri) {
      if (!has_sg(ri)) continue;
      const float4 a = pos[res_atom_off[ri] + kCysSgSlot];
      for (uint32_t rj = ri + 1; rj < r1; ++rj) {
        if (!has_sg(rj)) continue;
        const float4 b = pos[res_atom_off[rj] + kCysSgSlot];
        const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        const float d = sqrtf(dx * dx + dy * dy + dz * dz);
        if (d > kSsCutoff) continue;
        const float z = (d - kSsD0) / kSsSigma;
        e -= kSsDepth * expf(-0.5f * z * z);
      }
    }
  }
  out[c] = static_cast<double>(e);
}

// ---- host side ------------------------------------------------------------------

// Drops counts made for positions or a mask that no longer hold; the scan
// would otherwise consume them. Never part of a captured step loop.
int discard_pending_counts(dnab_ctx* ctx, dnab_population* pop) {
  if (!pop->counts_pending) return DNAB_OK;
  const uint64_t n = pop->sort_codes;
  if (n) {
    hipLaunchKernelGGL(k_fill_u32b, dim3(nblocks(n, 256) > 1024 ? 1024 : nblocks(n, 256)),
                       dim3(256), 0, ctx->stream, pop->sort_count.p, n, 0u);
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  pop->counts_pending = false;
  return DNAB_OK;
}

int population_count(dnab_ctx* ctx, dnab_population* pop) {
  if (pop->counts_pending) return DNAB_OK;
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  if (nc) {
    hipLaunchKernelGGL(k_sort_count, dim3(nblocks(nc, 4)), dim3(256), 0, ctx->stream,
                       pop->views.p, pop->cand_ag.p, pop->cand_atom_off.p, nc,
                       pop->masked ? pop->cand_mask.p : nullptr, pop->pos.p, pop->sort_count.p,
                       pop->sort_code.p, pop->sort_rank.p, pop->force.p, pop->e_atom.p,
                       pop->bd_atom.p, pop->f64_atom.p);
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  pop->counts_pending = true;
  return DNAB_OK;
}

int population_sort(dnab_ctx* ctx, dnab_population* pop) {
  const uint32_t n = static_cast<uint32_t>(pop->n_atoms);
  // pending counts were made by the step kernel for all candidates
  if (pop->masked) DNAB_RETURN_IF(discard_pending_counts(ctx, pop));
  DNAB_RETURN_IF(population_count(ctx, pop));
  hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, ctx->stream, pop->views.p,
                     static_cast<uint32_t>(pop->n_ag), pop->sort_count.p, pop->sort_start.p,
                     pop->perm.p, pop->ag_tile_start.p, pop->step_counter.p,
                     pop->ho_list.p + pop->tile_cap);
  pop->ho_dirty = false;
  if (n) {
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream,
                       pop->sort_code.p, pop->sort_rank.p, n, pop->sort_start.p, pop->perm.p,
                       pop->pos.p, pop->par.p, pop->spos.p, pop->spar.p);
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  pop->counts_pending = false;
  pop->sorted = true;
  return DNAB_OK;
}

int population_evaluate(dnab_ctx* ctx, dnab_population* pop, unsigned flags) {
  if (!pop->sorted) DNAB_RETURN_IF(population_sort(ctx, pop));
  DNAB_RETURN_IF(population_pair(ctx, pop, flags));
  return population_reduce_energies(ctx, pop, flags);
}

static PairHandover pair_handover(dnab_population* pop) {
  PairHandover h;
  h.rest = pop->ho_rest.p;
  h.list = pop->ho_list.p;
  h.n_list = pop->ho_list.p + pop->tile_cap;
  h.chunks = pop->ho_chunks.p;
  h.part = pop->ho_part.p;
  h.status = pop->ho_status.p;
  h.n_slots = static_cast<uint32_t>(pop->tile_cap * kWave);
  return h;
}

// CLASH: clash pairs in FP64 (evaluations whose results are reported);
// without it every pair term is FP32 (the per-step loop, as the fused one).
template <bool FORCES, bool COUNT, bool CLASH>
static void launch_pair(dnab_ctx* ctx, dnab_population* pop, const TileOut& o) {
  const uint32_t n_ag = static_cast<uint32_t>(pop->n_ag);
  const PairHandover h = pair_handover(pop);
  const uint32_t cap = static_cast<uint32_t>(pop->tile_cap ? pop->tile_cap : 1);
  if (pop->ho_dirty) hipLaunchKernelGGL(k_pair_list_reset, dim3(1), dim3(64), 0, ctx->stream, h.n_list);
  pop->ho_dirty = true;  // cleared by the next sort
  hipLaunchKernelGGL((k_pair_chunks<FORCES, COUNT, CLASH>), dim3((cap + kTileWaves - 1) / kTileWaves),
                     dim3(256), 0, ctx->stream, pop->views.p, n_ag, pop->ag_tile_start.p,
                     pop->perm.p, pop->spos.p, pop->spar.p, o, h, pop->tune);
  const dim3 rest_grid(static_cast<uint32_t>(ctx->cu_count * 2));
  switch (pop->tune.rest_waves) {
    case 4:
      hipLaunchKernelGGL((k_pair_rest<FORCES, COUNT, CLASH, 4>), rest_grid, dim3(4 * 64), 0, ctx->stream,
                         pop->views.p, n_ag, pop->ag_tile_start.p, pop->perm.p, pop->spos.p,
                         pop->spar.p, o, h);
      break;
    case 16:
      hipLaunchKernelGGL((k_pair_rest<FORCES, COUNT, CLASH, 16>), rest_grid, dim3(16 * 64), 0, ctx->stream,
                         pop->views.p, n_ag, pop->ag_tile_start.p, pop->perm.p, pop->spos.p,
                         pop->spar.p, o, h);
      break;
    default:
      hipLaunchKernelGGL((k_pair_rest<FORCES, COUNT, CLASH, 8>), rest_grid, dim3(8 * 64), 0, ctx->stream,
                         pop->views.p, n_ag, pop->ag_tile_start.p, pop->perm.p, pop->spos.p,
                         pop->spar.p, o, h);
  }
}

int population_pair(dnab_ctx* ctx, dnab_population* pop, unsigned flags) {
  const uint32_t n_ag = static_cast<uint32_t>(pop->n_ag);
  const dim3 grid(static_cast<uint32_t>(pop->tile_cap ? pop->tile_cap : 1)), block(256);
  TileOut o{pop->force.p, pop->e_atom.p, pop->bd_atom.p, pop->pair_stats.p};
  if (flags & DNAB_EVAL_FP64) {
    hipLaunchKernelGGL(k_pair_tile_f64, grid, block, 0, ctx->stream, pop->views.p, n_ag,
                       pop->ag_tile_start.p, pop->perm.p, pop->pos.p, pop->par.p,
                       pop->f64_atom.p, pop->bd_atom.p, pop->e_atom.p);
  } else if (flags & DNAB_EVAL_BREAKDOWN) {
    hipEvent_t ev_end;
    hipEvent_t ev = timer_begin(ctx, "pair_breakdown", &ev_end);
    hipLaunchKernelGGL((k_pair_tile<false, true, false>), grid, block, 0, ctx->stream,
                       pop->views.p, n_ag, pop->ag_tile_start.p, pop->perm.p, pop->pos.p,
                       pop->par.p, o);
    timer_end(ctx, "pair_breakdown", ev, ev_end);
  } else if (flags & kEvalCount) {
    if (flags & kEvalLoop) launch_pair<true, true, false>(ctx, pop, o);
    else launch_pair<true, true, true>(ctx, pop, o);
  } else if (flags & DNAB_EVAL_FORCES) {
    hipEvent_t ev_end;
    hipEvent_t ev = timer_begin(ctx, "pair_ef", &ev_end);
    if (flags & kEvalLoop) launch_pair<true, false, false>(ctx, pop, o);
    else launch_pair<true, false, true>(ctx, pop, o);
    timer_end(ctx, "pair_ef", ev, ev_end);
  } else {
    hipEvent_t ev_end;
    hipEvent_t ev = timer_begin(ctx, "pair_e", &ev_end);
    launch_pair<false, false, true>(ctx, pop, o);
    timer_end(ctx, "pair_e", ev, ev_end);
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  pop->last_flags = ((flags & kEvalCount) ? (flags | DNAB_EVAL_FORCES) : flags) & ~(kEvalCount | kEvalLoop);
  pop->evaluated = true;
  return DNAB_OK;
}

int population_reduce_energies(dnab_ctx* ctx, dnab_population* pop, unsigned flags) {
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  if (nc) {
    const int with_abs = (flags & (DNAB_EVAL_FP64 | DNAB_EVAL_BREAKDOWN)) ? 0 : 1;
    hipLaunchKernelGGL(k_reduce_energy, dim3(nblocks(nc, 4)), dim3(256), 0, ctx->stream,
                       pop->cand_atom_off.p, nc, pop->e_atom.p, pop->force.p, with_abs,
                       pop->masked ? pop->cand_mask.p : nullptr, pop->energy.p,
                       pop->err_bound.p);
    if (flags & (DNAB_EVAL_FP64 | DNAB_EVAL_BREAKDOWN)) {
      hipLaunchKernelGGL(k_reduce_breakdown, dim3(nblocks(nc, 4)), dim3(256), 0, ctx->stream,
                         pop->cand_atom_off.p, nc, pop->bd_atom.p, pop->breakdown.p);
    }
    if (flags & DNAB_EVAL_BREAKDOWN) {
      hipLaunchKernelGGL(k_disulfide, dim3(nblocks(nc, 256)), dim3(256), 0, ctx->stream,
                         pop->cand_res_off.p, pop->res_atom_off.p,
                         pop->has_res_aa ? pop->res_aa.p : nullptr, pop->pos.p, nc,
                         pop->disulfide.p);
    }
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  pop->last_flags = ((flags & kEvalCount) ? (flags | DNAB_EVAL_FORCES) : flags) & ~(kEvalCount | kEvalLoop);
  pop->evaluated = true;
  return DNAB_OK;
}

static int pair_count(dnab_ctx* ctx, dnab_population* pop, uint64_t* in_cutoff,
                      uint64_t* tested) {
  if (!pop->sorted) DNAB_RETURN_IF(population_sort(ctx, pop));
  unsigned long long zero[2] = {0, 0};
  DNAB_RETURN_IF(upload(ctx, pop->pair_stats.p, zero, sizeof(zero)));
  TileOut o{pop->force.p, pop->e_atom.p, pop->bd_atom.p, pop->pair_stats.p};
  launch_pair<true, true, true>(ctx, pop, o);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  unsigned long long h[2];
  DNAB_RETURN_IF(download(ctx, h, pop->pair_stats.p, sizeof(h)));
  if (in_cutoff) *in_cutoff = h[0];
  if (tested) *tested = h[1];
  DNAB_RETURN_IF(population_reduce_energies(ctx, pop, DNAB_EVAL_FORCES));
  return DNAB_OK;
}

}  // namespace dnab

using namespace dnab;

namespace {

bool atom_cloud_ok(const dnab_atom_cloud* c) {
  return c && (c->n == 0 || (c->x && c->y && c->z && c->charge && c->r_min_half && c->epsilon &&
                             c->hydrophobic));
}

}  // namespace

namespace dnab {
// Finite charges and non-negative finite r_min_half and epsilon: the pair
// kernels take sqrt(epsilon), and a clash pair's FP64 terms are finite for
// any such values, so a NaN sum can only come from a non-finite coordinate.
bool atom_params_ok(const dnab_atom_cloud* c) {
  for (size_t i = 0; i < c->n; ++i)
    if (!std::isfinite(c->charge[i]) || !(c->r_min_half[i] >= 0.0f) ||
        !std::isfinite(c->r_min_half[i]) || !(c->epsilon[i] >= 0.0f) ||
        !std::isfinite(c->epsilon[i]))
      return false;
  return true;
}

int require_finite(dnab_ctx* ctx, const dnab_population* pop, const char* what) {
  if (!pop->nonfinite) return DNAB_OK;
  return dnab_set_error(ctx, DNAB_ERR_INVALID,
                        "%s: the population has a non-finite coordinate (evaluation only)", what);
}
}  // namespace dnab

namespace {

// Per-candidate status of the last evaluation from its per-atom outputs:
// DNAB_STATUS_NAN when the energy sum is NaN (a non-finite coordinate),
// DNAB_STATUS_SATURATED when an atom's energy or force sum reached the
// fixed-point bound (written as +-inf).
__global__ void __launch_bounds__(256)
k_status(const uint32_t* __restrict__ cand_atom_off, uint32_t n_cand,
         const double* __restrict__ e_atom, const float4* __restrict__ force, int with_force,
         uint8_t* __restrict__ status) {
  const uint32_t c = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  if (c >= n_cand) return;
  bool nan = false, inf = false;
  for (uint32_t k = cand_atom_off[c] + lane; k < cand_atom_off[c + 1]; k += kWave) {
    const double e = e_atom[k];
    nan = nan || isnan(e);
    inf = inf || isinf(e);
    if (with_force) {
      const float4 f = force[k];
      inf = inf || isinf(f.x) || isinf(f.y) || isinf(f.z);
    }
  }
  const bool any_nan = ballot64(nan) != 0, any_inf = ballot64(inf) != 0;
  if (lane == 0)
    status[c] = static_cast<uint8_t>((any_nan ? DNAB_STATUS_NAN : 0) |
                                     (any_inf ? DNAB_STATUS_SATURATED : 0));
}

template <class T>
int up(dnab_ctx* ctx, DevBuf<T>& b, const T* src, size_t n) {
  DNAB_ALLOC_CHECK(ctx, b.alloc(n));
  return upload(ctx, b.p, src, n * sizeof(T));
}

int build_ca_grid(dnab_ctx* ctx, dnab_antigen* ag, const dnab_residue_cloud* ca,
                  const uint8_t* aa) {
  ag->n_ca = ca->n;
  size_t n = ca->n;
  DNAB_RETURN_IF(up(ctx, ag->ca_x, ca->x, n));
  DNAB_RETURN_IF(up(ctx, ag->ca_y, ca->y, n));
  DNAB_RETURN_IF(up(ctx, ag->ca_z, ca->z, n));
  DNAB_RETURN_IF(up(ctx, ag->ca_q, ca->charge, n));
  DNAB_RETURN_IF(up(ctx, ag->ca_eps, ca->epsilon, n));
  DNAB_RETURN_IF(up(ctx, ag->ca_sig, ca->sigma, n));
  DNAB_RETURN_IF(up(ctx, ag->ca_hyd, ca->hydrophobic, n));
  if (aa) {
    DNAB_RETURN_IF(up(ctx, ag->ca_aa, aa, n));
    ag->has_ca_aa = true;
  }
  ag->ca_grid.cell = 10.0f;  // src/diffusion.rs:1003
  ag->ca_grid.inv = 1.0f / 10.0f;
  return grid_build_device(ctx, ag->ca_grid, ag->ca_x.p, ag->ca_y.p, ag->ca_z.p, n);
}

}  // namespace

extern "C" {

int dnab_antigen_create(dnab_ctx* ctx, const dnab_antigen_desc* d, dnab_antigen** out) {
  if (!ctx || !d || !out || !atom_cloud_ok(&d->atoms)) return DNAB_ERR_INVALID;
  if (d->ca.n && !(d->ca.x && d->ca.y && d->ca.z && d->ca.charge && d->ca.epsilon &&
                   d->ca.sigma && d->ca.hydrophobic))
    return DNAB_ERR_INVALID;
  if (!(d->cell_size > 0.0f) || !std::isfinite(d->cell_size) || !(d->cutoff > 0.0f) ||
      !std::isfinite(d->cutoff))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "antigen: cell_size and cutoff must be > 0");
  *out = nullptr;
  (void)hipSetDevice(ctx->device);
  dnab_antigen* ag = new (std::nothrow) dnab_antigen();
  if (!ag) return DNAB_ERR_ALLOC;
  ag->ctx = ctx_retain(ctx);
  ag->cutoff = d->cutoff;
  ag->restraint_radius = d->restraint_radius;
  for (int a = 0; a < 3; ++a) ag->center[a] = d->center[a];
  const size_t n = d->atoms.n;
  const dnab_atom_cloud& c = d->atoms;
  auto fail = [&](int rc) {
    dnab_antigen_destroy(ag);
    return rc;
  };
  if (!atom_params_ok(&c))
    return fail(dnab_set_error(ctx, DNAB_ERR_INVALID,
                               "antigen: non-finite charge, or negative r_min_half or epsilon"));
  // bounding box of the atoms (host)
  if (n) {
    for (int a = 0; a < 3; ++a) {
      ag->bbox_lo[a] = INFINITY;
      ag->bbox_hi[a] = -INFINITY;
    }
    const float* p[3] = {c.x, c.y, c.z};
    for (size_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        float v = p[a][i];
        if (!std::isfinite(v))
          return fail(dnab_set_error(ctx, DNAB_ERR_INVALID, "antigen: non-finite coordinate"));
        ag->bbox_lo[a] = std::min(ag->bbox_lo[a], v);
        ag->bbox_hi[a] = std::max(ag->bbox_hi[a], v);
      }
  }
  GridDev& g = ag->grid;
  g.cell = d->cell_size;
  g.inv = 1.0f / d->cell_size;
  DevBuf<float> q, rmh, eps;
  DevBuf<uint8_t> hyd, type;
  if (c.atom_type)
    for (size_t i = 0; i < n; ++i)
      if (c.atom_type[i] >= kNumAtomTypesDev)
        return fail(dnab_set_error(ctx, DNAB_ERR_INVALID, "antigen: bad atom type"));
  int rc = DNAB_OK;
  if ((rc = up(ctx, g.x, c.x, n)) || (rc = up(ctx, g.y, c.y, n)) || (rc = up(ctx, g.z, c.z, n)) ||
      (rc = up(ctx, q, c.charge, n)) || (rc = up(ctx, rmh, c.r_min_half, n)) ||
      (rc = up(ctx, eps, c.epsilon, n)) || (rc = up(ctx, hyd, c.hydrophobic, n)) ||
      (c.atom_type && (rc = up(ctx, type, c.atom_type, n))))
    return fail(rc);
  if ((rc = grid_build_device(ctx, g, g.x.p, g.y.p, g.z.p, n))) return fail(rc);
  if (ag->atoms.alloc(n) != hipSuccess || ag->raw.alloc(n) != hipSuccess)
    return fail(dnab_set_error(ctx, DNAB_ERR_ALLOC, "antigen: allocation failed"));
  if (n) {
    hipLaunchKernelGGL(k_ag_gather, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream,
                       g.sorted_idx.p, static_cast<uint32_t>(n), g.x.p, g.y.p, g.z.p, q.p, rmh.p,
                       eps.p, hyd.p, c.atom_type ? type.p : nullptr, ag->atoms.p, ag->raw.p);
  }
  if (hipGetLastError() != hipSuccess)
    return fail(dnab_set_error(ctx, DNAB_ERR_HIP, "antigen: gather launch failed"));
  {
    const uint64_t nc = g.ncodes;  // (an empty grid has one empty cell)
    if (ag->cell_box.alloc(2 * nc) != hipSuccess)
      return fail(dnab_set_error(ctx, DNAB_ERR_ALLOC, "antigen: allocation failed"));
    hipLaunchKernelGGL(k_cell_box, dim3(static_cast<uint32_t>((nc + 255) / 256)), dim3(256), 0,
                       ctx->stream, ag->atoms.p, g.cell_start.p, nc, ag->cell_box.p);
    if (hipGetLastError() != hipSuccess)
      return fail(dnab_set_error(ctx, DNAB_ERR_HIP, "antigen: cell box launch failed"));
  }
  if (d->ca_amino_acid)
    for (size_t i = 0; i < d->ca.n; ++i)
      if (d->ca_amino_acid[i] >= 20)
        return fail(dnab_set_error(ctx, DNAB_ERR_INVALID, "antigen: bad residue type"));
  if ((rc = build_ca_grid(ctx, ag, &d->ca, d->ca_amino_acid))) return fail(rc);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail(dnab_set_error(ctx, DNAB_ERR_HIP, "antigen: synchronize failed"));
  *out = ag;
  return DNAB_OK;
}

int dnab_antigen_info(const dnab_antigen* ag, size_t* n_atoms, size_t* n_cells_nonempty,
                      float* cutoff, float* cell_size) {
  if (!ag) return DNAB_ERR_INVALID;
  if (n_atoms) *n_atoms = ag->grid.n;
  if (n_cells_nonempty) *n_cells_nonempty = ag->grid.n_nonempty;
  if (cutoff) *cutoff = ag->cutoff;
  if (cell_size) *cell_size = ag->grid.cell;
  return DNAB_OK;
}

void dnab_antigen_destroy(dnab_antigen* ag) {
  if (!ag) return;
  dnab_ctx* ctx = ag->ctx;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  delete ag;
  ctx_release(ctx);
}

int dnab_population_create(dnab_ctx* ctx, dnab_antigen* const* antigens, size_t n_antigens,
                           const dnab_population_desc* d, dnab_population** out) {
  if (!ctx || !out || !d || (n_antigens && !antigens) || n_antigens == 0)
    return DNAB_ERR_INVALID;
  *out = nullptr;
  const size_t nc = d->n_cand, nr = d->n_res;
  if (nc && (!d->cand_antigen || !d->cand_res_off)) return DNAB_ERR_INVALID;
  if (nr && (!d->res_atom_off || !d->res_ca_atom)) return DNAB_ERR_INVALID;
  if (!atom_cloud_ok(&d->atoms)) return DNAB_ERR_INVALID;
  for (size_t a = 0; a < n_antigens; ++a)
    if (!antigens[a] || antigens[a]->ctx != ctx)
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: antigen %zu invalid", a);
  const size_t na = d->atoms.n;
  if (na >= 0xFFFFFF00ull || nc >= 0xFFFFFF00ull || nr >= 0xFFFFFF00ull)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "population: too large");
  // validate ranges
  if (nc) {
    if (d->cand_res_off[0] != 0 || d->cand_res_off[nc] != nr)
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: cand_res_off must span 0..n_res");
    for (size_t c = 0; c < nc; ++c) {
      if (d->cand_res_off[c + 1] < d->cand_res_off[c])
        return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: cand_res_off not monotone");
      if (d->cand_antigen[c] >= n_antigens)
        return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: cand_antigen out of range");
    }
  } else if (nr) {
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: residues without candidates");
  }
  if (nr) {
    if (d->res_atom_off[0] != 0 || d->res_atom_off[nr] != na)
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: res_atom_off must span 0..n_atoms");
    for (size_t r = 0; r < nr; ++r) {
      if (d->res_atom_off[r + 1] < d->res_atom_off[r])
        return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: res_atom_off not monotone");
      if (d->res_ca_atom[r] < d->res_atom_off[r] || d->res_ca_atom[r] >= d->res_atom_off[r + 1])
        return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: residue %zu Calpha outside it", r);
    }
  } else if (na) {
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: atoms without residues");
  }
  if (!atom_params_ok(&d->atoms))
    return dnab_set_error(ctx, DNAB_ERR_INVALID,
                          "population: non-finite charge, or negative r_min_half or epsilon");
  bool nonfinite = false;  // allowed: such a candidate evaluates to NaN
  for (size_t i = 0; i < na; ++i)
    nonfinite = nonfinite || !std::isfinite(d->atoms.x[i]) || !std::isfinite(d->atoms.y[i]) ||
                !std::isfinite(d->atoms.z[i]);
  if (d->atoms.atom_type)
    for (size_t i = 0; i < na; ++i)
      if (d->atoms.atom_type[i] >= kNumAtomTypesDev)
        return dnab_set_error(ctx, DNAB_ERR_INVALID, "population: bad atom type");
  (void)hipSetDevice(ctx->device);
  dnab_population* pop = new (std::nothrow) dnab_population();
  if (!pop) return DNAB_ERR_ALLOC;
  pop->ctx = ctx_retain(ctx);
  pop->n_cand = nc;
  pop->nonfinite = nonfinite;
  pop->n_res = nr;
  pop->n_atoms = na;
  pop->n_ag = n_antigens;
  pop->antigens.assign(antigens, antigens + n_antigens);
  auto fail = [&](int rc) {
    dnab_population_destroy(pop);
    return rc;
  };
  auto t_phase = std::chrono::steady_clock::now();
  auto phase = [&](const char* name) {
    const auto now = std::chrono::steady_clock::now();
    timer_host(ctx, name, std::chrono::duration<double, std::milli>(now - t_phase).count());
    t_phase = now;
  };
  // host staging of the per-candidate and per-residue tables; the per-atom
  // arrays are packed on the device from the caller's SoA (k_pack_atoms)
  std::vector<uint32_t> hcand_atom_off(nc + 1), hres_cand(nr);
  std::vector<float4> hres_cg(nr);
  for (size_t c = 0; c < nc; ++c) {
    uint32_t r0 = d->cand_res_off[c], r1 = d->cand_res_off[c + 1];
    hcand_atom_off[c] = nr ? d->res_atom_off[r0] : 0;
    for (uint32_t r = r0; r < r1; ++r) hres_cand[r] = static_cast<uint32_t>(c);
  }
  if (nc) hcand_atom_off[nc] = static_cast<uint32_t>(na);
  else hcand_atom_off[0] = 0;
  for (size_t c = 0; c < nc; ++c) {
    pop->max_cand_atoms = std::max(pop->max_cand_atoms, hcand_atom_off[c + 1] - hcand_atom_off[c]);
    pop->max_cand_res = std::max(pop->max_cand_res, d->cand_res_off[c + 1] - d->cand_res_off[c]);
  }
  pop->h_cand_atom_off = hcand_atom_off;
  pop->h_cand_res_off.assign(d->cand_res_off, d->cand_res_off + nc + 1);
  if (nc) pop->h_cand_ag.assign(d->cand_antigen, d->cand_antigen + nc);
  for (size_t r = 0; r < nr; ++r) {
    float q = 0, e = 0, s = 0;
    uint8_t h = 0;
    if (d->res_amino_acid) {
      if (dnab_residue_cg_params(d->res_amino_acid[r], &q, &e, &s, &h) != DNAB_OK)
        return fail(dnab_set_error(ctx, DNAB_ERR_INVALID, "population: bad residue type"));
    }
    hres_cg[r] = make_float4(q, e, s, static_cast<float>(h));
  }
  // per-antigen views and the candidate sort grids
  pop->views_host.resize(n_antigens);
  uint64_t code_base = 0;
  for (size_t a = 0; a < n_antigens; ++a) {
    const dnab_antigen* ag = antigens[a];
    AgView& v = pop->views_host[a];
    std::memset(&v, 0, sizeof(v));
    v.atoms = ag->atoms.p;
    v.raw = ag->raw.p;
    v.cell_start = ag->grid.cell_start.p;
    v.cell_box = ag->cell_box.p;
    for (int k = 0; k < 3; ++k) {
      v.kmin[k] = ag->grid.kmin[k];
      v.kmax[k] = ag->grid.kmax[k];
      v.center[k] = ag->center[k];
    }
    v.bits = ag->grid.bits;
    v.inv = ag->grid.inv;
    v.rc = ag->cutoff;
    v.rc2 = ag->cutoff * ag->cutoff;
    v.restraint_radius = ag->restraint_radius;
    v.n_atoms = static_cast<uint32_t>(ag->grid.n);
    // sort grid: the antigen box + cutoff (atoms beyond are excluded), 4 A
    // cells (coarser when the code space would exceed 2^22 per antigen)
    const float margin = ag->cutoff + 1.0f;
    float cell = kSortCell;
    for (;;) {
      uint64_t codes = 64;
      for (int k = 0; k < 3; ++k) {
        float ext = (ag->grid.n ? ag->bbox_hi[k] - ag->bbox_lo[k] : 0.0f) + 2.0f * margin;
        int dim = static_cast<int>(std::ceil(ext / cell)) + 1;
        v.sort_dim[k] = dim;
        v.sort_nb[k] = (dim + 3) / 4;
        codes *= static_cast<uint64_t>(v.sort_nb[k]);
      }
      if (codes <= (uint64_t(1) << 22)) {
        v.sort_ncode = static_cast<uint32_t>(codes);
        break;
      }
      cell *= 2.0f;
    }
    for (int k = 0; k < 3; ++k) {
      v.sort_origin[k] = (ag->grid.n ? ag->bbox_lo[k] : 0.0f) - margin;
      v.excl_lo[k] = ag->grid.n ? ag->bbox_lo[k] : INFINITY;
      v.excl_hi[k] = ag->grid.n ? ag->bbox_hi[k] : -INFINITY;
    }
    v.sort_inv = 1.0f / cell;
    v.sort_code_base = static_cast<uint32_t>(code_base);
    code_base += v.sort_ncode;
    v.ca_x = ag->ca_x.p;
    v.ca_y = ag->ca_y.p;
    v.ca_z = ag->ca_z.p;
    v.ca_q = ag->ca_q.p;
    v.ca_eps = ag->ca_eps.p;
    v.ca_sig = ag->ca_sig.p;
    v.ca_hyd = ag->ca_hyd.p;
    v.ca_aa = ag->has_ca_aa ? ag->ca_aa.p : nullptr;
    v.ca_grid = make_view(ag->ca_grid);
  }
  if (code_base >= 0xFFFFFF00ull)
    return fail(dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "population: sort grid too large"));
  pop->sort_codes = static_cast<uint32_t>(code_base);
  // every non-empty 16 A brick adds at most one partly filled tile
  pop->tile_cap = (na + kWave - 1) / kWave +
                  std::min<size_t>(na, static_cast<size_t>(pop->sort_codes / kWave)) + 1;
  phase("population_create_pack");
  int rc = DNAB_OK;
  if ((rc = up(ctx, pop->views, pop->views_host.data(), n_antigens)) ||
      (rc = up(ctx, pop->cand_atom_off, hcand_atom_off.data(), nc + 1)) ||
      (rc = up(ctx, pop->res_cand, hres_cand.data(), nr)) ||
      (rc = up(ctx, pop->res_cg, hres_cg.data(), nr)))
    return fail(rc);
  if (nc && ((rc = up(ctx, pop->cand_res_off, d->cand_res_off, nc + 1)) ||
             (rc = up(ctx, pop->cand_ag, d->cand_antigen, nc))))
    return fail(rc);
  if (nr && ((rc = up(ctx, pop->res_atom_off, d->res_atom_off, nr + 1)) ||
             (rc = up(ctx, pop->res_ca, d->res_ca_atom, nr))))
    return fail(rc);
  // the per-atom and per-tile arrays in one allocation
  const size_t nf = (na + 63) / 64 * 64;
  {
    DevSlab& sl = pop->slab;
    sl.carve(pop->pack_soa, 6 * nf + (2 * nf + 3) / 4);
    sl.carve(pop->pos, na);
    sl.carve(pop->par, na);
    sl.carve(pop->atom_ag, na);
    sl.carve(pop->atom_cand, na);
    sl.carve(pop->atom_res, na);
    sl.carve(pop->sort_count, static_cast<size_t>(pop->sort_codes) + 4);
    sl.carve(pop->sort_start, static_cast<size_t>(pop->sort_codes) + 4);
    sl.carve(pop->sort_code, na);
    sl.carve(pop->sort_rank, na);
    sl.carve(pop->perm, pop->tile_cap * kWave);
    sl.carve(pop->spos, pop->tile_cap * kWave);
    sl.carve(pop->spar, pop->tile_cap * kWave);
    sl.carve(pop->ho_rest, pop->tile_cap);
    sl.carve(pop->ho_list, pop->tile_cap + 4);
    sl.carve(pop->ho_chunks, pop->tile_cap * kHandoverCap);
    sl.carve(pop->ho_part, 5 * pop->tile_cap * kWave);
    sl.carve(pop->ho_status, pop->tile_cap * kWave);
    sl.carve(pop->ag_tile_start, n_antigens + 1);
    sl.carve(pop->force, na);
    sl.carve(pop->e_atom, na);
    sl.carve(pop->bd_atom, kBdTerms * na);
    sl.carve(pop->f64_atom, 3 * na);
    sl.carve(pop->energy, nc);
    sl.carve(pop->err_bound, nc);
    sl.carve(pop->breakdown, kBdTerms * nc);
    sl.carve(pop->disulfide, nc);
    sl.carve(pop->res_delta, nr);
    sl.carve(pop->step_counter, 4);
    sl.carve(pop->pair_stats, 2);
    sl.carve(pop->cand_mask, nc);
    sl.carve(pop->cand_status, nc);
    if (sl.alloc(&ctx->spare_slab) != hipSuccess)
      return fail(dnab_set_error(ctx, DNAB_ERR_ALLOC, "population: device allocation failed"));
  }
  {
    // the caller's per-atom SoA, 26 bytes per atom, packed into pos / par
    // and the per-atom owner indices on the device
    DevBuf<float>& soa = pop->pack_soa;
    const float* src[6] = {d->atoms.x, d->atoms.y, d->atoms.z, d->atoms.charge,
                           d->atoms.r_min_half, d->atoms.epsilon};
    for (int k = 0; k < 6; ++k)
      if ((rc = upload(ctx, soa.p + k * nf, src[k], na * sizeof(float)))) return fail(rc);
    uint8_t* bytes = reinterpret_cast<uint8_t*>(soa.p + 6 * nf);
    if ((rc = upload(ctx, bytes, d->atoms.hydrophobic, na))) return fail(rc);
    if (d->atoms.atom_type && (rc = upload(ctx, bytes + nf, d->atoms.atom_type, na)))
      return fail(rc);
    if (na) {
      hipLaunchKernelGGL(k_pack_atoms, dim3(nblocks(na, 256)), dim3(256), 0, ctx->stream,
                         soa.p, nf, d->atoms.atom_type ? 1 : 0, static_cast<uint32_t>(na),
                         static_cast<uint32_t>(nr), pop->res_atom_off.p, pop->res_cand.p,
                         pop->cand_ag.p, pop->pos.p, pop->par.p, pop->atom_ag.p,
                         pop->atom_cand.p, pop->atom_res.p);
      DNAB_HIP_CHECK(ctx, hipGetLastError());
    }
  }
  if (nr && d->res_amino_acid) {  // types checked above (< 20)
    if ((rc = up(ctx, pop->res_aa, d->res_amino_acid, nr))) return fail(rc);
    pop->has_res_aa = true;
  }
  {
    std::vector<uint8_t> ones(nc ? nc : 1, 1);
    if ((rc = upload(ctx, pop->cand_mask.p, ones.data(), nc))) return fail(rc);
    std::vector<uint32_t> zeros(static_cast<size_t>(pop->sort_codes) + 4, 0u);
    if ((rc = upload(ctx, pop->sort_count.p, zeros.data(), zeros.size() * 4))) return fail(rc);
    uint32_t ctr[4] = {0, 0, 0, 0};
    if ((rc = upload(ctx, pop->step_counter.p, ctr, sizeof(ctr)))) return fail(rc);
    const unsigned long long stats0[2] = {0, 0};
    if ((rc = upload(ctx, pop->pair_stats.p, stats0, sizeof(stats0)))) return fail(rc);
    if (hipMemsetAsync(pop->ho_rest.p, 0, pop->tile_cap * sizeof(uint32_t), ctx->stream) != hipSuccess ||
        hipMemsetAsync(pop->ho_list.p, 0, (pop->tile_cap + 4) * sizeof(uint32_t), ctx->stream) != hipSuccess)
      return fail(dnab_set_error(ctx, DNAB_ERR_HIP, "population: clearing the handover table failed"));
  }
  phase("population_create_upload");
  if ((rc = population_sort(ctx, pop))) return fail(rc);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    return fail(dnab_set_error(ctx, DNAB_ERR_HIP, "population: synchronize failed"));
  phase("population_create_sort");
  *out = pop;
  return DNAB_OK;
}

void dnab_population_destroy(dnab_population* pop) {
  if (!pop) return;
  dnab_ctx* ctx = pop->ctx;
  if (ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    // its slab, idle now, kept for the context's next population (the
    // larger of it and the one kept so far, up to kSpareSlabMax)
    DevBuf<unsigned char>& m = pop->slab.mem;
    if (m.p && m.owned && m.n <= kSpareSlabMax && m.n > ctx->spare_slab.n) {
      ctx->spare_slab.release();
      ctx->spare_slab.p = m.p;
      ctx->spare_slab.n = m.n;
      m.p = nullptr;
      m.n = 0;
    }
  }
  delete pop;
  ctx_release(ctx);
}

int dnab_population_info(const dnab_population* pop, size_t* n_cand, size_t* n_res,
                         size_t* n_atoms) {
  if (!pop) return DNAB_ERR_INVALID;
  if (n_cand) *n_cand = pop->n_cand;
  if (n_res) *n_res = pop->n_res;
  if (n_atoms) *n_atoms = pop->n_atoms;
  return DNAB_OK;
}

int dnab_population_evaluate(dnab_ctx* ctx, dnab_population* pop, unsigned flags) {
  if (!ctx || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  if (flags & ~(DNAB_EVAL_FORCES | DNAB_EVAL_BREAKDOWN | DNAB_EVAL_FP64)) return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  return population_evaluate(ctx, pop, flags);
}

int dnab_population_read_energies(dnab_ctx* ctx, const dnab_population* pop, double* energy,
                                  float* err_bound) {
  if (!ctx || !pop || !energy) return DNAB_ERR_INVALID;
  if (!pop->evaluated) return dnab_set_error(ctx, DNAB_ERR_INVALID, "population not evaluated");
  DNAB_RETURN_IF(download(ctx, energy, pop->energy.p, pop->n_cand * sizeof(double)));
  if (err_bound) DNAB_RETURN_IF(download(ctx, err_bound, pop->err_bound.p, pop->n_cand * sizeof(float)));
  return DNAB_OK;
}

int dnab_population_read_status(dnab_ctx* ctx, dnab_population* pop, uint8_t* status) {
  if (!ctx || !pop || pop->ctx != ctx || (pop->n_cand && !status)) return DNAB_ERR_INVALID;
  if (!pop->evaluated) return dnab_set_error(ctx, DNAB_ERR_INVALID, "population not evaluated");
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  if (!nc) return DNAB_OK;
  const bool f64 = (pop->last_flags & DNAB_EVAL_FP64) != 0;
  const bool forces = (pop->last_flags & DNAB_EVAL_FORCES) != 0 && !f64;
  hipLaunchKernelGGL(k_status, dim3(nblocks(nc, 4)), dim3(256), 0, ctx->stream,
                     pop->cand_atom_off.p, nc, pop->e_atom.p, pop->force.p, forces ? 1 : 0,
                     pop->cand_status.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return download(ctx, status, pop->cand_status.p, nc);
}

int dnab_population_read_breakdown(dnab_ctx* ctx, const dnab_population* pop, double* lj,
                                   double* coulomb, double* hydrophobic) {
  if (!ctx || !pop) return DNAB_ERR_INVALID;
  if (!pop->evaluated || !(pop->last_flags & (DNAB_EVAL_BREAKDOWN | DNAB_EVAL_FP64)))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "no breakdown evaluation");
  std::vector<double> h(kBdTerms * pop->n_cand);
  DNAB_RETURN_IF(download(ctx, h.data(), pop->breakdown.p, h.size() * sizeof(double)));
  for (size_t c = 0; c < pop->n_cand; ++c) {
    if (lj) lj[c] = h[kBdTerms * c];
    if (coulomb) coulomb[c] = h[kBdTerms * c + 1];
    if (hydrophobic) hydrophobic[c] = h[kBdTerms * c + 2];
  }
  return DNAB_OK;
}

int dnab_population_read_terms(dnab_ctx* ctx, const dnab_population* pop, double* terms) {
  if (!ctx || !pop || (pop->n_cand && !terms)) return DNAB_ERR_INVALID;
  if (!pop->evaluated || !(pop->last_flags & DNAB_EVAL_BREAKDOWN))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "no breakdown evaluation");
  const size_t nc = pop->n_cand;
  std::vector<double> h(kBdTerms * nc), ss(nc);
  DNAB_RETURN_IF(download(ctx, h.data(), pop->breakdown.p, h.size() * sizeof(double)));
  DNAB_RETURN_IF(download(ctx, ss.data(), pop->disulfide.p, nc * sizeof(double)));
  for (size_t c = 0; c < nc; ++c) {
    for (int t = 0; t < kBdTerms; ++t) terms[6 * c + t] = h[kBdTerms * c + t];
    terms[6 * c + 5] = ss[c];
  }
  return DNAB_OK;
}

int dnab_population_read_forces(dnab_ctx* ctx, const dnab_population* pop, float* fx,
                                float* fy, float* fz) {
  if (!ctx || !pop || (pop->n_atoms && (!fx || !fy || !fz))) return DNAB_ERR_INVALID;
  if (!pop->evaluated) return dnab_set_error(ctx, DNAB_ERR_INVALID, "population not evaluated");
  const size_t n = pop->n_atoms;
  if (pop->last_flags & DNAB_EVAL_FP64) {
    std::vector<double> h(3 * n);
    DNAB_RETURN_IF(download(ctx, h.data(), pop->f64_atom.p, h.size() * sizeof(double)));
    for (size_t i = 0; i < n; ++i) {
      fx[i] = static_cast<float>(h[3 * i]);
      fy[i] = static_cast<float>(h[3 * i + 1]);
      fz[i] = static_cast<float>(h[3 * i + 2]);
    }
    return DNAB_OK;
  }
  if (!(pop->last_flags & DNAB_EVAL_FORCES))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "last evaluation had no forces");
  std::vector<float4> h(n);
  DNAB_RETURN_IF(download(ctx, h.data(), pop->force.p, n * sizeof(float4)));
  for (size_t i = 0; i < n; ++i) {
    fx[i] = h[i].x;
    fy[i] = h[i].y;
    fz[i] = h[i].z;
  }
  return DNAB_OK;
}

int dnab_population_read_forces_f64(dnab_ctx* ctx, const dnab_population* pop, double* fx,
                                    double* fy, double* fz) {
  if (!ctx || !pop || (pop->n_atoms && (!fx || !fy || !fz))) return DNAB_ERR_INVALID;
  if (!pop->evaluated || !(pop->last_flags & DNAB_EVAL_FP64))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "no FP64 evaluation");
  const size_t n = pop->n_atoms;
  std::vector<double> h(3 * n);
  DNAB_RETURN_IF(download(ctx, h.data(), pop->f64_atom.p, h.size() * sizeof(double)));
  for (size_t i = 0; i < n; ++i) {
    fx[i] = h[3 * i];
    fy[i] = h[3 * i + 1];
    fz[i] = h[3 * i + 2];
  }
  return DNAB_OK;
}

int dnab_population_read_coords(dnab_ctx* ctx, const dnab_population* pop, float* x, float* y,
                                float* z) {
  if (!ctx || !pop || (pop->n_atoms && (!x || !y || !z))) return DNAB_ERR_INVALID;
  const size_t n = pop->n_atoms;
  std::vector<float4> h(n);
  DNAB_RETURN_IF(download(ctx, h.data(), pop->pos.p, n * sizeof(float4)));
  for (size_t i = 0; i < n; ++i) {
    x[i] = h[i].x;
    y[i] = h[i].y;
    z[i] = h[i].z;
  }
  return DNAB_OK;
}

int dnab_population_write_coords(dnab_ctx* ctx, dnab_population* pop, const float* x,
                                 const float* y, const float* z) {
  if (!ctx || !pop || (pop->n_atoms && (!x || !y || !z))) return DNAB_ERR_INVALID;
  const size_t n = pop->n_atoms;
  std::vector<float4> h(n);
  DNAB_RETURN_IF(download(ctx, h.data(), pop->pos.p, n * sizeof(float4)));
  bool nonfinite = false;  // allowed: such a candidate evaluates to NaN
  for (size_t i = 0; i < n; ++i) {
    nonfinite = nonfinite || !std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]);
    h[i].x = x[i];
    h[i].y = y[i];
    h[i].z = z[i];
  }
  DNAB_RETURN_IF(discard_pending_counts(ctx, pop));
  DNAB_RETURN_IF(upload(ctx, pop->pos.p, h.data(), n * sizeof(float4)));
  pop->nonfinite = nonfinite;
  pop->evaluated = false;
  DNAB_RETURN_IF(population_sort(ctx, pop));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_population_get_tuning(const dnab_population* pop, dnab_pair_tuning* out) {
  if (!pop || !out) return DNAB_ERR_INVALID;
  out->direct_survivors = pop->tune.direct_survivors;
  out->direct_chunks = pop->tune.direct_chunks;
  out->rest_waves = pop->tune.rest_waves;
  return DNAB_OK;
}

int dnab_population_set_tuning(dnab_ctx* ctx, dnab_population* pop, const dnab_pair_tuning* t) {
  if (!ctx || !pop || !t || pop->ctx != ctx) return DNAB_ERR_INVALID;
  if (t->direct_survivors == 0 || t->direct_chunks == 0 ||
      (t->rest_waves != 4 && t->rest_waves != 8 && t->rest_waves != 16))
    return dnab_set_error(ctx, DNAB_ERR_INVALID,
                          "tuning: survivors and chunks >= 1, rest_waves 4, 8 or 16");
  pop->tune = PairTuning{t->direct_survivors, t->direct_chunks, t->rest_waves};
  return DNAB_OK;
}

int dnab_population_autotune(dnab_ctx* ctx, dnab_population* pop, uint32_t reps,
                             dnab_pair_tuning* best, double* best_ms) {
  if (!ctx || !pop || pop->ctx != ctx || reps == 0) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "autotune"));
  (void)hipSetDevice(ctx->device);
  pop->masked = false;
  if (!pop->sorted) DNAB_RETURN_IF(population_sort(ctx, pop));
  TileOut o{pop->force.p, pop->e_atom.p, pop->bd_atom.p, pop->pair_stats.p};
  hipEvent_t e0, e1;
  DNAB_HIP_CHECK(ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return dnab_set_error(ctx, DNAB_ERR_HIP, "autotune: event creation failed");
  }
  const uint32_t surv[] = {24, 48, 96, 192}, chunks[] = {8, 16, 32}, waves[] = {4, 8, 16};
  const PairTuning saved = pop->tune;
  PairTuning win = saved;
  double win_ms = 1e30;
  int rc = DNAB_OK;
  launch_pair<true, false, true>(ctx, pop, o);  // warm-up
  for (uint32_t w : waves)
    for (uint32_t c : chunks)
      for (uint32_t sv : surv) {
        pop->tune = PairTuning{sv, c, w};
        (void)hipEventRecord(e0, ctx->stream);
        for (uint32_t r = 0; r < reps; ++r) launch_pair<true, false, true>(ctx, pop, o);
        (void)hipEventRecord(e1, ctx->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
          rc = dnab_set_error(ctx, DNAB_ERR_HIP, "autotune: launch failed");
          break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = static_cast<double>(ms) / reps;
        if (per < win_ms) {
          win_ms = per;
          win = pop->tune;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pop->tune = rc ? saved : win;
  if (rc) return rc;
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(population_reduce_energies(ctx, pop, DNAB_EVAL_FORCES));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (best) {
    best->direct_survivors = win.direct_survivors;
    best->direct_chunks = win.direct_chunks;
    best->rest_waves = win.rest_waves;
  }
  if (best_ms) *best_ms = win_ms;
  return DNAB_OK;
}

int dnab_population_read_pair_stats(dnab_ctx* ctx, const dnab_population* pop,
                                    uint64_t* in_cutoff, uint64_t* tested) {
  if (!ctx || !pop) return DNAB_ERR_INVALID;
  unsigned long long h[2];
  DNAB_RETURN_IF(download(ctx, h, pop->pair_stats.p, sizeof(h)));
  if (in_cutoff) *in_cutoff = h[0];
  if (tested) *tested = h[1];
  return DNAB_OK;
}

int dnab_population_reset_pair_stats(dnab_ctx* ctx, dnab_population* pop) {
  if (!ctx || !pop) return DNAB_ERR_INVALID;
  unsigned long long zero[2] = {0, 0};
  DNAB_RETURN_IF(upload(ctx, pop->pair_stats.p, zero, sizeof(zero)));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_population_pair_count(dnab_ctx* ctx, dnab_population* pop, uint64_t* in_cutoff,
                               uint64_t* tested) {
  if (!ctx || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  return pair_count(ctx, pop, in_cutoff, tested);
}

}  // extern "C"
