// Antigen maps and per-candidate features for the learned prescreen (no
// reference counterpart: the reference has no learned component,
// Cargo.toml:4).
//
// The maps hold, on a regular grid around the antigen, the pair sums that the
// exact scorer (src/energy.rs:261-314: LJ with eps_ij = sqrt(eps_i eps_j) and
// R_ij = Rmin/2_i + Rmin/2_j, Coulomb 332 q_i q_j / r, -0.5 per hydrophobic
// pair within 6 A, pairs with 0.25 <= r^2 <= cutoff^2) would form for a probe
// atom at the grid point:
//   - LJ of a probe of each of the kMapLjClasses distinct (Rmin/2, eps)
//     parameter sets of the AMBER99SB table (src/amber.rs:46-63),
//   - the Coulomb potential sum 332 q_i / r (times the probe's charge),
//   - the hydrophobic sum -0.5 per hydrophobic antigen atom within 6 A (for a
//     hydrophobic probe),
//   - kMapRbf Gaussian radial densities sum exp(-(r - mu_k)^2 / (2 sigma^2)).
// A candidate atom's map energy is the trilinear interpolation of its class's
// LJ plane plus q times the potential plus (hydrophobic) the hydrophobic
// plane; summed over the candidate it approximates the exact energy except
// where the repulsive wall is steeper than the grid resolves.
//
// Features of a candidate (kPrescreenFeatures, bf16): per atom class, the sum
// of its atoms' LJ map energies and of their radial densities (atom-class x
// RBF distance histograms against the antigen's heavy atoms), the Coulomb sums
// over positive and negative atoms, the hydrophobic sum, the total map
// energy, and counts of clashing, contacting and in-map atoms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "dnab_internal.h"
#include "ff_tables.h"
#include "mlp_kernels.h"
#include "pair_device.h"

namespace dnab {

constexpr float kMapDefaultSpacing = 0.5f;
constexpr float kRbfMu0 = 2.0f, kRbfDmu = 1.0f, kRbfSigma = 0.75f;
constexpr float kMapClamp = 60000.0f;   // fp16 range; a wall value beyond it is saturated
constexpr size_t kMapMaxBytes = size_t(6) << 30;
constexpr int kFeatClasses = kMapLjClasses + 1;  // LJ classes + hydrophobic carbon
constexpr float kClashE = 1.0f;     // an atom's LJ map energy above this is a clash
constexpr float kContactE = -0.05f; // ... below this, a contact

// Feature layout (index ranges)
constexpr int kFLj = 0;            // [kFeatClasses] LJ map energy per atom class
constexpr int kFCoulPos = 8, kFCoulNeg = 9, kFHyd = 10, kFTotal = 11;
constexpr int kFClash = 12, kFContact = 13, kFInMap = 14;
constexpr int kFRbf = 16;          // [kFeatClasses][kMapRbf]
static_assert(kFRbf + kFeatClasses * kMapRbf <= kPrescreenFeatures, "feature layout");

struct MapClasses {
  float rmh[kMapLjClasses];
  float eps[kMapLjClasses];
  float sqe[kMapLjClasses];
};

// The distinct (Rmin/2, eps) pairs of the AtomType table, in AtomType order.
static MapClasses map_classes() {
  MapClasses c{};
  int n = 0;
  for (int t = 0; t < kNumAtomTypes; ++t) {
    bool seen = false;
    for (int k = 0; k < n; ++k)
      seen = seen || (c.rmh[k] == kAmberNb[t].r_min_half && c.eps[k] == kAmberNb[t].epsilon);
    if (!seen && n < kMapLjClasses) {
      c.rmh[n] = kAmberNb[t].r_min_half;
      c.eps[n] = kAmberNb[t].epsilon;
      c.sqe[n] = std::sqrt(kAmberNb[t].epsilon);
      ++n;
    }
  }
  return c;
}

__device__ inline uint16_t to_half_bits(float v) {
  v = fminf(fmaxf(v, -kMapClamp), kMapClamp);
  return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
}
__device__ inline float half_bits_to_f32(uint16_t u) {
  return static_cast<float>(__builtin_bit_cast(_Float16, u));
}

struct MapBuildArgs {
  const AgAtom* atoms;
  const uint32_t* cell_start;
  int32_t kmin[3], kmax[3];
  float inv;
  float rc, rc2;
  float origin[3];
  float h;
  int32_t n[3];
  MapClasses cls;
  uint16_t* pts;  // [plane][point] fp16
  uint16_t* rbf;  // [point][kMapRbf] fp16
  unsigned long long npts;
};

// One workgroup per brick of 8 x 8 x 4 grid points, one thread per point.
// The antigen atoms of the grid cells within the cutoff of the brick are
// streamed through LDS in slabs of 256 (cells enumerated 256 at a time, their
// atom ranges concatenated by a block scan).
__global__ void __launch_bounds__(256) k_map_points(MapBuildArgs a) {
  __shared__ AgAtom s_at[256];
  __shared__ uint32_t s_cs[256];
  __shared__ uint32_t s_pre[256];
  __shared__ uint32_t s_wsum[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const int32_t ix = static_cast<int32_t>(blockIdx.x * 8 + (tid & 7u));
  const int32_t iy = static_cast<int32_t>(blockIdx.y * 8 + ((tid >> 3) & 7u));
  const int32_t iz = static_cast<int32_t>(blockIdx.z * 4 + (tid >> 6));
  const bool valid = ix < a.n[0] && iy < a.n[1] && iz < a.n[2];
  const float px = a.origin[0] + a.h * static_cast<float>(ix);
  const float py = a.origin[1] + a.h * static_cast<float>(iy);
  const float pz = a.origin[2] + a.h * static_cast<float>(iz);
  // the brick's box and the grid cells within the cutoff of it (cell_key is
  // monotone: a slightly widened range is conservative)
  const float m = a.rc * 1.001f + 1e-3f;
  const float lox = a.origin[0] + a.h * static_cast<float>(blockIdx.x * 8);
  const float loy = a.origin[1] + a.h * static_cast<float>(blockIdx.y * 8);
  const float loz = a.origin[2] + a.h * static_cast<float>(blockIdx.z * 4);
  const int32_t kx0 = max(cell_key(lox - m, a.inv), a.kmin[0]);
  const int32_t ky0 = max(cell_key(loy - m, a.inv), a.kmin[1]);
  const int32_t kz0 = max(cell_key(loz - m, a.inv), a.kmin[2]);
  const int32_t kx1 = min(cell_key(lox + 7.0f * a.h + m, a.inv), a.kmax[0]);
  const int32_t ky1 = min(cell_key(loy + 7.0f * a.h + m, a.inv), a.kmax[1]);
  const int32_t kz1 = min(cell_key(loz + 3.0f * a.h + m, a.inv), a.kmax[2]);
  const int32_t nx = kx1 - kx0 + 1, ny = ky1 - ky0 + 1, nz = kz1 - kz0 + 1;
  const int32_t ncell = (nx > 0 && ny > 0 && nz > 0) ? nx * ny * nz : 0;

  float lj[kMapLjClasses];
  float rb[kMapRbf];
#pragma unroll
  for (int c = 0; c < kMapLjClasses; ++c) lj[c] = 0.0f;
#pragma unroll
  for (int k = 0; k < kMapRbf; ++k) rb[k] = 0.0f;
  float phi = 0.0f, hyd = 0.0f;
  const float inv2s2 = 1.0f / (2.0f * kRbfSigma * kRbfSigma);

  for (int32_t cb = 0; cb < ncell; cb += 256) {
    const int32_t c = cb + static_cast<int32_t>(tid);
    uint32_t s = 0, len = 0;
    if (c < ncell) {
      const int32_t cz = c / (nx * ny), rem = c - cz * nx * ny;
      const int32_t cy = rem / nx, cx = rem - cy * nx;
      const uint32_t code = morton3(static_cast<uint32_t>(kx0 + cx - a.kmin[0]),
                                    static_cast<uint32_t>(ky0 + cy - a.kmin[1]),
                                    static_cast<uint32_t>(kz0 + cz - a.kmin[2]));
      s = a.cell_start[code];
      len = a.cell_start[code + 1] - s;
    }
    uint32_t incl = len;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(incl, o, 64);
      if (lane >= static_cast<uint32_t>(o)) incl += u;
    }
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t wo = 0;
    for (uint32_t w = 0; w < wave; ++w) wo += s_wsum[w];
    const uint32_t total = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
    s_cs[tid] = s;
    s_pre[tid] = wo + incl - len;
    __syncthreads();
    for (uint32_t a0 = 0; a0 < total; a0 += 256) {
      const uint32_t q = a0 + tid;
      if (q < total) {
        // the last listed cell whose first concatenated index is <= q
        int lo = 0, hi = 255;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_pre[mid] <= q) lo = mid;
          else hi = mid - 1;
        }
        s_at[tid] = a.atoms[s_cs[lo] + (q - s_pre[lo])];
      }
      __syncthreads();
      const uint32_t cnt = min(256u, total - a0);
      for (uint32_t u = 0; u < cnt; ++u) {
        const AgAtom B = s_at[u];
        const float dx = px - B.x, dy = py - B.y, dz = pz - B.z;
        const float r2 = (dx * dx + dy * dy) + dz * dz;
        if (valid && r2 <= a.rc2 && r2 >= kMinRSq) {
          const float rinv = 1.0f / sqrtf(r2);
          const float rinv2 = rinv * rinv;
#pragma unroll
          for (int k = 0; k < kMapLjClasses; ++k) {
            const float R = a.cls.rmh[k] + B.rmh;
            const float s2 = (R * R) * rinv2;
            const float s6 = (s2 * s2) * s2;
            lj[k] += (a.cls.sqe[k] * B.sqe) * (s6 * s6 - 2.0f * s6);
          }
          phi += B.q332 * rinv;
          if (tag_hyd(B.tag) && r2 < kHydroSq) hyd -= 0.5f;
          const float r = r2 * rinv;
#pragma unroll
          for (int k = 0; k < kMapRbf; ++k) {
            const float u2 = r - (kRbfMu0 + kRbfDmu * static_cast<float>(k));
            rb[k] += __expf(-(u2 * u2) * inv2s2);
          }
        }
      }
      __syncthreads();
    }
  }
  if (!valid) return;
  const size_t pt = (static_cast<size_t>(iz) * a.n[1] + iy) * a.n[0] + ix;
#pragma unroll
  for (int k = 0; k < kMapLjClasses; ++k) a.pts[k * a.npts + pt] = to_half_bits(lj[k]);
  a.pts[kMapLjClasses * a.npts + pt] = to_half_bits(phi);
  a.pts[(kMapLjClasses + 1) * a.npts + pt] = to_half_bits(hyd);
  uint16_t r16[kMapRbf];
#pragma unroll
  for (int k = 0; k < kMapRbf; ++k) r16[k] = to_half_bits(rb[k]);
  *reinterpret_cast<uint4*>(a.rbf + pt * kMapRbf) = *reinterpret_cast<const uint4*>(r16);
}

// Cell records: for every plane and cell, its 8 corner values (x fastest).
__global__ void k_map_expand(const uint16_t* __restrict__ pts, unsigned long long npts, int32_t n0,
                             int32_t n1, int32_t n2, unsigned long long ncells,
                             uint16_t* __restrict__ erec) {
  const unsigned long long t = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x;
  if (t >= ncells * kMapPlanes) return;
  const unsigned long long p = t / ncells, cell = t - p * ncells;
  const int32_t cx = n0 - 1, cy = n1 - 1;
  const int32_t i = static_cast<int32_t>(cell % cx);
  const int32_t j = static_cast<int32_t>((cell / cx) % cy);
  const int32_t k = static_cast<int32_t>(cell / (static_cast<unsigned long long>(cx) * cy));
  const uint16_t* src = pts + p * npts;
  uint16_t v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const size_t q = (static_cast<size_t>(k + (c >> 2)) * n1 + (j + ((c >> 1) & 1))) * n0 + (i + (c & 1));
    v[c] = src[q];
  }
  *reinterpret_cast<uint4*>(erec + t * 8) = *reinterpret_cast<const uint4*>(v);
}

struct FeatArgs {
  const MapView* views;  // per antigen
  const uint32_t* cand_ag;
  const uint32_t* cand_atom_off;
  const float4* pos;  // (x, y, z, q)
  const float4* par;  // (Rmin/2, eps, hydrophobic, 0)
  uint32_t n_cand;
  MapClasses cls;
  uint16_t* feat;  // [n_cand][kPrescreenFeatures] bf16, nullable
  float* map_e;    // [n_cand], nullable
};

__device__ inline float trilerp(const uint16_t* rec, float fx, float fy, float fz) {
  const uint4 w = *reinterpret_cast<const uint4*>(rec);
  const uint16_t* v = reinterpret_cast<const uint16_t*>(&w);
  const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
  const float x00 = gx * half_bits_to_f32(v[0]) + fx * half_bits_to_f32(v[1]);
  const float x10 = gx * half_bits_to_f32(v[2]) + fx * half_bits_to_f32(v[3]);
  const float x01 = gx * half_bits_to_f32(v[4]) + fx * half_bits_to_f32(v[5]);
  const float x11 = gx * half_bits_to_f32(v[6]) + fx * half_bits_to_f32(v[7]);
  return gz * (gy * x00 + fy * x10) + fz * (gy * x01 + fy * x11);
}

// Per-lane sums: per feature class its LJ map energy and radial densities
// (kFeatClasses x (1 + kMapRbf)), then the Coulomb sums over positive and
// negative atoms, the hydrophobic sum, the total map energy and the clash,
// contact and in-map counts.
constexpr int kClassCols = kFeatClasses * (1 + kMapRbf);  // 72
constexpr int kSumCols = kClassCols + 7;                   // 79
constexpr int kTrCols = 40;                                // columns per transpose round
constexpr int kTrStride = kTrCols + 1;                     // odd: conflict-free rows and columns

// The feature a sum column feeds.
__device__ inline int feature_of_column(int col) {
  if (col < kClassCols) {
    const int cl = col / (1 + kMapRbf), v = col - cl * (1 + kMapRbf);
    return v == 0 ? kFLj + cl : kFRbf + cl * kMapRbf + (v - 1);
  }
  return kFCoulPos + (col - kClassCols);  // 8..14 in column order
}

// One wave per candidate, a lane per atom. Each lane sums its own atoms'
// values per column in registers (fixed order over the candidate's passes of
// 64 atoms); the 64 lanes' columns are then added through an LDS transpose
// (lane r's row, summed by the column's lane in row order), which keeps the
// sums deterministic.
__global__ void __launch_bounds__(256) k_map_features(FeatArgs g) {
  __shared__ float s_tr[4][kWave * kTrStride];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t c = blockIdx.x * 4 + wave;
  if (c >= g.n_cand) return;  // whole wave (no block barrier below)
  const MapView V = g.views[g.cand_ag[c]];
  const uint32_t a0 = g.cand_atom_off[c], a1 = g.cand_atom_off[c + 1];
  const int32_t cx = V.n[0] - 1, cy = V.n[1] - 1, cz = V.n[2] - 1;
  float acc[kFeatClasses][1 + kMapRbf];
#pragma unroll
  for (int cl = 0; cl < kFeatClasses; ++cl)
#pragma unroll
    for (int v = 0; v <= kMapRbf; ++v) acc[cl][v] = 0.0f;
  float cpos = 0.0f, cneg = 0.0f, shyd = 0.0f, stot = 0.0f, nclash = 0.0f, ncont = 0.0f, nin = 0.0f;
  for (uint32_t b = a0; b < a1; b += kWave) {
    const uint32_t a = b + lane;
    const bool live = a < a1;
    float4 P = make_float4(0.f, 0.f, 0.f, 0.f), Q = P;
    if (live) {
      P = g.pos[a];
      Q = g.par[a];
    }
    // the atom's LJ class: its (Rmin/2, eps) pair, else the nearest one
    int lc = -1;
    float best = 3.0e38f;
    int near = 0;
#pragma unroll
    for (int k = 0; k < kMapLjClasses; ++k) {
      if (Q.x == g.cls.rmh[k] && Q.y == g.cls.eps[k]) lc = k;
      const float d = fabsf(Q.x - g.cls.rmh[k]) + fabsf(Q.y - g.cls.eps[k]);
      if (d < best) {
        best = d;
        near = k;
      }
    }
    if (lc < 0) lc = near;
    const bool hydp = Q.z != 0.0f;
    const int fc = hydp ? kMapLjClasses : lc;
    float elj = 0.0f, phi = 0.0f, hh = 0.0f;
    float rb[kMapRbf];
#pragma unroll
    for (int k = 0; k < kMapRbf; ++k) rb[k] = 0.0f;
    bool inside = false;
    if (live && V.valid) {
      const float ux = (P.x - V.origin[0]) * V.inv_h;
      const float uy = (P.y - V.origin[1]) * V.inv_h;
      const float uz = (P.z - V.origin[2]) * V.inv_h;
      const float ffx = floorf(ux), ffy = floorf(uy), ffz = floorf(uz);
      // (a NaN coordinate fails these tests: outside)
      inside = ffx >= 0.0f && ffy >= 0.0f && ffz >= 0.0f && ffx < static_cast<float>(cx) &&
               ffy < static_cast<float>(cy) && ffz < static_cast<float>(cz);
      if (inside) {
        const int32_t i = static_cast<int32_t>(ffx), j = static_cast<int32_t>(ffy), k = static_cast<int32_t>(ffz);
        const float fx = ux - ffx, fy = uy - ffy, fz = uz - ffz;
        const size_t cell = (static_cast<size_t>(k) * cy + j) * cx + i;
        elj = trilerp(V.erec + (static_cast<size_t>(lc) * V.ncells + cell) * 8, fx, fy, fz);
        phi = trilerp(V.erec + (static_cast<size_t>(kMapLjClasses) * V.ncells + cell) * 8, fx, fy, fz);
        hh = trilerp(V.erec + (static_cast<size_t>(kMapLjClasses + 1) * V.ncells + cell) * 8, fx, fy, fz);
        // the nearest grid point's radial densities
        const int32_t ni = i + (fx >= 0.5f ? 1 : 0), nj = j + (fy >= 0.5f ? 1 : 0), nk = k + (fz >= 0.5f ? 1 : 0);
        const size_t pt = (static_cast<size_t>(nk) * V.n[1] + nj) * V.n[0] + ni;
        const uint4 w = *reinterpret_cast<const uint4*>(V.rbf + pt * kMapRbf);
        const uint16_t* v = reinterpret_cast<const uint16_t*>(&w);
#pragma unroll
        for (int q = 0; q < kMapRbf; ++q) rb[q] = half_bits_to_f32(v[q]);
      }
    }
    const float ec = P.w * phi;
    const float eh = hydp ? hh : 0.0f;
    cpos += P.w > 0.0f ? ec : 0.0f;
    cneg += P.w < 0.0f ? ec : 0.0f;
    shyd += eh;
    stot += elj + ec + eh;
    nclash += (inside && elj > kClashE) ? 1.0f : 0.0f;
    ncont += (inside && elj < kContactE) ? 1.0f : 0.0f;
    nin += inside ? 1.0f : 0.0f;
#pragma unroll
    for (int cl = 0; cl < kFeatClasses; ++cl) {
      const bool mine = inside && fc == cl;
      if (!ballot64(mine)) continue;  // uniform: no atom of the pass in this class
      acc[cl][0] += mine ? elj : 0.0f;
#pragma unroll
      for (int q = 0; q < kMapRbf; ++q) acc[cl][1 + q] += mine ? rb[q] : 0.0f;
    }
  }
  // the column sums over the lanes: two transpose rounds of kTrCols columns
  float* tr = s_tr[wave];
  float colsum[2];
#pragma unroll
  for (int round = 0; round < 2; ++round) {
#pragma unroll
    for (int q = 0; q < kTrCols; ++q) {
      const int col = round * kTrCols + q;
      float v = 0.0f;
      if (col < kClassCols) v = acc[col / (1 + kMapRbf)][col % (1 + kMapRbf)];
      else if (col == kClassCols + 0) v = cpos;
      else if (col == kClassCols + 1) v = cneg;
      else if (col == kClassCols + 2) v = shyd;
      else if (col == kClassCols + 3) v = stot;
      else if (col == kClassCols + 4) v = nclash;
      else if (col == kClassCols + 5) v = ncont;
      else if (col == kClassCols + 6) v = nin;
      tr[lane * kTrStride + q] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float sum = 0.0f;
    if (lane < static_cast<uint32_t>(kTrCols))
      for (int r = 0; r < kWave; ++r) sum += tr[r * kTrStride + lane];
    colsum[round] = sum;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // the total map energy: column kClassCols + 3 (round 1, lane kClassCols + 3 - kTrCols)
  const float total = readlane_f32(colsum[1], kClassCols + 3 - kTrCols);
  if (g.map_e && lane == 0) g.map_e[c] = total;
  if (g.feat) {
    uint16_t* row = g.feat + static_cast<size_t>(c) * kPrescreenFeatures;
    auto put = [&](int k, float v) {
      float t = 0.0f;
      if (k <= kFHyd) t = asinhf(v * 0.5f);
      else if (k == kFTotal) t = asinhf(v * 0.1f);
      else if (k >= kFClash && k <= kFInMap) t = log1pf(v);
      else if (k >= kFRbf && k < kFRbf + kFeatClasses * kMapRbf) t = log1pf(v);
      row[k] = f32_to_bf16(t);
    };
#pragma unroll
    for (int round = 0; round < 2; ++round) {
      const int col = round * kTrCols + static_cast<int>(lane);
      if (lane < static_cast<uint32_t>(kTrCols) && col < kSumCols) put(feature_of_column(col), colsum[round]);
    }
    // features no column feeds
    if (lane == 0) row[kFInMap + 1] = 0;
    for (int k = kFRbf + kFeatClasses * kMapRbf + static_cast<int>(lane); k < kPrescreenFeatures; k += kWave)
      row[k] = 0;
  }
}

// Map values at listed grid points (test hook): the energy planes from the
// cell records (the point as a corner of an adjacent cell) and the radial
// densities, as f32.
__global__ void k_map_gather(const uint16_t* __restrict__ erec, const uint16_t* __restrict__ rbf,
                             int32_t n0, int32_t n1, int32_t n2, unsigned long long ncells,
                             const int32_t* __restrict__ ijk, uint32_t n, float* __restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int32_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
  const int32_t cx = n0 - 1, cy = n1 - 1;
  const int32_t ci = min(i, cx - 1), cj = min(j, cy - 1), ck = min(k, n2 - 2);
  const int corner = (i - ci) + 2 * (j - cj) + 4 * (k - ck);
  const size_t cell = (static_cast<size_t>(ck) * cy + cj) * cx + ci;
  float* o = out + static_cast<size_t>(q) * (kMapPlanes + kMapRbf);
  for (int p = 0; p < kMapPlanes; ++p) o[p] = half_bits_to_f32(erec[(p * ncells + cell) * 8 + corner]);
  const size_t pt = (static_cast<size_t>(k) * n1 + j) * n0 + i;
  for (int t = 0; t < kMapRbf; ++t) o[kMapPlanes + t] = half_bits_to_f32(rbf[pt * kMapRbf + t]);
}

static inline uint32_t nblk(unsigned long long n, uint32_t t) {
  const unsigned long long b = (n + t - 1) / t;
  return static_cast<uint32_t>(b ? b : 1);
}

int antigen_build_maps(dnab_ctx* ctx, dnab_antigen* ag, float spacing) {
  if (!(spacing > 0.0f)) spacing = kMapDefaultSpacing;
  if (!std::isfinite(spacing) || spacing < 0.05f)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "maps: spacing must be >= 0.05 A");
  MapDev& M = ag->map;
  M.erec.release();
  M.rbf.release();
  M.npts = M.ncells = 0;
  M.h = spacing;
  M.inv_h = 1.0f / spacing;
  const float pad = ag->cutoff + 2.0f * spacing;
  const bool empty = ag->grid.n == 0;
  size_t npts = 1;
  for (int d = 0; d < 3; ++d) {
    const float lo = empty ? 0.0f : ag->bbox_lo[d], hi = empty ? 0.0f : ag->bbox_hi[d];
    M.origin[d] = lo - pad;
    M.n[d] = static_cast<int32_t>(std::ceil((hi - lo + 2.0f * pad) / spacing)) + 1;
    npts *= static_cast<size_t>(M.n[d]);
  }
  const size_t ncells = static_cast<size_t>(M.n[0] - 1) * (M.n[1] - 1) * (M.n[2] - 1);
  const size_t bytes = ncells * kMapPlanes * 16 + npts * kMapRbf * 2 + npts * kMapPlanes * 2;
  if (bytes > kMapMaxBytes)
    return dnab_set_error(ctx, DNAB_ERR_CAPACITY, "maps: %zu bytes at spacing %.3f A exceed %zu",
                          bytes, spacing, kMapMaxBytes);
  DevBuf<uint16_t> pts;
  if (pts.alloc(npts * kMapPlanes) != hipSuccess || M.rbf.alloc(npts * kMapRbf) != hipSuccess ||
      M.erec.alloc(ncells * kMapPlanes * 8) != hipSuccess) {
    M.erec.release();
    M.rbf.release();
    return dnab_set_error(ctx, DNAB_ERR_ALLOC, "maps: allocation of %zu bytes failed", bytes);
  }
  MapBuildArgs a{};
  a.atoms = ag->atoms.p;
  a.cell_start = ag->grid.cell_start.p;
  for (int d = 0; d < 3; ++d) {
    a.kmin[d] = ag->grid.kmin[d];
    a.kmax[d] = empty ? ag->grid.kmin[d] - 1 : ag->grid.kmax[d];
    a.origin[d] = M.origin[d];
    a.n[d] = M.n[d];
  }
  a.inv = ag->grid.inv;
  a.rc = ag->cutoff;
  a.rc2 = ag->cutoff * ag->cutoff;
  a.h = spacing;
  a.cls = map_classes();
  a.pts = pts.p;
  a.rbf = M.rbf.p;
  a.npts = npts;
  const dim3 grid(nblk(M.n[0], 8), nblk(M.n[1], 8), nblk(M.n[2], 4));
  hipLaunchKernelGGL(k_map_points, grid, dim3(256), 0, ctx->stream, a);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_map_expand, dim3(nblk(ncells * kMapPlanes, 256)), dim3(256), 0, ctx->stream,
                     pts.p, static_cast<unsigned long long>(npts), M.n[0], M.n[1], M.n[2],
                     static_cast<unsigned long long>(ncells), M.erec.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  // the temporary point planes are freed when this returns: wait for the expansion
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  M.npts = npts;
  M.ncells = ncells;
  return DNAB_OK;
}

int population_map_features(dnab_ctx* ctx, dnab_population* pop, uint16_t* feat, float* map_e) {
  std::vector<MapView> views(pop->antigens.size());
  for (size_t a = 0; a < pop->antigens.size(); ++a) {
    dnab_antigen* ag = pop->antigens[a];
    if (ag->map.ncells == 0) DNAB_RETURN_IF(antigen_build_maps(ctx, ag, 0.0f));
    const MapDev& M = ag->map;
    MapView& v = views[a];
    v.erec = M.erec.p;
    v.rbf = M.rbf.p;
    for (int d = 0; d < 3; ++d) {
      v.origin[d] = M.origin[d];
      v.n[d] = M.n[d];
    }
    v.inv_h = M.inv_h;
    v.valid = M.ncells > 0 && ag->grid.n > 0 ? 1u : 0u;
    v.ncells = M.ncells;
  }
  DNAB_ALLOC_CHECK(ctx, pop->map_views.ensure(views.size()));
  DNAB_RETURN_IF(upload(ctx, pop->map_views.p, views.data(), views.size() * sizeof(MapView)));
  if (pop->n_cand == 0) return DNAB_OK;
  FeatArgs g{};
  g.views = pop->map_views.p;
  g.cand_ag = pop->cand_ag.p;
  g.cand_atom_off = pop->cand_atom_off.p;
  g.pos = pop->pos.p;
  g.par = pop->par.p;
  g.n_cand = static_cast<uint32_t>(pop->n_cand);
  g.cls = map_classes();
  g.feat = feat;
  g.map_e = map_e;
  hipEvent_t ev_end;
  hipEvent_t ev = timer_begin(ctx, "map_features", &ev_end);
  hipLaunchKernelGGL(k_map_features, dim3(nblk(pop->n_cand, 4)), dim3(256), 0, ctx->stream, g);
  timer_end(ctx, "map_features", ev, ev_end);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return DNAB_OK;
}

}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_antigen_build_maps(dnab_ctx* ctx, dnab_antigen* ag, float spacing) {
  if (!ctx || !ag || ag->ctx != ctx) return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  return antigen_build_maps(ctx, ag, spacing);
}

int dnab_antigen_map_info(const dnab_antigen* ag, float* spacing, float origin[3], int32_t dims[3],
                          size_t* bytes) {
  if (!ag) return DNAB_ERR_INVALID;
  const MapDev& M = ag->map;
  if (spacing) *spacing = M.ncells ? M.h : 0.0f;
  for (int d = 0; d < 3; ++d) {
    if (origin) origin[d] = M.origin[d];
    if (dims) dims[d] = M.ncells ? M.n[d] : 0;
  }
  if (bytes) *bytes = M.ncells * kMapPlanes * 16 + M.npts * kMapRbf * 2;
  return DNAB_OK;
}

int dnab_antigen_map_points(dnab_ctx* ctx, const dnab_antigen* ag, const int32_t* ijk, size_t n,
                            float* out) {
  if (!ctx || !ag || ag->ctx != ctx || (n && (!ijk || !out))) return DNAB_ERR_INVALID;
  const MapDev& M = ag->map;
  if (M.ncells == 0) return dnab_set_error(ctx, DNAB_ERR_INVALID, "map_points: no maps built");
  if (M.n[0] < 2 || M.n[1] < 2 || M.n[2] < 2) return DNAB_ERR_INVALID;
  for (size_t q = 0; q < n; ++q)
    for (int d = 0; d < 3; ++d)
      if (ijk[3 * q + d] < 0 || ijk[3 * q + d] >= M.n[d])
        return dnab_set_error(ctx, DNAB_ERR_INVALID, "map_points: point %zu outside the grid", q);
  if (n == 0) return DNAB_OK;
  (void)hipSetDevice(ctx->device);
  DevBuf<int32_t> d_ijk;
  DevBuf<float> d_out;
  DNAB_ALLOC_CHECK(ctx, d_ijk.alloc(3 * n));
  DNAB_ALLOC_CHECK(ctx, d_out.alloc(n * (kMapPlanes + kMapRbf)));
  DNAB_RETURN_IF(upload(ctx, d_ijk.p, ijk, 3 * n * sizeof(int32_t)));
  hipLaunchKernelGGL(k_map_gather, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, M.erec.p, M.rbf.p,
                     M.n[0], M.n[1], M.n[2], static_cast<unsigned long long>(M.ncells), d_ijk.p,
                     static_cast<uint32_t>(n), d_out.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(download(ctx, out, d_out.p, n * (kMapPlanes + kMapRbf) * sizeof(float)));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_population_map_energy(dnab_ctx* ctx, dnab_population* pop, float* out) {
  if (!ctx || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "map_energy"));
  (void)hipSetDevice(ctx->device);
  DNAB_ALLOC_CHECK(ctx, pop->scores.ensure(pop->n_cand));
  DNAB_RETURN_IF(population_map_features(ctx, pop, nullptr, pop->scores.p));
  if (out) DNAB_RETURN_IF(download(ctx, out, pop->scores.p, pop->n_cand * sizeof(float)));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_prescreen_feature_count(size_t* n) {
  if (!n) return DNAB_ERR_INVALID;
  *n = kPrescreenFeatures;
  return DNAB_OK;
}

}  // extern "C"
