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
//   - kMapRbf Gaussian radial densities sum exp(-(r - mu_k)^2 / (2 sigma^2))
//     (on every second grid point).
// A candidate atom's map energy is the trilinear interpolation of its class's
// LJ plane plus q times the potential plus (hydrophobic) the hydrophobic
// plane; summed over the candidate it approximates the exact energy except
// where the repulsive wall is steeper than the grid resolves. The planes of a
// grid cell share one record, so an atom's three reads touch one or two cache
// lines.
//
// Features of a candidate (kPrescreenFeatures = 32, bf16): the LJ map energy
// summed per atom class (the 7 LJ parameter classes, hydrophobic atoms a class
// of their own), the Coulomb sums over positive and negative atoms, the
// hydrophobic sum, the total map energy, counts of clashing, contacting and
// near atoms (in a cell with an antigen atom within the cutoff of a corner),
// and the radial densities summed over polar (N, O) and over
// the other atoms: atom-class x RBF distance histograms against the antigen's
// heavy atoms.
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

constexpr float kMapDefaultSpacing = 0.625f;
constexpr float kRbfMu0 = 2.0f, kRbfDmu = 1.0f, kRbfSigma = 0.75f;
constexpr float kMapClamp = 60000.0f;   // fp16 range; a wall value beyond it is saturated
constexpr size_t kMapMaxBytes = size_t(6) << 30;
constexpr int kFeatClasses = kMapLjClasses + 1;  // LJ classes + hydrophobic atoms
constexpr float kClashE = 1.0f;     // an atom's LJ map energy above this is a clash
constexpr float kContactE = -0.05f; // ... below this, a contact

// Feature layout
constexpr int kFLj = 0;            // [kFeatClasses] LJ map energy per atom class
constexpr int kFCoulPos = 8, kFCoulNeg = 9, kFHyd = 10, kFTotal = 11;
constexpr int kFClash = 12, kFContact = 13, kFNear = 14;
constexpr int kFRbfPolar = 16;     // [kMapRbf] radial densities at N and O atoms
constexpr int kFRbfApolar = 24;    // [kMapRbf] ... at the other atoms
static_assert(kFRbfApolar + kMapRbf == kPrescreenFeatures, "feature layout");
static_assert(kFCoulNeg == kFCoulPos + 1 && kFHyd == kFCoulPos + 2 && kFTotal == kFCoulPos + 3 &&
                  kFClash == kFCoulPos + 4 && kFContact == kFCoulPos + 5 && kFNear == kFCoulPos + 6,
              "the physics features follow the columns in order");

// N* (class 2), O, O2 (3), OH (4), OS (5) in the class order of map_classes
__device__ inline bool polar_class(int fc) { return fc >= 2 && fc <= 5; }

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
  uint8_t* phot;  // [point]: an antigen atom passes the pair predicate
  uint16_t* rbf;  // [coarse point][kMapRbf] fp16
  int32_t nc[3];  // coarse points per axis (every second point)
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
  bool any = false;
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
          any = true;
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
  a.phot[pt] = any ? 1 : 0;
  if ((ix | iy | iz) & 1) return;  // radial densities on every second point
  uint16_t r16[kMapRbf];
#pragma unroll
  for (int k = 0; k < kMapRbf; ++k) r16[k] = to_half_bits(rb[k]);
  const size_t cpt = (static_cast<size_t>(iz >> 1) * a.nc[1] + (iy >> 1)) * a.nc[0] + (ix >> 1);
  *reinterpret_cast<uint4*>(a.rbf + cpt * kMapRbf) = *reinterpret_cast<const uint4*>(r16);
}

// Cell records: for every cell and plane, the plane's 8 corner values (x
// fastest), in the plane's slot of the cell's record or, for the sulfur
// class, in its own plane.
__global__ void k_map_expand(const uint16_t* __restrict__ pts, unsigned long long npts, int32_t n0,
                             int32_t n1, int32_t n2, unsigned long long ncells,
                             uint16_t* __restrict__ erec, uint16_t* __restrict__ esul) {
  const unsigned long long t = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x;
  if (t >= ncells * kMapPlanes) return;
  const unsigned long long cell = t / kMapPlanes;
  const int p = static_cast<int>(t - cell * kMapPlanes);
  const int32_t cx = n0 - 1, cy = n1 - 1;
  const int32_t i = static_cast<int32_t>(cell % cx);
  const int32_t j = static_cast<int32_t>((cell / cx) % cy);
  const int32_t k = static_cast<int32_t>(cell / (static_cast<unsigned long long>(cx) * cy));
  const uint16_t* src = pts + static_cast<size_t>(p) * npts;
  uint16_t v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const size_t q = (static_cast<size_t>(k + (c >> 2)) * n1 + (j + ((c >> 1) & 1))) * n0 + (i + (c & 1));
    v[c] = src[q];
  }
  const int slot = map_slot(p);
  uint16_t* dst = slot < 0 ? esul + cell * 8 : erec + (cell * kMapRecSlots + slot) * 8;
  *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(v);
}

// One bit per cell (32 cells a word, in cell order): some corner of the
// cell has an antigen atom within the cutoff. An atom in a cold cell reads
// only zeros (every corner, and its nearest radial-density point, which is
// a corner).
__global__ void k_map_hot(const uint8_t* __restrict__ phot, int32_t n0, int32_t n1,
                          unsigned long long ncells, uint32_t* __restrict__ hot) {
  const unsigned long long w = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x;
  if (w * 32 >= ncells) return;
  const int32_t cx = n0 - 1, cy = n1 - 1;
  uint32_t bits = 0;
  for (int b = 0; b < 32; ++b) {
    const unsigned long long cell = w * 32 + b;
    if (cell >= ncells) break;
    const int32_t i = static_cast<int32_t>(cell % cx);
    const int32_t j = static_cast<int32_t>((cell / cx) % cy);
    const int32_t k = static_cast<int32_t>(cell / (static_cast<unsigned long long>(cx) * cy));
    bool h = false;
    for (int c = 0; c < 8; ++c)
      h = h || phot[(static_cast<size_t>(k + (c >> 2)) * n1 + (j + ((c >> 1) & 1))) * n0 + (i + (c & 1))];
    bits |= h ? (1u << b) : 0u;
  }
  hot[w] = bits;
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

// The nearest point of the coarse grid (spacing 2h) to fine coordinate u >= 0.
__device__ inline int32_t coarse_nearest(float u, int32_t nc) {
  const float uc = u * 0.5f;
  const float f = floorf(uc);
  return min(static_cast<int32_t>(f) + (uc - f >= 0.5f ? 1 : 0), nc - 1);
}

__device__ inline float trilerp(uint4 w, float fx, float fy, float fz) {
  const uint16_t* v = reinterpret_cast<const uint16_t*>(&w);
  const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
  const float x00 = gx * half_bits_to_f32(v[0]) + fx * half_bits_to_f32(v[1]);
  const float x10 = gx * half_bits_to_f32(v[2]) + fx * half_bits_to_f32(v[3]);
  const float x01 = gx * half_bits_to_f32(v[4]) + fx * half_bits_to_f32(v[5]);
  const float x11 = gx * half_bits_to_f32(v[6]) + fx * half_bits_to_f32(v[7]);
  return gz * (gy * x00 + fy * x10) + fz * (gy * x01 + fy * x11);
}

// Per-lane sums, transposed through LDS into column sums: the LJ map energy
// per class, the radial densities at polar and at other atoms, then the
// Coulomb sums over positive and negative atoms, the hydrophobic sum, the
// total map energy and the clash, contact and near-atom counts.
constexpr int kCols = kFeatClasses + 2 * kMapRbf + 7;  // 31
constexpr int kTrStride = 33;                          // odd: conflict-free rows and columns
constexpr int kAtomsPerLane = 3;                       // atoms whose loads a lane has in flight

__device__ inline int feature_of_column(int col) {
  if (col < kFeatClasses) return kFLj + col;
  if (col < kFeatClasses + kMapRbf) return kFRbfPolar + (col - kFeatClasses);
  if (col < kFeatClasses + 2 * kMapRbf) return kFRbfApolar + (col - kFeatClasses - kMapRbf);
  return kFCoulPos + (col - kFeatClasses - 2 * kMapRbf);  // 8..14 in column order
}

// One wave per candidate, a lane per atom, kAtomsPerLane atoms a lane at a
// time (all their loads issued before any is used). Each lane sums its own
// atoms' values per column in registers in a fixed order; the lanes' columns
// are then added through an LDS transpose (lane r's row, summed by the
// column's lane in row order), which keeps the sums deterministic.
__global__ void __launch_bounds__(256) k_map_features(FeatArgs g) {
  __shared__ float s_tr[4][kWave * kTrStride];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t c = blockIdx.x * 4 + wave;
  if (c >= g.n_cand) return;  // whole wave (no block barrier below)
  const MapView V = g.views[g.cand_ag[c]];
  const uint32_t a0 = g.cand_atom_off[c], a1 = g.cand_atom_off[c + 1];
  const int32_t cx = V.n[0] - 1, cy = V.n[1] - 1, cz = V.n[2] - 1;
  float lj[kFeatClasses], rbp[kMapRbf], rba[kMapRbf];
#pragma unroll
  for (int k = 0; k < kFeatClasses; ++k) lj[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < kMapRbf; ++k) rbp[k] = rba[k] = 0.0f;
  float cpos = 0.0f, cneg = 0.0f, shyd = 0.0f, stot = 0.0f, nclash = 0.0f, ncont = 0.0f, nin = 0.0f;
  for (uint32_t b = a0; b < a1; b += kWave * kAtomsPerLane) {
    float4 P[kAtomsPerLane], Q[kAtomsPerLane];
#pragma unroll
    for (int q = 0; q < kAtomsPerLane; ++q) {
      const uint32_t a = b + q * kWave + lane;
      P[q] = make_float4(0.f, 0.f, 0.f, 0.f);
      Q[q] = P[q];
      if (a < a1) {
        P[q] = g.pos[a];
        Q[q] = g.par[a];
      }
    }
    int lc[kAtomsPerLane];
    bool in[kAtomsPerLane];
    float fx[kAtomsPerLane], fy[kAtomsPerLane], fz[kAtomsPerLane];
    uint4 rl[kAtomsPerLane], rp[kAtomsPerLane], rh[kAtomsPerLane], rr[kAtomsPerLane];
#pragma unroll
    for (int q = 0; q < kAtomsPerLane; ++q) {
      const uint32_t a = b + q * kWave + lane;
      // the atom's LJ class: its (Rmin/2, eps) pair, else the nearest one
      int cls = -1, near = 0;
      float best = 3.0e38f;
#pragma unroll
      for (int k = 0; k < kMapLjClasses; ++k) {
        if (Q[q].x == g.cls.rmh[k] && Q[q].y == g.cls.eps[k]) cls = k;
        const float d = fabsf(Q[q].x - g.cls.rmh[k]) + fabsf(Q[q].y - g.cls.eps[k]);
        if (d < best) {
          best = d;
          near = k;
        }
      }
      lc[q] = cls < 0 ? near : cls;
      const float ux = (P[q].x - V.origin[0]) * V.inv_h;
      const float uy = (P[q].y - V.origin[1]) * V.inv_h;
      const float uz = (P[q].z - V.origin[2]) * V.inv_h;
      const float ffx = floorf(ux), ffy = floorf(uy), ffz = floorf(uz);
      // (a NaN coordinate fails these tests: outside)
      in[q] = a < a1 && V.valid && ffx >= 0.0f && ffy >= 0.0f && ffz >= 0.0f &&
              ffx < static_cast<float>(cx) && ffy < static_cast<float>(cy) && ffz < static_cast<float>(cz);
      fx[q] = ux - ffx;
      fy[q] = uy - ffy;
      fz[q] = uz - ffz;
      rl[q] = rp[q] = rh[q] = rr[q] = make_uint4(0u, 0u, 0u, 0u);
      size_t cell = 0;
      if (in[q]) {
        const int32_t i = static_cast<int32_t>(ffx), j = static_cast<int32_t>(ffy), k = static_cast<int32_t>(ffz);
        cell = (static_cast<size_t>(k) * cy + j) * cx + i;
        in[q] = (V.hot[cell >> 5] >> (cell & 31)) & 1u;  // a cold cell reads zeros
      }
      if (in[q]) {
        const uint16_t* rec = V.erec + cell * (kMapRecSlots * 8);
        const int sl = map_slot(lc[q]);
        rl[q] = *reinterpret_cast<const uint4*>(sl < 0 ? V.esul + cell * 8 : rec + sl * 8);
        rp[q] = *reinterpret_cast<const uint4*>(rec + map_slot(kMapLjClasses) * 8);
        rh[q] = *reinterpret_cast<const uint4*>(rec + map_slot(kMapLjClasses + 1) * 8);
        // the radial densities of the nearest point of the coarse grid
        const int32_t ni = coarse_nearest(ux, V.nc[0]), nj = coarse_nearest(uy, V.nc[1]),
                      nk = coarse_nearest(uz, V.nc[2]);
        const size_t pt = (static_cast<size_t>(nk) * V.nc[1] + nj) * V.nc[0] + ni;
        rr[q] = *reinterpret_cast<const uint4*>(V.rbf + pt * kMapRbf);
      }
    }
#pragma unroll
    for (int q = 0; q < kAtomsPerLane; ++q) {
      const bool hydp = Q[q].z != 0.0f;
      const int fc = hydp ? kMapLjClasses : lc[q];
      const float elj = in[q] ? trilerp(rl[q], fx[q], fy[q], fz[q]) : 0.0f;
      const float phi = in[q] ? trilerp(rp[q], fx[q], fy[q], fz[q]) : 0.0f;
      const float hh = in[q] ? trilerp(rh[q], fx[q], fy[q], fz[q]) : 0.0f;
      const float ec = P[q].w * phi;
      const float eh = hydp ? hh : 0.0f;
      cpos += P[q].w > 0.0f ? ec : 0.0f;
      cneg += P[q].w < 0.0f ? ec : 0.0f;
      shyd += eh;
      stot += elj + ec + eh;
      nclash += (in[q] && elj > kClashE) ? 1.0f : 0.0f;
      ncont += (in[q] && elj < kContactE) ? 1.0f : 0.0f;
      nin += in[q] ? 1.0f : 0.0f;
#pragma unroll
      for (int k = 0; k < kFeatClasses; ++k) lj[k] += fc == k ? elj : 0.0f;
      const uint16_t* v = reinterpret_cast<const uint16_t*>(&rr[q]);
      const bool pol = polar_class(fc);
#pragma unroll
      for (int k = 0; k < kMapRbf; ++k) {
        const float r = half_bits_to_f32(v[k]);  // zero outside the map
        rbp[k] += pol ? r : 0.0f;
        rba[k] += pol ? 0.0f : r;
      }
    }
  }
  // the column sums over the lanes (one transpose)
  float* tr = s_tr[wave];
#pragma unroll
  for (int col = 0; col < kCols; ++col) {
    float v;
    if (col < kFeatClasses) v = lj[col];
    else if (col < kFeatClasses + kMapRbf) v = rbp[col - kFeatClasses];
    else if (col < kFeatClasses + 2 * kMapRbf) v = rba[col - kFeatClasses - kMapRbf];
    else if (col == kFeatClasses + 2 * kMapRbf + 0) v = cpos;
    else if (col == kFeatClasses + 2 * kMapRbf + 1) v = cneg;
    else if (col == kFeatClasses + 2 * kMapRbf + 2) v = shyd;
    else if (col == kFeatClasses + 2 * kMapRbf + 3) v = stot;
    else if (col == kFeatClasses + 2 * kMapRbf + 4) v = nclash;
    else if (col == kFeatClasses + 2 * kMapRbf + 5) v = ncont;
    else v = nin;
    tr[lane * kTrStride + col] = v;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float sum = 0.0f;
  if (lane < static_cast<uint32_t>(kCols))
    for (int r = 0; r < kWave; ++r) sum += tr[r * kTrStride + lane];
  const float total = readlane_f32(sum, kFeatClasses + 2 * kMapRbf + 3);
  if (g.map_e && lane == 0) g.map_e[c] = total;
  if (g.feat) {
    uint16_t* row = g.feat + static_cast<size_t>(c) * kPrescreenFeatures;
    if (lane < static_cast<uint32_t>(kCols)) {
      const int k = feature_of_column(static_cast<int>(lane));
      float t;
      if (k <= kFHyd) t = asinhf(sum * 0.5f);
      else if (k == kFTotal) t = asinhf(sum * 0.1f);
      else t = log1pf(sum);  // counts and radial densities
      row[k] = f32_to_bf16(t);
    } else if (lane == static_cast<uint32_t>(kCols)) {
      row[kFNear + 1] = 0;  // the one feature no column feeds
    }
  }
}

// Map values at listed grid points (test hook): the energy planes from the
// cell records (the point as a corner of an adjacent cell), as f32, and the
// radial densities where the point is on the coarse grid (NaN elsewhere).
__global__ void k_map_gather(const uint16_t* __restrict__ erec, const uint16_t* __restrict__ esul,
                             const uint16_t* __restrict__ rbf, const uint32_t* __restrict__ hot,
                             int32_t n0, int32_t n1, int32_t n2, int32_t nc0, int32_t nc1,
                             const int32_t* __restrict__ ijk, uint32_t n, float* __restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int32_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
  const int32_t cx = n0 - 1, cy = n1 - 1;
  const int32_t ci = min(i, cx - 1), cj = min(j, cy - 1), ck = min(k, n2 - 2);
  const int corner = (i - ci) + 2 * (j - cj) + 4 * (k - ck);
  const size_t cell = (static_cast<size_t>(ck) * cy + cj) * cx + ci;
  float* o = out + static_cast<size_t>(q) * (kMapPlanes + kMapRbf + 1);
  for (int p = 0; p < kMapPlanes; ++p) {
    const int sl = map_slot(p);
    o[p] = half_bits_to_f32(sl < 0 ? esul[cell * 8 + corner] : erec[cell * (kMapRecSlots * 8) + sl * 8 + corner]);
  }
  const bool coarse = ((i | j | k) & 1) == 0;
  const size_t pt = (static_cast<size_t>(k >> 1) * nc1 + (j >> 1)) * nc0 + (i >> 1);
  for (int t = 0; t < kMapRbf; ++t)
    o[kMapPlanes + t] = coarse ? half_bits_to_f32(rbf[pt * kMapRbf + t]) : __builtin_nanf("");
  // the occupancy bit of the cell whose lower corner is the point
  const bool lower = i < cx && j < cy && k < n2 - 1;
  const size_t own = (static_cast<size_t>(k) * cy + j) * cx + i;
  o[kMapPlanes + kMapRbf] = lower ? static_cast<float>((hot[own >> 5] >> (own & 31)) & 1u) : __builtin_nanf("");
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
  M.esul.release();
  M.rbf.release();
  M.hot.release();
  M.npts = M.ncells = M.ncpts = 0;
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
  size_t ncpts = 1;
  for (int d = 0; d < 3; ++d) {
    M.nc[d] = (M.n[d] - 1) / 2 + 1;
    ncpts *= static_cast<size_t>(M.nc[d]);
  }
  const size_t bytes = ncells * (kMapRecSlots + 1) * 16 + ncells / 8 + ncpts * kMapRbf * 2 +
                       npts * (kMapPlanes * 2 + 1);
  if (bytes > kMapMaxBytes)
    return dnab_set_error(ctx, DNAB_ERR_CAPACITY, "maps: %zu bytes at spacing %.3f A exceed %zu",
                          bytes, spacing, kMapMaxBytes);
  DevBuf<uint16_t> pts;
  DevBuf<uint8_t> phot;
  if (pts.alloc(npts * kMapPlanes) != hipSuccess || phot.alloc(npts) != hipSuccess ||
      M.hot.alloc((ncells + 31) / 32) != hipSuccess || M.rbf.alloc(ncpts * kMapRbf) != hipSuccess ||
      M.erec.alloc(ncells * kMapRecSlots * 8) != hipSuccess || M.esul.alloc(ncells * 8) != hipSuccess) {
    M.erec.release();
    M.esul.release();
    M.rbf.release();
    M.hot.release();
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
    a.nc[d] = M.nc[d];
  }
  a.inv = ag->grid.inv;
  a.rc = ag->cutoff;
  a.rc2 = ag->cutoff * ag->cutoff;
  a.h = spacing;
  a.cls = map_classes();
  a.pts = pts.p;
  a.phot = phot.p;
  a.rbf = M.rbf.p;
  a.npts = npts;
  const dim3 grid(nblk(M.n[0], 8), nblk(M.n[1], 8), nblk(M.n[2], 4));
  hipLaunchKernelGGL(k_map_points, grid, dim3(256), 0, ctx->stream, a);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_map_expand, dim3(nblk(ncells * kMapPlanes, 256)), dim3(256), 0, ctx->stream,
                     pts.p, static_cast<unsigned long long>(npts), M.n[0], M.n[1], M.n[2],
                     static_cast<unsigned long long>(ncells), M.erec.p, M.esul.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_map_hot, dim3(nblk((ncells + 31) / 32, 256)), dim3(256), 0, ctx->stream, phot.p,
                     M.n[0], M.n[1], static_cast<unsigned long long>(ncells), M.hot.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  // the temporary point planes are freed when this returns: wait for the expansion
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  M.npts = npts;
  M.ncells = ncells;
  M.ncpts = ncpts;
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
    v.esul = M.esul.p;
    v.rbf = M.rbf.p;
    v.hot = M.hot.p;
    for (int d = 0; d < 3; ++d) {
      v.origin[d] = M.origin[d];
      v.n[d] = M.n[d];
    }
    v.inv_h = M.inv_h;
    v.valid = M.ncells > 0 && ag->grid.n > 0 ? 1u : 0u;
    for (int d = 0; d < 3; ++d) v.nc[d] = M.nc[d];
    v.pad = 0;
  }
  const bool same = pop->map_views_host.size() == views.size() &&
                    std::memcmp(pop->map_views_host.data(), views.data(), views.size() * sizeof(MapView)) == 0;
  if (!same) {
    DNAB_ALLOC_CHECK(ctx, pop->map_views.ensure(views.size()));
    DNAB_RETURN_IF(upload(ctx, pop->map_views.p, views.data(), views.size() * sizeof(MapView)));
    pop->map_views_host = views;
  }
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
  if (bytes) *bytes = M.ncells * (kMapRecSlots + 1) * 16 + (M.ncells + 31) / 32 * 4 + M.ncpts * kMapRbf * 2;
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
  DNAB_ALLOC_CHECK(ctx, d_out.alloc(n * (kMapPlanes + kMapRbf + 1)));
  DNAB_RETURN_IF(upload(ctx, d_ijk.p, ijk, 3 * n * sizeof(int32_t)));
  hipLaunchKernelGGL(k_map_gather, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, M.erec.p, M.esul.p, M.rbf.p, M.hot.p,
                     M.n[0], M.n[1], M.n[2], M.nc[0], M.nc[1], d_ijk.p,
                     static_cast<uint32_t>(n), d_out.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(download(ctx, out, d_out.p, n * (kMapPlanes + kMapRbf + 1) * sizeof(float)));
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
