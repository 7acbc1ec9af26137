// Device-side grid lookup and residue-level pair force, shared by the grid,
// pair and step kernels.
#pragma once
#include "device_common.h"
#include "dnab_internal.h"

namespace dnab {

__device__ inline bool cell_range(const GridView& g, int32_t kx, int32_t ky, int32_t kz,
                                  uint32_t& b, uint32_t& e) {
  if (kx < g.kmin[0] || kx > g.kmax[0] || ky < g.kmin[1] || ky > g.kmax[1] || kz < g.kmin[2] ||
      kz > g.kmax[2])
    return false;
  uint32_t code = morton3(static_cast<uint32_t>(kx - g.kmin[0]), static_cast<uint32_t>(ky - g.kmin[1]),
                          static_cast<uint32_t>(kz - g.kmin[2]));
  b = g.cell_start[code];
  e = g.cell_start[code + 1];
  return e > b;
}

// ---- residue-level forces (src/energy.rs:154-234, 322-398) -------------------
// One thread per antibody point; the pair loop visits antigen points in the
// reference's order and evaluates every operation in the reference's order
// with IEEE f32 (correctly rounded division and sqrt, no contraction), so the
// accumulated forces reproduce the reference's f32 arithmetic.

struct CgCloud {
  const float *x, *y, *z, *q, *eps, *sig;
  const uint8_t* hyd;
};

// The force term of antigen point i on the antibody point b: false when the
// pair is outside the reference's predicate; else (tx, ty, tz) = f * d, the
// products the reference adds (src/energy.rs:353-391).
__device__ inline bool cg_term(const CgCloud& ag, uint32_t i, float bx, float by, float bz,
                               float bq, float be, float bs, uint8_t bh, float& tx, float& ty,
                               float& tz) {
  const float kCutoffSq = 100.0f, kMinRSq = 0.25f, kHydroSq = 36.0f, kCoulomb = 332.0f;
  float dx = bx - ag.x[i];
  float dy = by - ag.y[i];
  float dz = bz - ag.z[i];
  float r_sq = dx * dx + dy * dy + dz * dz;
  if (r_sq > kCutoffSq || r_sq < kMinRSq) return false;
  float eps_ij = sqrtf(be * ag.eps[i]);
  float sig_ij = 0.5f * (bs + ag.sig[i]);
  float sigma_sq_ij = sig_ij * sig_ij;
  float s2 = sigma_sq_ij / r_sq;
  float s6 = s2 * s2 * s2;
  float s12 = s6 * s6;
  float f_lj = (24.0f * eps_ij / r_sq) * (2.0f * s12 - s6);
  float q2 = ag.q[i];
  float f_coul = 0.0f;
  if (bq != 0.0f && q2 != 0.0f) {
    float r = sqrtf(r_sq);
    f_coul = kCoulomb * bq * q2 / (r_sq * r);
  }
  float f_hydro = 0.0f;
  if (bh == 1 && ag.hyd[i] == 1 && r_sq < kHydroSq) f_hydro = -0.1f / sqrtf(r_sq);
  float f_total = f_lj + f_coul + f_hydro;
  tx = f_total * dx;
  ty = f_total * dy;
  tz = f_total * dz;
  return true;
}

__device__ inline void cg_pair(const CgCloud& ag, uint32_t i, float bx, float by, float bz,
                               float bq, float be, float bs, uint8_t bh, float& ax, float& ay,
                               float& az) {
  float tx, ty, tz;
  if (!cg_term(ag, i, bx, by, bz, bq, be, bs, bh, tx, ty, tz)) return;
  ax += tx;
  ay += ty;
  az += tz;
}

// Device-side residue-level force for one Calpha (used by the gradient step).
__device__ inline void cg_force_point(const GridView& g, const CgCloud& ag, float bx, float by,
                               float bz, float bq, float be, float bs, uint8_t bh, float& ax,
                               float& ay, float& az) {
  ax = ay = az = 0.0f;
  if (g.n == 0) return;
  int32_t cx = cell_key(bx, g.inv), cy = cell_key(by, g.inv), cz = cell_key(bz, g.inv);
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        uint32_t b, e;
        if (!cell_range(g, wrap_add(cx, dx), wrap_add(cy, dy), wrap_add(cz, dz), b, e)) continue;
        for (uint32_t k = b; k < e; ++k)
          cg_pair(ag, g.sorted_idx[k], bx, by, bz, bq, be, bs, bh, ax, ay, az);
      }
}

}  // namespace dnab
