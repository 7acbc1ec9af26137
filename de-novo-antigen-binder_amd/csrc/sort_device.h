// Candidate-atom sort codes, shared by the sort kernels (pair.hip) and the
// step kernel (step.hip), which counts atoms for the next step's sort.
#pragma once
#include "device_common.h"
#include "dnab_internal.h"

namespace dnab {

// Sort code of a candidate atom, or kExcluded when it lies farther than the
// cutoff from the antigen's atom bounding box (then no antigen atom is within
// the cutoff: the f32 box distance never exceeds the f32 pair distance; an
// infinite coordinate lands here, its r^2 being +inf), or kNonFinite for a
// NaN coordinate (then every pair's r^2 is NaN, which the reference's skip
// test keeps, src/energy.rs:284: its terms and the atom's sums are NaN).
__device__ inline uint32_t sort_code_of(const AgView& V, float x, float y, float z) {
  if (isnan(x) || isnan(y) || isnan(z)) return kNonFinite;
  const float bx = fmaxf(fmaxf(V.excl_lo[0] - x, x - V.excl_hi[0]), 0.0f);
  const float by = fmaxf(fmaxf(V.excl_lo[1] - y, y - V.excl_hi[1]), 0.0f);
  const float bz = fmaxf(fmaxf(V.excl_lo[2] - z, z - V.excl_hi[2]), 0.0f);
  const float d2 = (bx * bx + by * by) + bz * bz;
  if (!(d2 <= V.rc2)) return kExcluded;
  const float p[3] = {x, y, z};
  uint32_t c[3];
  for (int a = 0; a < 3; ++a) {
    int32_t k = f32_to_i32_sat(floorf((p[a] - V.sort_origin[a]) * V.sort_inv));
    k = k < 0 ? 0 : (k >= V.sort_dim[a] ? V.sort_dim[a] - 1 : k);
    c[a] = static_cast<uint32_t>(k);
  }
  const uint32_t brick =
      ((c[2] >> 2) * static_cast<uint32_t>(V.sort_nb[1]) + (c[1] >> 2)) *
          static_cast<uint32_t>(V.sort_nb[0]) + (c[0] >> 2);
  return V.sort_code_base + brick * 64u + morton3(c[0] & 3u, c[1] & 3u, c[2] & 3u);
}

// The per-atom outputs of an atom left out of the tiles: zeros when it has no
// pair (kExcluded), NaN when a coordinate is NaN and the antigen has atoms
// (kNonFinite: its sums over the antigen are NaN).
__device__ inline void fill_atom_outputs(uint32_t j, uint32_t code, const AgView& V, float4* force,
                                         double* e_atom, double* bd_atom, double* f64_atom) {
  const bool nan = code == kNonFinite && V.n_atoms > 0;
  const float vf = nan ? __builtin_nanf("") : 0.0f;
  const double vd = nan ? __builtin_nan("") : 0.0;
  force[j] = make_float4(vf, vf, vf, vf);
  e_atom[j] = vd;
  const size_t kb = static_cast<size_t>(kBdTerms) * j;
  for (int t = 0; t < kBdTerms; ++t) bd_atom[kb + t] = vd;
  const size_t k = 3 * static_cast<size_t>(j);
  f64_atom[k] = f64_atom[k + 1] = f64_atom[k + 2] = vd;
}

}  // namespace dnab
