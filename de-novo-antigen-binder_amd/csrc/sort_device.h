// Candidate-atom sort codes, shared by the sort kernels (pair.hip) and the
// step kernel (step.hip), which counts atoms for the next step's sort.
#pragma once
#include "device_common.h"
#include "dnab_internal.h"

namespace dnab {

// Sort code of a candidate atom, or kExcluded when it lies farther than the
// cutoff from the antigen's atom bounding box (then no antigen atom is within
// the cutoff: the f32 box distance never exceeds the f32 pair distance).
__device__ inline uint32_t sort_code_of(const AgView& V, float x, float y, float z) {
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

// Zero an excluded atom's per-atom outputs (it has no pair).
__device__ inline void zero_atom_outputs(uint32_t j, float4* force, double* e_atom,
                                         double* bd_atom, double* f64_atom) {
  force[j] = make_float4(0.f, 0.f, 0.f, 0.f);
  e_atom[j] = 0.0;
  const size_t kb = static_cast<size_t>(kBdTerms) * j;
  for (int t = 0; t < kBdTerms; ++t) bd_atom[kb + t] = 0.0;
  const size_t k = 3 * static_cast<size_t>(j);
  f64_atom[k] = f64_atom[k + 1] = f64_atom[k + 2] = 0.0;
}

}  // namespace dnab
