// Residue geometry shared by the host structure builder (builder.cpp) and the
// device refinement kernel (refine.hip): the same f32 operations in the same
// order on both sides, so a residue placed on the device equals the host-built
// one bit for bit (both are compiled without contraction and with IEEE
// division and square root). Sines and cosines are passed in: the host takes
// them from its libm, the device from tables the host fills with the same
// calls.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define DNAB_HD __host__ __device__
#else
#define DNAB_HD
#endif

namespace dnab {
namespace geo {

struct V3 {
  float x, y, z;
};
DNAB_HD inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
DNAB_HD inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
// cross3 (src/allatom.rs:483-485)
DNAB_HD inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// norm3 (src/allatom.rs:486-489): the length floored at 1e-8
DNAB_HD inline V3 normalize(V3 v) {
  float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  if (l < 1e-8f) l = 1e-8f;
  return {v.x / l, v.y / l, v.z / l};
}
DNAB_HD inline float deg2rad(float d) { return d * (3.14159265358979323846f / 180.0f); }

// NERF (Parsons et al. 2005, J. Comput. Chem. 26(10)), the placement of
// place_atom (src/rotamer.rs:198-232): D with |CD| = bond_len, angle BCD of
// sine sa / cosine ca, dihedral ABCD of sine sd / cosine cd.
DNAB_HD inline V3 place(V3 A, V3 B, V3 C, float bond_len, float sa, float ca, float sd, float cd) {
  const V3 bc = normalize(sub(B, C));
  const V3 n = normalize(cross(sub(B, A), bc));
  const V3 m = cross(bc, n);
  return {C.x + bond_len * (-ca * bc.x + sa * cd * m.x + sa * sd * n.x),
          C.y + bond_len * (-ca * bc.y + sa * cd * m.y + sa * sd * n.y),
          C.z + bond_len * (-ca * bc.z + sa * cd * m.z + sa * sd * n.z)};
}

// build_ideal_backbone_around_ca (src/allatom.rs:502-517): N, CA, C, O, CB
// around a Calpha; returns 4 for Gly (no CB), else 5.
DNAB_HD inline int ideal_backbone(float cx, float cy, float cz, bool gly, V3 bb[5]) {
  const float kNCa = 1.46f, kCaC = 1.52f, kCO = 1.23f, kCaCb = 1.52f;
  bb[0] = {cx - kNCa, cy, cz};
  bb[1] = {cx, cy, cz};
  bb[2] = {cx + kCaC, cy, cz};
  bb[3] = {bb[2].x + 0.5f, bb[2].y + kCO * 0.9f, bb[2].z};
  bb[4] = {cx, cy - kCaCb, cz};
  return gly ? 4 : 5;
}

// rodrigues (src/allatom.rs:467-479): point rotated about the unit axis
// through origin by the angle of cosine ca and sine sa.
DNAB_HD inline V3 rodrigues(V3 p, V3 origin, V3 axis, float ca, float sa) {
  const V3 v = sub(p, origin);
  const float dot = v.x * axis.x + v.y * axis.y + v.z * axis.z;
  const V3 cx = cross(axis, v);
  const float omc = 1.0f - ca;
  const V3 r = {v.x * ca + cx.x * sa + axis.x * dot * omc, v.y * ca + cx.y * sa + axis.y * dot * omc,
                v.z * ca + cx.z * sa + axis.z * dot * omc};
  return add(r, origin);
}

}  // namespace geo
}  // namespace dnab
