// Residue update of the gradient step, shared by the per-step kernel
// (step.hip) and the fused trajectory kernel (traj.hip) so that both perform
// the same f32 operations in the same order.
//
// Restates src/diffusion.rs:300-339: the radial restraint toward the antigen
// centre, the clamped gradient displacement, optional uniform noise and
// set_ca_pos as a rigid translation of the residue (src/allatom.rs:279-293).
#pragma once
#include "device_common.h"
#include "dnab_internal.h"

namespace dnab {

struct StepParams {
  float step, max_disp, restraint_k, restraint_radius;
  int with_noise, force_model;
  uint32_t seed_lo, seed_hi;
};

// An atom's force component as the residue update reads it: a saturated
// component (+-inf, pair_device.h) counts as the saturation bound 2^39, so
// opposite saturated atoms of one residue do not make a NaN step; every
// other value passes unchanged.
__device__ inline float dyn_force(float v) {
  constexpr float kSat = 549755813888.0f;  // 2^59 units of 2^-20
  return fminf(fmaxf(v, -kSat), kSat);
}

__device__ inline float clamp_rust(float v, float lo, float hi) {
  // f32::clamp: NaN passes through.
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

// Displacement of residue `r` (global index) whose Calpha sits at `p`, given
// the summed force (fx, fy, fz) on its atoms; (cx, cy, cz) is the antigen
// centre and r0 the restraint radius.
__device__ inline float4 residue_delta(float cx, float cy, float cz, const StepParams& sp, float r0,
                                       float4 p, float fx, float fy, float fz, float sigma,
                                       uint32_t iter, uint32_t r) {
  // src/diffusion.rs:318-327
  const float dx_c = p.x - cx;
  const float dy_c = p.y - cy;
  const float dz_c = p.z - cz;
  const float d = fmaxf(sqrtf(dx_c * dx_c + dy_c * dy_c + dz_c * dz_c), 0.001f);
  const float excess = d - r0;
  const float f_r = -sp.restraint_k * excess / d;
  fx += f_r * dx_c;
  fy += f_r * dy_c;
  fz += f_r * dz_c;
  // src/diffusion.rs:329-337
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  if (sp.with_noise) {
    const Philox4 u = philox4x32_10(iter, r, 0x6e6f6973u, 0u, sp.seed_lo, sp.seed_hi);
    const float two = sigma + sigma;
    nx = u01_from_u32(u.v[0]) * two - sigma;
    ny = u01_from_u32(u.v[1]) * two - sigma;
    nz = u01_from_u32(u.v[2]) * two - sigma;
  }
  const float new_x = p.x + clamp_rust(sp.step * fx, -sp.max_disp, sp.max_disp) + nx;
  const float new_y = p.y + clamp_rust(sp.step * fy, -sp.max_disp, sp.max_disp) + ny;
  const float new_z = p.z + clamp_rust(sp.step * fz, -sp.max_disp, sp.max_disp) + nz;
  // set_ca_pos: the Calpha's displacement, added to every atom of the residue
  return make_float4(new_x - p.x, new_y - p.y, new_z - p.z, 0.0f);
}

// traj.hip: the fused trajectory kernel for the all-atom loop.
bool trajectory_supported(const dnab_population* pop, const dnab_langevin_params* p);
int trajectory_run(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                   const StepParams& sp);
// step.hip: the loop as per-step population kernels; top-k rows 0..n_mid
int step_kernel_loop(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                     const StepParams& sp, uint32_t nc, uint32_t k, uint32_t n_mid);

}  // namespace dnab
