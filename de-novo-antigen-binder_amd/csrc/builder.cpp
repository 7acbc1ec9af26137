// Host-side construction of all-atom clouds and synthetic inputs.
//
// protein_from_ca_trace semantics: src/allatom.rs:494-541 (ideal backbone
// around each Calpha, most probable rotamer, push_residue at
// src/allatom.rs:201-240); NERF placement: src/rotamer.rs:198-232,
// build_side_chain: src/rotamer.rs:439-462. The synthetic Calpha generators
// (ring layout of src/diffusion.rs:264-277, globular random walks) use this
// library's own SplitMix64 stream; the reference's rand::SmallRng streams are
// not reproduced.
#include <cmath>
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "dnab_internal.h"
#include "geometry.h"
#include "ff_tables.h"

namespace dnab {

using geo::V3;

// NERF (Parsons et al. 2005, J. Comput. Chem. 26(10)): place D so that
// |CD| = bond_len, angle BCD = angle_deg, dihedral ABCD = dihedral_deg.
void place_atom(const float a[3], const float b[3], const float c[3], float bond_len,
                float angle_deg, float dihedral_deg, float out[3]) {
  const float ang = geo::deg2rad(angle_deg), dih = geo::deg2rad(dihedral_deg);
  const V3 d = geo::place({a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}, bond_len,
                          std::sin(ang), std::cos(ang), std::sin(dih), std::cos(dih));
  out[0] = d.x;
  out[1] = d.y;
  out[2] = d.z;
}

// Axis-aligned ideal backbone around a Calpha, src/allatom.rs:502-517.
int ideal_backbone(const float ca[3], int aa, float bb[5][3]) {
  V3 v[5];
  const int n = geo::ideal_backbone(ca[0], ca[1], ca[2], aa == kGly, v);
  for (int k = 0; k < 5; ++k) {
    bb[k][0] = v[k].x;
    bb[k][1] = v[k].y;
    bb[k][2] = v[k].z;
  }
  return n;
}

// Append all heavy atoms of one residue; returns the number appended.
int build_residue_atoms(int aa, const float ca[3], const float chi[4], float* xyz_out,
                        int* n_backbone_out) {
  float atoms[16][3];
  int nb = ideal_backbone(ca, aa, reinterpret_cast<float(*)[3]>(atoms));
  const ResTopology& topo = kResidueTopology[aa];
  if (nb > topo.n_atoms) nb = topo.n_atoms;
  int n = nb;
  const BondDef* defs = kSidechainDefs[aa];
  for (int k = 0; k < kSidechainCount[aa] && n < topo.n_atoms; ++k) {
    const BondDef& d = defs[k];
    float dih = d.chi_idx == 255 ? d.dihedral_fixed : chi[d.chi_idx];
    place_atom(atoms[d.ref_a], atoms[d.ref_b], atoms[d.ref_c], d.bond_len, d.angle_deg, dih,
               atoms[n]);
    ++n;
  }
  std::memcpy(xyz_out, atoms, sizeof(float) * 3 * n);
  if (n_backbone_out) *n_backbone_out = nb;
  return n;
}

uint64_t SplitMix64::next() {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
float SplitMix64::uniform(float lo, float hi) {
  float u = static_cast<float>(next() >> 40) * (1.0f / 16777216.0f);
  return lo + (hi - lo) * u;
}
uint32_t SplitMix64::below(uint32_t n) {
  return static_cast<uint32_t>(((next() >> 32) * static_cast<uint64_t>(n)) >> 32);
}

}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_build_protein(const float* ca_x, const float* ca_y, const float* ca_z,
                       const uint8_t* amino_acid, const uint8_t* rotamer_idx, size_t n_res,
                       dnab_protein_out* out, size_t* n_atoms_out) {
  if (!out || (n_res && (!ca_x || !ca_y || !ca_z || !amino_acid))) return DNAB_ERR_INVALID;
  size_t need = 0;
  for (size_t r = 0; r < n_res; ++r) {
    if (amino_acid[r] >= kNumAminoAcids) return DNAB_ERR_INVALID;
    need += kResidueTopology[amino_acid[r]].n_atoms;
  }
  if (n_atoms_out) *n_atoms_out = need;
  if (need > out->atom_capacity) return DNAB_ERR_CAPACITY;
  size_t a = 0;
  if (out->res_start) out->res_start[0] = 0;
  for (size_t r = 0; r < n_res; ++r) {
    int aa = amino_acid[r];
    float chi[4] = {0, 0, 0, 0};
    if (kRotamerCount[aa]) {
      int k = rotamer_idx ? rotamer_idx[r] % kRotamerCount[aa] : 0;
      std::memcpy(chi, kRotamerLib[aa][k].chi, sizeof(chi));
    }
    float ca[3] = {ca_x[r], ca_y[r], ca_z[r]};
    float xyz[16][3];
    int n = build_residue_atoms(aa, ca, chi, &xyz[0][0], nullptr);
    const ResTopology& topo = kResidueTopology[aa];
    if (out->ca_atom_idx) out->ca_atom_idx[r] = static_cast<uint32_t>(a + topo.ca_idx);
    if (out->amino_acid) out->amino_acid[r] = static_cast<uint8_t>(aa);
    for (int k = 0; k < n; ++k, ++a) {
      const ResAtomDef& d = topo.atoms[k];
      const NbParams& nb = kAmberNb[d.atom_type];
      out->x[a] = xyz[k][0];
      out->y[a] = xyz[k][1];
      out->z[a] = xyz[k][2];
      if (out->charge) out->charge[a] = d.charge;
      if (out->r_min_half) out->r_min_half[a] = nb.r_min_half;
      if (out->epsilon) out->epsilon[a] = nb.epsilon;
      if (out->hydrophobic) out->hydrophobic[a] = d.hydrophobic;
      if (out->atom_type) out->atom_type[a] = d.atom_type;
      if (out->residue_idx) out->residue_idx[a] = static_cast<uint32_t>(r);
    }
    if (out->res_start) out->res_start[r + 1] = static_cast<uint32_t>(a);
  }
  return DNAB_OK;
}

int dnab_rebuild_side_chain(uint8_t amino_acid, uint8_t rotamer, float* xyz, size_t n_atoms) {
  if (!xyz || amino_acid >= kNumAminoAcids) return DNAB_ERR_INVALID;
  const int aa = amino_acid;
  const ResTopology& topo = kResidueTopology[aa];
  if (n_atoms != topo.n_atoms) return DNAB_ERR_INVALID;
  if (kRotamerCount[aa] ? rotamer >= kRotamerCount[aa] : rotamer != 0) return DNAB_ERR_INVALID;
  float chi[4] = {0, 0, 0, 0};
  if (kRotamerCount[aa]) std::memcpy(chi, kRotamerLib[aa][rotamer].chi, sizeof(chi));
  float atoms[16][3];
  std::memcpy(atoms, xyz, sizeof(float) * 3 * n_atoms);
  int n = aa == kGly ? 4 : 5;  // the backbone kept (src/allatom.rs:248)
  const BondDef* defs = kSidechainDefs[aa];
  for (int k = 0; k < kSidechainCount[aa] && n < topo.n_atoms; ++k) {
    const BondDef& d = defs[k];
    const float dih = d.chi_idx == 255 ? d.dihedral_fixed : chi[d.chi_idx];
    place_atom(atoms[d.ref_a], atoms[d.ref_b], atoms[d.ref_c], d.bond_len, d.angle_deg, dih,
               atoms[n]);
    ++n;
  }
  std::memcpy(xyz, atoms, sizeof(float) * 3 * n_atoms);
  return DNAB_OK;
}

int dnab_place_atom(const float a[3], const float b[3], const float c[3], float bond_len,
                    float angle_deg, float dihedral_deg, float out[3]) {
  if (!a || !b || !c || !out) return DNAB_ERR_INVALID;
  place_atom(a, b, c, bond_len, angle_deg, dihedral_deg, out);
  return DNAB_OK;
}

int dnab_residue_atom_count(uint8_t amino_acid) {
  return amino_acid < kNumAminoAcids ? kResidueTopology[amino_acid].n_atoms : -1;
}

int dnab_residue_cg_params(uint8_t amino_acid, float* charge, float* epsilon, float* sigma,
                           uint8_t* hydrophobic) {
  if (amino_acid >= kNumAminoAcids) return DNAB_ERR_INVALID;
  if (charge) *charge = aa_charge(amino_acid);
  if (epsilon) *epsilon = aa_lj_epsilon(amino_acid);
  if (sigma) *sigma = aa_lj_sigma(amino_acid);
  if (hydrophobic) *hydrophobic = aa_is_hydrophobic(amino_acid);
  return DNAB_OK;
}

int dnab_ff_type_params(uint8_t atom_type, float* dg_ref, float* volume, float* lambda,
                        uint8_t* hbond_donor, uint8_t* hbond_acceptor) {
  if (atom_type >= kNumAtomTypes) return DNAB_ERR_INVALID;
  if (dg_ref) *dg_ref = kEef1[atom_type].dg_ref;
  if (volume) *volume = kEef1[atom_type].vol;
  if (lambda) *lambda = kEef1[atom_type].lambda;
  if (hbond_donor) *hbond_donor = (kHbondDonorMask >> atom_type) & 1u;
  if (hbond_acceptor) *hbond_acceptor = (kHbondAcceptorMask >> atom_type) & 1u;
  return DNAB_OK;
}

int dnab_ff_table(uint8_t amino_acid, int slot, uint8_t* atom_type, float* charge,
                  uint8_t* hydrophobic, float* r_min_half, float* epsilon) {
  if (amino_acid >= kNumAminoAcids) return DNAB_ERR_INVALID;
  const ResTopology& t = kResidueTopology[amino_acid];
  if (slot < 0 || slot >= t.n_atoms) return DNAB_ERR_INVALID;
  const ResAtomDef& d = t.atoms[slot];
  if (atom_type) *atom_type = d.atom_type;
  if (charge) *charge = d.charge;
  if (hydrophobic) *hydrophobic = d.hydrophobic;
  if (r_min_half) *r_min_half = kAmberNb[d.atom_type].r_min_half;
  if (epsilon) *epsilon = kAmberNb[d.atom_type].epsilon;
  return DNAB_OK;
}

// Calpha random walk with 3.8 A steps, uniform directions, steps leaving the
// sphere of `radius` around the origin resampled; uniform residues and
// uniform rotamer index.
int dnab_synth_globule_trace(size_t n_res, float radius, uint64_t seed, float* x, float* y,
                             float* z, uint8_t* amino_acid, uint8_t* rotamer_idx) {
  if (n_res && (!x || !y || !z || !amino_acid)) return DNAB_ERR_INVALID;
  if (!(radius > 0.0f)) return DNAB_ERR_INVALID;
  SplitMix64 rng{seed};
  float px = 0, py = 0, pz = 0;
  for (size_t i = 0; i < n_res; ++i) {
    if (i > 0) {
      for (int tries = 0; tries < 1000; ++tries) {
        // uniform direction on the sphere (Marsaglia 1972)
        float u, v, s;
        do {
          u = rng.uniform(-1.0f, 1.0f);
          v = rng.uniform(-1.0f, 1.0f);
          s = u * u + v * v;
        } while (s >= 1.0f || s == 0.0f);
        float f = 2.0f * std::sqrt(1.0f - s);
        float dx = u * f, dy = v * f, dz = 1.0f - 2.0f * s;
        float nx = px + 3.8f * dx, ny = py + 3.8f * dy, nz = pz + 3.8f * dz;
        if (nx * nx + ny * ny + nz * nz <= radius * radius || tries == 999) {
          px = nx; py = ny; pz = nz;
          break;
        }
      }
    }
    x[i] = px; y[i] = py; z[i] = pz;
    amino_acid[i] = static_cast<uint8_t>(rng.below(kNumAminoAcids));
    if (rotamer_idx) rotamer_idx[i] = static_cast<uint8_t>(rng.below(5));
  }
  return DNAB_OK;
}

// Ring layout of random_allatom_antibody (src/diffusion.rs:264-277): residue i
// at golden angle i*2.399 rad on a ring of `ring_radius` in the x-z plane
// around `center`, jittered U(-5,5) in x and z and U(-R/2,R/2) in y, with a
// uniformly drawn residue.
int dnab_synth_binder_ring(size_t n_res, const float center[3], float ring_radius,
                           uint64_t seed, float* x, float* y, float* z, uint8_t* amino_acid) {
  if (n_res && (!x || !y || !z || !amino_acid || !center)) return DNAB_ERR_INVALID;
  SplitMix64 rng{seed};
  for (size_t i = 0; i < n_res; ++i) {
    float angle = static_cast<float>(i) * 2.399f;
    x[i] = center[0] + ring_radius * std::cos(angle) + rng.uniform(-5.0f, 5.0f);
    y[i] = center[1] + rng.uniform(-ring_radius * 0.5f, ring_radius * 0.5f);
    z[i] = center[2] + ring_radius * std::sin(angle) + rng.uniform(-5.0f, 5.0f);
    amino_acid[i] = static_cast<uint8_t>(rng.below(kNumAminoAcids));
  }
  return DNAB_OK;
}

// Compact binder next to a globule: a Calpha walk (3.8 A steps) that starts
// at a uniform point on the sphere of `shell_radius` around `center` and is
// kept outside `min_radius`; uniform residues.
int dnab_synth_binder_shell(size_t n_res, const float center[3], float shell_radius,
                            float min_radius, uint64_t seed, float* x, float* y, float* z,
                            uint8_t* amino_acid) {
  if (n_res && (!x || !y || !z || !amino_acid || !center)) return DNAB_ERR_INVALID;
  SplitMix64 rng{seed};
  auto unit = [&rng](float& dx, float& dy, float& dz) {
    float u, v, s;
    do {
      u = rng.uniform(-1.0f, 1.0f);
      v = rng.uniform(-1.0f, 1.0f);
      s = u * u + v * v;
    } while (s >= 1.0f || s == 0.0f);
    float f = 2.0f * std::sqrt(1.0f - s);
    dx = u * f; dy = v * f; dz = 1.0f - 2.0f * s;
  };
  float dx, dy, dz;
  unit(dx, dy, dz);
  float px = shell_radius * dx, py = shell_radius * dy, pz = shell_radius * dz;
  for (size_t i = 0; i < n_res; ++i) {
    if (i > 0) {
      for (int tries = 0; tries < 1000; ++tries) {
        unit(dx, dy, dz);
        float nx = px + 3.8f * dx, ny = py + 3.8f * dy, nz = pz + 3.8f * dz;
        if (nx * nx + ny * ny + nz * nz >= min_radius * min_radius || tries == 999) {
          px = nx; py = ny; pz = nz;
          break;
        }
      }
    }
    x[i] = center[0] + px;
    y[i] = center[1] + py;
    z[i] = center[2] + pz;
    amino_acid[i] = static_cast<uint8_t>(rng.below(kNumAminoAcids));
  }
  return DNAB_OK;
}

// A population of shell binders: candidate k is dnab_synth_binder_shell with
// seed seed_base + k, built by dnab_build_protein, all candidates n_res long.
// Traces are drawn first to size the atom arrays, then the candidates are
// built in parallel into their slices (same result for any thread count).
int dnab_synth_shell_population(size_t n_cand, size_t n_res, const float center[3],
                                float shell_radius, float min_radius, uint64_t seed_base,
                                int nthreads, dnab_protein_out* out, uint32_t* cand_res_off,
                                size_t* n_atoms_out) {
  if (!center || (out && !cand_res_off)) return DNAB_ERR_INVALID;
  const size_t nr = n_cand * n_res;
  std::vector<float> tx(nr), ty(nr), tz(nr);
  std::vector<uint8_t> taa(nr);
  std::vector<size_t> atom_off(n_cand + 1, 0);
  const int nt = std::max(1, std::min(nthreads > 0 ? nthreads : 1, 64));
  auto parallel = [&](auto fn) {
    std::vector<std::thread> th;
    const size_t per = (n_cand + nt - 1) / nt;
    for (int k = 0; k < nt; ++k) {
      const size_t c0 = std::min(n_cand, k * per), c1 = std::min(n_cand, c0 + per);
      if (c0 < c1) th.emplace_back(fn, c0, c1);
    }
    for (auto& t : th) t.join();
  };
  parallel([&](size_t c0, size_t c1) {
    for (size_t c = c0; c < c1; ++c)
      dnab_synth_binder_shell(n_res, center, shell_radius, min_radius, seed_base + c,
                              &tx[c * n_res], &ty[c * n_res], &tz[c * n_res], &taa[c * n_res]);
  });
  for (size_t c = 0; c < n_cand; ++c) {
    size_t n = 0;
    for (size_t r = 0; r < n_res; ++r) n += kResidueTopology[taa[c * n_res + r]].n_atoms;
    atom_off[c + 1] = atom_off[c] + n;
  }
  if (n_atoms_out) *n_atoms_out = atom_off[n_cand];
  if (!out) return DNAB_OK;
  if (!out->x || !out->y || !out->z) return DNAB_ERR_INVALID;
  if (atom_off[n_cand] > out->atom_capacity) return DNAB_ERR_CAPACITY;
  for (size_t c = 0; c <= n_cand; ++c) cand_res_off[c] = static_cast<uint32_t>(c * n_res);
  std::atomic<int> status{DNAB_OK};
  parallel([&](size_t c0, size_t c1) {
    std::vector<uint32_t> rs(n_res + 1), ca(n_res);
    for (size_t c = c0; c < c1; ++c) {
      const size_t a0 = atom_off[c], r0 = c * n_res;
      dnab_protein_out o{};
      o.atom_capacity = atom_off[c + 1] - a0;
      o.x = out->x + a0;
      o.y = out->y + a0;
      o.z = out->z + a0;
      o.charge = out->charge ? out->charge + a0 : nullptr;
      o.r_min_half = out->r_min_half ? out->r_min_half + a0 : nullptr;
      o.epsilon = out->epsilon ? out->epsilon + a0 : nullptr;
      o.hydrophobic = out->hydrophobic ? out->hydrophobic + a0 : nullptr;
      o.atom_type = out->atom_type ? out->atom_type + a0 : nullptr;
      o.residue_idx = nullptr;
      o.res_start = rs.data();
      o.ca_atom_idx = ca.data();
      o.amino_acid = out->amino_acid ? out->amino_acid + r0 : nullptr;
      size_t got = 0;
      const int st = dnab_build_protein(&tx[r0], &ty[r0], &tz[r0], &taa[r0], nullptr, n_res, &o,
                                        &got);
      if (st != DNAB_OK) {
        status = st;
        return;
      }
      for (size_t r = 0; r < n_res; ++r) {
        if (out->res_start) out->res_start[r0 + r] = static_cast<uint32_t>(a0 + rs[r]);
        if (out->ca_atom_idx) out->ca_atom_idx[r0 + r] = static_cast<uint32_t>(a0 + ca[r]);
        if (out->residue_idx)
          for (uint32_t k = rs[r]; k < rs[r + 1]; ++k)
            out->residue_idx[a0 + k] = static_cast<uint32_t>(r0 + r);
      }
    }
  });
  if (out->res_start) out->res_start[nr] = static_cast<uint32_t>(atom_off[n_cand]);
  return status.load();
}

}  // extern "C"
