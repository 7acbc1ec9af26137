// Host-buffer entry points: score_batch (src/gpu.rs:166-267) and the
// all-atom energy / force functions of src/energy.rs, each run as a one-call
// device-resident antigen + population evaluation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <vector>

#include "dnab_internal.h"

namespace {

struct AgGuard {
  dnab_antigen* p = nullptr;
  ~AgGuard() { dnab_antigen_destroy(p); }
};
struct PopGuard {
  dnab_population* p = nullptr;
  ~PopGuard() { dnab_population_destroy(p); }
};

dnab_antigen_desc antigen_desc(const dnab_atom_cloud* ag, float cutoff) {
  dnab_antigen_desc d{};
  d.atoms = *ag;
  d.cell_size = cutoff;
  d.cutoff = cutoff;
  return d;
}

// One candidate of the given residues (or, with n_res == 0 and no residue
// arrays, every antibody atom as a single residue).
int single_candidate(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                     const dnab_atom_cloud* antibody, float cutoff, AgGuard& ag, PopGuard& pop,
                     size_t n_res = 0, const uint32_t* res_atom_off = nullptr,
                     const uint32_t* res_ca_atom = nullptr, const uint8_t* amino_acid = nullptr) {
  if (!(cutoff > 0.0f) || !std::isfinite(cutoff))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "cutoff must be positive");
  if (n_res >= 0xFFFFFF00ull)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "antibody: too many residues");
  dnab_antigen_desc ad = antigen_desc(antigen, cutoff);
  DNAB_RETURN_IF(dnab_antigen_create(ctx, &ad, &ag.p));
  const uint32_t n = static_cast<uint32_t>(antibody->n);
  uint32_t cand_ag = 0, cand_res_off[2] = {0, n ? 1u : 0u}, one_res[2] = {0, n}, res_ca = 0;
  dnab_population_desc pd{};
  pd.n_cand = 1;
  pd.n_res = n ? 1 : 0;
  pd.cand_antigen = &cand_ag;
  pd.cand_res_off = cand_res_off;
  pd.res_atom_off = one_res;
  pd.res_ca_atom = &res_ca;
  if (res_atom_off) {
    cand_res_off[1] = static_cast<uint32_t>(n_res);
    pd.n_res = n_res;
    pd.res_atom_off = res_atom_off;
    pd.res_ca_atom = res_ca_atom;
    pd.res_amino_acid = amino_acid;
  }
  pd.atoms = *antibody;
  dnab_antigen* ags[1] = {ag.p};
  return dnab_population_create(ctx, ags, 1, &pd, &pop.p);
}

bool cloud_ok(const dnab_atom_cloud* c) {
  return c && (c->n == 0 || (c->x && c->y && c->z && c->charge && c->r_min_half && c->epsilon &&
                             c->hydrophobic));
}

}  // namespace

extern "C" {

int dnab_interaction_energy_atoms_breakdown(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                          const dnab_atom_cloud* antibody, float cutoff,
                          dnab_energy_breakdown* out) {
  if (!ctx || !cloud_ok(antigen) || !cloud_ok(antibody) || !out) return DNAB_ERR_INVALID;
  AgGuard ag;
  PopGuard pop;
  DNAB_RETURN_IF(single_candidate(ctx, antigen, antibody, cutoff, ag, pop));
  DNAB_RETURN_IF(dnab_population_evaluate(ctx, pop.p, DNAB_EVAL_BREAKDOWN));
  double t[6];
  DNAB_RETURN_IF(dnab_population_read_terms(ctx, pop.p, t));
  out->lj = t[0];
  out->coulomb = t[1];
  out->hydrophobic = t[2];
  out->hbond = t[3];
  out->solvation = t[4];
  return DNAB_OK;
}

int dnab_interaction_energy_atoms(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                                  const dnab_atom_cloud* antibody, float cutoff,
                                  double* energy) {
  // src/energy.rs:256-258: the breakdown's total(), summed in the order of
  // src/energy.rs:247-249.
  if (!energy) return DNAB_ERR_INVALID;
  dnab_energy_breakdown b{};
  DNAB_RETURN_IF(dnab_interaction_energy_atoms_breakdown(ctx, antigen, antibody, cutoff, &b));
  *energy = (((b.lj + b.coulomb) + b.hydrophobic) + b.hbond) + b.solvation;
  return DNAB_OK;
}

int dnab_allatom_energy(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                        const dnab_atom_cloud* antibody, const uint32_t* res_atom_off,
                        const uint32_t* res_ca_atom, const uint8_t* amino_acid, size_t n_res,
                        float cutoff, double* energy, double* disulfide) {
  // src/diffusion.rs:281-283: interaction_energy_atoms + the antibody's
  // disulfide_energy (src/allatom.rs:155-180), which needs its residues.
  if (!ctx || !cloud_ok(antigen) || !cloud_ok(antibody) || !energy) return DNAB_ERR_INVALID;
  if (!res_atom_off || (n_res && (!res_ca_atom || !amino_acid))) return DNAB_ERR_INVALID;
  AgGuard ag;
  PopGuard pop;
  DNAB_RETURN_IF(single_candidate(ctx, antigen, antibody, cutoff, ag, pop, n_res, res_atom_off,
                                  res_ca_atom, amino_acid));
  DNAB_RETURN_IF(dnab_population_evaluate(ctx, pop.p, DNAB_EVAL_BREAKDOWN));
  double t[6];
  DNAB_RETURN_IF(dnab_population_read_terms(ctx, pop.p, t));
  *energy = ((((t[0] + t[1]) + t[2]) + t[3]) + t[4]) + t[5];
  if (disulfide) *disulfide = t[5];
  return DNAB_OK;
}

int dnab_atom_forces(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                     const dnab_atom_cloud* antibody, float cutoff, float* fx, float* fy,
                     float* fz, double* energy) {
  if (!ctx || !cloud_ok(antigen) || !cloud_ok(antibody)) return DNAB_ERR_INVALID;
  if (antibody->n && (!fx || !fy || !fz)) return DNAB_ERR_INVALID;
  AgGuard ag;
  PopGuard pop;
  DNAB_RETURN_IF(single_candidate(ctx, antigen, antibody, cutoff, ag, pop));
  DNAB_RETURN_IF(dnab_population_evaluate(ctx, pop.p, DNAB_EVAL_FORCES));
  if (antibody->n) DNAB_RETURN_IF(dnab_population_read_forces(ctx, pop.p, fx, fy, fz));
  if (energy) DNAB_RETURN_IF(dnab_population_read_energies(ctx, pop.p, energy, nullptr));
  return DNAB_OK;
}

int dnab_score_batch(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                     const dnab_gpu_atom* candidates, size_t n_atoms, size_t n_ab,
                     float* energies_out, size_t capacity, size_t* n_cand_out) {
  if (!ctx || !cloud_ok(antigen)) return DNAB_ERR_INVALID;
  // src/gpu.rs:172-174: n_cand = len / n_ab, empty result for n_ab == 0.
  const size_t n_cand = n_ab == 0 ? 0 : n_atoms / n_ab;
  if (n_cand_out) *n_cand_out = n_cand;
  if (n_cand == 0) return DNAB_OK;
  if (!candidates) return DNAB_ERR_INVALID;
  if (!energies_out || capacity < n_cand) return DNAB_ERR_CAPACITY;
  const size_t na = n_cand * n_ab;
  std::vector<float> x(na), y(na), z(na), q(na), rmh(na), eps(na);
  std::vector<uint8_t> hyd(na);
  for (size_t i = 0; i < na; ++i) {
    const dnab_gpu_atom& a = candidates[i];
    x[i] = a.x;
    y[i] = a.y;
    z[i] = a.z;
    q[i] = a.q;
    rmh[i] = a.r_min_half;
    eps[i] = a.epsilon;
    hyd[i] = a.hydrophobic > 0.5f ? 1 : 0;  // shaders/energy.wgsl:91
  }
  std::vector<uint32_t> cand_ag(n_cand, 0), cand_res_off(n_cand + 1), res_atom_off(n_cand + 1),
      res_ca(n_cand);
  for (size_t c = 0; c <= n_cand; ++c) {
    cand_res_off[c] = static_cast<uint32_t>(c);
    res_atom_off[c] = static_cast<uint32_t>(c * n_ab);
  }
  for (size_t c = 0; c < n_cand; ++c) res_ca[c] = static_cast<uint32_t>(c * n_ab);
  AgGuard ag;
  PopGuard pop;
  dnab_antigen_desc ad = antigen_desc(antigen, 10.0f);  // shaders/energy.wgsl:16
  DNAB_RETURN_IF(dnab_antigen_create(ctx, &ad, &ag.p));
  dnab_population_desc pd{};
  pd.n_cand = n_cand;
  pd.n_res = n_cand;
  pd.cand_antigen = cand_ag.data();
  pd.cand_res_off = cand_res_off.data();
  pd.res_atom_off = res_atom_off.data();
  pd.res_ca_atom = res_ca.data();
  pd.atoms.n = na;
  pd.atoms.x = x.data();
  pd.atoms.y = y.data();
  pd.atoms.z = z.data();
  pd.atoms.charge = q.data();
  pd.atoms.r_min_half = rmh.data();
  pd.atoms.epsilon = eps.data();
  pd.atoms.hydrophobic = hyd.data();
  dnab_antigen* ags[1] = {ag.p};
  DNAB_RETURN_IF(dnab_population_create(ctx, ags, 1, &pd, &pop.p));
  DNAB_RETURN_IF(dnab_population_evaluate(ctx, pop.p, 0));
  std::vector<double> e(n_cand);
  DNAB_RETURN_IF(dnab_population_read_energies(ctx, pop.p, e.data(), nullptr));
  for (size_t c = 0; c < n_cand; ++c) energies_out[c] = static_cast<float>(e[c]);
  return DNAB_OK;
}

}  // extern "C"
