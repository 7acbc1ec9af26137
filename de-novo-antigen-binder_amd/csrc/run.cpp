// run_allatom (src/diffusion.rs:983-1107) as one call: the broad phase over
// n_broad binders, the refinement of the survivors with noise, their
// allatom_energy, the ascending sort and the truncation to max(top_n, 1).
// A host composition of the library's own calls (synthetic ring binders,
// dnab_refine, population scoring); the work runs in their kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "dnab_internal.h"
#include "ff_tables.h"

using namespace dnab;

namespace {

constexpr size_t kMaxResAtoms = 14;  // the largest residue topology (Trp)

// Candidates as the library's calls take them: concatenated atoms, residue
// ranges and candidate ranges, plus each residue's rotamer index.
struct Batch {
  std::vector<float> x, y, z, q, rmh, eps;
  std::vector<uint8_t> hyd, type, aa, rot;
  std::vector<uint32_t> res_start{0}, ca, cand_res_off{0}, cand_ag;

  size_t n_cand() const { return cand_res_off.size() - 1; }
  size_t n_res() const { return ca.size(); }
  size_t n_atoms() const { return x.size(); }

  dnab_population_desc desc() const {
    dnab_population_desc d{};
    d.n_cand = n_cand();
    d.n_res = n_res();
    d.cand_antigen = cand_ag.data();
    d.cand_res_off = cand_res_off.data();
    d.res_atom_off = res_start.data();
    d.res_ca_atom = ca.data();
    d.res_amino_acid = aa.data();
    d.atoms.n = n_atoms();
    d.atoms.x = x.data();
    d.atoms.y = y.data();
    d.atoms.z = z.data();
    d.atoms.charge = q.data();
    d.atoms.r_min_half = rmh.data();
    d.atoms.epsilon = eps.data();
    d.atoms.hydrophobic = hyd.data();
    d.atoms.atom_type = type.data();
    return d;
  }

  // buffers for n_res residues of at most kMaxResAtoms atoms each
  void reserve_out(size_t n_res, dnab_protein_out& o) {
    const size_t na = n_res * kMaxResAtoms;
    for (auto* v : {&x, &y, &z, &q, &rmh, &eps}) v->assign(std::max<size_t>(na, 1), 0.0f);
    hyd.assign(std::max<size_t>(na, 1), 0);
    type.assign(std::max<size_t>(na, 1), 0);
    res_start.assign(n_res + 1, 0);
    ca.assign(n_res, 0);
    aa.assign(n_res, 0);
    rot.assign(n_res, 0);
    o = dnab_protein_out{};
    o.atom_capacity = na;
    o.x = x.data();
    o.y = y.data();
    o.z = z.data();
    o.charge = q.data();
    o.r_min_half = rmh.data();
    o.epsilon = eps.data();
    o.hydrophobic = hyd.data();
    o.atom_type = type.data();
    o.res_start = res_start.data();
    o.ca_atom_idx = ca.data();
    o.amino_acid = aa.data();
  }
  void trim(size_t n_atoms) {
    for (auto* v : {&x, &y, &z, &q, &rmh, &eps}) v->resize(n_atoms);
    hyd.resize(n_atoms);
    type.resize(n_atoms);
  }

  // append candidate c of `src`
  void append(const Batch& src, size_t c) {
    const uint32_t r0 = src.cand_res_off[c], r1 = src.cand_res_off[c + 1];
    const uint32_t a0 = src.res_start[r0], a1 = src.res_start[r1];
    const uint32_t base = static_cast<uint32_t>(n_atoms());
    auto app = [&](auto& dst, const auto& s, size_t b, size_t e) {
      dst.insert(dst.end(), s.begin() + b, s.begin() + e);
    };
    app(x, src.x, a0, a1); app(y, src.y, a0, a1); app(z, src.z, a0, a1);
    app(q, src.q, a0, a1); app(rmh, src.rmh, a0, a1); app(eps, src.eps, a0, a1);
    app(hyd, src.hyd, a0, a1); app(type, src.type, a0, a1);
    for (uint32_t r = r0; r < r1; ++r) {
      res_start.push_back(base + (src.res_start[r + 1] - a0));
      ca.push_back(base + (src.ca[r] - a0));
      aa.push_back(src.aa[r]);
      rot.push_back(src.rot[r]);
    }
    cand_res_off.push_back(static_cast<uint32_t>(ca.size()));
    cand_ag.push_back(0);
  }
};

// random_allatom_antibody (src/diffusion.rs:264-277): a ring of ab_length
// residues around `center` with uniform residue types, built by
// protein_from_ca_trace; the library's SplitMix64 stream `seed`.
int ring_candidate(size_t ab_length, const float center[3], float radius, uint64_t seed,
                   Batch& b) {
  std::vector<float> x(ab_length), y(ab_length), z(ab_length);
  std::vector<uint8_t> aa(ab_length);
  DNAB_RETURN_IF(dnab_synth_binder_ring(ab_length, center, radius, seed, x.data(), y.data(),
                                        z.data(), aa.data()));
  Batch one;
  dnab_protein_out o;
  one.reserve_out(ab_length, o);
  size_t na = 0;
  DNAB_RETURN_IF(dnab_build_protein(x.data(), y.data(), z.data(), aa.data(), nullptr, ab_length,
                                    &o, &na));
  one.trim(na);
  one.cand_res_off = {0, static_cast<uint32_t>(ab_length)};
  one.cand_ag = {0};
  b.append(one, 0);
  return DNAB_OK;
}

// allatom_diffusion_step n_steps times for every candidate (dnab_refine);
// `out` receives the refined batch and, with energies, allatom_energy each.
int refine(dnab_ctx* ctx, dnab_antigen* ag, const Batch& in, const dnab_langevin_params& lp,
           const dnab_mc_params& mc, Batch& out, std::vector<double>* energies) {
  const dnab_population_desc d = in.desc();
  dnab_protein_out o;
  out.reserve_out(in.n_res(), o);
  if (energies) energies->assign(in.n_cand(), 0.0);
  size_t na = 0;
  DNAB_RETURN_IF(dnab_refine(ctx, ag, &d, in.rot.data(), &lp, &mc, &o, out.rot.data(),
                             energies ? energies->data() : nullptr, &na));
  out.trim(na);
  out.cand_res_off = in.cand_res_off;
  out.cand_ag = in.cand_ag;
  return DNAB_OK;
}

// score_batch's terms (LJ + Coulomb + hydrophobic, src/gpu.rs:166-267) of
// every candidate, as the reference's broad phase ranks them
// (src/diffusion.rs:1041-1049).
int score(dnab_ctx* ctx, dnab_antigen* ag, const Batch& b, std::vector<double>& e) {
  const dnab_population_desc d = b.desc();
  dnab_population* pop = nullptr;
  dnab_antigen* ags[1] = {ag};
  DNAB_RETURN_IF(dnab_population_create(ctx, ags, 1, &d, &pop));
  int rc = dnab_population_evaluate(ctx, pop, 0);
  e.assign(b.n_cand(), 0.0);
  if (rc == DNAB_OK) rc = dnab_population_read_energies(ctx, pop, e.data(), nullptr);
  dnab_population_destroy(pop);
  return rc;
}

// Ascending energy, NaN last, ties by index (a stable order; the reference's
// sort_by with partial_cmp().unwrap_or(Equal), src/diffusion.rs:1047,1097).
std::vector<uint32_t> ascending(const std::vector<double>& e) {
  std::vector<uint32_t> idx(e.size());
  std::iota(idx.begin(), idx.end(), 0u);
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
    const bool na = std::isnan(e[a]), nb = std::isnan(e[b]);
    if (na || nb) return !na && nb;
    return e[a] < e[b];
  });
  return idx;
}

uint64_t mix(uint64_t seed, uint64_t k) {
  SplitMix64 r{seed ^ (0x9E3779B97F4A7C15ull * (k + 1))};
  return r.next();
}

}  // namespace

extern "C" {

int dnab_run_allatom_default_params(dnab_run_allatom_params* p) {
  if (!p) return DNAB_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->ab_length = 0;        // the antigen's length, the CLI default (src/main.rs:145)
  p->population = 64;      // TOP_K, the CLI default (src/main.rs:308)
  p->cpu_iterations = 600;  // CPU_STEPS, src/diffusion.rs:38
  p->top_n = 1;             // src/main.rs:84
  p->n_broad = 1024;        // AA_POPULATION, src/diffusion.rs:34
  p->broad_steps = 200;     // GPU_STEPS, src/diffusion.rs:36
  p->broad_moves = 1;
  p->seed = 0;
  p->ring_radius = 20.0f;   // INIT_RADIUS, src/diffusion.rs:81
  return DNAB_OK;
}

int dnab_run_allatom(dnab_ctx* ctx, dnab_antigen* antigen, const dnab_run_allatom_params* p,
                     dnab_protein_out* out, double* out_energy, uint32_t* n_results,
                     size_t* n_atoms_out) {
  if (!ctx || !antigen || !p || antigen->ctx != ctx) return DNAB_ERR_INVALID;
  const size_t ab_length = p->ab_length ? p->ab_length : antigen->n_ca;
  if (ab_length == 0 || p->population == 0 || !(p->ring_radius > 0.0f))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "run_allatom: ab_length, population and radius > 0");
  if (p->n_broad > 1u << 20 || p->population > 1u << 20 || ab_length > 512)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "run_allatom: too large");
  dnab_langevin_params lp;
  dnab_langevin_default_params(&lp);
  dnab_mc_params mc;
  dnab_mc_default_params(&mc);
  const float* center = antigen->center;

  // the broad phase (src/diffusion.rs:1008-1062): n_broad ring binders,
  // broad_steps gradient steps without noise on the schedule of the whole
  // run, scored once at the end; the `population` lowest survive. Without
  // it, `population` ring binders go straight to refinement (the
  // reference's path without a GPU, src/diffusion.rs:1064-1071).
  Batch survivors;
  if (p->n_broad > 0) {
    Batch all;
    for (uint32_t k = 0; k < p->n_broad; ++k)
      DNAB_RETURN_IF(ring_candidate(ab_length, center, p->ring_radius, mix(p->seed, k), all));
    Batch moved;
    lp.n_steps = p->broad_steps;
    lp.with_noise = 0;
    lp.seed = mix(p->seed, 0xB40ADull);
    lp.schedule_offset = 0;
    lp.schedule_total = std::max<uint32_t>(p->broad_steps + p->cpu_iterations, 1);
    lp.restraint_radius = p->ring_radius;
    if (p->broad_moves) {
      // allatom_diffusion_step with its moves and no noise, as the
      // reference steps the broad phase (src/diffusion.rs:1021-1024)
      DNAB_RETURN_IF(refine(ctx, antigen, all, lp, mc, moved, nullptr));
    } else {
      // the gradient-only loop (no moves), one fused launch
      const dnab_population_desc d = all.desc();
      dnab_population* pop = nullptr;
      dnab_antigen* ags[1] = {antigen};
      DNAB_RETURN_IF(dnab_population_create(ctx, ags, 1, &d, &pop));
      lp.score_every = 0;
      lp.top_k = 1;
      uint32_t idx = 0, n = 0;
      double e = 0.0;
      int rc = dnab_langevin_run(ctx, pop, &lp, &idx, &e, &n);
      moved = all;
      if (rc == DNAB_OK)
        rc = dnab_population_read_coords(ctx, pop, moved.x.data(), moved.y.data(), moved.z.data());
      dnab_population_destroy(pop);
      DNAB_RETURN_IF(rc);
    }
    std::vector<double> e;
    DNAB_RETURN_IF(score(ctx, antigen, moved, e));
    const std::vector<uint32_t> order = ascending(e);
    const size_t keep = std::min<size_t>(p->population, order.size());
    for (size_t k = 0; k < keep; ++k) survivors.append(moved, order[k]);
  } else {
    for (uint32_t k = 0; k < p->population; ++k)
      DNAB_RETURN_IF(ring_candidate(ab_length, center, p->ring_radius, mix(p->seed, k),
                                    survivors));
  }

  // refinement (src/diffusion.rs:1083-1095): cpu_iterations steps with noise
  // and the moves, the schedule restarted, then allatom_energy of each
  Batch refined;
  std::vector<double> energy;
  lp.n_steps = p->cpu_iterations;
  lp.with_noise = 1;
  lp.seed = mix(p->seed, 0x5EF1E5ull);
  lp.schedule_offset = 0;
  lp.schedule_total = std::max<uint32_t>(p->cpu_iterations, 1);
  lp.restraint_radius = p->ring_radius;
  DNAB_RETURN_IF(refine(ctx, antigen, survivors, lp, mc, refined, &energy));

  // sort ascending and keep max(top_n, 1) (src/diffusion.rs:1097-1098)
  const std::vector<uint32_t> order = ascending(energy);
  const size_t n_keep = std::min<size_t>(std::max<uint32_t>(p->top_n, 1u), order.size());
  Batch res;
  for (size_t k = 0; k < n_keep; ++k) res.append(refined, order[k]);
  if (n_results) *n_results = static_cast<uint32_t>(n_keep);
  if (n_atoms_out) *n_atoms_out = res.n_atoms();
  if (out_energy)
    for (size_t k = 0; k < n_keep; ++k) out_energy[k] = energy[order[k]];
  if (!out) return DNAB_OK;
  if (out->atom_capacity < res.n_atoms()) return DNAB_ERR_CAPACITY;
  if (!out->x || !out->y || !out->z) return DNAB_ERR_INVALID;
  const size_t na = res.n_atoms(), nr = res.n_res();
  auto put = [&](auto* dst, const auto& src, size_t n) {
    if (dst) std::copy(src.begin(), src.begin() + n, dst);
  };
  put(out->x, res.x, na); put(out->y, res.y, na); put(out->z, res.z, na);
  put(out->charge, res.q, na); put(out->r_min_half, res.rmh, na);
  put(out->epsilon, res.eps, na); put(out->hydrophobic, res.hyd, na);
  put(out->atom_type, res.type, na);
  put(out->res_start, res.res_start, nr + 1);
  put(out->ca_atom_idx, res.ca, nr);
  put(out->amino_acid, res.aa, nr);
  if (out->residue_idx)
    for (size_t r = 0; r < nr; ++r)
      for (uint32_t a = res.res_start[r]; a < res.res_start[r + 1]; ++a)
        out->residue_idx[a] = static_cast<uint32_t>(r % ab_length);
  return DNAB_OK;
}

}  // extern "C"
