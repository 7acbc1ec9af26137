// binder.hpp — C++ host interface over libdnab's C ABI (include/dnab.h),
// mirroring the reference crate's API for the all-atom hot path so that its
// call sites read the same:
//
//   binder::AtomCloud                      src/allatom.rs:16-80
//   binder::AtomProtein, protein_from_ca_trace
//                                          src/allatom.rs:92-101, 523-541
//   binder::ResidueCloud                   src/atom.rs:185-257
//   binder::SpatialHashGrid                src/spatial.rs:12-68 (device-built)
//   binder::GpuContext::try_init           src/gpu.rs:55-57 (async body 59-155; std::nullopt = None)
//   binder::GpuContext::score_batch        src/gpu.rs:166-267
//   binder::interaction_energy_atoms[_breakdown], compute_forces[_with_grid]
//                                          src/energy.rs:154-398
//   binder::allatom_energy, disulfide_energy
//                                          src/diffusion.rs:281-283,
//                                          src/allatom.rs:155-180
//   binder::affinity_maturation_allatom    src/diffusion.rs:924-961
//   binder::parse_pdb, write_pdb_allatom, write_pdb_multi
//                                          src/pdb.rs:19-57, 121-191
//   binder::CandidateReport, write_fasta, write_json, write_csv
//                                          src/report.rs:8-113
//   binder::allatom_diffusion_steps        src/diffusion.rs:290-416 (with the moves,
//                                          a batch of antibodies in one call)
//   binder::run_allatom_gpu_phase          src/diffusion.rs:1008-1062 (GPU phase,
//                                          gradient steps + scoring, no MC moves)
//   binder::run_allatom, AllAtomResult     src/diffusion.rs:93-97, 983-1107
//
// Errors: functions that the reference cannot fail return values as the
// reference does; a failing C call throws binder::Error (the reference
// panics and aborts, Cargo.toml:27). Header-only; link with -ldnab.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/dnab.h"

namespace binder {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(int status, const dnab_ctx* ctx = nullptr) {
  if (status == DNAB_OK) return;
  std::string msg = dnab_status_string(status);
  if (ctx) msg += std::string(": ") + dnab_ctx_last_error(ctx);
  throw Error(status, msg);
}

enum class AminoAcid : uint8_t {
  Ala = 0, Arg, Asn, Asp, Cys, Gln, Glu, Gly, His, Ile,
  Leu, Lys, Met, Phe, Pro, Ser, Thr, Trp, Tyr, Val
};

struct AtomCloud {
  std::vector<float> x, y, z, charge, r_min_half, epsilon;
  std::vector<uint8_t> hydrophobic, atom_type;
  std::vector<uint32_t> residue_idx;
  size_t len() const { return x.size(); }
  bool is_empty() const { return x.empty(); }
  dnab_atom_cloud view() const {
    dnab_atom_cloud c{};
    c.n = len();
    c.x = x.data(); c.y = y.data(); c.z = z.data();
    c.charge = charge.data(); c.r_min_half = r_min_half.data(); c.epsilon = epsilon.data();
    c.hydrophobic = hydrophobic.data();
    c.atom_type = atom_type.empty() ? nullptr : atom_type.data();
    c.residue_idx = residue_idx.empty() ? nullptr : residue_idx.data();
    return c;
  }
};

struct ResidueCloud {
  std::vector<float> x, y, z, charge, epsilon, sigma;
  std::vector<uint8_t> hydrophobic;
  std::vector<AminoAcid> amino_acid;
  size_t len() const { return x.size(); }
  // ResidueCloud::push (src/atom.rs:226-235)
  void push(float px, float py, float pz, AminoAcid aa) {
    float q = 0, e = 0, s = 0;
    uint8_t h = 0;
    check(dnab_residue_cg_params(static_cast<uint8_t>(aa), &q, &e, &s, &h));
    x.push_back(px); y.push_back(py); z.push_back(pz);
    charge.push_back(q); epsilon.push_back(e); sigma.push_back(s);
    hydrophobic.push_back(h); amino_acid.push_back(aa);
  }
  dnab_residue_cloud view() const {
    dnab_residue_cloud c{};
    c.n = len();
    c.x = x.data(); c.y = y.data(); c.z = z.data();
    c.charge = charge.data(); c.epsilon = epsilon.data(); c.sigma = sigma.data();
    c.hydrophobic = hydrophobic.data();
    return c;
  }
};

struct AtomProtein {
  AtomCloud atoms;
  std::vector<uint32_t> res_start{0};
  std::vector<AminoAcid> amino_acid;
  std::vector<uint32_t> ca_atom_idx;
  size_t n_residues() const { return amino_acid.size(); }
  size_t n_atoms() const { return atoms.len(); }
  // src/allatom.rs:183-195
  std::array<float, 3> ca_center_of_mass() const {
    std::array<float, 3> c{0.f, 0.f, 0.f};
    if (n_residues() == 0) return c;
    for (uint32_t i : ca_atom_idx) {
      c[0] += atoms.x[i]; c[1] += atoms.y[i]; c[2] += atoms.z[i];
    }
    float n = static_cast<float>(n_residues());
    return {c[0] / n, c[1] / n, c[2] / n};
  }
};

// protein_from_ca_trace (src/allatom.rs:523-541)
inline AtomProtein protein_from_ca_trace(const std::vector<float>& x, const std::vector<float>& y,
                                         const std::vector<float>& z,
                                         const std::vector<AminoAcid>& aa) {
  const size_t n = aa.size();
  AtomProtein p;
  const size_t cap = 14 * n + 1;
  auto& a = p.atoms;
  a.x.resize(cap); a.y.resize(cap); a.z.resize(cap); a.charge.resize(cap);
  a.r_min_half.resize(cap); a.epsilon.resize(cap); a.hydrophobic.resize(cap);
  a.atom_type.resize(cap); a.residue_idx.resize(cap);
  p.res_start.assign(n + 1, 0);
  p.ca_atom_idx.assign(n, 0);
  std::vector<uint8_t> aa8(n), aa_out(n + 1);
  for (size_t i = 0; i < n; ++i) aa8[i] = static_cast<uint8_t>(aa[i]);
  dnab_protein_out out{cap, a.x.data(), a.y.data(), a.z.data(), a.charge.data(),
                       a.r_min_half.data(), a.epsilon.data(), a.hydrophobic.data(),
                       a.atom_type.data(), a.residue_idx.data(), p.res_start.data(),
                       p.ca_atom_idx.data(), aa_out.data()};
  size_t na = 0;
  check(dnab_build_protein(x.data(), y.data(), z.data(), aa8.data(), nullptr, n, &out, &na));
  for (auto* v : {&a.x, &a.y, &a.z, &a.charge, &a.r_min_half, &a.epsilon}) v->resize(na);
  a.hydrophobic.resize(na); a.atom_type.resize(na); a.residue_idx.resize(na);
  p.amino_acid = aa;
  return p;
}

struct AtomEnergyBreakdown {
  double lj = 0, coulomb = 0, hydrophobic = 0, hbond = 0, solvation = 0;
  // src/energy.rs:247-249
  double total() const { return lj + coulomb + hydrophobic + hbond + solvation; }
};

class GpuContext;

class SpatialHashGrid {
 public:
  // SpatialHashGrid::new (src/spatial.rs:23-28), on the context's device
  SpatialHashGrid(const GpuContext& ctx, float cell_size);
  ~SpatialHashGrid() { dnab_grid_destroy(g_); }
  SpatialHashGrid(const SpatialHashGrid&) = delete;
  SpatialHashGrid& operator=(const SpatialHashGrid&) = delete;
  // build (src/spatial.rs:31-39)
  void build(const std::vector<float>& x, const std::vector<float>& y,
             const std::vector<float>& z) {
    check(dnab_grid_build(ctx_, g_, x.data(), y.data(), z.data(), x.size()), ctx_);
  }
  // query_neighbors (src/spatial.rs:44-58): the callback sees each index in
  // the reference's order.
  template <class F>
  void query_neighbors(float px, float py, float pz, F&& callback) const {
    uint32_t off[2];
    size_t total = 0;
    int st = dnab_grid_query(ctx_, g_, &px, &py, &pz, 1, off, nullptr, 0, &total);
    if (st != DNAB_OK && st != DNAB_ERR_CAPACITY) check(st, ctx_);
    std::vector<uint32_t> idx(total ? total : 1);
    check(dnab_grid_query(ctx_, g_, &px, &py, &pz, 1, off, idx.data(), idx.size(), &total), ctx_);
    for (size_t k = 0; k < total; ++k) callback(idx[k]);
  }
  const dnab_grid* handle() const { return g_; }

 private:
  dnab_ctx* ctx_;
  dnab_grid* g_ = nullptr;
};

class GpuContext {
 public:
  // GpuContext::try_init (src/gpu.rs:55-57, async body 59-155): nullopt when no device.
  static std::optional<GpuContext> try_init(int device = 0) {
    dnab_ctx* c = nullptr;
    int st = dnab_ctx_create(device, &c);
    if (st == DNAB_ERR_NO_DEVICE) return std::nullopt;
    check(st);
    return GpuContext(c);
  }
  GpuContext(GpuContext&& o) noexcept : c_(o.c_) { o.c_ = nullptr; }
  GpuContext& operator=(GpuContext&& o) noexcept {
    std::swap(c_, o.c_);
    return *this;
  }
  GpuContext(const GpuContext&) = delete;
  ~GpuContext() { dnab_ctx_destroy(c_); }
  dnab_ctx* raw() const { return c_; }

  // score_batch (src/gpu.rs:166-267): n_cand = candidates.size() / n_ab.
  std::vector<float> score_batch(const AtomCloud& antigen,
                                 const std::vector<dnab_gpu_atom>& candidates,
                                 size_t n_ab) const {
    const size_t n_cand = n_ab ? candidates.size() / n_ab : 0;
    std::vector<float> e(n_cand);
    size_t got = 0;
    dnab_atom_cloud ag = antigen.view();
    check(dnab_score_batch(c_, &ag, candidates.data(), candidates.size(), n_ab, e.data(),
                           e.size(), &got), c_);
    e.resize(got);
    return e;
  }

 private:
  explicit GpuContext(dnab_ctx* c) : c_(c) {}
  dnab_ctx* c_ = nullptr;
};

inline SpatialHashGrid::SpatialHashGrid(const GpuContext& ctx, float cell_size)
    : ctx_(ctx.raw()) {
  check(dnab_grid_create(ctx_, cell_size, &g_), ctx_);
}

// interaction_energy_atoms_breakdown (src/energy.rs:261-314): LJ, Coulomb,
// hydrophobic, and H-bond and EEF1 solvation when both clouds carry atom
// types, at the reference cutoff of 10 A.
inline AtomEnergyBreakdown interaction_energy_atoms_breakdown(const GpuContext& gpu,
                                                              const AtomCloud& antigen,
                                                              const AtomCloud& antibody,
                                                              float cutoff = 10.0f) {
  dnab_atom_cloud a = antigen.view(), b = antibody.view();
  dnab_energy_breakdown o{};
  check(dnab_interaction_energy_atoms_breakdown(gpu.raw(), &a, &b, cutoff, &o), gpu.raw());
  return {o.lj, o.coulomb, o.hydrophobic, o.hbond, o.solvation};
}

// interaction_energy_atoms (src/energy.rs:256-258)
inline double interaction_energy_atoms(const GpuContext& gpu, const AtomCloud& antigen,
                                       const AtomCloud& antibody, float cutoff = 10.0f) {
  dnab_atom_cloud a = antigen.view(), b = antibody.view();
  double e = 0;
  check(dnab_interaction_energy_atoms(gpu.raw(), &a, &b, cutoff, &e), gpu.raw());
  return e;
}

// allatom_energy (src/diffusion.rs:281-283): interaction_energy_atoms plus the
// antibody's disulfide_energy (src/allatom.rs:155-180).
inline double allatom_energy(const GpuContext& gpu, const AtomProtein& ag, const AtomProtein& ab,
                             double* disulfide = nullptr, float cutoff = 10.0f) {
  dnab_atom_cloud a = ag.atoms.view(), b = ab.atoms.view();
  std::vector<uint8_t> aa(ab.amino_acid.size());
  for (size_t i = 0; i < aa.size(); ++i) aa[i] = static_cast<uint8_t>(ab.amino_acid[i]);
  double e = 0;
  check(dnab_allatom_energy(gpu.raw(), &a, &b, ab.res_start.data(), ab.ca_atom_idx.data(),
                            aa.data(), aa.size(), cutoff, &e, disulfide), gpu.raw());
  return e;
}

// AtomProtein::disulfide_energy (src/allatom.rs:155-180), evaluated on the
// device against an empty antigen.
inline double disulfide_energy(const GpuContext& gpu, const AtomProtein& p) {
  double ds = 0;
  allatom_energy(gpu, AtomProtein{}, p, &ds);
  return ds;
}

// affinity_maturation_allatom (src/diffusion.rs:955-961): greedy point
// mutations at the `mutable` residues (ascending), scored by allatom_energy;
// ab is replaced by the final protein and the best score returned. Residues
// carry the most probable rotamer (this interface keeps no chi angles).
inline double affinity_maturation_allatom(const GpuContext& gpu, AtomProtein& ab,
                                          const AtomProtein& ag,
                                          const std::vector<uint32_t>& mutable_res,
                                          float cutoff = 10.0f) {
  const size_t nr = ab.n_residues();
  std::vector<uint8_t> aa(nr);
  for (size_t i = 0; i < nr; ++i) aa[i] = static_cast<uint8_t>(ab.amino_acid[i]);
  uint32_t cand_ag = 0, cand_res_off[2] = {0, static_cast<uint32_t>(nr)};
  uint32_t mut_off[2] = {0, static_cast<uint32_t>(mutable_res.size())};
  dnab_population_desc d{};
  d.n_cand = 1;
  d.n_res = nr;
  d.cand_antigen = &cand_ag;
  d.cand_res_off = cand_res_off;
  d.res_atom_off = ab.res_start.data();
  d.res_ca_atom = ab.ca_atom_idx.data();
  d.res_amino_acid = aa.data();
  d.atoms = ab.atoms.view();
  dnab_antigen_desc ad{};
  ad.atoms = ag.atoms.view();
  ad.cell_size = cutoff;
  ad.cutoff = cutoff;
  dnab_antigen* dag = nullptr;
  check(dnab_antigen_create(gpu.raw(), &ad, &dag), gpu.raw());
  std::unique_ptr<dnab_antigen, void (*)(dnab_antigen*)> hold(dag, dnab_antigen_destroy);
  const size_t cap = 14 * nr + 1;
  AtomProtein out;
  auto& o = out.atoms;
  o.x.resize(cap); o.y.resize(cap); o.z.resize(cap); o.charge.resize(cap);
  o.r_min_half.resize(cap); o.epsilon.resize(cap); o.hydrophobic.resize(cap);
  o.atom_type.resize(cap); o.residue_idx.resize(cap);
  out.res_start.assign(nr + 1, 0);
  out.ca_atom_idx.assign(nr, 0);
  std::vector<uint8_t> aa_out(nr + 1);
  dnab_protein_out po{cap, o.x.data(), o.y.data(), o.z.data(), o.charge.data(),
                      o.r_min_half.data(), o.epsilon.data(), o.hydrophobic.data(),
                      o.atom_type.data(), o.residue_idx.data(), out.res_start.data(),
                      out.ca_atom_idx.data(), aa_out.data()};
  double e = 0;
  size_t na = 0;
  check(dnab_affinity_maturation(gpu.raw(), dag, &d, nullptr, mut_off,
                                 mutable_res.data(), &po, nullptr, nullptr, &e, &na),
        gpu.raw());
  for (auto* v : {&o.x, &o.y, &o.z, &o.charge, &o.r_min_half, &o.epsilon}) v->resize(na);
  o.hydrophobic.resize(na); o.atom_type.resize(na); o.residue_idx.resize(na);
  out.amino_acid.resize(nr);
  for (size_t i = 0; i < nr; ++i) out.amino_acid[i] = static_cast<AminoAcid>(aa_out[i]);
  ab = std::move(out);
  return e;
}

// parse_pdb (src/pdb.rs:19-57): the CA atoms of a PDB file as a residue cloud.
inline ResidueCloud parse_pdb(const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error(DNAB_ERR_INVALID, "cannot open " + path);
  std::string text;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  std::fclose(f);
  size_t n = 0;
  const int st = dnab_parse_pdb(text.data(), text.size(), nullptr, nullptr, nullptr, nullptr, 0, &n);
  if (st != DNAB_OK && st != DNAB_ERR_CAPACITY) check(st);
  std::vector<float> x(n), y(n), z(n);
  std::vector<uint8_t> aa(n);
  check(dnab_parse_pdb(text.data(), text.size(), x.data(), y.data(), z.data(), aa.data(), n, &n));
  ResidueCloud c;
  for (size_t i = 0; i < n; ++i) c.push(x[i], y[i], z[i], static_cast<AminoAcid>(aa[i]));
  return c;
}

// write_pdb_multi (src/pdb.rs:157-191); with one chain and ter = false,
// write_pdb_allatom (src/pdb.rs:121-155).
inline std::string write_pdb_multi(const std::vector<std::pair<const AtomProtein*, char>>& chains,
                                   bool ter = true) {
  std::vector<dnab_pdb_chain> cs;
  std::vector<std::vector<uint8_t>> aas;
  aas.reserve(chains.size());
  for (const auto& [p, id] : chains) {
    aas.emplace_back(p->amino_acid.size());
    for (size_t i = 0; i < p->amino_acid.size(); ++i) aas.back()[i] = static_cast<uint8_t>(p->amino_acid[i]);
    cs.push_back({p->atoms.x.data(), p->atoms.y.data(), p->atoms.z.data(), p->res_start.data(),
                  aas.back().data(), p->n_residues(), id});
  }
  size_t n = 0;
  check(dnab_write_pdb_chains(cs.data(), cs.size(), ter ? 1 : 0, nullptr, 0, &n));
  std::string out(n + 1, '\0');
  check(dnab_write_pdb_chains(cs.data(), cs.size(), ter ? 1 : 0, out.data(), out.size(), &n));
  out.resize(n);
  return out;
}
inline std::string write_pdb_allatom(const AtomProtein& p, char chain_id) {
  return write_pdb_multi({{&p, chain_id}}, false);
}

// CandidateReport (src/report.rs:8-20)
struct CandidateReport {
  size_t rank = 0;
  std::string sequence;
  std::optional<std::string> light_sequence;
  float energy = 0.0f, entropy_penalty = 0.0f, dg_corrected = 0.0f;
  int32_t net_charge = 0;
  bool aggregation_risk = false;
  size_t n_interface = 0;
  std::optional<size_t> clashes, rama_outliers;
};

inline std::string write_report(const std::vector<CandidateReport>& reports, int format) {
  std::vector<dnab_candidate_report> rs;
  for (const auto& r : reports)
    rs.push_back({r.rank, r.sequence.c_str(), r.light_sequence ? r.light_sequence->c_str() : nullptr,
                  r.energy, r.entropy_penalty, r.dg_corrected, r.net_charge,
                  r.aggregation_risk ? 1 : 0, r.n_interface,
                  r.clashes ? static_cast<int64_t>(*r.clashes) : -1,
                  r.rama_outliers ? static_cast<int64_t>(*r.rama_outliers) : -1});
  size_t n = 0;
  check(dnab_write_report(rs.data(), rs.size(), format, nullptr, 0, &n));
  std::string out(n + 1, '\0');
  check(dnab_write_report(rs.data(), rs.size(), format, out.data(), out.size(), &n));
  out.resize(n);
  return out;
}
// write_fasta / write_json / write_csv (src/report.rs:41-113)
inline std::string write_fasta(const std::vector<CandidateReport>& r) {
  return write_report(r, DNAB_REPORT_FASTA);
}
inline std::string write_json(const std::vector<CandidateReport>& r) {
  return write_report(r, DNAB_REPORT_JSON);
}
inline std::string write_csv(const std::vector<CandidateReport>& r) {
  return write_report(r, DNAB_REPORT_CSV);
}

// compute_forces_with_grid (src/energy.rs:322-398): outputs sized >= n_ab.
inline void compute_forces_with_grid(const GpuContext& gpu, const ResidueCloud& antigen,
                                     const SpatialHashGrid& grid, const ResidueCloud& antibody,
                                     std::vector<float>& fx, std::vector<float>& fy,
                                     std::vector<float>& fz) {
  dnab_residue_cloud a = antigen.view(), b = antibody.view();
  fx.resize(std::max(fx.size(), antibody.len()));
  fy.resize(fx.size());
  fz.resize(fx.size());
  check(dnab_cg_forces_with_grid(gpu.raw(), &a, grid.handle(), &b, fx.data(), fy.data(),
                                 fz.data()), gpu.raw());
}

// compute_forces (src/energy.rs:154-234)
inline void compute_forces(const GpuContext& gpu, const ResidueCloud& antigen,
                           const ResidueCloud& antibody, std::vector<float>& fx,
                           std::vector<float>& fy, std::vector<float>& fz) {
  dnab_residue_cloud a = antigen.view(), b = antibody.view();
  fx.resize(std::max(fx.size(), antibody.len()));
  fy.resize(fx.size());
  fz.resize(fx.size());
  check(dnab_cg_forces(gpu.raw(), &a, &b, fx.data(), fy.data(), fz.data()), gpu.raw());
}

struct GpuPhaseResult {
  uint32_t candidate;  // index into the candidates passed in
  double energy;       // LJ + Coulomb + hydrophobic at the final scoring
};

// The GPU phase of run_allatom (src/diffusion.rs:1008-1062) with every
// candidate moving for params.n_steps gradient steps, scoring every
// params.score_every steps and keeping params.top_k (default 64).
inline std::vector<GpuPhaseResult> run_allatom_gpu_phase(
    const GpuContext& gpu, const AtomProtein& antigen, const std::vector<AtomProtein>& cands,
    const dnab_langevin_params& params, float cutoff = 10.0f, float cell_size = 10.0f) {
  dnab_ctx* c = gpu.raw();
  ResidueCloud ca;
  for (size_t r = 0; r < antigen.n_residues(); ++r) {
    uint32_t i = antigen.ca_atom_idx[r];
    ca.push(antigen.atoms.x[i], antigen.atoms.y[i], antigen.atoms.z[i], antigen.amino_acid[r]);
  }
  dnab_antigen_desc ad{};
  ad.atoms = antigen.atoms.view();
  ad.ca = ca.view();
  auto ctr = antigen.ca_center_of_mass();
  std::memcpy(ad.center, ctr.data(), sizeof(ad.center));
  ad.cell_size = cell_size;
  ad.cutoff = cutoff;
  dnab_antigen* ag = nullptr;
  check(dnab_antigen_create(c, &ad, &ag), c);
  std::unique_ptr<dnab_antigen, void (*)(dnab_antigen*)> ag_guard(ag, dnab_antigen_destroy);
  // flatten the candidates
  AtomCloud all;
  std::vector<uint32_t> cand_ag(cands.size(), 0), cand_res_off{0}, res_atom_off{0}, res_ca;
  std::vector<uint8_t> res_aa;
  for (const auto& p : cands) {
    const uint32_t base = static_cast<uint32_t>(all.len());
    for (size_t r = 0; r < p.n_residues(); ++r) {
      res_atom_off.push_back(base + p.res_start[r + 1]);
      res_ca.push_back(base + p.ca_atom_idx[r]);
      res_aa.push_back(static_cast<uint8_t>(p.amino_acid[r]));
    }
    cand_res_off.push_back(static_cast<uint32_t>(res_ca.size()));
    auto app = [](auto& dst, const auto& src) { dst.insert(dst.end(), src.begin(), src.end()); };
    app(all.x, p.atoms.x); app(all.y, p.atoms.y); app(all.z, p.atoms.z);
    app(all.charge, p.atoms.charge); app(all.r_min_half, p.atoms.r_min_half);
    app(all.epsilon, p.atoms.epsilon); app(all.hydrophobic, p.atoms.hydrophobic);
  }
  dnab_population_desc pd{};
  pd.n_cand = cands.size();
  pd.n_res = res_ca.size();
  pd.cand_antigen = cand_ag.data();
  pd.cand_res_off = cand_res_off.data();
  pd.res_atom_off = res_atom_off.data();
  pd.res_ca_atom = res_ca.data();
  pd.res_amino_acid = res_aa.data();
  pd.atoms = all.view();
  dnab_population* pop = nullptr;
  dnab_antigen* ags[1] = {ag};
  check(dnab_population_create(c, ags, 1, &pd, &pop), c);
  std::unique_ptr<dnab_population, void (*)(dnab_population*)> pop_guard(pop,
                                                                         dnab_population_destroy);
  std::vector<uint32_t> idx(params.top_k ? params.top_k : 1);
  std::vector<double> e(idx.size());
  uint32_t n = 0;
  check(dnab_langevin_run(c, pop, &params, idx.data(), e.data(), &n), c);
  std::vector<GpuPhaseResult> out(n);
  for (uint32_t k = 0; k < n; ++k) out[k] = {idx[k], e[k]};
  return out;
}


// allatom_diffusion_step (src/diffusion.rs:290-416) lp.n_steps times for every
// antibody against the antigen, in one dnab_refine call (mc: the move rates,
// NULL for the reference's). Each antibody is replaced by its refined protein;
// rotamers (optional, one vector per antibody) carry each residue's rotamer
// index in and out. Returns each refined antibody's allatom_energy.
inline std::vector<double> allatom_diffusion_steps(
    const GpuContext& gpu, const AtomProtein& antigen, std::vector<AtomProtein>& abs,
    const dnab_langevin_params& lp, const dnab_mc_params* mc = nullptr,
    std::vector<std::vector<uint8_t>>* rotamers = nullptr, float cutoff = 10.0f) {
  dnab_ctx* c = gpu.raw();
  ResidueCloud ca;
  std::vector<uint8_t> ca_aa;
  for (size_t r = 0; r < antigen.n_residues(); ++r) {
    const uint32_t i = antigen.ca_atom_idx[r];
    ca.push(antigen.atoms.x[i], antigen.atoms.y[i], antigen.atoms.z[i], antigen.amino_acid[r]);
    ca_aa.push_back(static_cast<uint8_t>(antigen.amino_acid[r]));
  }
  dnab_antigen_desc ad{};
  ad.atoms = antigen.atoms.view();
  ad.ca = ca.view();
  ad.ca_amino_acid = ca_aa.data();
  const auto ctr = antigen.ca_center_of_mass();
  std::memcpy(ad.center, ctr.data(), sizeof(ad.center));
  ad.cell_size = cutoff;
  ad.cutoff = cutoff;
  dnab_antigen* ag = nullptr;
  check(dnab_antigen_create(c, &ad, &ag), c);
  std::unique_ptr<dnab_antigen, void (*)(dnab_antigen*)> ag_guard(ag, dnab_antigen_destroy);
  AtomCloud all;
  std::vector<uint32_t> cand_ag(abs.size(), 0), cand_res_off{0}, res_atom_off{0}, res_ca;
  std::vector<uint8_t> res_aa, res_rot;
  for (size_t a = 0; a < abs.size(); ++a) {
    const AtomProtein& p = abs[a];
    const uint32_t base = static_cast<uint32_t>(all.len());
    for (size_t r = 0; r < p.n_residues(); ++r) {
      res_atom_off.push_back(base + p.res_start[r + 1]);
      res_ca.push_back(base + p.ca_atom_idx[r]);
      res_aa.push_back(static_cast<uint8_t>(p.amino_acid[r]));
      res_rot.push_back(rotamers && a < rotamers->size() && r < (*rotamers)[a].size()
                            ? (*rotamers)[a][r]
                            : 0);
    }
    cand_res_off.push_back(static_cast<uint32_t>(res_ca.size()));
    auto app = [](auto& dst, const auto& src) { dst.insert(dst.end(), src.begin(), src.end()); };
    app(all.x, p.atoms.x); app(all.y, p.atoms.y); app(all.z, p.atoms.z);
    app(all.charge, p.atoms.charge); app(all.r_min_half, p.atoms.r_min_half);
    app(all.epsilon, p.atoms.epsilon); app(all.hydrophobic, p.atoms.hydrophobic);
  }
  const size_t nr = res_ca.size(), cap = 14 * nr + 1;
  dnab_population_desc pd{};
  pd.n_cand = abs.size();
  pd.n_res = nr;
  pd.cand_antigen = cand_ag.data();
  pd.cand_res_off = cand_res_off.data();
  pd.res_atom_off = res_atom_off.data();
  pd.res_ca_atom = res_ca.data();
  pd.res_amino_acid = res_aa.data();
  pd.atoms = all.view();
  AtomCloud o;
  o.x.resize(cap); o.y.resize(cap); o.z.resize(cap); o.charge.resize(cap);
  o.r_min_half.resize(cap); o.epsilon.resize(cap); o.hydrophobic.resize(cap);
  o.atom_type.resize(cap); o.residue_idx.resize(cap);
  std::vector<uint32_t> rs(nr + 1), cai(std::max<size_t>(nr, 1));
  std::vector<uint8_t> aa_out(nr + 1), rot_out(nr + 1);
  dnab_protein_out po{cap, o.x.data(), o.y.data(), o.z.data(), o.charge.data(),
                      o.r_min_half.data(), o.epsilon.data(), o.hydrophobic.data(),
                      o.atom_type.data(), o.residue_idx.data(), rs.data(), cai.data(),
                      aa_out.data()};
  dnab_mc_params mp{};
  if (mc) mp = *mc;
  else check(dnab_mc_default_params(&mp));
  std::vector<double> e(std::max<size_t>(abs.size(), 1));
  size_t na = 0;
  check(dnab_refine(c, ag, &pd, res_rot.data(), &lp, &mp, &po, rot_out.data(), e.data(), &na), c);
  for (size_t a = 0; a < abs.size(); ++a) {
    const uint32_t r0 = cand_res_off[a], r1 = cand_res_off[a + 1], a0 = rs[r0], a1 = rs[r1];
    AtomProtein q;
    auto cut = [&](auto& dst, const auto& src) { dst.assign(src.begin() + a0, src.begin() + a1); };
    cut(q.atoms.x, o.x); cut(q.atoms.y, o.y); cut(q.atoms.z, o.z); cut(q.atoms.charge, o.charge);
    cut(q.atoms.r_min_half, o.r_min_half); cut(q.atoms.epsilon, o.epsilon);
    cut(q.atoms.hydrophobic, o.hydrophobic); cut(q.atoms.atom_type, o.atom_type);
    cut(q.atoms.residue_idx, o.residue_idx);
    q.res_start.assign(1, 0);
    for (uint32_t r = r0; r < r1; ++r) {
      q.res_start.push_back(rs[r + 1] - a0);
      q.ca_atom_idx.push_back(cai[r] - a0);
      q.amino_acid.push_back(static_cast<AminoAcid>(aa_out[r]));
    }
    if (rotamers) {
      if ((*rotamers).size() < abs.size()) (*rotamers).resize(abs.size());
      (*rotamers)[a].assign(rot_out.begin() + r0, rot_out.begin() + r1);
    }
    abs[a] = std::move(q);
  }
  e.resize(abs.size());
  return e;
}

// AllAtomResult (src/diffusion.rs:93-97)
struct AllAtomResult {
  AtomProtein antibody;
  double energy;         // allatom_energy
  std::string sequence;  // one-letter codes
};

// run_allatom (src/diffusion.rs:983-1107) in one dnab_run_allatom call: the
// broad phase of `n_broad` binders (0: none, the reference's path without a
// GPU), the refinement of the `population` survivors over cpu_iterations
// steps, results in ascending allatom_energy, max(top_n, 1) of them. The
// antigen's Calpha centre of mass is the binders' ring centre.
inline std::vector<AllAtomResult> run_allatom(const GpuContext& gpu, const AtomProtein& antigen,
                                              size_t ab_length, size_t population,
                                              size_t cpu_iterations, size_t top_n,
                                              uint32_t n_broad = 1024, uint64_t seed = 0,
                                              float cutoff = 10.0f) {
  dnab_ctx* c = gpu.raw();
  ResidueCloud ca;
  std::vector<uint8_t> ca_aa;
  for (size_t r = 0; r < antigen.n_residues(); ++r) {
    const uint32_t i = antigen.ca_atom_idx[r];
    ca.push(antigen.atoms.x[i], antigen.atoms.y[i], antigen.atoms.z[i], antigen.amino_acid[r]);
    ca_aa.push_back(static_cast<uint8_t>(antigen.amino_acid[r]));
  }
  dnab_antigen_desc ad{};
  ad.atoms = antigen.atoms.view();
  ad.ca = ca.view();
  ad.ca_amino_acid = ca_aa.data();
  const auto ctr = antigen.ca_center_of_mass();
  std::memcpy(ad.center, ctr.data(), sizeof(ad.center));
  ad.cell_size = cutoff;
  ad.cutoff = cutoff;
  dnab_antigen* ag = nullptr;
  check(dnab_antigen_create(c, &ad, &ag), c);
  std::unique_ptr<dnab_antigen, void (*)(dnab_antigen*)> ag_guard(ag, dnab_antigen_destroy);
  dnab_run_allatom_params p;
  dnab_run_allatom_default_params(&p);
  p.ab_length = static_cast<uint32_t>(ab_length);
  p.population = static_cast<uint32_t>(population);
  p.cpu_iterations = static_cast<uint32_t>(cpu_iterations);
  p.top_n = static_cast<uint32_t>(top_n);
  p.n_broad = n_broad;
  p.seed = seed;
  const size_t pool = n_broad ? std::min<size_t>(population, n_broad) : population;
  const size_t n_keep = std::min<size_t>(std::max<size_t>(top_n, 1), pool);
  const size_t nr = n_keep * ab_length, cap = 14 * nr + 1;
  AtomCloud all;
  all.x.resize(cap); all.y.resize(cap); all.z.resize(cap); all.charge.resize(cap);
  all.r_min_half.resize(cap); all.epsilon.resize(cap); all.hydrophobic.resize(cap);
  all.atom_type.resize(cap); all.residue_idx.resize(cap);
  std::vector<uint32_t> rs(nr + 1), cai(std::max<size_t>(nr, 1));
  std::vector<uint8_t> aa(std::max<size_t>(nr, 1));
  dnab_protein_out o{cap, all.x.data(), all.y.data(), all.z.data(), all.charge.data(),
                     all.r_min_half.data(), all.epsilon.data(), all.hydrophobic.data(),
                     all.atom_type.data(), all.residue_idx.data(), rs.data(), cai.data(),
                     aa.data()};
  std::vector<double> e(std::max<size_t>(n_keep, 1));
  uint32_t n = 0;
  size_t na = 0;
  check(dnab_run_allatom(c, ag, &p, &o, e.data(), &n, &na), c);
  static const char kOne[] = "ARNDCQEGHILKMFPSTWYV";  // AminoAcid order
  std::vector<AllAtomResult> out;
  for (uint32_t k = 0; k < n; ++k) {
    AllAtomResult r;
    const size_t g0 = k * ab_length, g1 = g0 + ab_length;
    const uint32_t a0 = rs[g0], a1 = rs[g1];
    AtomCloud& b = r.antibody.atoms;
    auto cut = [&](auto& dst, const auto& src) { dst.assign(src.begin() + a0, src.begin() + a1); };
    cut(b.x, all.x); cut(b.y, all.y); cut(b.z, all.z); cut(b.charge, all.charge);
    cut(b.r_min_half, all.r_min_half); cut(b.epsilon, all.epsilon);
    cut(b.hydrophobic, all.hydrophobic); cut(b.atom_type, all.atom_type);
    cut(b.residue_idx, all.residue_idx);
    r.antibody.res_start.assign(ab_length + 1, 0);
    for (size_t g = g0; g < g1; ++g) {
      r.antibody.res_start[g - g0 + 1] = rs[g + 1] - a0;
      r.antibody.ca_atom_idx.push_back(cai[g] - a0);
      r.antibody.amino_acid.push_back(static_cast<AminoAcid>(aa[g]));
      r.sequence.push_back(kOne[aa[g]]);
    }
    r.energy = e[k];
    out.push_back(std::move(r));
  }
  return out;
}

}  // namespace binder
