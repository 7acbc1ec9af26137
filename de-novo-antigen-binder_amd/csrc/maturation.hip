// Batched affinity maturation: the greedy point-mutation scan of
// affinity_maturation_allatom (src/diffusion.rs:924-961) for many antibodies
// against one antigen, with every trial's score taken from one ragged
// scoring batch.
//
// The scan keeps, per mutable position r (in the given order), the first
// amino acid in ALL_AA order (src/atom.rs:29-50) whose mutant scores strictly
// below the best score so far; a non-improving trial is undone by
// mutate_residue(r, old, old_chi), which rebuilds the old residue on an ideal
// backbone (src/allatom.rs:299-384). So residue r only ever holds one of 21
// atom sets: its input atoms, the old type rebuilt, or one of 19 mutants, all
// of them built around the input Calpha. allatom_energy (src/diffusion.rs:
// 281-283) is a sum over binder atoms of antigen pair terms plus the Cys SG
// disulfide wells, so a trial's score is the state's score with residue r's
// pair terms and wells exchanged.
//
// Device work: one population whose candidates are single residues (every
// input residue, then 20 built residues per mutable position), evaluated once
// with the five-term breakdown kernel; then k_maturation_scan, one thread per
// antibody, walks the positions and trials in the reference's order with
// FP64 sums. The built residues come from the host builder (the same one
// protein_from_ca_trace uses); the caller receives the final proteins.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "dnab_internal.h"
#include "host_parallel.h"
#include "ff_tables.h"
#include "ff_terms.h"

namespace dnab {
namespace {

// Residue atom sets per mutable position: slot 0 is the old type rebuilt
// with its rotamer, slots 1..19 the other types in ALL_AA order, rotamer 0.
constexpr uint32_t kSlots = kNumAminoAcids;

struct PopGuard {
  dnab_population* p = nullptr;
  ~PopGuard() { dnab_population_destroy(p); }
};

// Candidate c's SG position (topology slot 5) when it is a Cys holding one.
__device__ inline bool sg_of(const uint32_t* __restrict__ res_atom_off,
                             const uint8_t* __restrict__ res_aa, const float4* __restrict__ pos,
                             uint32_t c, float3& p) {
  if (res_aa[c] != kCys || res_atom_off[c + 1] - res_atom_off[c] <= kCysSgSlot) return false;
  const float4 v = pos[res_atom_off[c] + kCysSgSlot];
  p = make_float3(v.x, v.y, v.z);
  return true;
}

// One disulfide well (src/allatom.rs:155-180): the pair set by the f32
// distance as the reference decides it, the value in FP64.
__device__ inline double ss_well(float3 a, float3 b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  if (sqrtf(dx * dx + dy * dy + dz * dz) > kSsCutoff) return 0.0;
  const double ex = static_cast<double>(a.x) - b.x, ey = static_cast<double>(a.y) - b.y,
               ez = static_cast<double>(a.z) - b.z;
  const double t = (sqrt(ex * ex + ey * ey + ez * ez) - static_cast<double>(kSsD0)) /
                   static_cast<double>(kSsSigma);
  return -static_cast<double>(kSsDepth) * exp(-0.5 * t * t);
}

// One thread per antibody. state[g] = candidate whose atoms residue g holds.
__global__ void __launch_bounds__(64)
k_maturation_scan(const uint32_t* __restrict__ ab_res_off, const uint32_t* __restrict__ mut_off,
                  const uint32_t* __restrict__ mut_res, const double* __restrict__ bd,
                  const uint32_t* __restrict__ res_atom_off, const uint8_t* __restrict__ res_aa,
                  const float4* __restrict__ pos, uint32_t n_res, uint32_t n_ab,
                  uint32_t* __restrict__ state, double* __restrict__ energy) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_ab) return;
  const uint32_t r0 = ab_res_off[a], r1 = ab_res_off[a + 1];
  // interaction_energy_atoms of one residue: total() order (src/energy.rs:247-249)
  auto T = [&](uint32_t c) {
    const double* t = bd + static_cast<size_t>(kBdTerms) * c;
    return (((t[0] + t[1]) + t[2]) + t[3]) + t[4];
  };
  double inter = 0.0;
  for (uint32_t g = r0; g < r1; ++g) {
    state[g] = g;
    inter += T(g);
  }
  double ss = 0.0;
  for (uint32_t i = r0; i < r1; ++i) {
    float3 pi;
    if (!sg_of(res_atom_off, res_aa, pos, i, pi)) continue;
    for (uint32_t j = i + 1; j < r1; ++j) {
      float3 pj;
      if (sg_of(res_atom_off, res_aa, pos, j, pj)) ss += ss_well(pi, pj);
    }
  }
  // change of the disulfide term when residue g's atoms go from c_old to c_new
  auto ss_swap = [&](uint32_t g, uint32_t c_old, uint32_t c_new) {
    float3 po, pn;
    const bool ho = sg_of(res_atom_off, res_aa, pos, c_old, po);
    const bool hn = sg_of(res_atom_off, res_aa, pos, c_new, pn);
    double d = 0.0;
    if (!ho && !hn) return d;
    for (uint32_t s = r0; s < r1; ++s) {
      float3 ps;
      if (s == g || !sg_of(res_atom_off, res_aa, pos, state[s], ps)) continue;
      if (ho) d -= ss_well(po, ps);
      if (hn) d += ss_well(pn, ps);
    }
    return d;
  };
  double best = inter + ss;
  for (uint32_t m = mut_off[a]; m < mut_off[a + 1]; ++m) {
    const uint32_t g = mut_res[m];
    const uint32_t base = n_res + kSlots * m;
    uint32_t cur = state[g];
    double tcur = T(cur);
    for (uint32_t k = 1; k < kSlots; ++k) {
      const uint32_t c = base + k;
      const double inter_c = (inter - tcur) + T(c);
      const double ss_c = ss + ss_swap(g, cur, c);
      const double e = inter_c + ss_c;
      if (e < best) {  // keep the first strict improvement (src/diffusion.rs:940-948)
        best = e;
        inter = inter_c;
        ss = ss_c;
        state[g] = c;
        break;
      }
      if (cur != base) {  // undo: the old type rebuilt (src/diffusion.rs:950)
        ss += ss_swap(g, cur, base);
        inter = (inter - tcur) + T(base);
        cur = base;
        tcur = T(base);
        state[g] = base;
      }
    }
  }
  energy[a] = best;
}

}  // namespace
}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_affinity_maturation(dnab_ctx* ctx, dnab_antigen* antigen,
                             const dnab_population_desc* ab, const uint8_t* res_rotamer,
                             const uint32_t* mutable_off, const uint32_t* mutable_res,
                             dnab_protein_out* out, uint8_t* out_rotamer, uint8_t* out_rebuilt,
                             double* out_energy, size_t* n_atoms_out) {
  if (!ctx || !antigen || !ab || !mutable_off || antigen->ctx != ctx) return DNAB_ERR_INVALID;
  const size_t n_ab = ab->n_cand, n_res = ab->n_res, n_in = ab->atoms.n;
  if (n_ab && (!ab->cand_res_off || !out_energy)) return DNAB_ERR_INVALID;
  if (n_res && (!ab->res_atom_off || !ab->res_ca_atom || !ab->res_amino_acid))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "maturation: residues need offsets, Calpha and types");
  if (n_ab >= 0xFFFFFF00ull || n_res >= 0xFFFFFF00ull)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "maturation: too large");
  if (n_ab && (ab->cand_res_off[0] != 0 || ab->cand_res_off[n_ab] != n_res))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "maturation: cand_res_off must span 0..n_res");
  if (mutable_off[0] != 0) return dnab_set_error(ctx, DNAB_ERR_INVALID, "maturation: mutable_off[0] != 0");
  for (size_t r = 0; r < n_res; ++r) {
    if (ab->res_amino_acid[r] >= kNumAminoAcids)
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "maturation: residue %zu type out of range", r);
    if (res_rotamer && kRotamerCount[ab->res_amino_acid[r]] &&
        res_rotamer[r] >= kRotamerCount[ab->res_amino_acid[r]])
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "maturation: residue %zu rotamer out of range", r);
  }
  // mutable positions: local residue indices, strictly ascending per antibody
  // (the reference's (0..n).filter(is_interface), src/main.rs:342-344)
  const size_t n_mut = n_ab ? mutable_off[n_ab] : 0;
  if (n_mut && !mutable_res) return DNAB_ERR_INVALID;
  if (n_mut >= (0xFFFFFF00ull - n_res) / kSlots)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "maturation: too many mutable positions");
  std::vector<uint32_t> mut_global(n_mut), mut_ab(n_mut);
  for (size_t a = 0; a < n_ab; ++a) {
    if (ab->cand_res_off[a + 1] < ab->cand_res_off[a] || mutable_off[a + 1] < mutable_off[a] ||
        mutable_off[a + 1] > n_mut)
      return dnab_set_error(ctx, DNAB_ERR_INVALID, "maturation: offsets not monotone");
    const uint32_t nr = ab->cand_res_off[a + 1] - ab->cand_res_off[a];
    for (uint32_t m = mutable_off[a]; m < mutable_off[a + 1]; ++m) {
      if (mutable_res[m] >= nr || (m > mutable_off[a] && mutable_res[m] <= mutable_res[m - 1]))
        return dnab_set_error(ctx, DNAB_ERR_INVALID,
                              "maturation: antibody %zu mutable positions must be ascending "
                              "residue indices within it", a);
      mut_global[m] = ab->cand_res_off[a] + mutable_res[m];
      mut_ab[m] = static_cast<uint32_t>(a);
    }
  }
  if (n_ab == 0) {
    if (n_atoms_out) *n_atoms_out = 0;
    if (out && out->res_start) out->res_start[0] = 0;
    return DNAB_OK;
  }
  if (n_res && (ab->res_atom_off[0] != 0 || ab->res_atom_off[n_res] != n_in))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "maturation: res_atom_off must span 0..n_atoms");
  if (n_in && (!ab->atoms.x || !ab->atoms.y || !ab->atoms.z || !ab->atoms.charge ||
               !ab->atoms.r_min_half || !ab->atoms.epsilon || !ab->atoms.hydrophobic))
    return DNAB_ERR_INVALID;

  // ---- the scoring batch: input residues, then kSlots built residues per position
  auto t0 = std::chrono::steady_clock::now();
  const size_t nc = n_res + kSlots * n_mut;
  std::vector<uint8_t> c_aa(nc), c_rot(nc, 0);
  std::vector<uint32_t> c_atom_off(nc + 1, 0);
  for (size_t g = 0; g < n_res; ++g) {
    c_aa[g] = ab->res_amino_acid[g];
    c_rot[g] = res_rotamer ? res_rotamer[g] : 0;
    c_atom_off[g + 1] = ab->res_atom_off[g + 1];
  }
  for (size_t m = 0; m < n_mut; ++m) {
    const uint8_t old = ab->res_amino_acid[mut_global[m]];
    size_t c = n_res + kSlots * m;
    c_aa[c] = old;
    c_rot[c] = res_rotamer ? res_rotamer[mut_global[m]] : 0;
    for (uint8_t t = 0; t < kNumAminoAcids; ++t)
      if (t != old) c_aa[++c] = t;
  }
  for (size_t c = n_res; c < nc; ++c)
    c_atom_off[c + 1] = c_atom_off[c] + kResidueTopology[c_aa[c]].n_atoms;
  const size_t na = c_atom_off[nc];
  // every element is written below (the input atoms copied, each built
  // residue's atoms by its build): no zero fill of ~30 B per atom first
  HostVec<float> x(na), y(na), z(na), q(na), rmh(na), eps(na);
  HostVec<uint8_t> hyd(na);
  std::vector<uint8_t> typ(na);  // zero where the input has no types
  std::vector<uint32_t> ca(nc), cand_res_off(nc + 1), cand_ag(nc, 0);
  const bool typed = ab->atoms.atom_type != nullptr;
  std::memcpy(x.data(), ab->atoms.x, n_in * sizeof(float));
  std::memcpy(y.data(), ab->atoms.y, n_in * sizeof(float));
  std::memcpy(z.data(), ab->atoms.z, n_in * sizeof(float));
  std::memcpy(q.data(), ab->atoms.charge, n_in * sizeof(float));
  std::memcpy(rmh.data(), ab->atoms.r_min_half, n_in * sizeof(float));
  std::memcpy(eps.data(), ab->atoms.epsilon, n_in * sizeof(float));
  std::memcpy(hyd.data(), ab->atoms.hydrophobic, n_in);
  if (typed) std::memcpy(typ.data(), ab->atoms.atom_type, n_in);
  for (size_t c = 0; c <= nc; ++c) cand_res_off[c] = static_cast<uint32_t>(c);
  for (size_t g = 0; g < n_res; ++g) ca[g] = ab->res_ca_atom[g];
  std::atomic<int> status{DNAB_OK};
  parallel_for(n_mut * kSlots, [&](size_t b, size_t e) {
    for (size_t k = b; k < e; ++k) {
      const size_t c = n_res + k;
      const uint32_t gca = ab->res_ca_atom[mut_global[k / kSlots]];
      const float cx = ab->atoms.x[gca], cy = ab->atoms.y[gca], cz = ab->atoms.z[gca];
      const size_t a0 = c_atom_off[c];
      uint32_t rs[2], cai;
      uint8_t aa_out;
      dnab_protein_out o{};
      o.atom_capacity = c_atom_off[c + 1] - a0;
      o.x = &x[a0];
      o.y = &y[a0];
      o.z = &z[a0];
      o.charge = &q[a0];
      o.r_min_half = &rmh[a0];
      o.epsilon = &eps[a0];
      o.hydrophobic = &hyd[a0];
      o.atom_type = &typ[a0];
      o.res_start = rs;
      o.ca_atom_idx = &cai;
      o.amino_acid = &aa_out;
      size_t got = 0;
      const int st = dnab_build_protein(&cx, &cy, &cz, &c_aa[c], &c_rot[c], 1, &o, &got);
      if (st != DNAB_OK || rs[1] != o.atom_capacity) {
        status = st != DNAB_OK ? st : DNAB_ERR_INVALID;
        return;
      }
      ca[c] = static_cast<uint32_t>(a0 + cai);
    }
  });
  if (status.load() != DNAB_OK)
    return dnab_set_error(ctx, status.load(), "maturation: residue build failed");
  timer_host(ctx, "maturation_host_build", ms_since(t0));
  t0 = std::chrono::steady_clock::now();

  PopGuard pop;
  dnab_population_desc pd{};
  pd.n_cand = nc;
  pd.n_res = nc;
  pd.cand_antigen = cand_ag.data();
  pd.cand_res_off = cand_res_off.data();
  pd.res_atom_off = c_atom_off.data();
  pd.res_ca_atom = ca.data();
  pd.res_amino_acid = c_aa.data();
  pd.atoms.n = na;
  pd.atoms.x = x.data();
  pd.atoms.y = y.data();
  pd.atoms.z = z.data();
  pd.atoms.charge = q.data();
  pd.atoms.r_min_half = rmh.data();
  pd.atoms.epsilon = eps.data();
  pd.atoms.hydrophobic = hyd.data();
  pd.atoms.atom_type = typed ? typ.data() : nullptr;
  dnab_antigen* ags[1] = {antigen};
  DNAB_RETURN_IF(dnab_population_create(ctx, ags, 1, &pd, &pop.p));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  timer_host(ctx, "maturation_host_upload", ms_since(t0));
  t0 = std::chrono::steady_clock::now();
  DNAB_RETURN_IF(dnab_population_evaluate(ctx, pop.p, DNAB_EVAL_BREAKDOWN));

  // ---- the scan
  DevBuf<uint32_t> d_ab_off, d_mut_off, d_mut, d_state;
  DevBuf<double> d_energy;
  DNAB_ALLOC_CHECK(ctx, d_ab_off.alloc(n_ab + 1));
  DNAB_ALLOC_CHECK(ctx, d_mut_off.alloc(n_ab + 1));
  DNAB_ALLOC_CHECK(ctx, d_mut.alloc(std::max<size_t>(n_mut, 1)));
  DNAB_ALLOC_CHECK(ctx, d_state.alloc(std::max<size_t>(n_res, 1)));
  DNAB_ALLOC_CHECK(ctx, d_energy.alloc(n_ab));
  DNAB_RETURN_IF(upload(ctx, d_ab_off.p, ab->cand_res_off, (n_ab + 1) * sizeof(uint32_t)));
  DNAB_RETURN_IF(upload(ctx, d_mut_off.p, mutable_off, (n_ab + 1) * sizeof(uint32_t)));
  if (n_mut) DNAB_RETURN_IF(upload(ctx, d_mut.p, mut_global.data(), n_mut * sizeof(uint32_t)));
  hipEvent_t ev_end = nullptr;
  hipEvent_t ev = timer_begin(ctx, "maturation_scan", &ev_end);
  hipLaunchKernelGGL(k_maturation_scan, dim3(static_cast<uint32_t>((n_ab + 63) / 64)), dim3(64),
                     0, ctx->stream, d_ab_off.p, d_mut_off.p, d_mut.p, pop.p->breakdown.p,
                     pop.p->res_atom_off.p, pop.p->res_aa.p, pop.p->pos.p,
                     static_cast<uint32_t>(n_res), static_cast<uint32_t>(n_ab), d_state.p,
                     d_energy.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  timer_end(ctx, "maturation_scan", ev, ev_end);
  std::vector<uint32_t> state(n_res);
  if (n_res) DNAB_RETURN_IF(download(ctx, state.data(), d_state.p, n_res * sizeof(uint32_t)));
  DNAB_RETURN_IF(download(ctx, out_energy, d_energy.p, n_ab * sizeof(double)));
  timer_host(ctx, "maturation_host_device", ms_since(t0));
  t0 = std::chrono::steady_clock::now();

  // ---- the final proteins: each residue's input atoms or its chosen built set
  size_t n_out = 0;
  for (size_t g = 0; g < n_res; ++g) n_out += c_atom_off[state[g] + 1] - c_atom_off[state[g]];
  if (n_atoms_out) *n_atoms_out = n_out;
  for (size_t g = 0; g < n_res; ++g) {
    if (out_rotamer) out_rotamer[g] = c_rot[state[g]];
    if (out_rebuilt) out_rebuilt[g] = state[g] != g;
  }
  if (!out) return DNAB_OK;
  if (!out->x || !out->y || !out->z) return DNAB_ERR_INVALID;
  if (n_out > out->atom_capacity) return DNAB_ERR_CAPACITY;
  // each residue's first output atom and its antibody, then the residues'
  // atoms copied in parallel
  std::vector<size_t> r_out(n_res + 1, 0);
  std::vector<uint32_t> r_ab(n_res);
  for (size_t a = 0; a < n_ab; ++a)
    for (uint32_t g = ab->cand_res_off[a]; g < ab->cand_res_off[a + 1]; ++g) {
      r_ab[g] = static_cast<uint32_t>(a);
      r_out[g + 1] = r_out[g] + (c_atom_off[state[g] + 1] - c_atom_off[state[g]]);
    }
  if (out->res_start) out->res_start[0] = 0;
  parallel_for(n_res, [&](size_t g0, size_t g1) {
    for (size_t g = g0; g < g1; ++g) {
      const uint32_t c = state[g], b = c_atom_off[c], e = c_atom_off[c + 1];
      const uint32_t local = static_cast<uint32_t>(g) - ab->cand_res_off[r_ab[g]];
      size_t o = r_out[g];
      if (out->ca_atom_idx) out->ca_atom_idx[g] = static_cast<uint32_t>(o + (ca[c] - b));
      if (out->amino_acid) out->amino_acid[g] = c_aa[c];
      for (uint32_t i = b; i < e; ++i, ++o) {
        out->x[o] = x[i];
        out->y[o] = y[i];
        out->z[o] = z[i];
        if (out->charge) out->charge[o] = q[i];
        if (out->r_min_half) out->r_min_half[o] = rmh[i];
        if (out->epsilon) out->epsilon[o] = eps[i];
        if (out->hydrophobic) out->hydrophobic[o] = hyd[i];
        if (out->atom_type) out->atom_type[o] = typ[i];
        if (out->residue_idx)
          out->residue_idx[o] = (c == g && ab->atoms.residue_idx) ? ab->atoms.residue_idx[i] : local;
      }
      if (out->res_start) out->res_start[g + 1] = static_cast<uint32_t>(o);
    }
  });
  timer_host(ctx, "maturation_host_assemble", ms_since(t0));
  return DNAB_OK;
}

}  // extern "C"
