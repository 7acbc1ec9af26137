// Internal host-side definitions of the dnab library objects.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/dnab.h"

namespace dnab {

// Host RNG for synthetic inputs: SplitMix64 (Steele, Lea, Flood 2014).
struct SplitMix64 {
  uint64_t state;
  uint64_t next();
  float uniform(float lo, float hi);  // [lo, hi)
  uint32_t below(uint32_t n);         // [0, n)
};

void place_atom(const float a[3], const float b[3], const float c[3], float bond_len,
                float angle_deg, float dihedral_deg, float out[3]);
int ideal_backbone(const float ca[3], int aa, float bb[5][3]);

// Owning device allocation.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  bool owned = true;  // false: a view into another buffer (DevSlab)
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = true;
  }
  hipError_t alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e == hipSuccess) n = count;
    else p = nullptr;
    return e;
  }
  hipError_t ensure(size_t count) { return count <= n ? hipSuccess : alloc(count); }
  void view(T* q, size_t count) {
    release();
    p = q;
    n = count;
    owned = false;
  }
};

// Many buffers in one device allocation: carve() plans each (256-byte
// aligned), alloc() makes the allocation and points every planned DevBuf
// into it. A free of a large allocation waits for the device and unmaps it
// (~0.1-0.2 ms each on MI355X), so one slab instead of dozens of buffers
// saves milliseconds per population. The DevBufs must not outlive the slab;
// one that later needs more room allocates its own memory.
struct DevSlab {
  DevBuf<unsigned char> mem;
  size_t bytes = 0;
  std::vector<std::function<void(unsigned char*)>> bind;
  template <class T>
  void carve(DevBuf<T>& b, size_t count) {
    if (count == 0) count = 1;
    const size_t off = (bytes + 255) & ~static_cast<size_t>(255);
    bytes = off + count * sizeof(T);
    DevBuf<T>* pb = &b;
    bind.push_back([pb, off, count](unsigned char* base) { pb->view(reinterpret_cast<T*>(base + off), count); });
  }
  // spare (nullable): an owned buffer to take instead of allocating when it
  // is large enough (its previous user's work must be complete)
  hipError_t alloc(DevBuf<unsigned char>* spare = nullptr) {
    if (spare && spare->p && spare->owned && spare->n >= bytes) {
      mem.release();
      mem.p = spare->p;
      mem.n = spare->n;
      spare->p = nullptr;
      spare->n = 0;
    } else {
      const hipError_t e = mem.alloc(bytes);
      if (e != hipSuccess) return e;
    }
    for (auto& f : bind) f(mem.p);
    bind.clear();
    return hipSuccess;
  }
};
constexpr size_t kSpareSlabMax = size_t(1) << 30;

// Scratch of a multi-pass top-k (step.hip): chunk lengths per pass and two
// ping-pong buffers; kept across calls, it keeps its buffers.
struct TopkPlan {
  uint32_t k = 0;
  uint32_t rows = 1;
  std::vector<uint32_t> lens;
  DevBuf<double> e[2];
  DevBuf<uint32_t> idx[2];
};

// A SpatialHashGrid laid out on the device: points sorted cell-major with
// cells in Morton order of (key - key_min); within a cell, ascending
// original index (the reference's append order, src/spatial.rs:35-38).
struct GridDev {
  float cell = 0.0f;
  float inv = 0.0f;  // fl(1/cell), src/spatial.rs:26
  int32_t kmin[3] = {0, 0, 0};
  int32_t kmax[3] = {-1, -1, -1};
  int bits = 0;       // Morton bits per axis
  uint64_t ncodes = 0;
  size_t n = 0;
  size_t n_nonempty = 0;
  DevBuf<uint32_t> cell_start;  // ncodes + 1
  DevBuf<uint32_t> sorted_idx;  // sorted slot -> original index
  DevBuf<int32_t> keys;         // 3 * n, per original point
  DevBuf<float> x, y, z;        // original order
};

// View of one grid for kernels (passed by value).
struct GridView {
  const uint32_t* cell_start;
  const uint32_t* sorted_idx;
  int32_t kmin[3];
  int32_t kmax[3];
  int bits;
  float inv;
  uint32_t n;
};

inline GridView make_view(const GridDev& g) {
  GridView v;
  v.cell_start = g.cell_start.p;
  v.sorted_idx = g.sorted_idx.p;
  for (int a = 0; a < 3; ++a) {
    v.kmin[a] = g.kmin[a];
    v.kmax[a] = g.kmax[a];
  }
  v.bits = g.bits;
  v.inv = g.inv;
  v.n = static_cast<uint32_t>(g.n);
  return v;
}

// Antigen atom record for the pair kernel: 32 bytes, read by one scalar
// load per atom. q332 = 332*q; sqe = sqrt(epsilon).
// Work a tile's own wave hands to a workgroup (tiles deep in the antigen):
// the chunks it did not take, as (first antigen atom, atom count), and its
// partial sums as 2^-20 fixed-point integers. Written and read with plain
// stores and loads: the kernel boundary orders them.
constexpr uint32_t kHandoverCap = 128;  // chunks per tile
struct PairHandover {
  uint32_t* rest;     // per tile: chunks handed over (0: none)
  uint32_t* list;     // tiles with chunks handed over (*n_list of them)
  uint32_t* n_list;   // cleared by the sort scan (or a reset) before each launch
  uint2* chunks;      // per tile: kHandoverCap entries
  long long* part;    // 5 x n_slots partial sums (energy, sum |e|, fx, fy, fz)
  uint32_t* status;   // n_slots: the partial sums' sticky status bits (pair_device.h)
  uint32_t n_slots;
};

// Work split of the pair kernels (dnab_pair_tuning); results do not depend
// on it, only the time does.
struct PairTuning {
  uint32_t direct_survivors;  // a tile's wave sums this many antigen atoms itself
  uint32_t direct_chunks;     // ... or this many chunks, then hands the rest on
  uint32_t rest_waves;        // waves per workgroup for handed tiles (4, 8, 16)
};

// Antigen atom record of the pair kernels, 32 bytes: position and 332 q;
// Rmin/2, sqrt(eps), the raw eps (clash pairs are evaluated from it in
// FP64), and a tag holding the AtomType and the hydrophobic flag (ag_tag,
// pair_device.h).
struct AgAtom {
  float x, y, z, q332;
  float rmh, sqe, eps, tag;
};
static_assert(sizeof(AgAtom) == 32, "AgAtom is two float4 loads");

constexpr int kNumAtomTypesDev = 16;  // AtomType values (src/amber.rs:13-32)
// Per-atom and per-candidate energy terms of a breakdown evaluation: LJ,
// Coulomb, hydrophobic, H-bond, EEF1 solvation.
constexpr int kBdTerms = 5;

// Per-antigen constants read by the pair kernels.
struct AgView {
  const AgAtom* atoms;         // grid-sorted
  const float2* raw;           // (q, epsilon), grid-sorted, for the FP64 kernel
  const uint32_t* cell_start;  // Morton cell ranges
  const float4* cell_box;      // per Morton code: its atoms' box (lo, hi)
  int32_t kmin[3];
  int32_t kmax[3];
  int bits;
  float inv;
  float rc;
  float rc2;
  // candidate-atom sort grid around this antigen: 4 x 4 x 4-cell bricks in
  // row-major order, Morton order inside a brick (64 codes per brick)
  float sort_origin[3];
  float sort_inv;
  int32_t sort_dim[3];
  int32_t sort_nb[3];
  uint32_t sort_ncode;
  uint32_t sort_code_base;
  // atoms farther than rc from this box (the antigen atoms' bounding box)
  // have no pair and are left out of the tiles
  float excl_lo[3];
  float excl_hi[3];
  uint32_t n_atoms;
  // residue-level model: antigen Calpha cloud and its grid
  const float* ca_x;
  const float* ca_y;
  const float* ca_z;
  const float* ca_q;
  const float* ca_eps;
  const float* ca_sig;
  const uint8_t* ca_hyd;
  const uint8_t* ca_aa;  // residue types (features); may be null
  GridView ca_grid;
  float center[3];
  float restraint_radius;  // 0 = use the step parameter
};

// Antigen maps for the learned prescreen's features: values on a regular
// grid of spacing h over the antigen's atom box widened by the cutoff and two
// spacings. Energy planes (fp16): the LJ energy of a probe atom of each of the
// kMapLjClasses parameter classes, the Coulomb potential sum 332 q_i / r and
// the hydrophobic sum (-0.5 per hydrophobic antigen atom within 6 A), kept per
// grid cell as one record of kMapRecSlots 16-byte slots, each the 8 corner
// values of one plane (x fastest): potential, hydrophobic, LJ classes 0..5;
// the rare sulfur class (6) in a plane of its own. An atom reads 3 slots; for
// the two carbon classes they share the record's first 64 bytes. kMapRbf
// Gaussian radial densities on the grid of spacing 2h (every second point),
// one 16-byte record per point.
constexpr int kMapLjClasses = 7;
constexpr int kMapPlanes = kMapLjClasses + 2;
constexpr int kMapRecSlots = 8;    // 128 bytes per cell
constexpr int kMapRbf = 8;
constexpr int kMapSulfur = 6;      // the LJ class kept outside the record
// the record slot of plane p (planes: LJ classes 0..6, potential,
// hydrophobic); -1 for the sulfur class
__host__ __device__ constexpr int map_slot(int p) {
  return p == kMapSulfur ? -1 : (p < kMapLjClasses ? 2 + p : p - kMapLjClasses);
}
struct MapDev {
  float h = 0.0f, inv_h = 0.0f;
  float origin[3] = {0, 0, 0};
  int32_t n[3] = {0, 0, 0};    // grid points per axis (cells: n - 1)
  int32_t nc[3] = {0, 0, 0};   // points of the radial-density grid (every second point)
  size_t npts = 0, ncells = 0, ncpts = 0;
  DevBuf<uint16_t> erec;       // [cell][kMapRecSlots][8] fp16
  DevBuf<uint16_t> esul;       // [cell][8] fp16: the sulfur class's LJ
  DevBuf<uint16_t> rbf;        // [coarse point][8] fp16
  DevBuf<uint32_t> hot;        // one bit per cell: some corner has an antigen atom within the
                               // cutoff (else every value the cell's atoms read is zero)
};
struct MapView {
  const uint16_t* erec;
  const uint16_t* esul;
  const uint16_t* rbf;
  const uint32_t* hot;
  float origin[3];
  float inv_h;
  int32_t n[3];
  uint32_t valid;
  int32_t nc[3];
  uint32_t pad;
};

struct KernelTimer {
  std::string name;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double total_ms = 0.0;
  uint64_t launches = 0;
};

}  // namespace dnab

struct dnab_ctx {
  // One reference for the caller (dropped by dnab_ctx_destroy) and one per
  // object made on the context (dropped when the object is destroyed): the
  // stream and the struct outlive every object, in any destroy order.
  int refs = 1;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  int cu_count = 0;
  int wave_size = 64;
  char arch[64] = {0};
  bool timing = false;
  std::vector<dnab::KernelTimer> timers;
  std::vector<hipEvent_t> event_pool;
  // pinned staging for host<->device copies (ctx.hip): uploads fill a ring
  // whose slots are reused behind stage_event; downloads land in dl_stage
  char* stage = nullptr;
  size_t stage_cap = 0, stage_off = 0;
  hipEvent_t stage_event = nullptr;
  char* dl_stage = nullptr;
  size_t dl_cap = 0;
  // the device slab of the last population destroyed on this context (up
  // to kSpareSlabMax bytes), kept for the next population that fits: a
  // large allocation and its free cost ~0.5 and ~0.2 ms
  dnab::DevBuf<unsigned char> spare_slab;
};

struct dnab_grid {
  dnab_ctx* ctx = nullptr;
  dnab::GridDev g;
};

struct dnab_antigen {
  dnab_ctx* ctx = nullptr;
  dnab::GridDev grid;
  float cutoff = 10.0f;
  float restraint_radius = 0.0f;
  float center[3] = {0, 0, 0};
  float bbox_lo[3] = {0, 0, 0};
  float bbox_hi[3] = {0, 0, 0};
  dnab::DevBuf<dnab::AgAtom> atoms;
  dnab::DevBuf<float2> raw;
  // per Morton code of the grid: the box of the cell's atoms (lo, hi; an
  // empty cell +inf, -inf)
  dnab::DevBuf<float4> cell_box;
  size_t n_ca = 0;
  dnab::DevBuf<float> ca_x, ca_y, ca_z, ca_q, ca_eps, ca_sig;
  dnab::DevBuf<uint8_t> ca_hyd, ca_aa;
  bool has_ca_aa = false;
  dnab::GridDev ca_grid;
  dnab::MapDev map;  // prescreen feature maps (built on first use)
};

struct dnab_population {
  dnab_ctx* ctx = nullptr;
  dnab::DevSlab slab;  // the per-atom and per-tile arrays below, in one allocation
  dnab::DevBuf<float> pack_soa;  // the caller's SoA, staged for k_pack_atoms
  size_t n_cand = 0, n_res = 0, n_atoms = 0, n_ag = 0;
  std::vector<dnab_antigen*> antigens;
  std::vector<dnab::AgView> views_host;
  dnab::DevBuf<dnab::AgView> views;
  // candidate atoms: pos = (x, y, z, q), par = (r_min_half, epsilon, hydrophobic,
  // AtomType as a float or -1)
  dnab::DevBuf<float4> pos, par;
  dnab::DevBuf<uint32_t> atom_ag;    // per atom
  dnab::DevBuf<uint32_t> atom_cand;  // per atom
  dnab::DevBuf<uint32_t> atom_res;   // per atom: its residue
  dnab::DevBuf<uint32_t> cand_atom_off, cand_res_off, cand_ag;
  dnab::DevBuf<uint32_t> res_atom_off, res_ca, res_cand;
  uint32_t max_cand_atoms = 0, max_cand_res = 0;  // largest candidate
  std::vector<uint32_t> h_cand_atom_off, h_cand_res_off, h_cand_ag;  // host copies
  // fused trajectory kernel: workgroup shape, candidate grouping (0: consecutive,
  // 1: by measured work, 2: by position, the group's atoms in compact tiles)
  uint32_t traj_waves = 16, traj_cpw = 4, traj_cells = 2, traj_chunks = 2;
  uint32_t traj_balance = 1;
  uint32_t traj_rebalance = 0;   // steps per launch between regroupings (0: one launch)
  uint32_t traj_engine = 0;      // the engine of step_kernels = 2 (0 fused, 1 per-step kernels)
  dnab::DevBuf<uint32_t> traj_work, traj_group_off, traj_group_cand;
  dnab::DevBuf<unsigned long long> traj_key;  // per candidate, for grouping by position
  std::vector<uint32_t> h_group_off, h_group_cand;  // sources of the group-table copies
  std::vector<float> h_sigma;                       // source of the noise-table copy
  dnab::DevBuf<uint8_t> res_aa;         // residue types, when given
  bool has_res_aa = false;
  dnab::DevBuf<float4> res_cg;  // residue-level params (q, eps, sigma, hydrophobic)
  // candidate-atom spatial sort
  uint32_t sort_codes = 0;
  size_t tile_cap = 0;
  dnab::DevBuf<uint32_t> sort_count;     // sort_codes, zero between sorts
  dnab::DevBuf<uint32_t> sort_start;     // sort_codes
  dnab::DevBuf<uint32_t> sort_rank;      // per atom
  dnab::DevBuf<uint32_t> sort_code;      // per atom (kExcluded when left out)
  bool counts_pending = false;           // counts filled by the step kernel
  dnab::DevBuf<uint32_t> perm;           // tile_cap * 64 slots -> atom (or ~0u)
  dnab::DevBuf<float4> spos, spar;       // per slot: the atom's pos and par
  dnab::PairTuning tune{48, 16, 8};
  dnab::DevBuf<uint32_t> ho_rest;         // handover: chunks left per tile
  dnab::DevBuf<uint32_t> ho_list;         // handover: tiles with chunks left, and the count
  bool ho_dirty = false;                  // count left set by the last pair launch
  dnab::DevBuf<uint2> ho_chunks;          // handover: tile_cap * kHandoverCap
  dnab::DevBuf<long long> ho_part;        // handover: 5 * tile_cap * 64
  dnab::DevBuf<uint32_t> ho_status;       // handover: tile_cap * 64 status words
  dnab::DevBuf<uint32_t> ag_tile_start;  // n_ag + 1
  bool sorted = false;
  // evaluation outputs
  dnab::DevBuf<float4> force;     // per atom (fx, fy, fz, sum|e|)
  dnab::DevBuf<double> e_atom;    // per atom energy (FP32 kernel, FP64 sum)
  dnab::DevBuf<double> bd_atom;   // per atom (lj, coulomb, hydrophobic)
  dnab::DevBuf<double> f64_atom;  // per atom FP64 oracle (fx, fy, fz, e)
  dnab::DevBuf<double> energy;    // per candidate
  dnab::DevBuf<float> err_bound;  // per candidate
  dnab::DevBuf<double> breakdown; // per candidate (lj, coulomb, hydrophobic, hbond, solvation)
  dnab::DevBuf<double> disulfide; // per candidate, the binder's S-S term (breakdown runs)
  unsigned last_flags = 0;
  bool evaluated = false;
  // gradient-step state
  dnab::DevBuf<float4> res_delta;  // per residue: this step's displacement
  dnab::DevBuf<float> sigma_table;
  dnab::DevBuf<uint32_t> step_counter;
  dnab::DevBuf<unsigned long long> pair_stats;  // (in cutoff, tested)
  dnab::DevBuf<uint32_t> topk_idx;
  dnab::DevBuf<double> topk_e;
  // score history of the last loop run: hist_rows x hist_k (top k per scoring)
  dnab::DevBuf<uint32_t> hist_idx;
  dnab::DevBuf<double> hist_e;
  uint32_t hist_rows = 0, hist_k = 0;
  dnab::DevBuf<double> score_e;  // fused loop: candidate energies per scoring
  dnab::TopkPlan run_topk;       // the loop's top-k scratch
  // candidate mask: only masked-in candidates are sorted and scored
  dnab::DevBuf<uint8_t> cand_mask;
  bool masked = false;
  // some candidate atom has a non-finite coordinate: the population can be
  // evaluated (its candidates score NaN, as the reference's sums do) but not
  // stepped, refined or screened
  bool nonfinite = false;
  dnab::DevBuf<uint8_t> cand_status;  // per candidate: DNAB_STATUS_* of the last evaluation
  // prescreen
  dnab::DevBuf<uint16_t> feat;   // n_cand x kPrescreenFeatures bf16
  dnab::DevBuf<dnab::MapView> map_views;  // per antigen, uploaded when they change
  std::vector<dnab::MapView> map_views_host;  // what map_views holds
  dnab::DevBuf<uint32_t> sel_scratch;     // select_smallest's stage lists
  dnab::DevBuf<float> scores;    // n_cand
  dnab::DevBuf<double> scores64;
  dnab::DevBuf<uint32_t> keep_idx;
  dnab::DevBuf<double> keep_e;
};

// Error helpers (host).
int dnab_set_error(dnab_ctx* ctx, int code, const char* fmt, ...);

#define DNAB_HIP_CHECK(ctx, expr)                                                     \
  do {                                                                                \
    hipError_t e__ = (expr);                                                          \
    if (e__ != hipSuccess)                                                            \
      return dnab_set_error((ctx), DNAB_ERR_HIP, "%s failed: %s (%s:%d)", #expr,     \
                            hipGetErrorString(e__), __FILE__, __LINE__);              \
  } while (0)

#define DNAB_ALLOC_CHECK(ctx, expr)                                                   \
  do {                                                                                \
    hipError_t e__ = (expr);                                                          \
    if (e__ != hipSuccess)                                                            \
      return dnab_set_error((ctx), DNAB_ERR_ALLOC, "%s failed: %s (%s:%d)", #expr,   \
                            hipGetErrorString(e__), __FILE__, __LINE__);              \
  } while (0)

#define DNAB_RETURN_IF(rc) \
  do {                     \
    int rc__ = (rc);       \
    if (rc__ != DNAB_OK) return rc__; \
  } while (0)

namespace dnab {
// internal evaluate flag: the counting variant of the E+F kernel
constexpr unsigned kEvalCount = 1u << 8;
// internal evaluate flag: a step of the loop (every pair term FP32, as in the
// fused loop); without it clash pairs are FP64 (reported evaluations)
constexpr unsigned kEvalLoop = 1u << 9;
constexpr uint32_t kExcluded = 0xFFFFFFFFu;   // sort code: no pair, zero outputs
constexpr uint32_t kNonFinite = 0xFFFFFFFEu;  // sort code: a NaN coordinate, NaN outputs
// grid.hip
int grid_build_device(dnab_ctx* ctx, GridDev& g, const float* d_x, const float* d_y,
                      const float* d_z, size_t n);
// ctx.hip
dnab_ctx* ctx_retain(dnab_ctx* ctx);   // an object made on ctx keeps it alive
void ctx_release(dnab_ctx* ctx);       // the last release frees the stream and ctx
hipEvent_t timer_begin(dnab_ctx* ctx, const char* name, hipEvent_t* end_out);
void timer_end(dnab_ctx* ctx, const char* name, hipEvent_t start, hipEvent_t end);
void timer_host(dnab_ctx* ctx, const char* name, double ms);  // host-clock phase, when timing
int upload(dnab_ctx* ctx, void* dst, const void* src, size_t bytes);
int download(dnab_ctx* ctx, void* dst, const void* src, size_t bytes);
int download2(dnab_ctx* ctx, void* dst1, const void* src1, size_t b1, void* dst2, const void* src2,
              size_t b2);
// step.hip: multi-pass top-k (ascending, NaN last, ties by index)
int topk_plan(dnab_ctx* ctx, TopkPlan& pl, uint32_t n, uint32_t k, uint32_t rows = 1);
// first_idx (nullable): the entries' indices (else their positions), for
// the output and for breaking ties
void topk_launch(dnab_ctx* ctx, TopkPlan& pl, const double* energy, double* final_e,
                 uint32_t* final_idx, const uint32_t* counter = nullptr, uint32_t row_div = 1,
                 uint32_t row_fixed = 0, const uint32_t* first_idx = nullptr);
// step.hip: the min(m, n) smallest scores (ties by index) as indices in
// index order, 1 <= m <= 2048
int select_smallest(dnab_ctx* ctx, const float* scores, uint32_t n, uint32_t m, uint32_t* out_idx,
                    DevBuf<uint32_t>& scratch);
// pair.hip
int population_sort(dnab_ctx* ctx, dnab_population* pop);
int population_evaluate(dnab_ctx* ctx, dnab_population* pop, unsigned flags);
int population_pair(dnab_ctx* ctx, dnab_population* pop, unsigned flags);
int discard_pending_counts(dnab_ctx* ctx, dnab_population* pop);
int population_count(dnab_ctx* ctx, dnab_population* pop);
int population_reduce_energies(dnab_ctx* ctx, dnab_population* pop, unsigned flags);
// DNAB_ERR_INVALID naming `what` when the population has a non-finite coordinate
int require_finite(dnab_ctx* ctx, const dnab_population* pop, const char* what);
// parameters of an atom cloud: finite charge, r_min_half >= 0 and epsilon >= 0
bool atom_params_ok(const dnab_atom_cloud* c);
// noalloc.hip: the no-alloc entry's scratch address on `device` (current)
int noalloc_init_device(int device);
// maps.hip: the prescreen's antigen maps and per-candidate features
constexpr int kPrescreenFeatures = 32;
int antigen_build_maps(dnab_ctx* ctx, dnab_antigen* ag, float spacing);
// features (bf16 rows of kPrescreenFeatures, nullable) and the map energy of
// every candidate (nullable); builds missing antigen maps first
int population_map_features(dnab_ctx* ctx, dnab_population* pop, uint16_t* feat, float* map_e);
// traj.hip: final-state evaluation (energies, forces; clash pairs in FP64) of
// the n candidates listed in d_idx (device), one workgroup each
// (score_scratch: n_cand doubles the kernel may write)
int trajectory_eval_subset(dnab_ctx* ctx, dnab_population* pop, const uint32_t* d_idx, uint32_t n,
                           double* score_scratch);
}  // namespace dnab
