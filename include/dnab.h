/* dnab — MI355X (gfx950) batched all-atom binder scoring: C ABI.
 *
 * One shared library (libdnab.so) exposes the hot path of the reference's
 * all-atom engine as plain C entry points: plain pointers and sizes, status
 * codes, every output buffer owned by the caller. Each entry point names the
 * reference interface it replaces (reference file:line).
 *
 * Conventions
 *  - Every function returns DNAB_OK (0) or a DNAB_ERR_* code; no exceptions or
 *    aborts cross the boundary. dnab_ctx_last_error() gives the message.
 *  - A context owns one HIP stream on one device. A context is not
 *    thread-safe: use one context per host thread.
 *  - Host-pointer entry points (score_batch, energy/force/grid calls) copy in,
 *    run, and copy back before returning. Device-resident objects
 *    (dnab_antigen, dnab_population) keep their data in HBM between calls.
 *  - Units: Angstrom, kcal/mol, elementary charge.
 */
#ifndef DNAB_H_
#define DNAB_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DNAB_OK 0
#define DNAB_ERR_INVALID 1     /* bad argument */
#define DNAB_ERR_NO_DEVICE 2   /* no usable GPU (reference: try_init -> None) */
#define DNAB_ERR_HIP 3         /* a HIP runtime call failed */
#define DNAB_ERR_ALLOC 4       /* device or host allocation failed */
#define DNAB_ERR_CAPACITY 5    /* caller buffer too small; required size reported */
#define DNAB_ERR_UNSUPPORTED 6 /* valid input outside what this build handles */

typedef struct dnab_ctx dnab_ctx;
typedef struct dnab_grid dnab_grid;
typedef struct dnab_antigen dnab_antigen;
typedef struct dnab_population dnab_population;

/* ---- data layouts ------------------------------------------------------- */

/* 32-byte AoS atom, field-for-field the reference GpuAtom
 * (src/gpu.rs:18-29, shaders/energy.wgsl:21-30). */
typedef struct dnab_gpu_atom {
  float x, y, z, q, r_min_half, epsilon, hydrophobic, _pad;
} dnab_gpu_atom;

/* The layout is the ABI: 32 bytes, fields at 4-byte steps in the order of
 * the reference's #[repr(C)] struct. */
#ifdef __cplusplus
#define DNAB_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define DNAB_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
DNAB_STATIC_ASSERT(sizeof(dnab_gpu_atom) == 32, "dnab_gpu_atom must be 32 bytes");
DNAB_STATIC_ASSERT(offsetof(dnab_gpu_atom, x) == 0 && offsetof(dnab_gpu_atom, y) == 4 &&
                       offsetof(dnab_gpu_atom, z) == 8 && offsetof(dnab_gpu_atom, q) == 12 &&
                       offsetof(dnab_gpu_atom, r_min_half) == 16 &&
                       offsetof(dnab_gpu_atom, epsilon) == 20 &&
                       offsetof(dnab_gpu_atom, hydrophobic) == 24 &&
                       offsetof(dnab_gpu_atom, _pad) == 28,
                   "dnab_gpu_atom field offsets follow src/gpu.rs:18-29");

/* Read-only SoA view of an AtomCloud (src/allatom.rs:16-31). atom_type and
 * residue_idx are not used by the LJ/Coulomb/hydrophobic terms and may be
 * NULL. */
typedef struct dnab_atom_cloud {
  size_t n;
  const float* x;
  const float* y;
  const float* z;
  const float* charge;
  const float* r_min_half;
  const float* epsilon;
  const uint8_t* hydrophobic; /* 0 or 1 */
  const uint8_t* atom_type;
  const uint32_t* residue_idx;
} dnab_atom_cloud;

/* Read-only SoA view of a ResidueCloud (src/atom.rs:185-195). */
typedef struct dnab_residue_cloud {
  size_t n;
  const float* x;
  const float* y;
  const float* z;
  const float* charge;
  const float* epsilon;
  const float* sigma;
  const uint8_t* hydrophobic;
} dnab_residue_cloud;

/* AtomEnergyBreakdown (src/energy.rs:238-250). hbond and solvation need the
 * atom types of both clouds (atom_type); without them they are 0. */
typedef struct dnab_energy_breakdown {
  double lj, coulomb, hydrophobic, hbond, solvation;
} dnab_energy_breakdown;
DNAB_STATIC_ASSERT(sizeof(dnab_energy_breakdown) == 40, "five doubles");

/* Output buffers for dnab_build_protein (caller-allocated). Per-atom arrays
 * hold atom_capacity entries; per-residue arrays n_res (res_start n_res+1).
 * Any pointer except x/y/z may be NULL. */
typedef struct dnab_protein_out {
  size_t atom_capacity;
  float *x, *y, *z, *charge, *r_min_half, *epsilon;
  uint8_t *hydrophobic, *atom_type;
  uint32_t* residue_idx;
  uint32_t* res_start;
  uint32_t* ca_atom_idx;
  uint8_t* amino_acid;
} dnab_protein_out;

/* ---- context (replaces GpuContext, src/gpu.rs:46-57) ---------------------
 * dnab_ctx_create replaces GpuContext::try_init (src/gpu.rs:55-57, whose
 * async body src/gpu.rs:59-155 returns None when no adapter exists): a
 * non-zero status (DNAB_ERR_NO_DEVICE) takes the place of `None`. */
int dnab_ctx_create(int device, dnab_ctx** out);
/* Drops the caller's hold on the context. Objects made on it (grids,
 * antigens, populations, networks, datasets) keep it alive until they are
 * destroyed, so objects and context may be destroyed in any order (as Rust
 * drops or a garbage collector's finalizers order them); the context must
 * not be passed to any other call after this one. */
void dnab_ctx_destroy(dnab_ctx* ctx);
const char* dnab_ctx_last_error(const dnab_ctx* ctx);
const char* dnab_status_string(int status);
int dnab_ctx_synchronize(dnab_ctx* ctx);
int dnab_device_count(int* count);
/* Device properties of the context's GPU: CU count, wave size, arch name. */
int dnab_ctx_device_info(const dnab_ctx* ctx, int* cu_count, int* wave_size, char* arch,
                         size_t arch_len);
/* Kernel timing with HIP events on the context stream (no reference
 * counterpart; the reference times whole runs with Instant, src/main.rs:153).
 * When enabled, every launch of the pair kernel is bracketed by events;
 * dnab_ctx_kernel_time returns the summed milliseconds and launch count. */
int dnab_ctx_set_timing(dnab_ctx* ctx, int enable);
int dnab_ctx_kernel_time(dnab_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches);

/* ---- structure building (host only, no GPU needed) ------------------------
 * protein_from_ca_trace (src/allatom.rs:523-541) with an optional rotamer
 * index per residue (NULL = most probable rotamer, as the reference). */
int dnab_build_protein(const float* ca_x, const float* ca_y, const float* ca_z,
                       const uint8_t* amino_acid, const uint8_t* rotamer_idx, size_t n_res,
                       dnab_protein_out* out, size_t* n_atoms_out);
/* rebuild_side_chain (src/allatom.rs:244-271) of one residue: xyz holds the
 * residue's atoms in topology order (3 floats each); the side chain is placed
 * from its backbone atoms (N, CA, C, O and, except Gly, CB) with the chi of
 * rotamer `rotamer` (0 for types without rotamers) and written after them. */
int dnab_rebuild_side_chain(uint8_t amino_acid, uint8_t rotamer, float* xyz, size_t n_atoms);
/* ---- PDB text I/O (host only) ---------------------------------------------
 * parse_pdb (src/pdb.rs:19-57) on a PDB file's text: one residue per CA atom
 * of an ATOM/HETATM record, in file order; x, y, z from columns 31-54 (0
 * where a field is missing or not a number), the residue from columns 18-20
 * (from_three_letter: HID/HIE/HIP are His, unknown names Gly). Writes up to
 * `capacity` residues; n_out = the count. No CA atom: DNAB_ERR_INVALID (the
 * reference's EmptyInput); more than capacity: DNAB_ERR_CAPACITY. */
int dnab_parse_pdb(const char* text, size_t len, float* x, float* y, float* z,
                   uint8_t* amino_acid, size_t capacity, size_t* n_out);
/* write_pdb (src/pdb.rs:94-119): CA-only ATOM records of a residue cloud and
 * END. The writers fill `out` (capacity bytes, NUL-terminated) and set
 * n_chars to the text length; out = NULL only queries the length. */
int dnab_write_pdb_ca(const float* x, const float* y, const float* z, const uint8_t* amino_acid,
                      size_t n, char chain_id, char* out, size_t capacity, size_t* n_chars);
/* One all-atom chain: atoms in residue order (res_start, n_res + 1), each
 * residue's atoms in topology order. */
typedef struct dnab_pdb_chain {
  const float *x, *y, *z;
  const uint32_t* res_start;
  const uint8_t* amino_acid;
  size_t n_res;
  char chain_id;
} dnab_pdb_chain;
/* write_pdb_allatom (src/pdb.rs:121-155, one chain, ter = 0) and
 * write_pdb_multi (src/pdb.rs:157-191, ter = 1: a TER record after each
 * chain): ATOM records with the topology's atom names, serials continuing
 * across chains, then END. */
int dnab_write_pdb_chains(const dnab_pdb_chain* chains, size_t n_chains, int ter, char* out,
                          size_t capacity, size_t* n_chars);
/* One designed candidate's summary (CandidateReport, src/report.rs:8-20).
 * Strings are NUL-terminated; light_sequence = NULL for a single chain;
 * clashes / rama_outliers < 0 for none. */
typedef struct dnab_candidate_report {
  size_t rank;
  const char* sequence;
  const char* light_sequence;
  float energy, entropy_penalty, dg_corrected;
  int32_t net_charge;
  int aggregation_risk;
  size_t n_interface;
  int64_t clashes;
  int64_t rama_outliers;
} dnab_candidate_report;
#define DNAB_REPORT_FASTA 0
#define DNAB_REPORT_JSON 1
#define DNAB_REPORT_CSV 2
/* write_fasta / write_json / write_csv (src/report.rs:41-113) of n reports
 * in order, into `out` as the PDB writers do (out = NULL queries the
 * length). */
int dnab_write_report(const dnab_candidate_report* reports, size_t n, int format, char* out,
                      size_t capacity, size_t* n_chars);
/* place_atom (src/rotamer.rs:198-232), NERF. */
int dnab_place_atom(const float a[3], const float b[3], const float c[3], float bond_len,
                    float angle_deg, float dihedral_deg, float out[3]);
/* Heavy-atom count of one residue type (src/amber.rs:357-378); -1 if invalid. */
int dnab_residue_atom_count(uint8_t amino_acid);
/* Residue-level parameters (src/atom.rs:55-104). */
int dnab_residue_cg_params(uint8_t amino_acid, float* charge, float* epsilon, float* sigma,
                           uint8_t* hydrophobic);
/* One topology slot (src/amber.rs:126-351) with its AMBER_NB parameters
 * (src/amber.rs:46-63). */
int dnab_ff_table(uint8_t amino_acid, int slot, uint8_t* atom_type, float* charge,
                  uint8_t* hydrophobic, float* r_min_half, float* epsilon);
/* Per AtomType (0..15): EEF1 (dg_ref, volume, lambda), src/amber.rs:406-431,
 * and the H-bond donor / acceptor classes, src/amber.rs:72-84. */
int dnab_ff_type_params(uint8_t atom_type, float* dg_ref, float* volume, float* lambda,
                        uint8_t* hbond_donor, uint8_t* hbond_acceptor);
/* Synthetic Calpha traces (no reference counterpart except the ring layout,
 * which follows random_allatom_antibody, src/diffusion.rs:264-277). */
int dnab_synth_globule_trace(size_t n_res, float radius, uint64_t seed, float* x, float* y,
                             float* z, uint8_t* amino_acid, uint8_t* rotamer_idx);
int dnab_synth_binder_ring(size_t n_res, const float center[3], float ring_radius,
                           uint64_t seed, float* x, float* y, float* z, uint8_t* amino_acid);
int dnab_synth_binder_shell(size_t n_res, const float center[3], float shell_radius,
                            float min_radius, uint64_t seed, float* x, float* y, float* z,
                            uint8_t* amino_acid);
/* n_cand shell binders of n_res residues (candidate k: seed seed_base + k),
 * built into one population's arrays with nthreads host threads: per-atom
 * arrays (atom_capacity), amino_acid / ca_atom_idx / res_start (global
 * indices, n_cand * n_res (+1) entries), cand_res_off (n_cand + 1). With out
 * NULL only *n_atoms_out is set (sizing). */
int dnab_synth_shell_population(size_t n_cand, size_t n_res, const float center[3],
                                float shell_radius, float min_radius, uint64_t seed_base,
                                int nthreads, dnab_protein_out* out, uint32_t* cand_res_off,
                                size_t* n_atoms_out);

/* ---- spatial hash grid (replaces SpatialHashGrid, src/spatial.rs:12-68) ---
 * dnab_grid_create  = SpatialHashGrid::new   (src/spatial.rs:23-28)
 * dnab_grid_build   = SpatialHashGrid::build (src/spatial.rs:31-39)
 * dnab_grid_query   = query_neighbors        (src/spatial.rs:44-58), batched:
 *   probe p's list is indices[offsets[p] .. offsets[p+1]) in the reference's
 *   emission order (dx, dy, dz nested, ascending index within a cell).
 * dnab_grid_cell_of = cell_of                (src/spatial.rs:61-67), keys as
 *   (kx, ky, kz) triples.
 * If `indices` is NULL or `capacity` is too small, the offsets are still
 * written, *n_total is set and DNAB_ERR_CAPACITY is returned. */
int dnab_grid_create(dnab_ctx* ctx, float cell_size, dnab_grid** out);
int dnab_grid_build(dnab_ctx* ctx, dnab_grid* grid, const float* x, const float* y,
                    const float* z, size_t n);
int dnab_grid_query(dnab_ctx* ctx, const dnab_grid* grid, const float* px, const float* py,
                    const float* pz, size_t n_probe, uint32_t* offsets, uint32_t* indices,
                    size_t capacity, size_t* n_total);
int dnab_grid_cell_of(dnab_ctx* ctx, const dnab_grid* grid, const float* x, const float* y,
                      const float* z, size_t n, int32_t* keys);
/* Grid introspection: points, non-empty cells, key bounding box. */
int dnab_grid_info(const dnab_grid* grid, size_t* n_points, size_t* n_cells_nonempty,
                   int32_t key_min[3], int32_t key_max[3]);
void dnab_grid_destroy(dnab_grid* grid);

/* ---- residue-level forces -------------------------------------------------
 * compute_forces_with_grid (src/energy.rs:322-398): `grid` must have been
 * built over antigen->x/y/z. Outputs are overwritten (the reference zeroes
 * them first, src/energy.rs:332-334); length antibody->n. */
int dnab_cg_forces_with_grid(dnab_ctx* ctx, const dnab_residue_cloud* antigen,
                             const dnab_grid* grid, const dnab_residue_cloud* antibody,
                             float* fx, float* fy, float* fz);
/* compute_forces (src/energy.rs:154-234), brute force over all antigen points. */
int dnab_cg_forces(dnab_ctx* ctx, const dnab_residue_cloud* antigen,
                   const dnab_residue_cloud* antibody, float* fx, float* fy, float* fz);

/* ---- all-atom energies and forces (host buffers) ---------------------------
 * interaction_energy_atoms_breakdown (src/energy.rs:261-314): LJ + Coulomb +
 * hydrophobic + H-bond with the reference pair predicate 0.25 <= r^2 <=
 * cutoff^2 (cutoff 10 in the reference, src/energy.rs:16), and EEF1 solvation
 * over 0.25 <= r^2 < 81 (src/solvation.rs:52-84, counted within the cutoff).
 * H-bond and solvation need both clouds' atom_type. */
int dnab_interaction_energy_atoms_breakdown(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                          const dnab_atom_cloud* antibody, float cutoff,
                          dnab_energy_breakdown* out);
/* interaction_energy_atoms (src/energy.rs:256-258), the total of those terms. */
int dnab_interaction_energy_atoms(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                                  const dnab_atom_cloud* antibody, float cutoff,
                                  double* energy);
/* allatom_energy (src/diffusion.rs:281-283), the score run_allatom ranks its
 * refined candidates by (src/diffusion.rs:1092): interaction_energy_atoms plus
 * the antibody's disulfide_energy (src/allatom.rs:155-180) over its Cys SG
 * atoms (slot 5 of each Cys residue). The antibody is given as residues:
 * res_atom_off (n_res + 1, spanning 0..antibody->n), res_ca_atom (n_res) and
 * amino_acid (n_res). disulfide (nullable) receives the disulfide term. */
int dnab_allatom_energy(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                        const dnab_atom_cloud* antibody, const uint32_t* res_atom_off,
                        const uint32_t* res_ca_atom, const uint8_t* amino_acid, size_t n_res,
                        float cutoff, double* energy, double* disulfide);
/* All-atom AMBER forces on the antibody atoms (no reference counterpart: the
 * reference has residue-level forces only, src/energy.rs:154-234). Force on
 * atom j is sum_i [12 eps (s^12 - s^6)/r^2 + 332 q_i q_j / r^3 + f_h] (p_j -
 * p_i), s = R_ij/r, the exact negative gradient of the LJ and Coulomb terms;
 * f_h = -0.1/r for hydrophobic pairs with r^2 < 36, the reference's force
 * convention for the hydrophobic term (src/energy.rs:382-386). */
int dnab_atom_forces(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                     const dnab_atom_cloud* antibody, float cutoff, float* fx, float* fy,
                     float* fz, double* energy);

/* score_batch (src/gpu.rs:166-267): candidates are n_cand = n_atoms / n_ab
 * uniform-length blocks of GpuAtom; energies (LJ + Coulomb + hydrophobic,
 * cutoff 10 as shaders/energy.wgsl:16) are written to energies_out[0..n_cand).
 * n_ab == 0 or no candidates gives *n_cand = 0 and DNAB_OK, as the
 * reference's empty Vec (src/gpu.rs:172-174). */
int dnab_score_batch(dnab_ctx* ctx, const dnab_atom_cloud* antigen,
                     const dnab_gpu_atom* candidates, size_t n_atoms, size_t n_ab,
                     float* energies_out, size_t capacity, size_t* n_cand);
/* dnab_score_batch without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). The antigen is a
 * dnab_antigen (its cutoff applies; dnab_score_batch's is 10 A with 10 A
 * cells); candidates_dev (n_atoms GpuAtom records) and energies_dev
 * (capacity floats) are device memory owned by the caller. The work is
 * queued on the context's stream and the call returns without waiting:
 * energies_dev is ready when the stream has completed (dnab_ctx_synchronize).
 * Energies equal dnab_score_batch's on the same atoms and antigen; a NaN
 * coordinate gives NaN, a saturated sum +inf. The per-candidate sums live
 * in one fixed device array per GPU: at most 65,536 candidates per call
 * (DNAB_ERR_UNSUPPORTED beyond); calls that overlap on one GPU (from two
 * contexts or threads) race on that array. */
int dnab_score_batch_noalloc(dnab_ctx* ctx, const dnab_antigen* antigen,
                             const dnab_gpu_atom* candidates_dev, size_t n_atoms, size_t n_ab,
                             float* energies_dev, size_t capacity, size_t* n_cand);

/* ---- device-resident antigen ---------------------------------------------
 * The antigen's atoms are sorted into a Morton-ordered, cell-major SoA over
 * the SpatialHashGrid cells of `cell_size` (src/spatial.rs semantics) and
 * kept in HBM. `center` is the restraint centre of the gradient step (the
 * antigen's Calpha centre of mass, src/diffusion.rs:991). `ca` is the
 * antigen Calpha ResidueCloud used by the residue-level force model
 * (src/diffusion.rs:994-1004); it may be empty when that model is unused. */
typedef struct dnab_antigen_desc {
  dnab_atom_cloud atoms;
  dnab_residue_cloud ca;
  float center[3];
  float cell_size; /* grid cell edge; the reference uses 10 (src/diffusion.rs:1003) */
  float cutoff;    /* pair cutoff; the reference uses 10 (src/energy.rs:16) */
  float restraint_radius; /* gradient-step restraint radius around `center`;
                             0 = the step parameter restraint_radius */
  const uint8_t* ca_amino_acid; /* residue types of `ca` (AminoAcid order,
                                   src/atom.rs:4-25); needed by the prescreen
                                   features, may be NULL otherwise */
} dnab_antigen_desc;
int dnab_antigen_create(dnab_ctx* ctx, const dnab_antigen_desc* desc, dnab_antigen** out);
int dnab_antigen_info(const dnab_antigen* ag, size_t* n_atoms, size_t* n_cells_nonempty,
                      float* cutoff, float* cell_size);
void dnab_antigen_destroy(dnab_antigen* ag);

/* ---- device-resident ragged population ------------------------------------
 * No reference counterpart: score_batch needs one atom count for every
 * candidate (src/gpu.rs:166-174). Candidates here have their own residue and
 * atom ranges and each is bound to one antigen. Atoms of one candidate are
 * contiguous; residues of one candidate are contiguous. */
typedef struct dnab_population_desc {
  size_t n_cand, n_res;
  const uint32_t* cand_antigen; /* n_cand, index into the antigens array */
  const uint32_t* cand_res_off; /* n_cand + 1, residue ranges */
  const uint32_t* res_atom_off; /* n_res + 1, atom ranges (global atom index) */
  const uint32_t* res_ca_atom;  /* n_res, global atom index of each Calpha */
  const uint8_t* res_amino_acid;/* n_res, residue types (residue-level model) */
  dnab_atom_cloud atoms;        /* all candidate atoms, n = res_atom_off[n_res] */
} dnab_population_desc;
/* Atom parameters must be finite with r_min_half >= 0 and epsilon >= 0
 * (DNAB_ERR_INVALID otherwise; the same holds for dnab_antigen_create, whose
 * coordinates must also be finite). A candidate coordinate may be NaN or
 * infinite, as the reference's AtomCloud allows, and such a population can be
 * evaluated with the reference's results: an atom with a NaN coordinate has
 * NaN distances, whose pairs the reference's skip test keeps
 * (src/energy.rs:284), so its outputs and its candidate's energy are NaN
 * (dnab_population_read_status reports it); an atom with an infinite and no
 * NaN coordinate has r^2 = +inf, which the test skips, so it has no pair.
 * Every call that steps or screens such a population returns
 * DNAB_ERR_INVALID. */
int dnab_population_create(dnab_ctx* ctx, dnab_antigen* const* antigens, size_t n_antigens,
                           const dnab_population_desc* desc, dnab_population** out);
void dnab_population_destroy(dnab_population* pop);
/* Affinity maturation of a batch of antibodies against one device antigen
 * (its cutoff applies):
 * affinity_maturation_allatom (src/diffusion.rs:924-961), called by the
 * reference once per result (src/main.rs:341-347), here for all of them in one
 * call. Per antibody, for each mutable position r in order, the amino acids
 * are tried in ALL_AA order (src/atom.rs:29-50), each mutant built as
 * mutate_residue does (ideal backbone around r's Calpha, the type's first
 * rotamer, src/allatom.rs:299-384); the first whose allatom_energy is strictly
 * below the best so far is kept, and a non-improving trial is undone by
 * rebuilding the old type on an ideal backbone with its rotamer.
 *   ab: the antibodies as a population description (n_cand antibodies,
 *       cand_antigen ignored); res_amino_acid and res_ca_atom required;
 *       atom_type needed for the H-bond and solvation terms.
 *   res_rotamer (nullable, n_res): each residue's rotamer index, whose chi the
 *       undo rebuild uses (0, the most probable, when NULL).
 *   mutable_off (n_cand + 1), mutable_res: per antibody, strictly ascending
 *       residue indices within it (the reference's interface filter,
 *       src/main.rs:342-344).
 * Outputs: out (nullable: then only sizes and per-residue outputs) receives
 * the final antibodies concatenated, residues in input order, res_start and
 * ca_atom_idx as global atom indices; out_rotamer and out_rebuilt (nullable,
 * n_res) the rotamer index and whether the residue was rebuilt; out_energy
 * (n_cand) the returned score of the reference, the best allatom_energy
 * accepted (FP64); n_atoms_out the final atom count. Scores are taken from one
 * scoring batch of single residues: the trial scores equal full re-scoring to
 * FP64 rounding. */
int dnab_affinity_maturation(dnab_ctx* ctx, dnab_antigen* antigen,
                             const dnab_population_desc* ab, const uint8_t* res_rotamer,
                             const uint32_t* mutable_off, const uint32_t* mutable_res,
                             dnab_protein_out* out, uint8_t* out_rotamer, uint8_t* out_rebuilt,
                             double* out_energy, size_t* n_atoms_out);

#define DNAB_EVAL_FORCES 1u    /* also compute per-atom forces */
#define DNAB_EVAL_BREAKDOWN 2u /* energy-only, the reference's five terms kept
                                  separately: LJ, Coulomb, hydrophobic, and, when
                                  both sides carry atom types, H-bond and EEF1
                                  solvation (src/energy.rs:261-314,
                                  src/solvation.rs:52-84); the candidate energy
                                  is their total; plus the binder's disulfide
                                  term (src/allatom.rs:155-180) when residue
                                  types are given */
#define DNAB_EVAL_FP64 4u      /* all-FP64 oracle kernel instead of FP32 */
/* One evaluation of every candidate against its antigen: per-candidate
 * energies (FP32 pair terms, TwoSum-compensated per atom, FP64 total) and,
 * with DNAB_EVAL_FORCES, per-atom forces. Results stay on the device. */
int dnab_population_evaluate(dnab_ctx* ctx, dnab_population* pop, unsigned flags);
/* Per-candidate energies of the last evaluation; err_bound (nullable) is the
 * a-priori bound on |E - E_fp64| from the FP32 pair terms. */
int dnab_population_read_energies(dnab_ctx* ctx, const dnab_population* pop, double* energy,
                                  float* err_bound);
/* Per-candidate status of the last evaluation (no reference counterpart: the
 * reference's f32 sums carry NaN and overflow in the value itself).
 * DNAB_STATUS_NAN: the energy is NaN (a non-finite coordinate).
 * DNAB_STATUS_SATURATED: an atom's energy or force sum reached the exact
 * accumulator's bound, +-2^39 kcal/mol (or kcal/mol/A) (~5.5e11; only
 * reachable through dozens of sub-Angstrom clashes on one atom); the value is
 * reported as +-inf. 0: the results are finite and within the tolerance. */
#define DNAB_STATUS_NAN 1u
#define DNAB_STATUS_SATURATED 2u
int dnab_population_read_status(dnab_ctx* ctx, dnab_population* pop, uint8_t* status);
/* Per candidate, 6 values of the last DNAB_EVAL_BREAKDOWN run: LJ, Coulomb,
 * hydrophobic, H-bond, solvation (their sum is the candidate energy,
 * interaction_energy_atoms) and the binder's disulfide term; allatom_energy
 * (src/diffusion.rs:281-283) is the sum of all six. */
int dnab_population_read_terms(dnab_ctx* ctx, const dnab_population* pop, double* terms);
/* Per-candidate LJ, Coulomb, hydrophobic of the last DNAB_EVAL_BREAKDOWN run. */
int dnab_population_read_breakdown(dnab_ctx* ctx, const dnab_population* pop, double* lj,
                                   double* coulomb, double* hydrophobic);
/* Per-atom forces of the last evaluation (FP64 runs are rounded to float;
 * use dnab_population_read_forces_f64 for the oracle values). */
int dnab_population_read_forces(dnab_ctx* ctx, const dnab_population* pop, float* fx,
                                float* fy, float* fz);
int dnab_population_read_forces_f64(dnab_ctx* ctx, const dnab_population* pop, double* fx,
                                    double* fy, double* fz);
int dnab_population_read_coords(dnab_ctx* ctx, const dnab_population* pop, float* x,
                                float* y, float* z);
int dnab_population_write_coords(dnab_ctx* ctx, dnab_population* pop, const float* x,
                                 const float* y, const float* z);
/* Pair statistics at the current coordinates: pairs inside the FP32
 * predicate 0.25 <= r^2 <= cutoff^2, and pairs the kernel tested. */
/* Work split of the pair kernels and its autotuner (no reference
 * counterpart; north_star (b)). Per-chunk sums are added as exact fixed-point
 * integers, so every setting gives bit-identical results: only the time
 * changes. */
typedef struct dnab_pair_tuning {
  uint32_t direct_survivors; /* antigen atoms a tile's own wave sums (>= 1) */
  uint32_t direct_chunks;    /* ... or chunks (>= 1); the rest goes to a workgroup */
  uint32_t rest_waves;       /* waves per workgroup for handed-on tiles: 4, 8 or 16 */
} dnab_pair_tuning;
int dnab_population_get_tuning(const dnab_population* pop, dnab_pair_tuning* out);
int dnab_population_set_tuning(dnab_ctx* ctx, dnab_population* pop, const dnab_pair_tuning* t);
/* Times `reps` energy+force evaluations of the population's current state
 * (HIP events on the context stream) for each setting of a fixed grid
 * (survivors 24/48/96/192 x chunks 8/16/32 x waves 4/8/16), keeps the
 * fastest and returns it with its mean time per evaluation. */
int dnab_population_autotune(dnab_ctx* ctx, dnab_population* pop, uint32_t reps,
                             dnab_pair_tuning* best, double* best_ms);
int dnab_population_pair_count(dnab_ctx* ctx, dnab_population* pop, uint64_t* in_cutoff,
                               uint64_t* tested);
int dnab_population_info(const dnab_population* pop, size_t* n_cand, size_t* n_res,
                         size_t* n_atoms);
/* Pair statistics accumulated by runs with count_pairs set (zero when the
 * population is created; reset by dnab_population_reset_pair_stats). */
int dnab_population_read_pair_stats(dnab_ctx* ctx, const dnab_population* pop,
                                    uint64_t* in_cutoff, uint64_t* tested);
int dnab_population_reset_pair_stats(dnab_ctx* ctx, dnab_population* pop);

/* ---- gradient step and population loop ------------------------------------
 * The step restates allatom_diffusion_step lines src/diffusion.rs:300-339:
 * per residue, force F (all-atom: sum of the residue's atom forces; or the
 * residue-level grid force of src/energy.rs:322-398 on the Calpha), plus the
 * restraint -k (d - R0)/d (ca - center) (src/diffusion.rs:317-327), then
 * ca' = ca + clamp(step*F, +-max_disp) + U(-s, s) with s = noise_base*sqrt(T)
 * when with_noise (src/diffusion.rs:315, 329-337), applied as a rigid
 * translation of the residue's atoms (set_ca_pos, src/allatom.rs:289-293).
 * T(iter) = t_hot (t_cold/t_hot)^((iter + schedule_offset)/schedule_total)
 * (src/diffusion.rs:102-104). Monte-Carlo moves (src/diffusion.rs:341-418)
 * are not part of this path. */
#define DNAB_FORCE_ALLATOM 0 /* all-atom AMBER forces summed per residue */
#define DNAB_FORCE_CG 1      /* residue-level Calpha grid forces (reference step) */
typedef struct dnab_langevin_params {
  uint32_t n_steps;         /* gradient steps (reference GPU phase: 200) */
  uint32_t score_every;     /* score every k steps (0 = only at the end) */
  uint32_t top_k;           /* survivors kept per scoring (reference: 64) */
  uint64_t seed;
  int32_t with_noise;       /* reference GPU phase: 0, CPU phase: 1 */
  int32_t force_model;      /* DNAB_FORCE_ALLATOM or DNAB_FORCE_CG */
  int32_t use_graph;        /* replay the step loop from a captured hipGraph */
  float step;               /* 0.08  (src/diffusion.rs:66) */
  float max_disp;           /* 2.0   (src/diffusion.rs:82) */
  float noise_base;         /* 0.25  (src/diffusion.rs:67) */
  float t_hot, t_cold;      /* 3.0, 0.02 (src/diffusion.rs:68-69) */
  uint32_t schedule_offset; /* iter offset into the temperature schedule */
  uint32_t schedule_total;  /* schedule length (reference: GPU_STEPS + cpu_iterations) */
  float restraint_k;        /* 0.02  (src/diffusion.rs:83) */
  float restraint_radius;   /* 20.0  (INIT_RADIUS, src/diffusion.rs:81) */
  int32_t count_pairs;      /* accumulate pair statistics over the run (same
                               results, slower); read with
                               dnab_population_read_pair_stats */
  int32_t step_kernels;     /* 0: run the all-atom loop as one fused kernel
                               (a workgroup per group of candidates, every
                               step in LDS) when a candidate fits one
                               workgroup (up to ~1,500 atoms);
                               1: one set of population kernels per step
                               (hipGraph replay with use_graph);
                               2: the engine the population's trajectory
                               tuning names (dnab_traj_tuning.engine, chosen
                               by dnab_population_autotune_trajectory).
                               Results are bit-identical; pair statistics
                               agree on pairs in the cutoff, not on pairs
                               tested. */
} dnab_langevin_params;
int dnab_langevin_default_params(dnab_langevin_params* p);

/* ---- refinement with Monte-Carlo moves (allatom_diffusion_step) ----------
 * Rates of the Metropolis moves of allatom_diffusion_step
 * (src/diffusion.rs:353-415); the defaults are the reference's. */
#define DNAB_MC_SIDECHAIN 1u /* rotamer moves and amino-acid mutations */
#define DNAB_MC_BACKBONE 2u  /* phi/psi torsion moves of the downstream chain */
typedef struct dnab_mc_params {
  float mutation_p;      /* 0.08 (src/diffusion.rs:41); x3 at the interface, x0.4 elsewhere */
  float rotamer_move_p;  /* 0.12 (src/diffusion.rs:43); x2 at the interface */
  float backbone_move_p; /* 0.04 (src/diffusion.rs:45) */
  float max_torsion_deg; /* 15 (src/diffusion.rs:47) */
  float iface_cutoff_sq; /* 64, an 8 A Calpha-Calpha interface (src/diffusion.rs:51) */
  uint32_t moves;        /* DNAB_MC_SIDECHAIN | DNAB_MC_BACKBONE (the reference) */
  uint32_t waves;        /* waves per candidate: 0 = chosen by batch size (8 for batches up
                            to the CU count, 4 up to twice it, else 2), or 1 / 2 / 4 / 8; no
                            effect on results beyond the order of FP64 energy sums */
} dnab_mc_params;
int dnab_mc_default_params(dnab_mc_params* p);
/* Every iteration of allatom_diffusion_step (src/diffusion.rs:290-416) for a
 * batch of candidates, all iterations in one launch: residue-level forces from
 * the antigen Calpha grid, restraint, clamp, noise (with_noise) and rigid
 * residue moves, then Metropolis rotamer/mutation moves and phi/psi torsion
 * moves scored by residue_contribution_allatom plus disulfide_energy. This is
 * the per-candidate work of run_allatom's refinement phase
 * (src/diffusion.rs:1083-1095, with_noise = 1) and, with with_noise = 0, of
 * its broad-sampling steps (src/diffusion.rs:1017-1024). lp supplies n_steps,
 * seed, with_noise, step, max_disp, noise_base, t_hot, t_cold,
 * schedule_offset/total (temperature(offset + s, total)), restraint_k and
 * restraint_radius (or the antigen's own). The antigen needs its Calpha cloud
 * and residue types. Each input residue holds its topology's atoms in
 * topology order (as protein_from_ca_trace builds them); res_rotamer
 * (nullable) gives each residue's rotamer index. Random draws are
 * Philox4x32-10 by (iteration, residue, kind): runs repeat exactly, but the
 * reference's SmallRng streams are not reproduced. Outputs as
 * dnab_affinity_maturation; out_energy (nullable, n_cand) receives each
 * refined candidate's allatom_energy. Candidates of up to 512 residues. */
int dnab_refine(dnab_ctx* ctx, dnab_antigen* antigen, const dnab_population_desc* ab,
                const uint8_t* res_rotamer, const dnab_langevin_params* lp,
                const dnab_mc_params* mc, dnab_protein_out* out, uint8_t* out_rotamer,
                double* out_energy, size_t* n_atoms_out);

/* ---- run_allatom (src/diffusion.rs:983-1107) in one call -------------------
 * The broad phase: n_broad binders built as random_allatom_antibody builds
 * them (a ring of ab_length residues around the antigen's centre,
 * src/diffusion.rs:264-277), broad_steps of allatom_diffusion_step without
 * noise on the temperature schedule of the whole run (broad_moves = 1: with
 * its Monte-Carlo moves, as src/diffusion.rs:1021-1024 steps them; 0: the
 * gradient-only loop, one fused launch), then the `population` candidates
 * with the lowest LJ + Coulomb + hydrophobic energy (score_batch's terms,
 * src/diffusion.rs:1041-1049) survive; n_broad = 0 starts from `population`
 * ring binders instead (the reference's path without a GPU,
 * src/diffusion.rs:1064-1071). Then cpu_iterations steps with noise and the
 * moves, the schedule restarted (src/diffusion.rs:1083-1095), the
 * allatom_energy of each, the ascending sort (NaN last, ties by index) and
 * the first max(top_n, 1) (src/diffusion.rs:1097-1098). Outputs: the result
 * proteins concatenated in rank order into `out` (nullable: then only the
 * sizes; res_start and ca_atom_idx are global atom indices), their
 * allatom_energy in out_energy (nullable, max(top_n, 1) entries), the count
 * in n_results and the atom count in n_atoms_out. The antigen needs its
 * Calpha cloud and residue types; its centre is the ring's centre. Random
 * streams are the library's (SplitMix64 for construction, Philox for the
 * steps), seeded from `seed`: runs repeat exactly, the reference's SmallRng
 * trajectories are not reproduced. */
typedef struct dnab_run_allatom_params {
  uint32_t ab_length;      /* binder residues (0: the antigen's residue count, the CLI default) */
  uint32_t population;     /* survivors of the broad phase (TOP_K = 64) */
  uint32_t cpu_iterations; /* refinement iterations (CPU_STEPS = 600) */
  uint32_t top_n;          /* results kept (at least 1) */
  uint32_t n_broad;        /* broad-phase binders (AA_POPULATION = 1024; 0 = no broad phase) */
  uint32_t broad_steps;    /* broad-phase steps (GPU_STEPS = 200) */
  uint32_t broad_moves;    /* 1: the broad phase with the Monte-Carlo moves; 0: without */
  uint32_t _reserved;
  uint64_t seed;
  float ring_radius;       /* the binders' ring (INIT_RADIUS = 20 A) */
  float _pad;
} dnab_run_allatom_params;
DNAB_STATIC_ASSERT(sizeof(dnab_run_allatom_params) == 48, "dnab_run_allatom_params layout");
int dnab_run_allatom_default_params(dnab_run_allatom_params* p);
int dnab_run_allatom(dnab_ctx* ctx, dnab_antigen* antigen, const dnab_run_allatom_params* p,
                     dnab_protein_out* out, double* out_energy, uint32_t* n_results,
                     size_t* n_atoms_out);
/* One step at iteration `iter` (forces from a fresh evaluation). */
int dnab_langevin_step(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                       uint32_t iter);
/* The step loop of run_allatom's GPU phase (src/diffusion.rs:1018-1057), all
 * candidates kept moving, with scoring every score_every steps after the
 * update; the top_k (ascending energy, ties by index) of the final scoring
 * are written to topk_idx / topk_energy (each top_k entries; fewer if the
 * population is smaller, count in *n_topk). */
int dnab_langevin_run(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                      uint32_t* topk_idx, double* topk_energy, uint32_t* n_topk);
/* Workgroup shape of the fused trajectory kernel (no reference counterpart):
 * `waves` (4, 8 or 16) waves run the trajectories of up to `cand_per_group`
 * (1..16) candidates together, sharing their pair work. `balance` 0: groups
 * are consecutive candidates, each candidate's atoms in tiles of their own;
 * 1: a first evaluation measures each candidate's pair work and candidates
 * are grouped so that groups carry similar work (a candidate in deep contact
 * runs alone), heaviest group first; 2: candidates of one antigen are grouped
 * by position (the Morton order of their centroids) and each launch cuts the
 * group's atoms into compact tiles of 64 by recursive bisection, so a tile
 * sees fewer antigen atoms. Groups that do not fit one workgroup's LDS hold
 * fewer candidates.
 * Each step lists, per tile of 64 atoms, the antigen chunks (up to 64 atoms
 * of one grid cell) of the cells within the cutoff of the tile's box; a work
 * item is a run of `chunks_per_item` of them (or of `cells_per_item` cells
 * when the list overflows). Default 16 waves, 4 candidates, balance on,
 * 2 cells, 2 chunks, one launch (no regrouping). Results do not depend on
 * it. */
typedef struct dnab_traj_tuning {
  uint32_t waves;
  uint32_t cand_per_group;
  uint32_t balance;
  uint32_t cells_per_item;  /* grid cells per work item, for a tile whose
                               chunk list overflowed (1..64) */
  uint32_t chunks_per_item; /* listed chunks per work item (1..4) */
  uint32_t rebalance_steps; /* with balance 1 or 2: the loop runs in launches
                               of this many steps, regrouping between them by
                               the work each candidate had in the last, or by
                               the candidates' new positions (0: one launch) */
  uint32_t engine;          /* the engine a loop with step_kernels = 2 runs:
                               0 the fused kernel, 1 the per-step kernels */
} dnab_traj_tuning;
int dnab_population_get_trajectory_tuning(const dnab_population* pop, dnab_traj_tuning* t);
int dnab_population_set_trajectory_tuning(dnab_ctx* ctx, dnab_population* pop,
                                          const dnab_traj_tuning* t);
/* Picks the fused kernel's shape for this population and these loop
 * parameters (no reference counterpart): times trial_steps (1..10000) steps
 * of the fused loop, wall clock with the host's regrouping, from the
 * population's coordinates for each shape of a coordinate descent over waves
 * {4, 8, 16}, the pair (cand_per_group {1, 2, 4, 8}, balance {0, 1, 2}),
 * chunks_per_item {1, 2, 4} and cells_per_item {2, 8} (rebalance_steps kept), and keeps
 * the fastest in the population (also written to `chosen`, nullable). It then
 * times the per-step kernels on the same loop (graph replay when the
 * parameters ask for it) and sets `engine` to the faster of the two. The
 * coordinates are put back afterwards, no evaluation is current and the score
 * history is empty. The shape and the engine change only the speed: every
 * choice gives the same results. */
int dnab_population_autotune_trajectory(dnab_ctx* ctx, dnab_population* pop,
                                        const dnab_langevin_params* p, uint32_t trial_steps,
                                        dnab_traj_tuning* chosen);
/* Score history of the last dnab_langevin_run: one row per scoring (the
 * states at steps 0, score_every, 2 score_every, ... before n_steps, then the
 * final state), each its k best (ascending energy, ties by index). The last
 * row is what the run returned. Arrays of n_scores * k entries; NULL arrays
 * only report n_scores and k. (The reference scores once, at the end of the
 * GPU phase, src/diffusion.rs:1026; this keeps every scoring of the loop.) */
int dnab_population_read_scores(dnab_ctx* ctx, const dnab_population* pop, uint32_t* n_scores,
                                uint32_t* k, uint32_t* topk_idx, double* topk_energy);

/* ---- learned prescreen (no reference counterpart; the reference has no
 * learned components, Cargo.toml:4) -------------------------------------------
 * Antigen maps: on a grid of spacing h (default 0.625 A) over the antigen's atom
 * box widened by cutoff + 2h, the sums the exact scorer forms for a probe atom
 * at a grid point, over antigen atoms with 0.25 <= r^2 <= cutoff^2: LJ
 * eps_ij (s^12 - 2 s^6), s = R_ij / r, for each of the 7 distinct (Rmin/2, eps)
 * classes of the AMBER99SB table; 332 q_i / r; -0.5 per hydrophobic atom with
 * r^2 < 36; and 8 radial densities exp(-(r - mu_k)^2 / (2 * 0.75^2)),
 * mu_k = 2 + k A (on every second grid point). Stored in fp16 (saturated at
 * +-60000).
 * Features (32 per candidate, bf16): a candidate atom reads its class's LJ
 * plane, the potential and the hydrophobic plane by trilinear interpolation
 * and the radial densities at the nearest grid point; per candidate the LJ
 * map energy summed per atom class (7 LJ classes, hydrophobic atoms their
 * own), the Coulomb sums over positive and negative atoms and the
 * hydrophobic sum (asinh(x/2)), the total map energy (asinh(x/10)), log1p
 * counts of clashing (LJ > 1), contacting (LJ < -0.05) and near atoms (in a
 * cell with an antigen atom within the cutoff of a corner; the others read
 * only zeros and are skipped), and
 * the radial densities summed over N/O atoms and over the others (log1p).
 * Model: 32 -> 256 -> 256 -> 256 -> 1, SiLU, bf16 weights and activations,
 * FP32 accumulation (MFMA), fp32 master weights, Adam. Target: asinh(E / s)
 * with E the exact LJ + Coulomb + hydrophobic energy, s = energy_scale. */
#define DNAB_FEATURE_DIM 32
#define DNAB_MLP_HIDDEN 256
int dnab_prescreen_feature_count(size_t* n);
/* Builds (or rebuilds) the antigen's maps; spacing <= 0 selects 0.625 A.
 * Features build missing maps with the default spacing. */
int dnab_antigen_build_maps(dnab_ctx* ctx, dnab_antigen* ag, float spacing);
/* spacing 0 and dims 0 when no maps are built; bytes of device memory held. */
int dnab_antigen_map_info(const dnab_antigen* ag, float* spacing, float origin[3], int32_t dims[3],
                          size_t* bytes);
/* Map values at n grid points ijk[3 n] (test hook): out[n][18] = 7 LJ
 * classes, potential, hydrophobic, 8 radial densities (NaN off the grid of
 * every second point), and the occupancy bit of the cell whose lower corner
 * the point is (NaN on the upper faces). */
int dnab_antigen_map_points(dnab_ctx* ctx, const dnab_antigen* ag, const int32_t* ijk, size_t n,
                            float* out);
/* Per candidate the map energy (the sum the features' total carries): the
 * ranking score of the method without the network. */
int dnab_population_map_energy(dnab_ctx* ctx, dnab_population* pop, float* out);
typedef struct dnab_mlp dnab_mlp;
typedef struct dnab_dataset dnab_dataset;

/* Features of every candidate; copied to host_bf16 (n_cand * 32) when not
 * NULL. */
int dnab_population_features(dnab_ctx* ctx, dnab_population* pop, uint16_t* host_bf16);
int dnab_mlp_create(dnab_ctx* ctx, uint64_t seed, dnab_mlp** out);
void dnab_mlp_destroy(dnab_mlp* mlp);
/* Flat parameter vector: W1[32][256], b1[256], W2[256][256], b2, W3[256][256],
 * b3, w4[256], b4[1] (140,289 floats, fp32 master copy). */
int dnab_mlp_param_count(size_t* n);
int dnab_mlp_get_params(dnab_ctx* ctx, const dnab_mlp* mlp, float* out, size_t n);
int dnab_mlp_set_params(dnab_ctx* ctx, dnab_mlp* mlp, const float* in, size_t n);
/* Loss and gradient of mean((f(x_b) - t_b)^2) on one batch of host rows (the
 * oracle check of the device forward and backward passes). */
int dnab_mlp_loss_grad(dnab_ctx* ctx, dnab_mlp* mlp, const uint16_t* x_bf16, const float* t,
                       uint32_t batch, float* grads, float* loss, float* pred);
/* Training rows get the target t = asinh(min(E, energy_cap) / energy_scale):
 * candidates at or above energy_cap (clashes, no contact) share one value, so
 * the fit spends its capacity on the low-energy tail that screening ranks. */
int dnab_dataset_create(dnab_ctx* ctx, size_t capacity, float energy_scale, float energy_cap,
                        dnab_dataset** out);
void dnab_dataset_destroy(dnab_dataset* ds);
/* Appends one row per candidate: its features and the target computed from
 * the exact FP32 kernel's energy of the candidate. */
int dnab_dataset_add_population(dnab_ctx* ctx, dnab_dataset* ds, dnab_population* pop);
int dnab_dataset_size(const dnab_dataset* ds, size_t* n);
typedef struct dnab_train_params {
  uint32_t n_steps, batch;   /* batch: multiple of 64 */
  float lr, beta1, beta2, eps;
  uint64_t seed;
} dnab_train_params;
int dnab_train_default_params(dnab_train_params* p);
/* Adam training on random minibatches; loss_history (nullable) gets the mean
 * squared error of every step. */
int dnab_mlp_train(dnab_ctx* ctx, dnab_mlp* mlp, dnab_dataset* ds, const dnab_train_params* p,
                   float* loss_history);
/* Predicted target for every candidate (features + forward pass). */
int dnab_mlp_predict(dnab_ctx* ctx, dnab_mlp* mlp, dnab_population* pop, float* scores);
/* Screening: features + MLP for every candidate, the n_keep best predictions
 * re-scored exactly (one workgroup per kept candidate; the same energies as a
 * full evaluation), the top_k of those by exact energy (ascending, ties by
 * index) written to idx / energy (*n_out entries). Other candidates' energy
 * outputs are left stale. 1 <= top_k <= n_keep <= 2048. */
int dnab_prescreen(dnab_ctx* ctx, dnab_mlp* mlp, dnab_population* pop, uint32_t n_keep,
                   uint32_t top_k, uint32_t* idx, double* energy, uint32_t* n_out);
/* The min(m, n) smallest of n host scores (ties by index, NaN last) as
 * indices in ascending index order: the prescreen's keep selection.
 * 1 <= m <= 2048. */
int dnab_select_smallest(dnab_ctx* ctx, const float* scores, uint32_t n, uint32_t m,
                         uint32_t* idx, uint32_t* n_out);
/* Exact top-k of every candidate, for comparison: evaluate all + top-k. */
int dnab_exact_topk(dnab_ctx* ctx, dnab_population* pop, uint32_t top_k, uint32_t* idx,
                    double* energy, uint32_t* n_out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* DNAB_H_ */
