/* dnab oracle — CPU restatement of the reference's hot-path arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker and as the
 * timed CPU baseline; the product (libdnab.so) never links or calls it.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fno-fast-math). The
 * reference is Rust and compiles without contraction or reassociation;
 * ffp-contract=off gives the same IEEE f32 evaluation here.
 *
 * Functions and the reference lines they restate:
 *   or_cell_of                  SpatialHashGrid::cell_of      src/spatial.rs:61-67
 *   or_grid_query               build + query_neighbors       src/spatial.rs:31-58
 *   or_energy_f32               interaction_energy_atoms_breakdown, LJ/Coulomb/
 *                               hydrophobic terms, f32 in the reference's order
 *                                                             src/energy.rs:261-301
 *   or_energy_f64               the same pair set (f32 predicate) with every term
 *                               and sum in double, plus the analytic all-atom
 *                               forces (negative gradient of LJ + Coulomb, the
 *                               hydrophobic pull of src/energy.rs:382-386)
 *   or_population_f64           or_energy_f64 over a ragged batch, threaded
 *   or_hbond_energy_f32/_f64    hbond_energy                  src/energy.rs:32-39
 *   or_pair_solvation_f32/_f64  pair_solvation (EEF1)         src/solvation.rs:34-44
 *   or_energy_full_f32          interaction_energy_atoms_breakdown, all five terms
 *                               (LJ, Coulomb, hydrophobic, H-bond in the j-i loop,
 *                               solvation_interaction after it) in f32 and the
 *                               reference's order   src/energy.rs:261-314,
 *                                                   src/solvation.rs:52-84
 *   or_energy_full_f64          the same pair sets (f32 predicates), every term and
 *                               sum in double
 *   or_disulfide_f32/_f64       disulfide_energy over Cys SG positions
 *                                                   src/allatom.rs:155-180
 *   or_cg_forces(_with_grid)    compute_forces(_with_grid)    src/energy.rs:154-234, 322-398
 *   or_langevin_step            allatom_diffusion_step lines  src/diffusion.rs:300-339
 *                               with set_ca_pos               src/allatom.rs:279-293
 *   or_temperature, or_sigma    temperature + noise scale     src/diffusion.rs:102-104, 315
 *   or_philox4x32_10            counter RNG of the step's noise (Salmon et al. 2011)
 *   or_baseline_energy_brute    timed CPU baseline: src/energy.rs:269-301 pattern,
 *                               parallel over candidates like the Rayon loop of
 *                               src/diffusion.rs:1083-1095
 *   or_baseline_grid_ef         timed CPU baseline: grid E+F, src/energy.rs:322-398
 *                               pattern with AMBER per-atom parameters
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define COULOMB_K 332.0f   /* src/energy.rs:13 */
#define MIN_R_SQ 0.25f     /* src/energy.rs:17 */
#define HYDRO_SQ 36.0f     /* src/energy.rs:20 */
#define HYDRO_BONUS -0.5f  /* src/energy.rs:21 */

typedef struct {
  size_t n;
  const float *x, *y, *z, *q, *rmh, *eps;
  const uint8_t* hyd;
} or_cloud;

/* ---- SpatialHashGrid ---------------------------------------------------- */

static int32_t f32_to_i32_sat(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}

int32_t or_cell_of(float x, float inv_cell) {
  volatile float p = x * inv_cell;
  return f32_to_i32_sat(floorf(p));
}

float or_inv_cell(float cell) { return 1.0f / cell; }

static int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

/* Neighbour lists by definition: for each probe, the 27 cells in dx, dy, dz
 * nested order; within a cell every point whose key equals the cell, in
 * ascending index (the order `build` appends them). O(27 n) per probe. */
int or_grid_query(const float* x, const float* y, const float* z, size_t n, float cell,
                  const float* px, const float* py, const float* pz, size_t n_probe,
                  uint32_t* offsets, uint32_t* indices, size_t capacity) {
  float inv = 1.0f / cell;
  int32_t* keys = (int32_t*)malloc(3 * (n ? n : 1) * sizeof(int32_t));
  if (!keys) return -1;
  for (size_t i = 0; i < n; ++i) {
    keys[3 * i] = or_cell_of(x[i], inv);
    keys[3 * i + 1] = or_cell_of(y[i], inv);
    keys[3 * i + 2] = or_cell_of(z[i], inv);
  }
  size_t o = 0;
  for (size_t p = 0; p < n_probe; ++p) {
    offsets[p] = (uint32_t)o;
    int32_t cx = or_cell_of(px[p], inv), cy = or_cell_of(py[p], inv), cz = or_cell_of(pz[p], inv);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          int32_t kx = wrap_add(cx, dx), ky = wrap_add(cy, dy), kz = wrap_add(cz, dz);
          for (size_t i = 0; i < n; ++i)
            if (keys[3 * i] == kx && keys[3 * i + 1] == ky && keys[3 * i + 2] == kz) {
              if (indices && o < capacity) indices[o] = (uint32_t)i;
              ++o;
            }
        }
  }
  offsets[n_probe] = (uint32_t)o;
  free(keys);
  return o <= capacity ? 0 : 1;
}

/* ---- all-atom energies --------------------------------------------------- */

/* f32 in the reference's order (src/energy.rs:269-301): antibody outer,
 * antigen inner, three sequential f32 accumulators. */
void or_energy_f32(const or_cloud* ag, const or_cloud* ab, float cutoff, float out[3]) {
  const float cut2 = cutoff * cutoff;
  float lj = 0.0f, coul = 0.0f, hyd = 0.0f;
  for (size_t j = 0; j < ab->n; ++j) {
    const float bx = ab->x[j], by = ab->y[j], bz = ab->z[j], bq = ab->q[j];
    const float brm = ab->rmh[j], be = ab->eps[j];
    const uint8_t bh = ab->hyd[j];
    for (size_t i = 0; i < ag->n; ++i) {
      float dx = bx - ag->x[i];
      float dy = by - ag->y[i];
      float dz = bz - ag->z[i];
      float r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq > cut2 || r_sq < MIN_R_SQ) continue;
      float r_ij = brm + ag->rmh[i];
      float eps = sqrtf(be * ag->eps[i]);
      float r2 = (r_ij * r_ij) / r_sq;
      float r6 = r2 * r2 * r2;
      float r12 = r6 * r6;
      lj += eps * (r12 - 2.0f * r6);
      float q2 = ag->q[i];
      if (bq != 0.0f && q2 != 0.0f) coul += COULOMB_K * bq * q2 / sqrtf(r_sq);
      if (bh == 1 && ag->hyd[i] == 1 && r_sq < HYDRO_SQ) hyd += HYDRO_BONUS;
    }
  }
  out[0] = lj;
  out[1] = coul;
  out[2] = hyd;
}

/* FP64 oracle for one antibody atom against the whole antigen. The pair set
 * is decided by the reference's f32 predicate; terms, mixing and sums are in
 * double from the raw f32 parameters. */
static void atom_f64(const or_cloud* ag, const or_cloud* ab, size_t j, float cut2,
                     double acc[6], uint64_t* pairs) {
  const float bx = ab->x[j], by = ab->y[j], bz = ab->z[j];
  const double bq = ab->q[j], brm = ab->rmh[j], be = ab->eps[j];
  const uint8_t bh = ab->hyd[j];
  double lj = 0, co = 0, hy = 0, fx = 0, fy = 0, fz = 0;
  uint64_t np = 0;
  for (size_t i = 0; i < ag->n; ++i) {
    float dx = bx - ag->x[i];
    float dy = by - ag->y[i];
    float dz = bz - ag->z[i];
    float r_sq = dx * dx + dy * dy + dz * dz;
    if (r_sq > cut2 || r_sq < MIN_R_SQ) continue;
    ++np;
    double ddx = (double)bx - (double)ag->x[i];
    double ddy = (double)by - (double)ag->y[i];
    double ddz = (double)bz - (double)ag->z[i];
    double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
    double r = sqrt(r2);
    double R = (double)ag->rmh[i] + brm;
    double eps = sqrt((double)ag->eps[i] * be);
    double s2 = (R * R) / r2;
    double s6 = s2 * s2 * s2;
    double s12 = s6 * s6;
    lj += eps * (s12 - 2.0 * s6);
    double f = 12.0 * eps * (s12 - s6) / r2;
    double ec = 332.0 * bq * (double)ag->q[i] / r;
    co += ec;
    f += ec / r2;
    if (bh == 1 && ag->hyd[i] == 1 && r_sq < HYDRO_SQ) {
      hy += -0.5;
      f += -0.1 / r;
    }
    fx += f * ddx;
    fy += f * ddy;
    fz += f * ddz;
  }
  acc[0] = lj;
  acc[1] = co;
  acc[2] = hy;
  acc[3] = fx;
  acc[4] = fy;
  acc[5] = fz;
  if (pairs) *pairs += np;
}

void or_energy_f64(const or_cloud* ag, const or_cloud* ab, float cutoff, double out[3],
                   double* fx, double* fy, double* fz, uint64_t* n_pairs) {
  const float cut2 = cutoff * cutoff;
  double lj = 0, co = 0, hy = 0;
  uint64_t np = 0;
  for (size_t j = 0; j < ab->n; ++j) {
    double a[6];
    atom_f64(ag, ab, j, cut2, a, &np);
    lj += a[0];
    co += a[1];
    hy += a[2];
    if (fx) fx[j] = a[3];
    if (fy) fy[j] = a[4];
    if (fz) fz[j] = a[5];
  }
  out[0] = lj;
  out[1] = co;
  out[2] = hy;
  if (n_pairs) *n_pairs = np;
}

/* ---- the other two interaction terms and the disulfide term ---------------- */

#define HB_R0 2.9f         /* src/energy.rs:24-27 */
#define HB_SIGMA 0.35f
#define HB_DEPTH 1.5f
#define HB_CUTOFF_SQ 25.0f
#define PI32 5.568328f     /* pi^(3/2), src/solvation.rs:22 */
#define SOLV_CUTOFF_SQ 81.0f  /* (9 A)^2, src/solvation.rs:25-26 */

/* Per-atom EEF1 parameters and H-bond classes (src/amber.rs:72-84, 414-431),
 * passed in by the caller (the tests take them from the reference's tables). */
typedef struct {
  const float *dg, *vol, *lam;
  const uint8_t *don, *acc;
} or_ext;

float or_hbond_energy_f32(float r_sq) {
  if (r_sq > HB_CUTOFF_SQ) return 0.0f;
  float r = sqrtf(r_sq);
  float z = (r - HB_R0) / HB_SIGMA;
  return -HB_DEPTH * expf(-0.5f * z * z);
}

/* the double oracles take the reference's f32 constants, widened */
double or_hbond_energy_f64(double r_sq) {
  double r = sqrt(r_sq);
  double z = (r - (double)HB_R0) / (double)HB_SIGMA;
  return -(double)HB_DEPTH * exp(-0.5 * z * z);
}

/* powi(2) and powi(3) as repeated products, in the reference's order */
float or_pair_solvation_f32(float dgi, float voli, float lami, float dgj, float volj, float lamj,
                            float r_sq) {
  float r = sqrtf(r_sq);
  float ti = r / lami, tj = r / lamj;
  float fi = expf(-(ti * ti)) / (PI32 * (lami * lami * lami));
  float fj = expf(-(tj * tj)) / (PI32 * (lamj * lamj * lamj));
  return -(dgi * volj * fi + dgj * voli * fj);
}

double or_pair_solvation_f64(double dgi, double voli, double lami, double dgj, double volj,
                             double lamj, double r_sq) {
  const double pi32 = (double)PI32;
  double r = sqrt(r_sq);
  double fi = exp(-(r / lami) * (r / lami)) / (pi32 * lami * lami * lami);
  double fj = exp(-(r / lamj) * (r / lamj)) / (pi32 * lamj * lamj * lamj);
  return -(dgi * volj * fi + dgj * voli * fj);
}

static int hb_pair(const or_ext* ga, size_t i, const or_ext* gb, size_t j) {
  return (gb->don[j] && ga->acc[i]) || (gb->acc[j] && ga->don[i]);
}

/* All five terms in f32, the reference's order: the j-i loop with four
 * accumulators, then the solvation loop summed per antibody atom. out =
 * (lj, coulomb, hydrophobic, hbond, solvation). */
void or_energy_full_f32(const or_cloud* ag, const or_cloud* ab, const or_ext* ga,
                        const or_ext* gb, float cutoff, float out[5]) {
  float three[3];
  or_energy_f32(ag, ab, cutoff, three);
  const float cut2 = cutoff * cutoff;
  float hb = 0.0f, solv = 0.0f;
  for (size_t j = 0; j < ab->n; ++j) {
    for (size_t i = 0; i < ag->n; ++i) {
      float dx = ab->x[j] - ag->x[i];
      float dy = ab->y[j] - ag->y[i];
      float dz = ab->z[j] - ag->z[i];
      float r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq > cut2 || r_sq < MIN_R_SQ) continue;
      if (hb_pair(ga, i, gb, j)) hb += or_hbond_energy_f32(r_sq);
    }
  }
  for (size_t j = 0; j < ab->n; ++j) {
    float local = 0.0f;
    for (size_t i = 0; i < ag->n; ++i) {
      float dx = ab->x[j] - ag->x[i];
      float dy = ab->y[j] - ag->y[i];
      float dz = ab->z[j] - ag->z[i];
      float r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq >= SOLV_CUTOFF_SQ || r_sq < MIN_R_SQ) continue;
      local += or_pair_solvation_f32(ga->dg[i], ga->vol[i], ga->lam[i], gb->dg[j], gb->vol[j],
                                     gb->lam[j], r_sq);
    }
    solv += local;
  }
  out[0] = three[0];
  out[1] = three[1];
  out[2] = three[2];
  out[3] = hb;
  out[4] = solv;
}

/* The same pair sets (f32 predicates), terms and sums in double; per
 * antibody atom into e5[5 * j .. 5 * j + 4] when e5 is given. */
void or_energy_full_f64(const or_cloud* ag, const or_cloud* ab, const or_ext* ga,
                        const or_ext* gb, float cutoff, double out[5], double* e5) {
  const float cut2 = cutoff * cutoff;
  double tot[5] = {0, 0, 0, 0, 0};
  for (size_t j = 0; j < ab->n; ++j) {
    double a[6], hb = 0.0, solv = 0.0;
    atom_f64(ag, ab, j, cut2, a, NULL);
    for (size_t i = 0; i < ag->n; ++i) {
      float dx = ab->x[j] - ag->x[i];
      float dy = ab->y[j] - ag->y[i];
      float dz = ab->z[j] - ag->z[i];
      float r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq < MIN_R_SQ) continue;
      double ddx = (double)ab->x[j] - (double)ag->x[i];
      double ddy = (double)ab->y[j] - (double)ag->y[i];
      double ddz = (double)ab->z[j] - (double)ag->z[i];
      double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
      if (r_sq <= cut2 && r_sq <= HB_CUTOFF_SQ && hb_pair(ga, i, gb, j))
        hb += or_hbond_energy_f64(r2);
      if (r_sq < SOLV_CUTOFF_SQ)
        solv += or_pair_solvation_f64(ga->dg[i], ga->vol[i], ga->lam[i], gb->dg[j], gb->vol[j],
                                      gb->lam[j], r2);
    }
    const double t[5] = {a[0], a[1], a[2], hb, solv};
    for (int k = 0; k < 5; ++k) {
      tot[k] += t[k];
      if (e5) e5[5 * j + k] = t[k];
    }
  }
  for (int k = 0; k < 5; ++k) out[k] = tot[k];
}

/* disulfide_energy over the given SG positions (src/allatom.rs:155-180):
 * pairs i < j in order, d = sqrt(dx*dx + dy*dy + dz*dz), skipped beyond 4 A,
 * e -= 4 exp(-z^2 / 2), z = (d - 2.05) / 0.4. */
float or_disulfide_f32(const float* x, const float* y, const float* z, size_t n) {
  float e = 0.0f;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = i + 1; j < n; ++j) {
      float dx = x[i] - x[j];
      float dy = y[i] - y[j];
      float dz = z[i] - z[j];
      float d = sqrtf(dx * dx + dy * dy + dz * dz);
      if (d > 4.0f) continue;
      float t = (d - 2.05f) / 0.4f;
      e -= 4.0f * expf(-0.5f * t * t);
    }
  return e;
}

double or_disulfide_f64(const float* x, const float* y, const float* z, size_t n) {
  double e = 0.0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = i + 1; j < n; ++j) {
      float dx = x[i] - x[j];
      float dy = y[i] - y[j];
      float dz = z[i] - z[j];
      float d32 = sqrtf(dx * dx + dy * dy + dz * dz);
      if (d32 > 4.0f) continue;  /* the reference's f32 cutoff decides the pair set */
      double ddx = (double)x[i] - x[j], ddy = (double)y[i] - y[j], ddz = (double)z[i] - z[j];
      double t = (sqrt(ddx * ddx + ddy * ddy + ddz * ddz) - (double)2.05f) / (double)0.4f;
      e -= 4.0 * exp(-0.5 * t * t);
    }
  return e;
}

/* ---- threaded batch helpers ----------------------------------------------- */

typedef struct {
  const or_cloud* ag;
  const or_cloud* atoms;   /* all candidate atoms */
  const uint32_t* cand_off;
  size_t n_cand;
  float cutoff;
  float cell;
  double *e, *lj, *co, *hy, *fx, *fy, *fz;
  float* e32;
  float *f32x, *f32y, *f32z;
  uint64_t* pairs;         /* per candidate */
  size_t next;             /* work counter */
  pthread_mutex_t mu;
  int kind;
  /* grid baseline */
  const int32_t* ckey;     /* dense grid */
  const uint32_t* cstart;
  const uint32_t* cidx;
  int32_t kmin[3], kdim[3];
  float inv;
} or_job;

static size_t take(or_job* J, size_t chunk) {
  pthread_mutex_lock(&J->mu);
  size_t c = J->next;
  J->next += chunk;
  pthread_mutex_unlock(&J->mu);
  return c;
}

static or_cloud sub_cloud(const or_cloud* a, size_t b, size_t e) {
  or_cloud c;
  c.n = e - b;
  c.x = a->x + b;
  c.y = a->y + b;
  c.z = a->z + b;
  c.q = a->q + b;
  c.rmh = a->rmh + b;
  c.eps = a->eps + b;
  c.hyd = a->hyd + b;
  return c;
}

static void* worker_f64(void* arg) {
  or_job* J = (or_job*)arg;
  const float cut2 = J->cutoff * J->cutoff;
  for (;;) {
    size_t c = take(J, 1);
    if (c >= J->n_cand) break;
    size_t b = J->cand_off[c], e = J->cand_off[c + 1];
    double lj = 0, co = 0, hy = 0;
    uint64_t np = 0;
    for (size_t j = b; j < e; ++j) {
      double a[6];
      atom_f64(J->ag, J->atoms, j, cut2, a, &np);
      lj += a[0];
      co += a[1];
      hy += a[2];
      if (J->fx) {
        J->fx[j] = a[3];
        J->fy[j] = a[4];
        J->fz[j] = a[5];
      }
    }
    if (J->lj) {
      J->lj[c] = lj;
      J->co[c] = co;
      J->hy[c] = hy;
    }
    J->e[c] = (lj + co) + hy;
    if (J->pairs) J->pairs[c] = np;
  }
  return NULL;
}

static void run_threads(or_job* J, int nthreads, void* (*fn)(void*)) {
  if (nthreads < 1) nthreads = 1;
  pthread_mutex_init(&J->mu, NULL);
  J->next = 0;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  int started = 0;
  for (int t = 1; t < nthreads; ++t)
    if (pthread_create(&th[t], NULL, fn, J) == 0) ++started;
    else break;
  fn(J);
  for (int t = 1; t <= started; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&J->mu);
}

/* Per candidate: E = LJ + Coulomb + hydrophobic in FP64 (and the terms), per
 * atom FP64 forces (nullable), per candidate in-cutoff pair counts. */
void or_population_f64(const or_cloud* ag, const or_cloud* atoms, const uint32_t* cand_off,
                       size_t n_cand, float cutoff, double* e, double* lj, double* co,
                       double* hy, double* fx, double* fy, double* fz, uint64_t* pairs,
                       int nthreads) {
  or_job J;
  memset(&J, 0, sizeof(J));
  J.ag = ag;
  J.atoms = atoms;
  J.cand_off = cand_off;
  J.n_cand = n_cand;
  J.cutoff = cutoff;
  J.e = e;
  J.lj = lj;
  J.co = co;
  J.hy = hy;
  J.fx = fx;
  J.fy = fy;
  J.fz = fz;
  J.pairs = pairs;
  run_threads(&J, nthreads, worker_f64);
}

/* ---- timed CPU baselines ---------------------------------------------------- */

/* Reference pattern (src/energy.rs:269-301): per candidate a brute-force f32
 * breakdown over all antigen atoms. */
static void* worker_brute(void* arg) {
  or_job* J = (or_job*)arg;
  for (;;) {
    size_t c = take(J, 1);
    if (c >= J->n_cand) break;
    or_cloud ab = sub_cloud(J->atoms, J->cand_off[c], J->cand_off[c + 1]);
    float out[3];
    or_energy_f32(J->ag, &ab, J->cutoff, out);
    J->e32[c] = out[0] + out[1] + out[2];
  }
  return NULL;
}

void or_baseline_energy_brute(const or_cloud* ag, const or_cloud* atoms,
                              const uint32_t* cand_off, size_t n_cand, float cutoff, float* e,
                              int nthreads) {
  or_job J;
  memset(&J, 0, sizeof(J));
  J.ag = ag;
  J.atoms = atoms;
  J.cand_off = cand_off;
  J.n_cand = n_cand;
  J.cutoff = cutoff;
  J.e32 = e;
  run_threads(&J, nthreads, worker_brute);
}

/* Grid E+F baseline: the antigen in a dense cell grid (cell = cutoff), each
 * candidate atom probes the 27 neighbouring cells (src/spatial.rs:44-58) and
 * accumulates AMBER energy and force in f32 (src/energy.rs:336-397 pattern). */
static void* worker_grid(void* arg) {
  or_job* J = (or_job*)arg;
  const float cut2 = J->cutoff * J->cutoff;
  const or_cloud* ag = J->ag;
  for (;;) {
    size_t c = take(J, 1);
    if (c >= J->n_cand) break;
    float etot = 0.0f;
    for (size_t j = J->cand_off[c]; j < J->cand_off[c + 1]; ++j) {
      const float bx = J->atoms->x[j], by = J->atoms->y[j], bz = J->atoms->z[j];
      const float bq = J->atoms->q[j], brm = J->atoms->rmh[j], be = J->atoms->eps[j];
      const uint8_t bh = J->atoms->hyd[j];
      int32_t k[3] = {or_cell_of(bx, J->inv), or_cell_of(by, J->inv), or_cell_of(bz, J->inv)};
      float ax = 0, ay = 0, az = 0, e = 0;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            int32_t gx = k[0] + dx - J->kmin[0], gy = k[1] + dy - J->kmin[1],
                    gz = k[2] + dz - J->kmin[2];
            if (gx < 0 || gy < 0 || gz < 0 || gx >= J->kdim[0] || gy >= J->kdim[1] ||
                gz >= J->kdim[2])
              continue;
            size_t cell = ((size_t)gz * J->kdim[1] + gy) * J->kdim[0] + gx;
            for (uint32_t s = J->cstart[cell]; s < J->cstart[cell + 1]; ++s) {
              uint32_t i = J->cidx[s];
              float ddx = bx - ag->x[i], ddy = by - ag->y[i], ddz = bz - ag->z[i];
              float r_sq = ddx * ddx + ddy * ddy + ddz * ddz;
              if (r_sq > cut2 || r_sq < MIN_R_SQ) continue;
              float r_ij = brm + ag->rmh[i];
              float eps = sqrtf(be * ag->eps[i]);
              float s2 = (r_ij * r_ij) / r_sq;
              float s6 = s2 * s2 * s2;
              float s12 = s6 * s6;
              e += eps * (s12 - 2.0f * s6);
              float f = (12.0f * eps / r_sq) * (s12 - s6);
              float q2 = ag->q[i];
              if (bq != 0.0f && q2 != 0.0f) {
                float r = sqrtf(r_sq);
                e += COULOMB_K * bq * q2 / r;
                f += COULOMB_K * bq * q2 / (r_sq * r);
              }
              if (bh == 1 && ag->hyd[i] == 1 && r_sq < HYDRO_SQ) {
                e += HYDRO_BONUS;
                f += -0.1f / sqrtf(r_sq);
              }
              ax += f * ddx;
              ay += f * ddy;
              az += f * ddz;
            }
          }
      if (J->f32x) {
        J->f32x[j] = ax;
        J->f32y[j] = ay;
        J->f32z[j] = az;
      }
      etot += e;
    }
    J->e32[c] = etot;
  }
  return NULL;
}

int or_baseline_grid_ef(const or_cloud* ag, const or_cloud* atoms, const uint32_t* cand_off,
                        size_t n_cand, float cutoff, float* e, float* fx, float* fy,
                        float* fz, int nthreads) {
  or_job J;
  memset(&J, 0, sizeof(J));
  J.ag = ag;
  J.atoms = atoms;
  J.cand_off = cand_off;
  J.n_cand = n_cand;
  J.cutoff = cutoff;
  J.e32 = e;
  J.f32x = fx;
  J.f32y = fy;
  J.f32z = fz;
  J.inv = 1.0f / cutoff;
  int32_t kmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, kmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  int32_t* keys = (int32_t*)malloc(3 * (ag->n ? ag->n : 1) * sizeof(int32_t));
  if (!keys) return -1;
  for (size_t i = 0; i < ag->n; ++i) {
    const float p[3] = {ag->x[i], ag->y[i], ag->z[i]};
    for (int a = 0; a < 3; ++a) {
      keys[3 * i + a] = or_cell_of(p[a], J.inv);
      if (keys[3 * i + a] < kmin[a]) kmin[a] = keys[3 * i + a];
      if (keys[3 * i + a] > kmax[a]) kmax[a] = keys[3 * i + a];
    }
  }
  size_t ncell = 1;
  for (int a = 0; a < 3; ++a) {
    J.kmin[a] = ag->n ? kmin[a] : 0;
    J.kdim[a] = ag->n ? kmax[a] - kmin[a] + 1 : 0;
    ncell *= (size_t)(J.kdim[a] > 0 ? J.kdim[a] : 0);
  }
  uint32_t* cstart = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
  uint32_t* cidx = (uint32_t*)malloc((ag->n ? ag->n : 1) * sizeof(uint32_t));
  if (!cstart || !cidx) {
    free(keys);
    free(cstart);
    free(cidx);
    return -1;
  }
  for (size_t i = 0; i < ag->n; ++i) {
    size_t cell = ((size_t)(keys[3 * i + 2] - J.kmin[2]) * J.kdim[1] + (keys[3 * i + 1] - J.kmin[1])) *
                      J.kdim[0] + (keys[3 * i] - J.kmin[0]);
    cstart[cell + 1]++;
  }
  for (size_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
  uint32_t* fill = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
  for (size_t i = 0; i < ag->n; ++i) {
    size_t cell = ((size_t)(keys[3 * i + 2] - J.kmin[2]) * J.kdim[1] + (keys[3 * i + 1] - J.kmin[1])) *
                      J.kdim[0] + (keys[3 * i] - J.kmin[0]);
    cidx[cstart[cell] + fill[cell]++] = (uint32_t)i;
  }
  J.cstart = cstart;
  J.cidx = cidx;
  run_threads(&J, nthreads, worker_grid);
  free(fill);
  free(keys);
  free(cstart);
  free(cidx);
  return 0;
}

/* ---- residue-level forces ---------------------------------------------------- */

typedef struct {
  size_t n;
  const float *x, *y, *z, *q, *eps, *sig;
  const uint8_t* hyd;
} or_rcloud;

static void cg_pair(const or_rcloud* ag, size_t i, float bx, float by, float bz, float bq,
                    float be, float bs, uint8_t bh, float* ax, float* ay, float* az) {
  float dx = bx - ag->x[i];
  float dy = by - ag->y[i];
  float dz = bz - ag->z[i];
  float r_sq = dx * dx + dy * dy + dz * dz;
  if (r_sq > 100.0f || r_sq < MIN_R_SQ) return;  /* src/energy.rs:358 */
  float eps_ij = sqrtf(be * ag->eps[i]);
  float sig_ij = 0.5f * (bs + ag->sig[i]);
  float sigma_sq_ij = sig_ij * sig_ij;
  float s2 = sigma_sq_ij / r_sq;
  float s6 = s2 * s2 * s2;
  float s12 = s6 * s6;
  float f_lj = (24.0f * eps_ij / r_sq) * (2.0f * s12 - s6);
  float f_coul = 0.0f;
  float q2 = ag->q[i];
  if (bq != 0.0f && q2 != 0.0f) {
    float r = sqrtf(r_sq);
    f_coul = COULOMB_K * bq * q2 / (r_sq * r);
  }
  float f_hydro = 0.0f;
  if (bh == 1 && ag->hyd[i] == 1 && r_sq < HYDRO_SQ) f_hydro = -0.1f / sqrtf(r_sq);
  float f_total = f_lj + f_coul + f_hydro;
  *ax += f_total * dx;
  *ay += f_total * dy;
  *az += f_total * dz;
}

/* compute_forces (src/energy.rs:154-234) */
void or_cg_forces(const or_rcloud* ag, const or_rcloud* ab, float* fx, float* fy, float* fz) {
  for (size_t j = 0; j < ab->n; ++j) {
    float ax = 0, ay = 0, az = 0;
    for (size_t i = 0; i < ag->n; ++i)
      cg_pair(ag, i, ab->x[j], ab->y[j], ab->z[j], ab->q[j], ab->eps[j], ab->sig[j], ab->hyd[j],
              &ax, &ay, &az);
    fx[j] = ax;
    fy[j] = ay;
    fz[j] = az;
  }
}

/* compute_forces_with_grid (src/energy.rs:322-398) with a grid of cell 10 and
 * the query order of src/spatial.rs:44-58. */
static void cg_point_grid(const or_rcloud* ag, const int32_t* keys, float inv, float bx,
                          float by, float bz, float bq, float be, float bs, uint8_t bh,
                          float* fx, float* fy, float* fz) {
  float ax = 0, ay = 0, az = 0;
  int32_t cx = or_cell_of(bx, inv), cy = or_cell_of(by, inv), cz = or_cell_of(bz, inv);
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        int32_t kx = wrap_add(cx, dx), ky = wrap_add(cy, dy), kz = wrap_add(cz, dz);
        for (size_t i = 0; i < ag->n; ++i)
          if (keys[3 * i] == kx && keys[3 * i + 1] == ky && keys[3 * i + 2] == kz)
            cg_pair(ag, i, bx, by, bz, bq, be, bs, bh, &ax, &ay, &az);
      }
  *fx = ax;
  *fy = ay;
  *fz = az;
}

int or_cg_forces_with_grid(const or_rcloud* ag, float cell, const or_rcloud* ab, float* fx,
                           float* fy, float* fz) {
  float inv = 1.0f / cell;
  int32_t* keys = (int32_t*)malloc(3 * (ag->n ? ag->n : 1) * sizeof(int32_t));
  if (!keys) return -1;
  for (size_t i = 0; i < ag->n; ++i) {
    keys[3 * i] = or_cell_of(ag->x[i], inv);
    keys[3 * i + 1] = or_cell_of(ag->y[i], inv);
    keys[3 * i + 2] = or_cell_of(ag->z[i], inv);
  }
  for (size_t j = 0; j < ab->n; ++j)
    cg_point_grid(ag, keys, inv, ab->x[j], ab->y[j], ab->z[j], ab->q[j], ab->eps[j], ab->sig[j],
                  ab->hyd[j], &fx[j], &fy[j], &fz[j]);
  free(keys);
  return 0;
}

/* ---- the gradient step --------------------------------------------------------- */

typedef struct { uint32_t v[4]; } or_philox4;

static uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t h0 = mulhi32(0xD2511F53u, c0), l0 = 0xD2511F53u * c0;
    uint32_t h1 = mulhi32(0xCD9E8D57u, c2), l1 = 0xCD9E8D57u * c2;
    uint32_t n0 = h1 ^ c1 ^ k0, n1 = l1, n2 = h0 ^ c3 ^ k1, n3 = l0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* noise scale at iteration it: NOISE_BASE * sqrt(T_HOT (T_COLD/T_HOT)^(it/total)) */
/* temperature (src/diffusion.rs:102-104) in f32. */
float or_temperature(uint32_t it, uint32_t total, float t_hot, float t_cold) {
  return t_hot * powf(t_cold / t_hot, (float)it / (float)total);
}

float or_sigma(uint32_t it, uint32_t total, float t_hot, float t_cold, float noise_base) {
  float temp = t_hot * powf(t_cold / t_hot, (float)it / (float)total);
  return noise_base * sqrtf(temp);
}

static float clampf_rust(float v, float lo, float hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

typedef struct {
  float step, max_disp, restraint_k, restraint_radius, sigma;
  int with_noise, force_model; /* 0 all-atom (forces given), 1 residue level */
  uint32_t iter;
  uint32_t seed_lo, seed_hi;
  float center[3];
} or_step_params;

/* One step for a set of residues. x/y/z are all atoms (updated in place);
 * res_off[n_res+1] atom ranges, res_ca[n_res] Calpha atom index; afx/afy/afz
 * per-atom forces (all-atom model) or the residue-level inputs (cg model:
 * ag_ca, res_q/eps/sig/hyd). res_id0 is the global index of residue 0 (the
 * noise counter). */
void or_langevin_step(float* x, float* y, float* z, size_t n_res, const uint32_t* res_off,
                      const uint32_t* res_ca, const float* afx, const float* afy,
                      const float* afz, const or_rcloud* ag_ca, const float* res_q,
                      const float* res_eps, const float* res_sig, const uint8_t* res_hyd,
                      uint32_t res_id0, const or_step_params* p) {
  int32_t* keys = NULL;
  float inv = 1.0f / 10.0f;
  if (p->force_model == 1 && ag_ca && ag_ca->n) {
    keys = (int32_t*)malloc(3 * ag_ca->n * sizeof(int32_t));
    for (size_t i = 0; i < ag_ca->n; ++i) {
      keys[3 * i] = or_cell_of(ag_ca->x[i], inv);
      keys[3 * i + 1] = or_cell_of(ag_ca->y[i], inv);
      keys[3 * i + 2] = or_cell_of(ag_ca->z[i], inv);
    }
  }
  for (size_t r = 0; r < n_res; ++r) {
    uint32_t ca = res_ca[r];
    float px = x[ca], py = y[ca], pz = z[ca];
    float fx = 0, fy = 0, fz = 0;
    if (p->force_model == 1) {
      if (keys)
        cg_point_grid(ag_ca, keys, inv, px, py, pz, res_q[r], res_eps[r], res_sig[r], res_hyd[r],
                      &fx, &fy, &fz);
    } else {
      for (uint32_t k = res_off[r]; k < res_off[r + 1]; ++k) {
        fx += afx[k];
        fy += afy[k];
        fz += afz[k];
      }
    }
    float dx_c = px - p->center[0], dy_c = py - p->center[1], dz_c = pz - p->center[2];
    float d = sqrtf(dx_c * dx_c + dy_c * dy_c + dz_c * dz_c);
    if (d < 0.001f) d = 0.001f;
    float excess = d - p->restraint_radius;
    float f_r = -p->restraint_k * excess / d;
    fx += f_r * dx_c;
    fy += f_r * dy_c;
    fz += f_r * dz_c;
    float nx = 0, ny = 0, nz = 0;
    if (p->with_noise) {
      uint32_t ctr[4] = {p->iter, res_id0 + (uint32_t)r, 0x6e6f6973u, 0u};
      uint32_t key[2] = {p->seed_lo, p->seed_hi}, u[4];
      or_philox4x32_10(ctr, key, u);
      float two = p->sigma + p->sigma;
      nx = (float)(u[0] >> 8) * (1.0f / 16777216.0f) * two - p->sigma;
      ny = (float)(u[1] >> 8) * (1.0f / 16777216.0f) * two - p->sigma;
      nz = (float)(u[2] >> 8) * (1.0f / 16777216.0f) * two - p->sigma;
    }
    float new_x = px + clampf_rust(p->step * fx, -p->max_disp, p->max_disp) + nx;
    float new_y = py + clampf_rust(p->step * fy, -p->max_disp, p->max_disp) + ny;
    float new_z = pz + clampf_rust(p->step * fz, -p->max_disp, p->max_disp) + nz;
    float ddx = new_x - px, ddy = new_y - py, ddz = new_z - pz;
    for (uint32_t k = res_off[r]; k < res_off[r + 1]; ++k) {
      x[k] += ddx;
      y[k] += ddy;
      z[k] += ddz;
    }
  }
  free(keys);
}

/* ---- allatom_diffusion_step with Monte-Carlo moves, C port -----------------
 * The same algorithm as oracle/refine_oracle.py (src/diffusion.rs:290-416),
 * written for speed as the CPU baseline of the refinement: candidates in
 * parallel (the Rayon loop of src/diffusion.rs:1083-1095), residues in 14
 * atom slots (the splice of mutate_residue, src/allatom.rs:299-384), the
 * antigen Calpha grid as cell lists (src/spatial.rs:12-68), NERF from the
 * tables the caller passes (src/rotamer.rs), Philox draws by (iteration,
 * residue, kind), Metropolis energies as FP64 sums of f32 terms. */

#define RF_SLOTS 14
#define RF_DEFS 12

typedef struct {
  const uint8_t* natoms;  /* 20: atoms per residue type */
  const uint8_t* ndef;    /* 20: side-chain placements */
  const float* defs;      /* 20 x 12 x 7: a, b, c, bond, angle, fixed dihedral, chi (255 fixed) */
  const uint8_t* nrot;    /* 20 */
  const float* chi;       /* 20 x 5 x 4 degrees */
  const float* ca_par;    /* 20 x 5: Calpha r_min_half, epsilon, charge, hydrophobic, type */
  const float* cg_par;    /* 20 x 4: residue-level charge, epsilon, sigma, hydrophobic */
  const uint8_t* hb_don;  /* 16 */
  const uint8_t* hb_acc;  /* 16 */
} or_rf_tables;

typedef struct {
  float step, max_disp, restraint_k, restraint_radius;
  int with_noise;
  uint32_t seed_lo, seed_hi;
  float mut_p, rot_p, bb_p, max_torsion, iface_sq;
  uint32_t moves;
  float center[3];
} or_rf_params;

typedef struct {
  float x, y, z;
} rf_v3;

static rf_v3 rf_sub(rf_v3 a, rf_v3 b) { rf_v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static rf_v3 rf_cross(rf_v3 a, rf_v3 b) {
  rf_v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}
static rf_v3 rf_norm(rf_v3 v) {
  float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  if (l < 1e-8f) l = 1e-8f;
  rf_v3 r = {v.x / l, v.y / l, v.z / l};
  return r;
}
static float rf_rad(float d) { return d * (3.14159265358979323846f / 180.0f); }

/* place_atom (src/rotamer.rs:198-232) */
static rf_v3 rf_place(rf_v3 a, rf_v3 b, rf_v3 c, float bond, float ang_deg, float dih_deg) {
  float ang = rf_rad(ang_deg), dih = rf_rad(dih_deg);
  rf_v3 bc = rf_norm(rf_sub(b, c));
  rf_v3 n = rf_norm(rf_cross(rf_sub(b, a), bc));
  rf_v3 m = rf_cross(bc, n);
  float sa = sinf(ang), ca = cosf(ang), sd = sinf(dih), cd = cosf(dih);
  rf_v3 d = {c.x + bond * (-ca * bc.x + sa * cd * m.x + sa * sd * n.x),
             c.y + bond * (-ca * bc.y + sa * cd * m.y + sa * sd * n.y),
             c.z + bond * (-ca * bc.z + sa * cd * m.z + sa * sd * n.z)};
  return d;
}

static void rf_side_chain(const or_rf_tables* T, int aa, int rot, int nb, rf_v3* s) {
  int n = nb;
  for (int k = 0; k < T->ndef[aa] && n < T->natoms[aa]; ++k) {
    const float* d = T->defs + ((size_t)aa * RF_DEFS + k) * 7;
    int ci = (int)d[6];
    float dih = ci == 255 ? d[5] : (T->nrot[aa] ? T->chi[((size_t)aa * 5 + rot) * 4 + ci] : 0.0f);
    s[n] = rf_place(s[(int)d[0]], s[(int)d[1]], s[(int)d[2]], d[3], d[4], dih);
    ++n;
  }
}

/* mutate_residue (src/allatom.rs:299-384) on one residue's slots */
static void rf_mutate(const or_rf_tables* T, uint8_t* aa, uint8_t* rot, int r, int new_aa,
                      int new_rot, rf_v3* slots) {
  rf_v3* s = slots + (size_t)r * RF_SLOTS;
  if (aa[r] == new_aa) {
    rot[r] = (uint8_t)new_rot;
    rf_side_chain(T, new_aa, new_rot, new_aa == 7 ? 4 : 5, s);
    return;
  }
  const float cx = s[1].x, cy = s[1].y, cz = s[1].z;
  rf_v3 bb[5] = {{cx - 1.46f, cy, cz}, {cx, cy, cz}, {cx + 1.52f, cy, cz}, {0, 0, 0}, {cx, cy - 1.52f, cz}};
  bb[3].x = bb[2].x + 0.5f;
  bb[3].y = bb[2].y + 1.23f * 0.9f;
  bb[3].z = bb[2].z;
  int nb = new_aa == 7 ? 4 : 5;
  if (nb > T->natoms[new_aa]) nb = T->natoms[new_aa];
  for (int k = 0; k < nb; ++k) s[k] = bb[k];
  aa[r] = (uint8_t)new_aa;
  rot[r] = (uint8_t)new_rot;
  rf_side_chain(T, new_aa, new_rot, nb, s);
}

/* the antigen Calpha grid as cell lists over the key bounding box */
typedef struct {
  int32_t kmin[3], dim[3];
  uint32_t* start; /* cells + 1 */
  uint32_t* idx;
  float inv;
} rf_grid;

static int rf_grid_build(rf_grid* g, const or_rcloud* ag) {
  g->inv = 1.0f / 10.0f;
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  int32_t* keys = (int32_t*)malloc(3 * (ag->n ? ag->n : 1) * sizeof(int32_t));
  if (!keys) return -1;
  for (size_t i = 0; i < ag->n; ++i) {
    const float* c[3] = {ag->x, ag->y, ag->z};
    for (int k = 0; k < 3; ++k) {
      keys[3 * i + k] = or_cell_of(c[k][i], g->inv);
      if (keys[3 * i + k] < lo[k]) lo[k] = keys[3 * i + k];
      if (keys[3 * i + k] > hi[k]) hi[k] = keys[3 * i + k];
    }
  }
  size_t cells = 1;
  for (int k = 0; k < 3; ++k) {
    g->kmin[k] = ag->n ? lo[k] : 0;
    g->dim[k] = ag->n ? hi[k] - lo[k] + 1 : 1;
    cells *= (size_t)g->dim[k];
  }
  g->start = (uint32_t*)calloc(cells + 1, sizeof(uint32_t));
  g->idx = (uint32_t*)malloc((ag->n ? ag->n : 1) * sizeof(uint32_t));
  if (!g->start || !g->idx) { free(keys); return -1; }
  for (size_t i = 0; i < ag->n; ++i) {
    size_t c = ((size_t)(keys[3 * i] - g->kmin[0]) * g->dim[1] + (keys[3 * i + 1] - g->kmin[1])) *
                   g->dim[2] + (keys[3 * i + 2] - g->kmin[2]);
    g->start[c + 1]++;
  }
  for (size_t c = 0; c < cells; ++c) g->start[c + 1] += g->start[c];
  uint32_t* fill = (uint32_t*)calloc(cells, sizeof(uint32_t));
  for (size_t i = 0; i < ag->n; ++i) {  /* ascending index within a cell (src/spatial.rs:35-38) */
    size_t c = ((size_t)(keys[3 * i] - g->kmin[0]) * g->dim[1] + (keys[3 * i + 1] - g->kmin[1])) *
                   g->dim[2] + (keys[3 * i + 2] - g->kmin[2]);
    g->idx[g->start[c] + fill[c]++] = (uint32_t)i;
  }
  free(fill);
  free(keys);
  return 0;
}

/* compute_forces_with_grid for one Calpha (src/energy.rs:322-398), the
 * reference's 27-cell order */
static void rf_cg_force(const rf_grid* g, const or_rcloud* ag, float bx, float by, float bz,
                        const float* cg, float* fx, float* fy, float* fz) {
  float ax = 0, ay = 0, az = 0;
  int32_t c0 = or_cell_of(bx, g->inv), c1 = or_cell_of(by, g->inv), c2 = or_cell_of(bz, g->inv);
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        int64_t k0 = (int64_t)wrap_add(c0, dx) - g->kmin[0], k1 = (int64_t)wrap_add(c1, dy) - g->kmin[1],
                k2 = (int64_t)wrap_add(c2, dz) - g->kmin[2];
        if (k0 < 0 || k1 < 0 || k2 < 0 || k0 >= g->dim[0] || k1 >= g->dim[1] || k2 >= g->dim[2]) continue;
        size_t c = ((size_t)k0 * g->dim[1] + (size_t)k1) * g->dim[2] + (size_t)k2;
        for (uint32_t k = g->start[c]; k < g->start[c + 1]; ++k)
          cg_pair(ag, g->idx[k], bx, by, bz, cg[0], cg[1], cg[2], (uint8_t)cg[3], &ax, &ay, &az);
      }
  *fx = ax;
  *fy = ay;
  *fz = az;
}

/* residue_contribution_allatom (src/diffusion.rs:422-456) */
static double rf_contribution(const or_rf_tables* T, const or_rcloud* ag, const uint8_t* ag_aa,
                              int aa, rf_v3 p) {
  const float* b = T->ca_par + (size_t)aa * 5;
  const float brm = b[0], be = b[1], bq = b[2];
  const int bh = (int)b[3], bt = (int)b[4];
  double e = 0.0;
  for (size_t i = 0; i < ag->n; ++i) {
    float dx = p.x - ag->x[i], dy = p.y - ag->y[i], dz = p.z - ag->z[i];
    float r_sq = dx * dx + dy * dy + dz * dz;
    if (r_sq > 100.0f || r_sq < 0.25f) continue;
    const float* a = T->ca_par + (size_t)ag_aa[i] * 5;
    float r_ij = brm + a[0];
    float eps = sqrtf(be * a[1]);
    float r2 = (r_ij * r_ij) / r_sq;
    float r6 = r2 * r2 * r2;
    e += (double)(eps * (r6 * r6 - 2.0f * r6));
    if (bq != 0.0f && a[2] != 0.0f) e += (double)(332.0f * bq * a[2] / sqrtf(r_sq));
    if (bh == 1 && (int)a[3] == 1 && r_sq < 36.0f) e -= 0.5;
    int at = (int)a[4];
    if ((T->hb_don[bt] && T->hb_acc[at]) || (T->hb_acc[bt] && T->hb_don[at])) {
      if (r_sq <= 25.0f) {
        float z = (sqrtf(r_sq) - 2.9f) / 0.35f;
        e += (double)(-1.5f * expf(-0.5f * z * z));
      }
    }
  }
  return e;
}

static double rf_disulfide(const uint8_t* aa, const rf_v3* s, int nr) {
  double e = 0.0;
  for (int i = 0; i < nr; ++i) {
    if (aa[i] != 4) continue;
    rf_v3 a = s[(size_t)i * RF_SLOTS + 5];
    for (int j = i + 1; j < nr; ++j) {
      if (aa[j] != 4) continue;
      rf_v3 b = s[(size_t)j * RF_SLOTS + 5];
      float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
      if (sqrtf(dx * dx + dy * dy + dz * dz) > 4.0f) continue;
      double ex = (double)a.x - b.x, ey = (double)a.y - b.y, ez = (double)a.z - b.z;
      double t = (sqrt(ex * ex + ey * ey + ez * ez) - (double)2.05f) / (double)0.4f;
      e -= 4.0 * exp(-0.5 * t * t);
    }
  }
  return e;
}

static void rf_u4(uint32_t it, uint32_t r, uint32_t tag, const or_rf_params* p, float u[4]) {
  uint32_t ctr[4] = {it, r, tag, 0u}, key[2] = {p->seed_lo, p->seed_hi}, o[4];
  or_philox4x32_10(ctr, key, o);
  for (int k = 0; k < 4; ++k) u[k] = (float)(o[k] >> 8) * (1.0f / 16777216.0f);
}

/* perturb_phi / perturb_psi (src/allatom.rs:392-456) with rodrigues (467-479) */
static void rf_torsion(const or_rf_tables* T, const uint8_t* aa, rf_v3* s, int r, int nr,
                       float rad, int phi) {
  const rf_v3* b = s + (size_t)r * RF_SLOTS;
  rf_v3 p0 = phi ? b[0] : b[1], p1 = phi ? b[1] : b[2];
  rf_v3 axis = rf_norm(rf_sub(p1, p0));
  float cs = (float)cos((double)rad), sn = (float)sin((double)rad);
  for (int rr = r; rr < nr; ++rr)
    for (int k = 0; k < T->natoms[aa[rr]]; ++k) {
      if (rr == r && (phi ? k < 2 : k != 3)) continue;
      rf_v3* q = s + (size_t)rr * RF_SLOTS + k;
      rf_v3 v = rf_sub(*q, p1);
      float dot = v.x * axis.x + v.y * axis.y + v.z * axis.z;
      rf_v3 cx = rf_cross(axis, v);
      float omc = 1.0f - cs;
      q->x = (v.x * cs + cx.x * sn + axis.x * dot * omc) + p1.x;
      q->y = (v.y * cs + cx.y * sn + axis.y * dot * omc) + p1.y;
      q->z = (v.z * cs + cx.z * sn + axis.z * dot * omc) + p1.z;
    }
}

typedef struct {
  const or_rf_tables* T;
  const or_rcloud* ag;
  const uint8_t* ag_aa;
  const rf_grid* grid;
  const uint32_t* cand_res_off;
  uint8_t* aa;
  uint8_t* rot;
  rf_v3* slots;
  const float* temp;
  const float* sigma;
  uint32_t n_steps, iter0;
  const or_rf_params* p;
  size_t c0, c1;
} rf_job;

static void rf_candidate(const rf_job* J, size_t c) {
  const or_rf_tables* T = J->T;
  const or_rf_params* p = J->p;
  const uint32_t g0 = J->cand_res_off[c];
  const int nr = (int)(J->cand_res_off[c + 1] - g0);
  uint8_t* aa = J->aa + g0;
  uint8_t* rot = J->rot + g0;
  rf_v3* s = J->slots + (size_t)g0 * RF_SLOTS;
  uint8_t* iface = (uint8_t*)malloc(nr ? nr : 1);
  float* F = (float*)malloc(3 * (nr ? nr : 1) * sizeof(float));
  for (uint32_t st = 0; st < J->n_steps; ++st) {
    const uint32_t it = J->iter0 + st;
    const float temp = J->temp[st], sigma = J->sigma[st];
    for (int r = 0; r < nr; ++r) {
      rf_v3 ca = s[(size_t)r * RF_SLOTS + 1];
      rf_cg_force(J->grid, J->ag, ca.x, ca.y, ca.z, T->cg_par + (size_t)aa[r] * 4, &F[3 * r],
                  &F[3 * r + 1], &F[3 * r + 2]);
    }
    for (int r = 0; r < nr; ++r) {  /* src/diffusion.rs:318-338 */
      rf_v3 ca = s[(size_t)r * RF_SLOTS + 1];
      float fx = F[3 * r], fy = F[3 * r + 1], fz = F[3 * r + 2];
      float dx_c = ca.x - p->center[0], dy_c = ca.y - p->center[1], dz_c = ca.z - p->center[2];
      float d = sqrtf(dx_c * dx_c + dy_c * dy_c + dz_c * dz_c);
      if (d < 0.001f) d = 0.001f;
      float f_r = -p->restraint_k * (d - p->restraint_radius) / d;
      fx += f_r * dx_c;
      fy += f_r * dy_c;
      fz += f_r * dz_c;
      float nx = 0, ny = 0, nz = 0;
      if (p->with_noise) {
        float u[4];
        rf_u4(it, g0 + (uint32_t)r, 0x6e6f6973u, p, u);
        float two = sigma + sigma;
        nx = u[0] * two - sigma;
        ny = u[1] * two - sigma;
        nz = u[2] * two - sigma;
      }
      float ddx = (ca.x + clampf_rust(p->step * fx, -p->max_disp, p->max_disp) + nx) - ca.x;
      float ddy = (ca.y + clampf_rust(p->step * fy, -p->max_disp, p->max_disp) + ny) - ca.y;
      float ddz = (ca.z + clampf_rust(p->step * fz, -p->max_disp, p->max_disp) + nz) - ca.z;
      for (int k = 0; k < T->natoms[aa[r]]; ++k) {
        rf_v3* q = s + (size_t)r * RF_SLOTS + k;
        q->x += ddx;
        q->y += ddy;
        q->z += ddz;
      }
    }
    for (int r = 0; r < nr; ++r) {  /* src/diffusion.rs:341-351 */
      rf_v3 ca = s[(size_t)r * RF_SLOTS + 1];
      int in = 0;
      for (size_t i = 0; i < J->ag->n && !in; ++i) {
        float dx = ca.x - J->ag->x[i], dy = ca.y - J->ag->y[i], dz = ca.z - J->ag->z[i];
        in = dx * dx + dy * dy + dz * dz < p->iface_sq;
      }
      iface[r] = (uint8_t)in;
    }
    if (p->moves & 1u) {  /* src/diffusion.rs:353-398 */
      for (int r = 0; r < nr; ++r) {
        float u[4];
        rf_u4(it, g0 + (uint32_t)r, 0x6d636d76u, p, u);
        float eff_mut = iface[r] ? p->mut_p * 3.0f : p->mut_p * 0.4f;
        float eff_rot = iface[r] ? p->rot_p * 2.0f : p->rot_p;
        if (!(u[0] < eff_mut)) continue;
        int old_aa = aa[r], old_rot = rot[r];
        rf_v3* sr = s + (size_t)r * RF_SLOTS;
        if (u[1] < eff_rot) {
          int nrot = T->nrot[old_aa];
          if (!nrot) continue;
          int nrt = (int)(u[2] * (float)nrot);
          if (nrt > nrot - 1) nrt = nrot - 1;
          double old_e = rf_contribution(T, J->ag, J->ag_aa, old_aa, sr[1]) + rf_disulfide(aa, s, nr);
          rot[r] = (uint8_t)nrt;
          rf_side_chain(T, old_aa, nrt, old_aa == 7 ? 4 : 5, sr);
          double new_e = rf_contribution(T, J->ag, J->ag_aa, old_aa, sr[1]) + rf_disulfide(aa, s, nr);
          double delta = new_e - old_e;
          if (!(delta <= 0.0 || (double)u[3] < exp(-delta / temp))) {
            rot[r] = (uint8_t)old_rot;
            rf_side_chain(T, old_aa, old_rot, old_aa == 7 ? 4 : 5, sr);
          }
        } else {
          int new_aa = (int)(u[2] * 20.0f);
          if (new_aa > 19) new_aa = 19;
          double old_e = rf_contribution(T, J->ag, J->ag_aa, old_aa, sr[1]) + rf_disulfide(aa, s, nr);
          rf_mutate(T, aa, rot, r, new_aa, 0, s);
          double new_e = rf_contribution(T, J->ag, J->ag_aa, aa[r], sr[1]) + rf_disulfide(aa, s, nr);
          double delta = new_e - old_e;
          if (!(delta <= 0.0 || (double)u[3] < exp(-delta / temp)))
            rf_mutate(T, aa, rot, r, old_aa, old_rot, s);
        }
      }
    }
    if (p->moves & 2u) {  /* src/diffusion.rs:400-415 */
      for (int r = 0; r < nr; ++r) {
        float u[4];
        rf_u4(it, g0 + (uint32_t)r, 0x62626d76u, p, u);
        if (!(u[0] < p->bb_p)) continue;
        float rad = rf_rad(-p->max_torsion + (p->max_torsion + p->max_torsion) * u[1]);
        int phi = u[2] < 0.5f;
        double e[2];
        for (int pass = 0; pass < 2; ++pass) {
          double sum = 0.0;
          for (int i = r; i < nr; ++i)
            sum += rf_contribution(T, J->ag, J->ag_aa, aa[i], s[(size_t)i * RF_SLOTS + 1]);
          e[pass] = sum + rf_disulfide(aa, s, nr);
          if (pass == 0) rf_torsion(T, aa, s, r, nr, rad, phi);
        }
        double delta = e[1] - e[0];
        if (!(delta <= 0.0 || (double)u[3] < exp(-delta / temp))) rf_torsion(T, aa, s, r, nr, -rad, phi);
      }
    }
  }
  free(iface);
  free(F);
}

static void* rf_worker(void* arg) {
  const rf_job* J = (const rf_job*)arg;
  for (size_t c = J->c0; c < J->c1; ++c) rf_candidate(J, c);
  return NULL;
}

/* Refines n_cand candidates in place: slots (n_res x 14 x 3 floats), residue
 * types and rotamers; temp/sigma per step. Returns 0 or -1. */
int or_refine_run(const or_rf_tables* T, const or_rcloud* ag, const uint8_t* ag_aa,
                  size_t n_cand, const uint32_t* cand_res_off, uint8_t* res_aa, uint8_t* res_rot,
                  float* slots, const float* temp, const float* sigma, uint32_t n_steps,
                  uint32_t iter0, const or_rf_params* p, int nthreads) {
  rf_grid g;
  memset(&g, 0, sizeof(g));
  if (rf_grid_build(&g, ag)) return -1;
  if (nthreads < 1) nthreads = 1;
  if ((size_t)nthreads > n_cand) nthreads = n_cand ? (int)n_cand : 1;
  pthread_t th[256];
  rf_job jobs[256];
  if (nthreads > 256) nthreads = 256;
  size_t per = (n_cand + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    rf_job J = {T, ag, ag_aa, &g, cand_res_off, res_aa, res_rot, (rf_v3*)slots, temp, sigma,
                n_steps, iter0, p, 0, 0};
    J.c0 = (size_t)k * per < n_cand ? (size_t)k * per : n_cand;
    J.c1 = J.c0 + per < n_cand ? J.c0 + per : n_cand;
    jobs[k] = J;
    pthread_create(&th[k], NULL, rf_worker, &jobs[k]);
  }
  for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
  free(g.start);
  free(g.idx);
  return 0;
}
