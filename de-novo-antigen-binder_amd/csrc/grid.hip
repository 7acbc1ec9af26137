// SpatialHashGrid on the device (src/spatial.rs:12-68).
//
// build: cell keys with the reference's f32 arithmetic (x * fl(1/cell),
// floor, saturating cast), a counting sort over Morton codes of
// (key - key_min), and a per-cell insertion sort so each cell lists its
// points in ascending original index, the order in which the reference's
// `build` appends them (src/spatial.rs:35-38).
// query: the 27 cells dx -> dy -> dz nested (src/spatial.rs:46-56), every
// index of every existing cell, no distance test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include <cmath>

#include "device_common.h"
#include "dnab_internal.h"
#include "grid_device.h"

namespace dnab {

constexpr int kMaxMortonBits = 8;

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_bbox_init(int32_t* bb) {
  if (threadIdx.x < 3) bb[threadIdx.x] = INT32_MAX;
  else if (threadIdx.x < 6) bb[threadIdx.x] = INT32_MIN;
}

__global__ void k_grid_keys(const float* __restrict__ x, const float* __restrict__ y,
                            const float* __restrict__ z, uint32_t n, float inv,
                            int32_t* __restrict__ keys, int32_t* __restrict__ bb) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t k0 = cell_key(x[i], inv), k1 = cell_key(y[i], inv), k2 = cell_key(z[i], inv);
  keys[3 * i] = k0;
  keys[3 * i + 1] = k1;
  keys[3 * i + 2] = k2;
  atomicMin(&bb[0], k0);
  atomicMin(&bb[1], k1);
  atomicMin(&bb[2], k2);
  atomicMax(&bb[3], k0);
  atomicMax(&bb[4], k1);
  atomicMax(&bb[5], k2);
}

__device__ inline uint32_t grid_code(const int32_t k[3], const int32_t kmin[3]) {
  return morton3(static_cast<uint32_t>(k[0] - kmin[0]), static_cast<uint32_t>(k[1] - kmin[1]),
                 static_cast<uint32_t>(k[2] - kmin[2]));
}

__global__ void k_grid_count(const int32_t* __restrict__ keys, uint32_t n, int32_t kmin0,
                             int32_t kmin1, int32_t kmin2, uint32_t* __restrict__ count,
                             uint32_t* __restrict__ rank) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t kmin[3] = {kmin0, kmin1, kmin2};
  uint32_t code = grid_code(&keys[3 * i], kmin);
  rank[i] = atomicAdd(&count[code], 1u);
}

// Exclusive scan of data[0..n) in place by one workgroup of 1024 threads;
// data[n] receives the total; *nonzero the number of non-zero entries.
__global__ void __launch_bounds__(1024) k_scan_excl(uint32_t* data, uint64_t n,
                                                    uint32_t* nonzero) {
  __shared__ uint32_t sums[1024];
  __shared__ uint32_t nz[1024];
  const uint32_t t = threadIdx.x;
  const uint64_t chunk = (n + 1023) / 1024;
  const uint64_t b = t * chunk, e = min(n, b + chunk);
  uint32_t s = 0, c = 0;
  for (uint64_t i = b; i < e; ++i) {
    uint32_t v = data[i];
    s += v;
    c += v != 0;
  }
  sums[t] = s;
  nz[t] = c;
  __syncthreads();
  // Hillis-Steele inclusive scan over the 1024 partial sums.
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    uint32_t v = t >= off ? sums[t - off] : 0u;
    uint32_t w = t >= off ? nz[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    nz[t] += w;
    __syncthreads();
  }
  uint32_t run = t ? sums[t - 1] : 0u;
  for (uint64_t i = b; i < e; ++i) {
    uint32_t v = data[i];
    data[i] = run;
    run += v;
  }
  if (t == 1023) {
    data[n] = sums[1023];
    if (nonzero) *nonzero = nz[1023];
  }
}

__global__ void k_grid_scatter(const int32_t* __restrict__ keys, const uint32_t* __restrict__ rank,
                               uint32_t n, int32_t kmin0, int32_t kmin1, int32_t kmin2,
                               const uint32_t* __restrict__ start, uint32_t* __restrict__ sorted) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t kmin[3] = {kmin0, kmin1, kmin2};
  uint32_t code = grid_code(&keys[3 * i], kmin);
  sorted[start[code] + rank[i]] = i;
}

// Restore ascending original index inside each cell (atomic ranks are in
// arrival order).
__global__ void k_grid_cell_sort(const uint32_t* __restrict__ start, uint64_t ncodes,
                                 uint32_t* __restrict__ sorted) {
  uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (c >= ncodes) return;
  uint32_t b = start[c], e = start[c + 1];
  for (uint32_t i = b + 1; i < e; ++i) {
    uint32_t v = sorted[i];
    uint32_t j = i;
    while (j > b && sorted[j - 1] > v) {
      sorted[j] = sorted[j - 1];
      --j;
    }
    sorted[j] = v;
  }
}

static inline uint32_t blocks_for(uint64_t n, uint32_t t) {
  uint64_t b = (n + t - 1) / t;
  return static_cast<uint32_t>(b ? b : 1);
}

int grid_build_device(dnab_ctx* ctx, GridDev& g, const float* d_x, const float* d_y,
                      const float* d_z, size_t n) {
  if (n > 0xFFFFFFF0ull) return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "grid: too many points");
  g.n = n;
  g.n_nonempty = 0;
  DevBuf<int32_t> bb;
  DNAB_ALLOC_CHECK(ctx, bb.alloc(6));
  DNAB_ALLOC_CHECK(ctx, g.keys.alloc(3 * n));
  hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, ctx->stream, bb.p);
  if (n) {
    hipLaunchKernelGGL(k_grid_keys, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, d_x,
                       d_y, d_z, static_cast<uint32_t>(n), g.inv, g.keys.p, bb.p);
  }
  int32_t hbb[6];
  DNAB_RETURN_IF(download(ctx, hbb, bb.p, sizeof(hbb)));
  if (n == 0) {
    for (int a = 0; a < 3; ++a) {
      g.kmin[a] = 0;
      g.kmax[a] = -1;
    }
    g.bits = 0;
    g.ncodes = 1;
    DNAB_ALLOC_CHECK(ctx, g.cell_start.alloc(2));
    DNAB_ALLOC_CHECK(ctx, g.sorted_idx.alloc(1));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, g.cell_start.p,
                       static_cast<uint64_t>(2), 0u);
    DNAB_HIP_CHECK(ctx, hipGetLastError());
    return DNAB_OK;
  }
  int64_t ext_max = 1;
  for (int a = 0; a < 3; ++a) {
    g.kmin[a] = hbb[a];
    g.kmax[a] = hbb[3 + a];
    int64_t ext = static_cast<int64_t>(hbb[3 + a]) - hbb[a] + 1;
    ext_max = std::max(ext_max, ext);
  }
  int bits = 0;
  while ((int64_t(1) << bits) < ext_max) ++bits;
  if (bits > kMaxMortonBits)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED,
                          "grid: key extent %lld cells exceeds 2^%d per axis (cell %g)",
                          static_cast<long long>(ext_max), kMaxMortonBits, g.cell);
  g.bits = bits;
  g.ncodes = uint64_t(1) << (3 * bits);
  DNAB_ALLOC_CHECK(ctx, g.cell_start.alloc(g.ncodes + 1));
  DNAB_ALLOC_CHECK(ctx, g.sorted_idx.alloc(n));
  DevBuf<uint32_t> rank, nz;
  DNAB_ALLOC_CHECK(ctx, rank.alloc(n));
  DNAB_ALLOC_CHECK(ctx, nz.alloc(1));
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(g.ncodes + 1, 256)), dim3(256), 0, ctx->stream,
                     g.cell_start.p, g.ncodes + 1, 0u);
  hipLaunchKernelGGL(k_grid_count, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, g.keys.p,
                     static_cast<uint32_t>(n), g.kmin[0], g.kmin[1], g.kmin[2], g.cell_start.p,
                     rank.p);
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, ctx->stream, g.cell_start.p, g.ncodes,
                     nz.p);
  hipLaunchKernelGGL(k_grid_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                     g.keys.p, rank.p, static_cast<uint32_t>(n), g.kmin[0], g.kmin[1], g.kmin[2],
                     g.cell_start.p, g.sorted_idx.p);
  hipLaunchKernelGGL(k_grid_cell_sort, dim3(blocks_for(g.ncodes, 256)), dim3(256), 0,
                     ctx->stream, g.cell_start.p, g.ncodes, g.sorted_idx.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  uint32_t hnz = 0;
  DNAB_RETURN_IF(download(ctx, &hnz, nz.p, sizeof(hnz)));
  g.n_nonempty = hnz;
  return DNAB_OK;
}

// ---- query -------------------------------------------------------------------

__global__ void k_grid_query_count(GridView g, const float* __restrict__ px,
                                   const float* __restrict__ py, const float* __restrict__ pz,
                                   uint32_t n, uint32_t* __restrict__ count) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t cx = cell_key(px[p], g.inv), cy = cell_key(py[p], g.inv), cz = cell_key(pz[p], g.inv);
  uint32_t c = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        uint32_t b, e;
        if (cell_range(g, wrap_add(cx, dx), wrap_add(cy, dy), wrap_add(cz, dz), b, e)) c += e - b;
      }
  count[p] = c;
}

__global__ void k_grid_query_fill(GridView g, const float* __restrict__ px,
                                  const float* __restrict__ py, const float* __restrict__ pz,
                                  uint32_t n, const uint32_t* __restrict__ offsets,
                                  uint32_t* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t cx = cell_key(px[p], g.inv), cy = cell_key(py[p], g.inv), cz = cell_key(pz[p], g.inv);
  uint32_t o = offsets[p];
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        uint32_t b, e;
        if (cell_range(g, wrap_add(cx, dx), wrap_add(cy, dy), wrap_add(cz, dz), b, e))
          for (uint32_t k = b; k < e; ++k) out[o++] = g.sorted_idx[k];
      }
}

__global__ void k_cell_of(const float* __restrict__ x, const float* __restrict__ y,
                          const float* __restrict__ z, uint32_t n, float inv,
                          int32_t* __restrict__ keys) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[3 * i] = cell_key(x[i], inv);
  keys[3 * i + 1] = cell_key(y[i], inv);
  keys[3 * i + 2] = cell_key(z[i], inv);
}

// ---- residue-level forces (src/energy.rs:154-234, 322-398) -------------------
// One thread per antibody point; the pair loop visits antigen points in the
// reference's order and evaluates every operation in the reference's order
// with IEEE f32 (correctly rounded division and sqrt, no contraction), so the
// accumulated forces reproduce the reference's f32 arithmetic.

__global__ void k_cg_forces_grid(GridView g, CgCloud ag, const float* __restrict__ bx_,
                                 const float* __restrict__ by_, const float* __restrict__ bz_,
                                 const float* __restrict__ bq_, const float* __restrict__ be_,
                                 const float* __restrict__ bs_, const uint8_t* __restrict__ bh_,
                                 uint32_t n, float* __restrict__ fx, float* __restrict__ fy,
                                 float* __restrict__ fz) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float bx = bx_[j], by = by_[j], bz = bz_[j], bq = bq_[j], be = be_[j], bs = bs_[j];
  uint8_t bh = bh_[j];
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  int32_t cx = cell_key(bx, g.inv), cy = cell_key(by, g.inv), cz = cell_key(bz, g.inv);
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        uint32_t b, e;
        if (!cell_range(g, wrap_add(cx, dx), wrap_add(cy, dy), wrap_add(cz, dz), b, e)) continue;
        for (uint32_t k = b; k < e; ++k)
          cg_pair(ag, g.sorted_idx[k], bx, by, bz, bq, be, bs, bh, ax, ay, az);
      }
  fx[j] = ax;
  fy[j] = ay;
  fz[j] = az;
}

__global__ void k_cg_forces_brute(CgCloud ag, uint32_t n_ag, const float* __restrict__ bx_,
                                  const float* __restrict__ by_, const float* __restrict__ bz_,
                                  const float* __restrict__ bq_, const float* __restrict__ be_,
                                  const float* __restrict__ bs_, const uint8_t* __restrict__ bh_,
                                  uint32_t n, float* __restrict__ fx, float* __restrict__ fy,
                                  float* __restrict__ fz) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float bx = bx_[j], by = by_[j], bz = bz_[j], bq = bq_[j], be = be_[j], bs = bs_[j];
  uint8_t bh = bh_[j];
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  for (uint32_t i = 0; i < n_ag; ++i) cg_pair(ag, i, bx, by, bz, bq, be, bs, bh, ax, ay, az);
  fx[j] = ax;
  fy[j] = ay;
  fz[j] = az;
}

}  // namespace dnab

using namespace dnab;

namespace {

struct Staged {
  DevBuf<float> x, y, z;
};

int stage_xyz(dnab_ctx* ctx, Staged& s, const float* x, const float* y, const float* z,
              size_t n) {
  DNAB_ALLOC_CHECK(ctx, s.x.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.y.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.z.alloc(n));
  DNAB_RETURN_IF(upload(ctx, s.x.p, x, n * sizeof(float)));
  DNAB_RETURN_IF(upload(ctx, s.y.p, y, n * sizeof(float)));
  DNAB_RETURN_IF(upload(ctx, s.z.p, z, n * sizeof(float)));
  return DNAB_OK;
}

struct StagedCg {
  DevBuf<float> x, y, z, q, e, s;
  DevBuf<uint8_t> h;
  CgCloud view() const { return CgCloud{x.p, y.p, z.p, q.p, e.p, s.p, h.p}; }
};

int stage_cg(dnab_ctx* ctx, StagedCg& s, const dnab_residue_cloud* c) {
  size_t n = c->n;
  DNAB_ALLOC_CHECK(ctx, s.x.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.y.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.z.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.q.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.e.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.s.alloc(n));
  DNAB_ALLOC_CHECK(ctx, s.h.alloc(n));
  DNAB_RETURN_IF(upload(ctx, s.x.p, c->x, n * 4));
  DNAB_RETURN_IF(upload(ctx, s.y.p, c->y, n * 4));
  DNAB_RETURN_IF(upload(ctx, s.z.p, c->z, n * 4));
  DNAB_RETURN_IF(upload(ctx, s.q.p, c->charge, n * 4));
  DNAB_RETURN_IF(upload(ctx, s.e.p, c->epsilon, n * 4));
  DNAB_RETURN_IF(upload(ctx, s.s.p, c->sigma, n * 4));
  DNAB_RETURN_IF(upload(ctx, s.h.p, c->hydrophobic, n));
  return DNAB_OK;
}

bool cg_cloud_ok(const dnab_residue_cloud* c) {
  return c && (c->n == 0 || (c->x && c->y && c->z && c->charge && c->epsilon && c->sigma &&
                             c->hydrophobic));
}

}  // namespace

extern "C" {

int dnab_grid_create(dnab_ctx* ctx, float cell_size, dnab_grid** out) {
  if (!ctx || !out) return DNAB_ERR_INVALID;
  if (!(cell_size > 0.0f) || !std::isfinite(cell_size))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "grid: cell_size must be positive and finite");
  dnab_grid* g = new (std::nothrow) dnab_grid();
  if (!g) return DNAB_ERR_ALLOC;
  g->ctx = ctx_retain(ctx);
  g->g.cell = cell_size;
  g->g.inv = 1.0f / cell_size;  // src/spatial.rs:26, IEEE f32 division on the host
  *out = g;
  return DNAB_OK;
}

void dnab_grid_destroy(dnab_grid* g) {
  if (!g) return;
  dnab_ctx* ctx = g->ctx;
  delete g;
  ctx_release(ctx);
}

int dnab_grid_build(dnab_ctx* ctx, dnab_grid* grid, const float* x, const float* y,
                    const float* z, size_t n) {
  if (!ctx || !grid || (n && (!x || !y || !z))) return DNAB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  GridDev& g = grid->g;
  DNAB_ALLOC_CHECK(ctx, g.x.alloc(n));
  DNAB_ALLOC_CHECK(ctx, g.y.alloc(n));
  DNAB_ALLOC_CHECK(ctx, g.z.alloc(n));
  DNAB_RETURN_IF(upload(ctx, g.x.p, x, n * 4));
  DNAB_RETURN_IF(upload(ctx, g.y.p, y, n * 4));
  DNAB_RETURN_IF(upload(ctx, g.z.p, z, n * 4));
  return grid_build_device(ctx, g, g.x.p, g.y.p, g.z.p, n);
}

int dnab_grid_info(const dnab_grid* grid, size_t* n_points, size_t* n_cells_nonempty,
                   int32_t key_min[3], int32_t key_max[3]) {
  if (!grid) return DNAB_ERR_INVALID;
  if (n_points) *n_points = grid->g.n;
  if (n_cells_nonempty) *n_cells_nonempty = grid->g.n_nonempty;
  for (int a = 0; a < 3; ++a) {
    if (key_min) key_min[a] = grid->g.kmin[a];
    if (key_max) key_max[a] = grid->g.kmax[a];
  }
  return DNAB_OK;
}

int dnab_grid_cell_of(dnab_ctx* ctx, const dnab_grid* grid, const float* x, const float* y,
                      const float* z, size_t n, int32_t* keys) {
  if (!ctx || !grid || (n && (!x || !y || !z || !keys))) return DNAB_ERR_INVALID;
  if (n == 0) return DNAB_OK;
  (void)hipSetDevice(ctx->device);
  Staged s;
  DNAB_RETURN_IF(stage_xyz(ctx, s, x, y, z, n));
  DevBuf<int32_t> k;
  DNAB_ALLOC_CHECK(ctx, k.alloc(3 * n));
  hipLaunchKernelGGL(k_cell_of, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, s.x.p, s.y.p,
                     s.z.p, static_cast<uint32_t>(n), grid->g.inv, k.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return download(ctx, keys, k.p, 3 * n * sizeof(int32_t));
}

int dnab_grid_query(dnab_ctx* ctx, const dnab_grid* grid, const float* px, const float* py,
                    const float* pz, size_t n_probe, uint32_t* offsets, uint32_t* indices,
                    size_t capacity, size_t* n_total) {
  if (!ctx || !grid || !offsets || (n_probe && (!px || !py || !pz))) return DNAB_ERR_INVALID;
  if (n_probe > 0xFFFFFFF0ull) return DNAB_ERR_UNSUPPORTED;
  (void)hipSetDevice(ctx->device);
  if (n_probe == 0) {
    offsets[0] = 0;
    if (n_total) *n_total = 0;
    return DNAB_OK;
  }
  Staged s;
  DNAB_RETURN_IF(stage_xyz(ctx, s, px, py, pz, n_probe));
  DevBuf<uint32_t> off;
  DNAB_ALLOC_CHECK(ctx, off.alloc(n_probe + 1));
  GridView gv = make_view(grid->g);
  if (grid->g.n == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n_probe + 1, 256)), dim3(256), 0, ctx->stream,
                       off.p, static_cast<uint64_t>(n_probe + 1), 0u);
  } else {
    hipLaunchKernelGGL(k_grid_query_count, dim3(blocks_for(n_probe, 256)), dim3(256), 0,
                       ctx->stream, gv, s.x.p, s.y.p, s.z.p, static_cast<uint32_t>(n_probe),
                       off.p);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, ctx->stream, off.p,
                       static_cast<uint64_t>(n_probe), nullptr);
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(download(ctx, offsets, off.p, (n_probe + 1) * sizeof(uint32_t)));
  size_t total = offsets[n_probe];
  if (n_total) *n_total = total;
  if (!indices || capacity < total) return DNAB_ERR_CAPACITY;
  if (total == 0) return DNAB_OK;
  DevBuf<uint32_t> out;
  DNAB_ALLOC_CHECK(ctx, out.alloc(total));
  hipLaunchKernelGGL(k_grid_query_fill, dim3(blocks_for(n_probe, 256)), dim3(256), 0, ctx->stream,
                     gv, s.x.p, s.y.p, s.z.p, static_cast<uint32_t>(n_probe), off.p, out.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return download(ctx, indices, out.p, total * sizeof(uint32_t));
}

int dnab_cg_forces_with_grid(dnab_ctx* ctx, const dnab_residue_cloud* antigen,
                             const dnab_grid* grid, const dnab_residue_cloud* antibody,
                             float* fx, float* fy, float* fz) {
  if (!ctx || !grid || !cg_cloud_ok(antigen) || !cg_cloud_ok(antibody)) return DNAB_ERR_INVALID;
  if (antibody->n && (!fx || !fy || !fz)) return DNAB_ERR_INVALID;
  if (grid->g.n != antigen->n)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "grid has %zu points, antigen %zu", grid->g.n,
                          antigen->n);
  size_t n = antibody->n;
  if (n == 0) return DNAB_OK;
  (void)hipSetDevice(ctx->device);
  StagedCg ag, ab;
  DNAB_RETURN_IF(stage_cg(ctx, ag, antigen));
  DNAB_RETURN_IF(stage_cg(ctx, ab, antibody));
  DevBuf<float> ox, oy, oz;
  DNAB_ALLOC_CHECK(ctx, ox.alloc(n));
  DNAB_ALLOC_CHECK(ctx, oy.alloc(n));
  DNAB_ALLOC_CHECK(ctx, oz.alloc(n));
  hipLaunchKernelGGL(k_cg_forces_grid, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                     make_view(grid->g), ag.view(), ab.x.p, ab.y.p, ab.z.p, ab.q.p, ab.e.p, ab.s.p,
                     ab.h.p, static_cast<uint32_t>(n), ox.p, oy.p, oz.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(download(ctx, fx, ox.p, n * 4));
  DNAB_RETURN_IF(download(ctx, fy, oy.p, n * 4));
  return download(ctx, fz, oz.p, n * 4);
}

int dnab_cg_forces(dnab_ctx* ctx, const dnab_residue_cloud* antigen,
                   const dnab_residue_cloud* antibody, float* fx, float* fy, float* fz) {
  if (!ctx || !cg_cloud_ok(antigen) || !cg_cloud_ok(antibody)) return DNAB_ERR_INVALID;
  if (antibody->n && (!fx || !fy || !fz)) return DNAB_ERR_INVALID;
  size_t n = antibody->n;
  if (n == 0) return DNAB_OK;
  (void)hipSetDevice(ctx->device);
  StagedCg ag, ab;
  DNAB_RETURN_IF(stage_cg(ctx, ag, antigen));
  DNAB_RETURN_IF(stage_cg(ctx, ab, antibody));
  DevBuf<float> ox, oy, oz;
  DNAB_ALLOC_CHECK(ctx, ox.alloc(n));
  DNAB_ALLOC_CHECK(ctx, oy.alloc(n));
  DNAB_ALLOC_CHECK(ctx, oz.alloc(n));
  hipLaunchKernelGGL(k_cg_forces_brute, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                     ag.view(), static_cast<uint32_t>(antigen->n), ab.x.p, ab.y.p, ab.z.p, ab.q.p,
                     ab.e.p, ab.s.p, ab.h.p, static_cast<uint32_t>(n), ox.p, oy.p, oz.p);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  DNAB_RETURN_IF(download(ctx, fx, ox.p, n * 4));
  DNAB_RETURN_IF(download(ctx, fy, oy.p, n * 4));
  return download(ctx, fz, oz.p, n * 4);
}

}  // extern "C"
