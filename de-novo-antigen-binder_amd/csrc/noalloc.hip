// dnab_score_batch_noalloc: score_batch on device-resident inputs with no
// allocation, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The per-candidate sums live in `scratch`, a file-scope device array of a
// fixed size (kNfMaxCand candidates); the library takes its address once per
// device with hipGetSymbolAddress when a context is created on that device
// (noalloc_init_device), and the entry point uses only that address. The
// array is shared by every call on one GPU and keeps the last call's
// contents: each call clears the part it uses with a fill kernel (no
// memset, so the call can be captured in a graph). Calls that overlap on
// one GPU (two contexts, or two host threads) race on it.
//
// Work: tiles of 64 consecutive input atoms, one wave each, walk the antigen
// grid cells within the cutoff of the tile's box and sum every (atom, chunk)
// pair set with chunk_pairs (pair_device.h), the evaluation path with its
// FP64 clash pairs. An atom's integer sum is therefore the one the
// population kernels form (a chunk is fixed by the antigen), and a
// candidate's exact integer total gives the same double as the population
// reduction whenever it is below 2^53 units (8.6e9 kcal/mol).
#include <hip/hip_runtime.h>

#include <mutex>

#include "device_common.h"
#include "dnab_internal.h"
#include "pair_device.h"

namespace dnab {

constexpr uint32_t kNfMaxCand = 65536;
constexpr size_t kNfScratchBytes = static_cast<size_t>(kNfMaxCand) * (8 + 4);

// [kNfMaxCand] int64 energy sums (units of 2^-20), then [kNfMaxCand] uint32
// status words (pair_device.h kSt*)
static __device__ __align__(256) unsigned char scratch[kNfScratchBytes];

constexpr int kNfMaxDevices = 64;
static std::mutex g_nf_mu;
static void* g_nf_scratch[kNfMaxDevices] = {};

// The antigen fields the kernel reads.
struct NfView {
  const AgAtom* atoms;
  const uint32_t* cell_start;
  int32_t kmin[3], kmax[3];
  float inv, rc, rc2;
  uint32_t n_atoms;
};

__global__ void k_nf_clear(long long* __restrict__ sum, uint32_t* __restrict__ st, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    sum[i] = 0;
    st[i] = 0;
  }
}

__device__ inline long long wave_sum_i64(long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline uint32_t wave_or_u32(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// One wave per tile of 64 consecutive input atoms (candidate-major GpuAtom
// records); four tiles per workgroup.
__global__ void __launch_bounds__(256)
k_nf_pairs(NfView V, const dnab_gpu_atom* __restrict__ cands, uint32_t n_used, uint32_t n_ab,
           long long* __restrict__ sum, uint32_t* __restrict__ status) {
  __shared__ float4 s_atoms[4][kWave][2];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t t = blockIdx.x * 4 + wave;
  if (t * kWave >= n_used) return;  // whole wave; no workgroup barriers below
  const uint32_t i = t * kWave + lane;
  const bool live = i < n_used;
  float4 A = make_float4(0.f, 0.f, 0.f, 0.f), B = A;
  if (live) {
    A = *reinterpret_cast<const float4*>(&cands[i].x);          // x, y, z, q
    B = *reinterpret_cast<const float4*>(&cands[i].r_min_half);  // rmh, eps, hydrophobic, pad
  }
  const bool nan = live && (isnan(A.x) || isnan(A.y) || isnan(A.z));
  const bool valid = live && isfinite(A.x) && isfinite(A.y) && isfinite(A.z);
  // an atom with an infinite coordinate has r^2 = +inf with every antigen
  // atom and no pair; one with a NaN coordinate has NaN pairs (below)
  const float4 P = valid ? A : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 Q = valid ? B : make_float4(0.f, 0.f, 0.f, 0.f);
  const float xb = P.x, yb = P.y, zb = P.z, qb = P.w;
  const float rmhb = Q.x, epsb = Q.y;
  const float sqeb = sqrtf(epsb);
  const float sqe12b = 12.0f * sqeb;
  const bool hydb = Q.z > 0.5f;  // shaders/energy.wgsl:91
  const float rc2_lane = valid ? V.rc2 : -1.0f;

  const float inf = __builtin_inff();
  const float lox = wave_reduce_f32<OpMin>(valid ? xb : inf);
  const float loy = wave_reduce_f32<OpMin>(valid ? yb : inf);
  const float loz = wave_reduce_f32<OpMin>(valid ? zb : inf);
  const float hix = wave_reduce_f32<OpMax>(valid ? xb : -inf);
  const float hiy = wave_reduce_f32<OpMax>(valid ? yb : -inf);
  const float hiz = wave_reduce_f32<OpMax>(valid ? zb : -inf);
  const bool any_valid = ballot64(valid) != 0;
  const float margin = V.rc * 1.001f + 1e-3f;
  const int32_t kx0 = max(cell_key(lox - margin, V.inv), V.kmin[0]);
  const int32_t ky0 = max(cell_key(loy - margin, V.inv), V.kmin[1]);
  const int32_t kz0 = max(cell_key(loz - margin, V.inv), V.kmin[2]);
  const int32_t kx1 = min(cell_key(hix + margin, V.inv), V.kmax[0]);
  const int32_t ky1 = min(cell_key(hiy + margin, V.inv), V.kmax[1]);
  const int32_t kz1 = min(cell_key(hiz + margin, V.inv), V.kmax[2]);
  const int32_t nx = kx1 - kx0 + 1, ny = ky1 - ky0 + 1, nz = kz1 - kz0 + 1;
  const int32_t ncell = (any_valid && nx > 0 && ny > 0 && nz > 0) ? nx * ny * nz : 0;

  FixAcc acc;
  uint32_t n_in = 0, n_tested = 0;
  for (int32_t cb0 = 0; cb0 < ncell; cb0 += kWave) {
    // lane = a cell of this batch, in (kz, ky, kx) order
    const int32_t c = cb0 + static_cast<int32_t>(lane);
    uint32_t cb = 0, ce = 0;
    if (c < ncell) {
      const int32_t iz = c / (nx * ny), rem = c - iz * (nx * ny);
      const int32_t iy = rem / nx, ix = rem - iy * nx;
      const uint32_t code = morton3(static_cast<uint32_t>(kx0 + ix - V.kmin[0]),
                                    static_cast<uint32_t>(ky0 + iy - V.kmin[1]),
                                    static_cast<uint32_t>(kz0 + iz - V.kmin[2]));
      cb = V.cell_start[code];
      ce = V.cell_start[code + 1];
    }
    for (uint64_t cm = ballot64(ce > cb); cm; cm &= cm - 1) {
      const int cl = static_cast<int>(__builtin_ctzll(cm));
      const uint32_t c0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cb), cl));
      const uint32_t c1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ce), cl));
      for (uint32_t f = c0; f < c1; f += kWave) {  // the cell's chunks
        const uint32_t e = min(f + kWave, c1);
        float4 L0 = make_float4(0.f, 0.f, 0.f, 0.f), L1 = L0;
        bool near = false;
        if (f + lane < e) {
          L0 = *reinterpret_cast<const float4*>(&V.atoms[f + lane].x);
          L1 = *reinterpret_cast<const float4*>(&V.atoms[f + lane].rmh);
          const float bx = fmaxf(fmaxf(lox - L0.x, L0.x - hix), 0.0f);
          const float by = fmaxf(fmaxf(loy - L0.y, L0.y - hiy), 0.0f);
          const float bz = fmaxf(fmaxf(loz - L0.z, L0.z - hiz), 0.0f);
          near = ((bx * bx + by * by) + bz * bz) <= V.rc2;
        }
        const uint64_t mask = ballot64(near);
        if (!mask) continue;
        stage_chunk(s_atoms[wave], mask, lane, L0, L1);
        chunk_pairs<false, false, true>(s_atoms[wave], static_cast<uint32_t>(__builtin_popcountll(mask)),
                                        xb, yb, zb, qb, rmhb, sqeb, sqe12b, epsb, hydb, rc2_lane,
                                        valid, acc, n_in, n_tested);
      }
    }
  }
  if (nan && V.n_atoms > 0) acc.st |= kStNaN;

  // per candidate in the tile: one exact integer sum and one status word
  const uint32_t cand = live ? i / n_ab : 0u;
  for (uint64_t todo = ballot64(live); todo;) {
    const int l = static_cast<int>(__builtin_ctzll(todo));
    const uint32_t cc = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cand), l));
    const bool mine = live && cand == cc;
    const uint64_t m = ballot64(mine);
    const long long se = wave_sum_i64(mine ? acc.e : 0);
    const uint32_t so = wave_or_u32(mine ? acc.st : 0u);
    if (lane == static_cast<uint32_t>(l)) {
      if (se) atomicAdd(reinterpret_cast<unsigned long long*>(&sum[cc]), static_cast<unsigned long long>(se));
      if (so) atomicOr(&status[cc], so);
    }
    todo &= ~m;
  }
}

// Energy per candidate as f32 (the reference's Vec<f32>): NaN with a NaN
// coordinate, +inf when a sum saturated, else the exact total.
__global__ void k_nf_finish(const long long* __restrict__ sum, const uint32_t* __restrict__ status,
                            uint32_t n, float* __restrict__ energy) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const uint32_t st = status[c];
  double e = static_cast<double>(sum[c]) * kFixUnit;
  if (st & kStNaN) e = __builtin_nan("");
  else if (st & kStE) e = __builtin_huge_val();
  energy[c] = static_cast<float>(e);
}

// Takes the scratch array's address on `device` (current): once per device,
// at context creation.
int noalloc_init_device(int device) {
  if (device < 0 || device >= kNfMaxDevices) return DNAB_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_nf_mu);
  if (g_nf_scratch[device]) return DNAB_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DNAB_ERR_HIP;
  g_nf_scratch[device] = p;
  return DNAB_OK;
}

}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_score_batch_noalloc(dnab_ctx* ctx, const dnab_antigen* antigen,
                             const dnab_gpu_atom* candidates_dev, size_t n_atoms, size_t n_ab,
                             float* energies_dev, size_t capacity, size_t* n_cand_out) {
  if (!ctx || !antigen || antigen->ctx != ctx) return DNAB_ERR_INVALID;
  // src/gpu.rs:172-174: n_cand = len / n_ab, empty result for n_ab == 0.
  const size_t n_cand = n_ab == 0 ? 0 : n_atoms / n_ab;
  if (n_cand_out) *n_cand_out = n_cand;
  if (n_cand == 0) return DNAB_OK;
  if (!candidates_dev) return DNAB_ERR_INVALID;
  if (!energies_dev || capacity < n_cand) return DNAB_ERR_CAPACITY;
  if (n_cand > kNfMaxCand || n_cand * n_ab > 0xFFFFFF00ull)
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED,
                          "score_batch_noalloc: at most %u candidates and 2^32 atoms (got %zu, %zu)",
                          kNfMaxCand, n_cand, n_cand * n_ab);
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_nf_mu);
    if (ctx->device >= 0 && ctx->device < kNfMaxDevices) base = g_nf_scratch[ctx->device];
  }
  if (!base)
    return dnab_set_error(ctx, DNAB_ERR_HIP, "score_batch_noalloc: scratch address not available");
  long long* sum = static_cast<long long*>(base);
  uint32_t* status = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(base) +
                                                 static_cast<size_t>(kNfMaxCand) * 8);
  NfView V;
  V.atoms = antigen->atoms.p;
  V.cell_start = antigen->grid.cell_start.p;
  for (int k = 0; k < 3; ++k) {
    V.kmin[k] = antigen->grid.kmin[k];
    V.kmax[k] = antigen->grid.kmax[k];
  }
  V.inv = antigen->grid.inv;
  V.rc = antigen->cutoff;
  V.rc2 = antigen->cutoff * antigen->cutoff;
  V.n_atoms = static_cast<uint32_t>(antigen->grid.n);
  const uint32_t nc = static_cast<uint32_t>(n_cand);
  const uint32_t n_used = static_cast<uint32_t>(n_cand * n_ab);
  const uint32_t tiles = (n_used + kWave - 1) / kWave;
  (void)hipSetDevice(ctx->device);
  hipLaunchKernelGGL(k_nf_clear, dim3((nc + 255) / 256), dim3(256), 0, ctx->stream, sum, status, nc);
  hipLaunchKernelGGL(k_nf_pairs, dim3((tiles + 3) / 4), dim3(256), 0, ctx->stream, V, candidates_dev,
                     n_used, static_cast<uint32_t>(n_ab), sum, status);
  hipLaunchKernelGGL(k_nf_finish, dim3((nc + 255) / 256), dim3(256), 0, ctx->stream, sum, status, nc,
                     energies_dev);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return DNAB_OK;
}

}  // extern "C"
