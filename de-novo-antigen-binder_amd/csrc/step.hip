// Gradient (Langevin) step of the population loop and device-side top-k.
//
// The per-residue update restates src/diffusion.rs:300-339 (forces, the
// radial restraint, clamp, uniform noise, set_ca_pos as a rigid translation,
// src/allatom.rs:279-293) in the reference's operation order. The loop
// restates the GPU phase of run_allatom (src/diffusion.rs:1018-1057) with all
// candidates kept moving and scoring every `score_every` steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "dnab_internal.h"
#include "grid_device.h"
#include "sort_device.h"
#include "step_device.h"

namespace dnab {

constexpr int kTopkChunk = 4096;
constexpr uint32_t kNoIdx = 0xFFFFFFFFu;

// One wave per candidate: (1) the candidate's FP64 energy from its atoms'
// energies (DPP wave sum; the loop needs no separate reduction), (2) the
// gradient step of each residue (lanes = residues), whose displacement goes
// to res_delta for k_atom_move.
__global__ void __launch_bounds__(256)
k_candidate_update(const AgView* __restrict__ views, const uint32_t* __restrict__ cand_ag,
                   const uint32_t* __restrict__ cand_res_off,
                   const uint32_t* __restrict__ cand_atom_off,
                   const uint32_t* __restrict__ res_atom_off, const uint32_t* __restrict__ res_ca,
                   const float4* __restrict__ res_cg, uint32_t n_cand,
                   const float4* __restrict__ force, const double* __restrict__ e_atom,
                   const float4* __restrict__ pos, StepParams sp,
                   const float* __restrict__ sigma_table, const uint32_t* __restrict__ counter,
                   int do_energy, double* __restrict__ energy, float* __restrict__ err,
                   float4* __restrict__ res_delta) {
  const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (c >= n_cand) return;  // whole wave
  const AgView& V = views[cand_ag[c]];
  if (do_energy) {
    double s = 0.0, a = 0.0;
    for (uint32_t k = cand_atom_off[c] + lane; k < cand_atom_off[c + 1]; k += kWave) {
      s += e_atom[k];
      a += static_cast<double>(force[k].w);
    }
    s = wave_sum_f64(s);
    a = wave_sum_f64(a);
    if (lane == 0) {
      energy[c] = s;
      err[c] = static_cast<float>(a) * (32.0f / 16777216.0f);
    }
  }
  const uint32_t iter = counter[1];
  const float r0 = V.restraint_radius > 0.0f ? V.restraint_radius : sp.restraint_radius;
  for (uint32_t r = cand_res_off[c] + lane; r < cand_res_off[c + 1]; r += kWave) {
    const uint32_t ca = res_ca[r];
    const float4 p = pos[ca];
    const uint32_t a0 = res_atom_off[r], a1 = res_atom_off[r + 1];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (sp.force_model == DNAB_FORCE_CG) {
      const float4 g = res_cg[r];
      CgCloud ag{V.ca_x, V.ca_y, V.ca_z, V.ca_q, V.ca_eps, V.ca_sig, V.ca_hyd};
      cg_force_point(V.ca_grid, ag, p.x, p.y, p.z, g.x, g.y, g.z, g.w != 0.0f ? 1 : 0, fx, fy, fz);
    } else {
      // sum in atom order; four loads in flight at a time
      uint32_t k = a0;
      for (; k + 4 <= a1; k += 4) {
        const float4 f0 = force[k], f1 = force[k + 1], f2 = force[k + 2], f3 = force[k + 3];
        fx += dyn_force(f0.x); fy += dyn_force(f0.y); fz += dyn_force(f0.z);
        fx += dyn_force(f1.x); fy += dyn_force(f1.y); fz += dyn_force(f1.z);
        fx += dyn_force(f2.x); fy += dyn_force(f2.y); fz += dyn_force(f2.z);
        fx += dyn_force(f3.x); fy += dyn_force(f3.y); fz += dyn_force(f3.z);
      }
      for (; k < a1; ++k) {
        const float4 f = force[k];
        fx += dyn_force(f.x);
        fy += dyn_force(f.y);
        fz += dyn_force(f.z);
      }
    }
    const float sigma = sp.with_noise ? sigma_table[iter] : 0.0f;
    res_delta[r] = residue_delta(V.center[0], V.center[1], V.center[2], sp, r0, p, fx, fy, fz,
                                 sigma, iter, r);
  }
}

// One thread per atom: move by its residue's displacement; optionally count
// the atom for the next step's sort (or zero the outputs of an atom that left
// the cutoff box of its antigen). Lanes of a wave with the same sort code
// share one counter atomic (neighbouring atoms of a candidate often fall in
// the same 4 A cell), which cuts the traffic on contended counters.
__global__ void __launch_bounds__(256)
k_atom_move(const AgView* __restrict__ views, const uint32_t* __restrict__ atom_ag,
            const uint32_t* __restrict__ atom_res, const float4* __restrict__ res_delta,
            uint32_t n, float4* __restrict__ pos, int do_count, uint32_t* __restrict__ count,
            uint32_t* __restrict__ code_out, uint32_t* __restrict__ rank_out,
            float4* __restrict__ force, double* __restrict__ e_atom, double* __restrict__ bd_atom,
            double* __restrict__ f64_atom) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = k < n;
  uint32_t code = kExcluded;
  if (live) {
    const float4 d = res_delta[atom_res[k]];
    float4 q = pos[k];
    q.x += d.x;
    q.y += d.y;
    q.z += d.z;
    pos[k] = q;
    if (do_count) {
      const AgView& V = views[atom_ag[k]];
      code = sort_code_of(V, q.x, q.y, q.z);
      code_out[k] = code;
      if (code >= kNonFinite) fill_atom_outputs(k, code, V, force, e_atom, bd_atom, f64_atom);
    }
  }
  if (!do_count) return;  // uniform
  const bool need = live && code < kNonFinite;
  // group the wave's lanes by code: a leader per distinct code
  uint64_t rem = ballot64(need);
  uint32_t leader = 0, below = 0, size = 0;
  while (rem) {
    const int l = static_cast<int>(__builtin_ctzll(rem));
    const uint32_t lc = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(code), l));
    const uint64_t grp = ballot64(need && code == lc);
    if (need && code == lc) {
      leader = static_cast<uint32_t>(l);
      below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(grp >> 32),
                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(grp), 0u));
      size = static_cast<uint32_t>(__builtin_popcountll(grp));
    }
    rem &= ~grp;
  }
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t base = 0;
  if (need && lane == leader) base = atomicAdd(&count[code], size);
  base = static_cast<uint32_t>(__shfl(static_cast<int>(base), static_cast<int>(leader), 64));
  if (need) rank_out[k] = base + below;
}

// counter[1] = step index in use, counter[0] = the next one
__global__ void k_advance(uint32_t* counter) {
  if (threadIdx.x == 0) {
    const uint32_t next = counter[0];
    counter[1] = next;
    counter[0] = next + 1u;
  }
}

__global__ void k_set_counter(uint32_t* counter, uint32_t v) {
  if (threadIdx.x == 0) counter[0] = v;
}

// Ordering used for ranking: ascending energy, NaN last, ties by index.
__device__ inline bool rank_less(double ea, uint32_t ia, double eb, uint32_t ib) {
  const bool na = ea != ea, nb = eb != eb;
  if (na != nb) return nb;
  if (!na && ea != eb) return ea < eb;
  return ia < ib;
}

// Each block sorts one chunk of up to kTopkChunk (energy, index) entries with
// a bitonic network in LDS and writes its first k to out[block * k ...].
// blockIdx.y selects a row (an independent list; rows are in_stride apart in
// the input and out_stride apart in the output). The output starts
// out_row * out_stride further on, where out_row = counter[1] / row_div when
// `counter` is given (the step index in use: the scoring's row), else
// row_fixed.
__global__ void __launch_bounds__(1024)
k_topk_chunk(const double* __restrict__ in_e, const uint32_t* __restrict__ in_idx, uint32_t n,
             uint32_t k, double* __restrict__ out_e, uint32_t* __restrict__ out_idx,
             size_t in_stride, size_t out_stride, const uint32_t* __restrict__ counter,
             uint32_t row_div, uint32_t row_fixed) {
  __shared__ double se[kTopkChunk];
  __shared__ uint32_t si[kTopkChunk];
  const size_t out_row = (counter ? counter[1] / row_div : row_fixed) + blockIdx.y;
  in_e += blockIdx.y * in_stride;
  if (in_idx) in_idx += blockIdx.y * in_stride;
  out_e += out_row * out_stride;
  out_idx += out_row * out_stride;
  const uint32_t base = blockIdx.x * kTopkChunk;
  const uint32_t len = min(static_cast<uint32_t>(kTopkChunk), n - base);
  uint32_t m = 1;
  while (m < len) m <<= 1;
  for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
    if (t < len) {
      se[t] = in_e[base + t];
      si[t] = in_idx ? in_idx[base + t] : base + t;
    } else {
      se[t] = __builtin_huge_val();
      si[t] = kNoIdx;
    }
  }
  __syncthreads();
  for (uint32_t size = 2; size <= m; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = threadIdx.x; t < m / 2; t += blockDim.x) {
        const uint32_t i = 2 * t - (t & (stride - 1));
        const uint32_t j = i + stride;
        const bool up = (i & size) == 0;
        const double ea = se[i], eb = se[j];
        const uint32_t ia = si[i], ib = si[j];
        const bool swap = up ? rank_less(eb, ib, ea, ia) : rank_less(ea, ia, eb, ib);
        if (swap) {
          se[i] = eb;
          se[j] = ea;
          si[i] = ib;
          si[j] = ia;
        }
      }
      __syncthreads();
    }
  }
  const uint32_t kk = min(k, len);
  for (uint32_t t = threadIdx.x; t < kk; t += blockDim.x) {
    out_e[blockIdx.x * k + t] = se[t];
    out_idx[blockIdx.x * k + t] = si[t];
  }
  // chunks shorter than k pad their slots so later passes ignore them
  for (uint32_t t = kk + threadIdx.x; t < k; t += blockDim.x) {
    out_e[blockIdx.x * k + t] = __builtin_huge_val();
    out_idx[blockIdx.x * k + t] = kNoIdx;
  }
}

static inline uint32_t nblk(uint64_t n, uint32_t t) {
  uint64_t b = (n + t - 1) / t;
  return static_cast<uint32_t>(b ? b : 1);
}

// Multi-pass top-k: chunks of 4096 reduce n to ceil(n/4096)*k entries per pass.
int topk_plan(dnab_ctx* ctx, TopkPlan& pl, uint32_t n, uint32_t k, uint32_t rows) {
  pl.k = k;
  pl.rows = rows;
  pl.lens.clear();
  uint32_t len = n;
  size_t cap = 1;
  for (;;) {
    pl.lens.push_back(len);
    uint32_t chunks = nblk(len, kTopkChunk);
    size_t out = static_cast<size_t>(chunks) * k;
    cap = std::max(cap, out);
    if (chunks == 1) break;
    if (out >= len)
      return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED,
                            "top-%u of %u entries: k must stay below half of a %d-entry chunk", k,
                            n, kTopkChunk);
    len = static_cast<uint32_t>(out);
  }
  cap *= std::max<uint32_t>(rows, 1);
  DNAB_ALLOC_CHECK(ctx, pl.e[0].ensure(cap));  // a plan kept across calls keeps its buffers
  DNAB_ALLOC_CHECK(ctx, pl.e[1].ensure(cap));
  DNAB_ALLOC_CHECK(ctx, pl.idx[0].ensure(cap));
  DNAB_ALLOC_CHECK(ctx, pl.idx[1].ensure(cap));
  return DNAB_OK;
}

// Rows of pl.rows lists of n entries, n apart; row r's top k go to
// final_e / final_idx + (out_row + r) * k, with out_row as in k_topk_chunk.
void topk_launch(dnab_ctx* ctx, TopkPlan& pl, const double* energy, double* final_e,
                 uint32_t* final_idx, const uint32_t* counter, uint32_t row_div,
                 uint32_t row_fixed, const uint32_t* first_idx) {
  const double* in_e = energy;
  const uint32_t* in_i = first_idx;
  for (size_t p = 0; p < pl.lens.size(); ++p) {
    const uint32_t len = pl.lens[p];
    const uint32_t chunks = nblk(len, kTopkChunk);
    const bool last = p + 1 == pl.lens.size();
    double* oe = last ? final_e : pl.e[p & 1].p;
    uint32_t* oi = last ? final_idx : pl.idx[p & 1].p;
    hipLaunchKernelGGL(k_topk_chunk, dim3(chunks, std::max<uint32_t>(pl.rows, 1)), dim3(1024), 0,
                       ctx->stream, in_e, in_i, len, pl.k, oe, oi, static_cast<size_t>(len),
                       static_cast<size_t>(chunks) * pl.k, last ? counter : nullptr,
                       std::max<uint32_t>(row_div, 1), last ? row_fixed : 0u);
    in_e = oe;
    in_i = oi;
  }
}

// ---- selection of the m smallest scores (unordered, in index order) -----------
//
// A block of 1024 threads holds up to 1024 * KPT entries (thread t the
// consecutive slots t * KPT ...). It finds the m-th smallest key digit by
// digit (four 8-bit digits from the top: a block histogram of the digit in
// LDS among the keys that match the digits found so far, then the digit
// where the running count reaches the rank still needed), then writes, in
// slot order, every entry with a key below it and the first ties with it
// until m are taken. Keys are the scores' order-preserving integer images
// (NaN last). Stage inputs are scores (slot = index) or the previous stage's
// per-block lists (block b's entries at b * m_in, count cnt_in[b]).
constexpr uint32_t kSelBlock = 1024;
constexpr uint32_t kSelNone = 0xFFFFFFFFu;

__device__ inline uint32_t score_key(float f) {
  if (f != f) return kSelNone;
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Block-wide sum (all threads call; result in every thread).
__device__ inline uint32_t sel_block_sum(uint32_t v, uint32_t* s_w) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  v = wave_sum_u32(v);
  __syncthreads();
  if (lane == 0) s_w[wave] = v;
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t w = 0; w < kSelBlock / kWave; ++w) t += s_w[w];
  return t;
}

// Block-wide exclusive scan (all threads call); *total = the sum.
__device__ inline uint32_t sel_block_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(incl, o, 64);
    if (lane >= static_cast<uint32_t>(o)) incl += u;
  }
  __syncthreads();
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < kSelBlock / kWave; ++w) {
    before += w < wave ? s_w[w] : 0u;
    all += s_w[w];
  }
  *total = all;
  return before + incl - v;
}

template <int KPT, bool FROM_SCORES>
__global__ void __launch_bounds__(kSelBlock)
k_select_smallest(const float* __restrict__ scores, const uint32_t* __restrict__ keys_in,
                  const uint32_t* __restrict__ idx_in, const uint32_t* __restrict__ cnt_in,
                  uint32_t m_in, uint32_t n_slots, uint32_t m, uint32_t* __restrict__ keys_out,
                  uint32_t* __restrict__ idx_out, uint32_t* __restrict__ cnt_out) {
  __shared__ uint32_t s_w[kSelBlock / kWave];
  const uint32_t t = threadIdx.x;
  const uint32_t base = blockIdx.x * kSelBlock * KPT + t * KPT;
  uint32_t key[KPT], idx[KPT];
  uint32_t nv = 0;
#pragma unroll
  for (int q = 0; q < KPT; ++q) {
    const uint32_t sl = base + q;
    key[q] = kSelNone;
    idx[q] = kSelNone;
    if (sl < n_slots) {
      if (FROM_SCORES) {
        key[q] = score_key(scores[sl]);
        idx[q] = sl;
      } else {
        const uint32_t b = sl / m_in, j = sl - b * m_in;
        if (j < cnt_in[b]) {
          key[q] = keys_in[sl];
          idx[q] = idx_in[sl];
        }
      }
    }
    nv += idx[q] != kSelNone ? 1u : 0u;
  }
  const uint32_t mm = min(m, sel_block_sum(nv, s_w));
  if (mm > 0) {
    // the key of rank mm, 8 bits at a time from the top: a block histogram
    // of the digit over the keys that match the digits found so far
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_sel[2];
    const uint32_t lane = t & 63u, wave = t >> 6;
    uint32_t T = 0, need = mm;  // the digits found so far; the rank still needed among their keys
    for (int sh = 24; sh >= 0; sh -= 8) {
      const uint32_t hi = sh == 24 ? 0u : (0xFFFFFFFFu << (sh + 8));
      if (t < 256) s_hist[t] = 0;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < KPT; ++q)
        if (idx[q] != kSelNone && (key[q] & hi) == (T & hi)) atomicAdd(&s_hist[(key[q] >> sh) & 255u], 1u);
      __syncthreads();
      if (wave == 0) {
        // lane: digits 4 lane .. 4 lane + 3, running counts within the lane,
        // then across the lanes
        const uint4 h = *reinterpret_cast<const uint4*>(&s_hist[4 * lane]);
        const uint32_t c0 = h.x, c1 = c0 + h.y, c2 = c1 + h.z, c3 = c2 + h.w;
        uint32_t incl = c3;
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t u = __shfl_up(incl, o, 64);
          if (lane >= static_cast<uint32_t>(o)) incl += u;
        }
        const uint32_t before = incl - c3;  // keys in digits below 4 lane
        // the first digit whose running count reaches the rank needed
        const uint64_t reach = ballot64(incl >= need);
        const int L = __builtin_ctzll(reach);  // exists: the keys matching T number >= need
        if (lane == static_cast<uint32_t>(L)) {
          uint32_t d = 3, below = before + c2;
          if (before + c0 >= need) {
            d = 0;
            below = before;
          } else if (before + c1 >= need) {
            d = 1;
            below = before + c0;
          } else if (before + c2 >= need) {
            d = 2;
            below = before + c1;
          }
          s_sel[0] = T | ((4u * lane + d) << sh);
          s_sel[1] = need - below;
        }
      }
      __syncthreads();
      T = s_sel[0];
      need = s_sel[1];
    }
    // T is the mm-th smallest key and `need` its rank among the keys equal to
    // it: all keys below T, then the first `need` ties in slot order
    uint32_t neq = 0;
#pragma unroll
    for (int q = 0; q < KPT; ++q) neq += (idx[q] != kSelNone && key[q] == T) ? 1u : 0u;
    uint32_t eq_all;
    uint32_t eq_rank = sel_block_scan(neq, s_w, &eq_all);
    uint32_t ntake = 0;
    bool take[KPT];
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
      const bool v = idx[q] != kSelNone;
      const bool eq = v && key[q] == T;
      take[q] = (v && key[q] < T) || (eq && eq_rank < need);
      eq_rank += eq ? 1u : 0u;
      ntake += take[q] ? 1u : 0u;
    }
    uint32_t all;
    uint32_t pos = sel_block_scan(ntake, s_w, &all);
    const size_t out0 = static_cast<size_t>(blockIdx.x) * m;
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
      if (take[q]) {
        if (keys_out) keys_out[out0 + pos] = key[q];
        idx_out[out0 + pos] = idx[q];
        ++pos;
      }
    }
  }
  if (t == 0 && cnt_out) cnt_out[blockIdx.x] = mm;
}

// The min(m, n) smallest of n scores (ties by index) into out_idx, in index
// order. 1 <= m <= 2048.
int select_smallest(dnab_ctx* ctx, const float* scores, uint32_t n, uint32_t m, uint32_t* out_idx,
                    DevBuf<uint32_t>& scratch) {
  if (m == 0 || m > 2048) return dnab_set_error(ctx, DNAB_ERR_INVALID, "select: m must be 1..2048");
  if (n == 0) return DNAB_OK;
  constexpr uint32_t kChunk = kSelBlock * 8;  // stage inputs per block
  const uint32_t nb1 = (n + kChunk - 1) / kChunk;
  if (nb1 == 1) {
    hipLaunchKernelGGL((k_select_smallest<8, true>), dim3(1), dim3(kSelBlock), 0, ctx->stream, scores,
                       nullptr, nullptr, nullptr, 1u, n, m, nullptr, out_idx, nullptr);
    DNAB_HIP_CHECK(ctx, hipGetLastError());
    return DNAB_OK;
  }
  // scratch: two ping-pong lists of (keys, idx, counts)
  const size_t cap = static_cast<size_t>(nb1) * m;
  DNAB_ALLOC_CHECK(ctx, scratch.ensure(4 * cap + 2 * nb1 + 2));
  uint32_t* kbuf[2] = {scratch.p, scratch.p + cap};
  uint32_t* ibuf[2] = {scratch.p + 2 * cap, scratch.p + 3 * cap};
  uint32_t* cbuf[2] = {scratch.p + 4 * cap, scratch.p + 4 * cap + nb1 + 1};
  hipLaunchKernelGGL((k_select_smallest<8, true>), dim3(nb1), dim3(kSelBlock), 0, ctx->stream, scores,
                     nullptr, nullptr, nullptr, 1u, n, m, kbuf[0], ibuf[0], cbuf[0]);
  uint32_t slots = nb1 * m, cur = 0;
  while (slots > kChunk) {  // more chunked stages (m <= kChunk / 4: each at least quarters)
    const uint32_t nb = (slots + kChunk - 1) / kChunk;
    hipLaunchKernelGGL((k_select_smallest<8, false>), dim3(nb), dim3(kSelBlock), 0, ctx->stream,
                       nullptr, kbuf[cur], ibuf[cur], cbuf[cur], m, slots, m, kbuf[cur ^ 1],
                       ibuf[cur ^ 1], cbuf[cur ^ 1]);
    cur ^= 1;
    slots = nb * m;
  }
  hipLaunchKernelGGL((k_select_smallest<8, false>), dim3(1), dim3(kSelBlock), 0, ctx->stream,
                     nullptr, kbuf[cur], ibuf[cur], cbuf[cur], m, slots, m, nullptr, out_idx, nullptr);
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  return DNAB_OK;
}

static StepParams make_step_params(const dnab_langevin_params* p) {
  StepParams sp;
  sp.step = p->step;
  sp.max_disp = p->max_disp;
  sp.restraint_k = p->restraint_k;
  sp.restraint_radius = p->restraint_radius;
  sp.with_noise = p->with_noise ? 1 : 0;
  sp.force_model = p->force_model;
  sp.seed_lo = static_cast<uint32_t>(p->seed);
  sp.seed_hi = static_cast<uint32_t>(p->seed >> 32);
  return sp;
}

static int validate_params(dnab_ctx* ctx, const dnab_langevin_params* p) {
  if (p->force_model != DNAB_FORCE_ALLATOM && p->force_model != DNAB_FORCE_CG)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "langevin: unknown force model");
  if (!(p->schedule_total > 0)) return dnab_set_error(ctx, DNAB_ERR_INVALID, "langevin: schedule_total must be > 0");
  if (!std::isfinite(p->step) || !std::isfinite(p->max_disp) || !std::isfinite(p->restraint_k) ||
      !std::isfinite(p->restraint_radius) || !(p->t_hot > 0) || !(p->t_cold > 0))
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "langevin: non-finite parameter");
  if (p->step_kernels < 0 || p->step_kernels > 2)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "langevin: step_kernels must be 0, 1 or 2");
  return DNAB_OK;
}

// sigma(iter) = noise_base * sqrt(T(iter)), T = t_hot (t_cold/t_hot)^(i/total)
// in f32 as src/diffusion.rs:102-104 and 315, computed on the host.
static int upload_sigma(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                        uint32_t n_iters) {
  // kept in the population: the asynchronous copy reads it after this returns
  std::vector<float>& s = pop->h_sigma;
  s.resize(n_iters + 1);
  for (uint32_t i = 0; i <= n_iters; ++i) {
    float it = static_cast<float>(i + p->schedule_offset);
    float temp = p->t_hot * std::pow(p->t_cold / p->t_hot, it / static_cast<float>(p->schedule_total));
    s[i] = p->noise_base * std::sqrt(temp);
  }
  DNAB_ALLOC_CHECK(ctx, pop->sigma_table.ensure(s.size()));
  return upload(ctx, pop->sigma_table.p, s.data(), s.size() * sizeof(float));
}

static int launch_update(dnab_ctx* ctx, dnab_population* pop, const StepParams& sp,
                         bool do_energy, bool do_count) {
  // counts of the positions before the move would go stale
  if (!do_count) DNAB_RETURN_IF(discard_pending_counts(ctx, pop));
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  const uint32_t na = static_cast<uint32_t>(pop->n_atoms);
  if (nc) {
    hipLaunchKernelGGL(k_candidate_update, dim3(nblk(nc, 4)), dim3(256), 0, ctx->stream,
                       pop->views.p, pop->cand_ag.p, pop->cand_res_off.p, pop->cand_atom_off.p,
                       pop->res_atom_off.p, pop->res_ca.p, pop->res_cg.p, nc, pop->force.p,
                       pop->e_atom.p, pop->pos.p, sp, pop->sigma_table.p, pop->step_counter.p,
                       do_energy ? 1 : 0, pop->energy.p, pop->err_bound.p, pop->res_delta.p);
  }
  if (na) {
    hipLaunchKernelGGL(k_atom_move, dim3(nblk(na, 256)), dim3(256), 0, ctx->stream,
                       pop->views.p, pop->atom_ag.p, pop->atom_res.p, pop->res_delta.p, na,
                       pop->pos.p, do_count ? 1 : 0, pop->sort_count.p, pop->sort_code.p,
                       pop->sort_rank.p, pop->force.p, pop->e_atom.p, pop->bd_atom.p,
                       pop->f64_atom.p);
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  pop->counts_pending = do_count;
  pop->sorted = false;
  return DNAB_OK;
}

// One chunk of the loop: L steps; when `score`, the first step's energies
// (the state the chunk starts from) go through top-k into the scoring's row
// of the score history (row = step index / score_every, read on the device,
// so a captured chunk replays into successive rows).
static int enqueue_chunk(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                         const StepParams& sp, uint32_t L, TopkPlan* tk, bool score) {
  const bool allatom = p->force_model == DNAB_FORCE_ALLATOM;
  for (uint32_t l = 0; l < L; ++l) {
    const bool scoring = score && l == 0;
    const bool eval = allatom || scoring;
    if (eval) {
      DNAB_RETURN_IF(population_sort(ctx, pop));  // also advances the step counter
      DNAB_RETURN_IF(population_pair(ctx, pop, kEvalLoop | (p->count_pairs ? kEvalCount : DNAB_EVAL_FORCES)));
    } else {
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, ctx->stream, pop->step_counter.p);
    }
    // the next step evaluates (and so needs counts) in the all-atom model
    DNAB_RETURN_IF(launch_update(ctx, pop, sp, eval, allatom));
    if (scoring && tk)
      topk_launch(ctx, *tk, pop->energy.p, pop->hist_e.p, pop->hist_idx.p, pop->step_counter.p,
                  p->score_every);
  }
  return DNAB_OK;
}

}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_langevin_default_params(dnab_langevin_params* p) {
  if (!p) return DNAB_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->n_steps = 200;      // GPU_STEPS, src/diffusion.rs:36
  p->score_every = 10;
  p->top_k = 64;         // TOP_K, src/diffusion.rs:40
  p->seed = 0;
  p->with_noise = 0;     // GPU phase, src/diffusion.rs:1023
  p->force_model = DNAB_FORCE_ALLATOM;
  p->use_graph = 0;
  p->step = 0.08f;       // src/diffusion.rs:66
  p->max_disp = 2.0f;    // src/diffusion.rs:82
  p->noise_base = 0.25f; // src/diffusion.rs:67
  p->t_hot = 3.0f;       // src/diffusion.rs:68
  p->t_cold = 0.02f;     // src/diffusion.rs:69
  p->schedule_offset = 0;
  p->schedule_total = 200 + 600;  // GPU_STEPS + CPU_STEPS, src/diffusion.rs:1019
  p->restraint_k = 0.02f;         // src/diffusion.rs:83
  p->restraint_radius = 20.0f;    // src/diffusion.rs:81
  return DNAB_OK;
}

int dnab_langevin_step(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                       uint32_t iter) {
  if (!ctx || !pop || !p || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "langevin_step"));
  DNAB_RETURN_IF(validate_params(ctx, p));
  (void)hipSetDevice(ctx->device);
  DNAB_RETURN_IF(upload_sigma(ctx, pop, p, iter));
  pop->masked = false;
  hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(64), 0, ctx->stream, pop->step_counter.p, iter);
  const bool allatom = p->force_model == DNAB_FORCE_ALLATOM;
  if (allatom) {
    DNAB_RETURN_IF(population_sort(ctx, pop));  // counter[1] = iter
    DNAB_RETURN_IF(population_pair(ctx, pop, kEvalLoop | DNAB_EVAL_FORCES));
  } else {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, ctx->stream, pop->step_counter.p);
  }
  DNAB_RETURN_IF(launch_update(ctx, pop, make_step_params(p), allatom, false));
  if (allatom) {
    pop->last_flags = DNAB_EVAL_FORCES;
    pop->evaluated = true;  // energies and forces of the state before the step
  } else {
    pop->evaluated = false;
  }
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_langevin_run(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                      uint32_t* topk_idx, double* topk_energy, uint32_t* n_topk) {
  if (!ctx || !pop || !p || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "langevin_run"));
  DNAB_RETURN_IF(validate_params(ctx, p));
  if (p->top_k == 0 || p->top_k > 4096)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "langevin: top_k must be 1..4096");
  (void)hipSetDevice(ctx->device);
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  const uint32_t k = std::min<uint32_t>(p->top_k, nc);
  if (n_topk) *n_topk = k;
  if (nc == 0) return DNAB_OK;
  DNAB_RETURN_IF(upload_sigma(ctx, pop, p, p->n_steps));
  // score history: one row per scoring (steps 0, L, 2L, ... before the
  // last, then the final state), each the top k of that scoring
  const uint32_t n_mid = p->score_every ? (p->n_steps + p->score_every - 1) / p->score_every : 0;
  const uint32_t rows = n_mid + 1;
  DNAB_ALLOC_CHECK(ctx, pop->hist_e.ensure(static_cast<size_t>(rows) * k));
  DNAB_ALLOC_CHECK(ctx, pop->hist_idx.ensure(static_cast<size_t>(rows) * k));
  pop->hist_rows = 0;  // until this run completes
  pop->hist_k = k;
  pop->masked = false;
  hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(64), 0, ctx->stream, pop->step_counter.p, 0u);
  const StepParams sp = make_step_params(p);
  const bool fused = p->step_kernels == 0 || (p->step_kernels == 2 && pop->traj_engine == 0);
  if (fused && trajectory_supported(pop, p)) {
    // every step in one launch, which leaves each scoring's candidate
    // energies in a row of score_e; then the top k of every row at once
    DNAB_ALLOC_CHECK(ctx, pop->score_e.ensure(static_cast<size_t>(rows) * nc));
    TopkPlan& tr = pop->run_topk;
    DNAB_RETURN_IF(topk_plan(ctx, tr, nc, k, rows));
    DNAB_RETURN_IF(discard_pending_counts(ctx, pop));
    DNAB_RETURN_IF(trajectory_run(ctx, pop, p, sp));
    pop->sorted = false;
    pop->last_flags = DNAB_EVAL_FORCES;
    pop->evaluated = true;
    topk_launch(ctx, tr, pop->score_e.p, pop->hist_e.p, pop->hist_idx.p);
  } else {
    DNAB_RETURN_IF(step_kernel_loop(ctx, pop, p, sp, nc, k, n_mid));
  }
  DNAB_HIP_CHECK(ctx, hipGetLastError());
  const size_t last = static_cast<size_t>(n_mid) * k;
  DNAB_RETURN_IF(download2(ctx, topk_idx, pop->hist_idx.p + last, k * sizeof(uint32_t), topk_energy,
                           pop->hist_e.p + last, k * sizeof(double)));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  pop->hist_rows = rows;
  return DNAB_OK;
}

int dnab_population_autotune_trajectory(dnab_ctx* ctx, dnab_population* pop,
                                        const dnab_langevin_params* p, uint32_t trial_steps,
                                        dnab_traj_tuning* chosen) {
  if (!ctx || !pop || !p || pop->ctx != ctx) return DNAB_ERR_INVALID;
  DNAB_RETURN_IF(require_finite(ctx, pop, "autotune_trajectory"));
  DNAB_RETURN_IF(validate_params(ctx, p));
  if (trial_steps == 0 || trial_steps > 10000)
    return dnab_set_error(ctx, DNAB_ERR_INVALID, "autotune_trajectory: trial_steps must be 1..10000");
  (void)hipSetDevice(ctx->device);
  dnab_langevin_params lp = *p;
  lp.n_steps = trial_steps;
  lp.step_kernels = 0;
  const uint32_t nc = static_cast<uint32_t>(pop->n_cand);
  if (nc == 0 || !trajectory_supported(pop, &lp))
    return dnab_set_error(ctx, DNAB_ERR_UNSUPPORTED, "autotune_trajectory: no fused loop for this population");
  DNAB_RETURN_IF(upload_sigma(ctx, pop, &lp, lp.n_steps));
  const uint32_t rows = (lp.score_every ? (lp.n_steps + lp.score_every - 1) / lp.score_every : 0) + 1;
  DNAB_ALLOC_CHECK(ctx, pop->score_e.ensure(static_cast<size_t>(rows) * nc));
  DNAB_RETURN_IF(discard_pending_counts(ctx, pop));
  // every trial starts from the caller's coordinates, which are put back at the end
  DevBuf<float4> saved;
  DNAB_ALLOC_CHECK(ctx, saved.alloc(pop->n_atoms));
  const size_t pos_bytes = pop->n_atoms * sizeof(float4);
  DNAB_HIP_CHECK(ctx, hipMemcpyAsync(saved.p, pop->pos.p, pos_bytes, hipMemcpyDeviceToDevice, ctx->stream));
  const StepParams sp = make_step_params(&lp);
  dnab_traj_tuning best;
  DNAB_RETURN_IF(dnab_population_get_trajectory_tuning(pop, &best));
  // wall clock of one loop (host regrouping included), best of two runs; a
  // shape whose groups do not fit the LDS runs as the loop runs it then
  // (one candidate per group of 4 waves)
  auto trial = [&](const dnab_traj_tuning& t, double* sec) -> int {
    *sec = INFINITY;
    DNAB_RETURN_IF(dnab_population_set_trajectory_tuning(ctx, pop, &t));
    for (int rep = 0; rep < 2; ++rep) {
      DNAB_HIP_CHECK(ctx, hipMemcpyAsync(pop->pos.p, saved.p, pos_bytes, hipMemcpyDeviceToDevice, ctx->stream));
      hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(64), 0, ctx->stream, pop->step_counter.p, 0u);
      DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      const auto t0 = std::chrono::steady_clock::now();
      DNAB_RETURN_IF(trajectory_run(ctx, pop, &lp, sp));
      DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      *sec = std::min(*sec, s);
    }
    return DNAB_OK;
  };
  double best_s = INFINITY;
  DNAB_RETURN_IF(trial(best, &best_s));
  // coordinate descent over the shape, one dimension at a time, until a
  // pass changes nothing (at most three passes). Candidates per group and
  // the grouping mode form one dimension of 12 pairs: they interact (larger
  // groups pay off when grouped by position), and one at a time the descent
  // stops short of such pairs.
  static const uint32_t waves[] = {4, 8, 16}, cpw[] = {1, 2, 4, 8}, chunks[] = {1, 2, 4},
                        cells[] = {2, 8}, bal[] = {0, 1, 2};
  struct Dim { const uint32_t* v; int n; uint32_t dnab_traj_tuning::*f; const uint32_t* v2;
               int n2; uint32_t dnab_traj_tuning::*f2; };
  const Dim dims[] = {{waves, 3, &dnab_traj_tuning::waves, nullptr, 1, nullptr},
                      {cpw, 4, &dnab_traj_tuning::cand_per_group, bal, 3, &dnab_traj_tuning::balance},
                      {chunks, 3, &dnab_traj_tuning::chunks_per_item, nullptr, 1, nullptr},
                      {cells, 2, &dnab_traj_tuning::cells_per_item, nullptr, 1, nullptr}};
  int rc = DNAB_OK;
  for (int pass = 0; pass < 3 && rc == DNAB_OK; ++pass) {
    bool changed = false;
    for (const Dim& d : dims) {
      for (int i = 0; i < d.n * d.n2 && rc == DNAB_OK; ++i) {
        dnab_traj_tuning t = best;
        t.*d.f = d.v[i % d.n];
        if (d.f2) t.*d.f2 = d.v2[i / d.n];
        if (t.*d.f == best.*d.f && (!d.f2 || t.*d.f2 == best.*d.f2)) continue;
        double s;
        rc = trial(t, &s);
        if (rc == DNAB_OK && s < best_s) {
          best_s = s;
          best = t;
          changed = true;
        }
      }
    }
    if (!changed) break;
  }
  // the per-step kernels on the same loop: the engine of step_kernels = 2 is
  // the faster of the two (the per-step run, best of two, leaves its score
  // history, which is cleared below)
  if (rc == DNAB_OK) {
    dnab_langevin_params sk = lp;
    sk.step_kernels = 1;
    sk.count_pairs = 0;
    double step_s = INFINITY;
    for (int rep = 0; rep < 2 && rc == DNAB_OK; ++rep) {
      DNAB_HIP_CHECK(ctx, hipMemcpyAsync(pop->pos.p, saved.p, pos_bytes, hipMemcpyDeviceToDevice, ctx->stream));
      DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      pop->sorted = false;
      pop->evaluated = false;
      const auto t0 = std::chrono::steady_clock::now();
      rc = dnab_langevin_run(ctx, pop, &sk, nullptr, nullptr, nullptr);
      DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      step_s = std::min(step_s, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    best.engine = step_s < best_s ? 1u : 0u;
    pop->hist_rows = 0;
  }
  // the winner stays in the population; the coordinates are the caller's
  // again and no evaluation is current
  const int rc2 = dnab_population_set_trajectory_tuning(ctx, pop, &best);
  DNAB_HIP_CHECK(ctx, hipMemcpyAsync(pop->pos.p, saved.p, pos_bytes, hipMemcpyDeviceToDevice, ctx->stream));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  pop->sorted = false;
  pop->evaluated = false;
  pop->masked = false;
  if (rc != DNAB_OK) return rc;
  if (rc2 != DNAB_OK) return rc2;
  if (chosen) *chosen = best;
  return DNAB_OK;
}

int dnab_population_read_scores(dnab_ctx* ctx, const dnab_population* pop, uint32_t* n_scores,
                                uint32_t* k, uint32_t* topk_idx, double* topk_energy) {
  if (!ctx || !pop || pop->ctx != ctx) return DNAB_ERR_INVALID;
  if (n_scores) *n_scores = pop->hist_rows;
  if (k) *k = pop->hist_k;
  const size_t n = static_cast<size_t>(pop->hist_rows) * pop->hist_k;
  if (n == 0) return DNAB_OK;
  DNAB_RETURN_IF(download2(ctx, topk_idx, pop->hist_idx.p, n * sizeof(uint32_t), topk_energy, pop->hist_e.p,
                           n * sizeof(double)));
  return DNAB_OK;
}

}  // extern "C"

namespace dnab {

// The loop as per-step population kernels, eager or replayed from a graph.
int step_kernel_loop(dnab_ctx* ctx, dnab_population* pop, const dnab_langevin_params* p,
                     const StepParams& sp, uint32_t nc, uint32_t k, uint32_t n_mid) {
  TopkPlan& tk = pop->run_topk;
  DNAB_RETURN_IF(topk_plan(ctx, tk, nc, k));
  const uint32_t L = p->score_every ? p->score_every : std::max<uint32_t>(p->n_steps, 1);
  const bool graph = p->use_graph && p->n_steps >= L;
  // A chunk must start in the state it ends in: in the all-atom model every
  // update counts the moved atoms for the next sort, so counts are pending.
  if (p->force_model == DNAB_FORCE_ALLATOM) DNAB_RETURN_IF(population_count(ctx, pop));
  else DNAB_RETURN_IF(discard_pending_counts(ctx, pop));
  const bool saved_timing = ctx->timing;
  uint32_t done = 0;
  if (graph) {
    ctx->timing = false;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    DNAB_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_chunk(ctx, pop, p, sp, L, &tk, p->score_every != 0);
    hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    ctx->timing = saved_timing;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ce != hipSuccess) return dnab_set_error(ctx, DNAB_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      return dnab_set_error(ctx, DNAB_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
    }
    for (; done + L <= p->n_steps; done += L) {
      hipError_t le = hipGraphLaunch(ge, ctx->stream);
      if (le != hipSuccess) {
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        return dnab_set_error(ctx, DNAB_ERR_HIP, "graph launch failed: %s", hipGetErrorString(le));
      }
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
  }
  for (; done < p->n_steps; done += L) {
    const uint32_t len = std::min(L, p->n_steps - done);
    DNAB_RETURN_IF(enqueue_chunk(ctx, pop, p, sp, len, &tk, p->score_every != 0));
  }
  // final scoring after the last update (clash pairs in FP64: the reported
  // energies and forces)
  DNAB_RETURN_IF(population_sort(ctx, pop));
  DNAB_RETURN_IF(population_pair(ctx, pop, p->count_pairs ? kEvalCount : DNAB_EVAL_FORCES));
  DNAB_RETURN_IF(population_reduce_energies(ctx, pop, DNAB_EVAL_FORCES));
  topk_launch(ctx, tk, pop->energy.p, pop->hist_e.p, pop->hist_idx.p, nullptr, 1, n_mid);
  return DNAB_OK;
}

}  // namespace dnab
