// Context, errors, transfers and kernel timing.
//
// dnab_ctx_create replaces GpuContext::try_init (src/gpu.rs:55-57; async body 59-155): where
// the reference returns None when no adapter exists (src/gpu.rs:65-71) and
// the caller falls back to the CPU (src/main.rs:319-327), this returns
// DNAB_ERR_NO_DEVICE and leaves the fallback decision to the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstring>

#include "dnab_internal.h"

int dnab_set_error(dnab_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->last_error = buf;
  }
  return code;
}

namespace dnab {

// Host<->device copies. A copy from pageable memory is staged by the runtime
// through its own buffer, synchronously and with a large fixed cost per call;
// copies here go through pinned buffers of the context instead. An upload
// copies the source into the next slot of a pinned ring (so the caller may
// reuse its buffer at once, as before) and queues the DMA; the ring restarts
// only after the event recorded behind the last queued copy. Very large
// copies, and copies while the stream is being captured into a graph (whose
// nodes would keep pointing at the ring), go direct.
constexpr size_t kStageMax = size_t(64) << 20;
constexpr size_t kStageAlign = 256;

static bool stream_capturing(dnab_ctx* ctx) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

static bool ensure_pinned(char*& buf, size_t& cap, size_t need) {
  if (need <= cap) return true;
  if (buf) (void)hipHostFree(buf);
  buf = nullptr;
  cap = 0;
  size_t sz = std::max<size_t>(need, size_t(1) << 20);
  if (hipHostMalloc(reinterpret_cast<void**>(&buf), sz, hipHostMallocDefault) != hipSuccess) {
    buf = nullptr;
    return false;
  }
  cap = sz;
  return true;
}

int upload(dnab_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return DNAB_OK;
  const size_t need = (bytes + kStageAlign - 1) / kStageAlign * kStageAlign;
  bool staged = bytes <= kStageMax && !stream_capturing(ctx);
  if (staged && !ctx->stage_event &&
      hipEventCreateWithFlags(&ctx->stage_event, hipEventDisableTiming) != hipSuccess) {
    ctx->stage_event = nullptr;
    staged = false;
  }
  if (staged && ctx->stage_off + need > ctx->stage_cap) {
    // every copy queued from the ring so far has finished reading it
    DNAB_HIP_CHECK(ctx, hipEventSynchronize(ctx->stage_event));
    ctx->stage_off = 0;
    if (need > ctx->stage_cap)  // grow (to at most twice the largest staged copy)
      staged = ensure_pinned(ctx->stage, ctx->stage_cap,
                             std::max(need, std::min(2 * ctx->stage_cap, kStageMax)));
  }
  if (!staged) {
    DNAB_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return DNAB_OK;
  }
  char* slot = ctx->stage + ctx->stage_off;
  std::memcpy(slot, src, bytes);
  DNAB_HIP_CHECK(ctx, hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, ctx->stream));
  DNAB_HIP_CHECK(ctx, hipEventRecord(ctx->stage_event, ctx->stream));
  ctx->stage_off += need;
  return DNAB_OK;
}

int download(dnab_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return DNAB_OK;
  if (bytes > kStageMax || stream_capturing(ctx) || !ensure_pinned(ctx->dl_stage, ctx->dl_cap, bytes)) {
    DNAB_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return DNAB_OK;
  }
  DNAB_HIP_CHECK(ctx, hipMemcpyAsync(ctx->dl_stage, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(dst, ctx->dl_stage, bytes);
  return DNAB_OK;
}

// Two device-to-host copies (either may be empty: a null destination) with
// one wait for the stream.
int download2(dnab_ctx* ctx, void* dst1, const void* src1, size_t b1, void* dst2, const void* src2,
              size_t b2) {
  if (!dst1) b1 = 0;
  if (!dst2) b2 = 0;
  const size_t off2 = (b1 + 15) & ~size_t(15);
  if (b1 + b2 == 0) return DNAB_OK;
  if (off2 + b2 > kStageMax || stream_capturing(ctx) || !ensure_pinned(ctx->dl_stage, ctx->dl_cap, off2 + b2)) {
    DNAB_RETURN_IF(download(ctx, dst1, src1, b1));
    return download(ctx, dst2, src2, b2);
  }
  if (b1) DNAB_HIP_CHECK(ctx, hipMemcpyAsync(ctx->dl_stage, src1, b1, hipMemcpyDeviceToHost, ctx->stream));
  if (b2) DNAB_HIP_CHECK(ctx, hipMemcpyAsync(ctx->dl_stage + off2, src2, b2, hipMemcpyDeviceToHost, ctx->stream));
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (b1) std::memcpy(dst1, ctx->dl_stage, b1);
  if (b2) std::memcpy(dst2, ctx->dl_stage + off2, b2);
  return DNAB_OK;
}

static hipEvent_t take_event(dnab_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

hipEvent_t timer_begin(dnab_ctx* ctx, const char* name, hipEvent_t* end_out) {
  (void)name;
  *end_out = nullptr;
  if (!ctx->timing) return nullptr;
  hipEvent_t s = take_event(ctx), e = take_event(ctx);
  if (!s || !e) return nullptr;
  if (hipEventRecord(s, ctx->stream) != hipSuccess) return nullptr;
  *end_out = e;
  return s;
}

void timer_end(dnab_ctx* ctx, const char* name, hipEvent_t start, hipEvent_t end) {
  if (!start || !end) return;
  if (hipEventRecord(end, ctx->stream) != hipSuccess) return;
  for (auto& t : ctx->timers) {
    if (t.name == name) {
      t.pending.emplace_back(start, end);
      return;
    }
  }
  KernelTimer t;
  t.name = name;
  t.pending.emplace_back(start, end);
  ctx->timers.push_back(std::move(t));
}

void timer_host(dnab_ctx* ctx, const char* name, double ms) {
  if (!ctx->timing) return;
  for (auto& t : ctx->timers) {
    if (t.name == name) {
      t.total_ms += ms;
      t.launches += 1;
      return;
    }
  }
  KernelTimer t;
  t.name = name;
  t.total_ms = ms;
  t.launches = 1;
  ctx->timers.push_back(std::move(t));
}

dnab_ctx* ctx_retain(dnab_ctx* ctx) {
  if (ctx) ++ctx->refs;
  return ctx;
}

void ctx_release(dnab_ctx* ctx) {
  if (!ctx || --ctx->refs > 0) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->timers)
    for (auto& pr : t.pending) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->stage_event) (void)hipEventDestroy(ctx->stage_event);
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  if (ctx->dl_stage) (void)hipHostFree(ctx->dl_stage);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

}  // namespace dnab

using namespace dnab;

extern "C" {

const char* dnab_status_string(int status) {
  switch (status) {
    case DNAB_OK: return "ok";
    case DNAB_ERR_INVALID: return "invalid argument";
    case DNAB_ERR_NO_DEVICE: return "no usable GPU";
    case DNAB_ERR_HIP: return "HIP runtime error";
    case DNAB_ERR_ALLOC: return "allocation failed";
    case DNAB_ERR_CAPACITY: return "output buffer too small";
    case DNAB_ERR_UNSUPPORTED: return "unsupported input";
    default: return "unknown status";
  }
}

int dnab_device_count(int* count) {
  if (!count) return DNAB_ERR_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return DNAB_OK;
}

int dnab_ctx_create(int device, dnab_ctx** out) {
  if (!out) return DNAB_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DNAB_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return DNAB_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return DNAB_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DNAB_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DNAB_ERR_NO_DEVICE;
  dnab_ctx* ctx = new (std::nothrow) dnab_ctx();
  if (!ctx) return DNAB_ERR_ALLOC;
  ctx->device = device;
  ctx->cu_count = prop.multiProcessorCount;
  ctx->wave_size = prop.warpSize;
  std::strncpy(ctx->arch, prop.gcnArchName, sizeof(ctx->arch) - 1);
  // Default-priority stream; stream priorities are left untouched.
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return DNAB_ERR_HIP;
  }
  // the no-alloc entry's scratch address on this device (once per device; on
  // failure that entry reports it and everything else works)
  (void)dnab::noalloc_init_device(device);
  *out = ctx;
  return DNAB_OK;
}


// Drops the caller's reference: the context is freed once the objects made on
// it are destroyed too, whatever the order.
void dnab_ctx_destroy(dnab_ctx* ctx) { dnab::ctx_release(ctx); }

const char* dnab_ctx_last_error(const dnab_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

int dnab_ctx_synchronize(dnab_ctx* ctx) {
  if (!ctx) return DNAB_ERR_INVALID;
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DNAB_OK;
}

int dnab_ctx_device_info(const dnab_ctx* ctx, int* cu_count, int* wave_size, char* arch,
                         size_t arch_len) {
  if (!ctx) return DNAB_ERR_INVALID;
  if (cu_count) *cu_count = ctx->cu_count;
  if (wave_size) *wave_size = ctx->wave_size;
  if (arch && arch_len) {
    std::strncpy(arch, ctx->arch, arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return DNAB_OK;
}

int dnab_ctx_set_timing(dnab_ctx* ctx, int enable) {
  if (!ctx) return DNAB_ERR_INVALID;
  ctx->timing = enable != 0;
  return DNAB_OK;
}

int dnab_ctx_kernel_time(dnab_ctx* ctx, const char* kernel, double* total_ms,
                         uint64_t* launches) {
  if (!ctx || !kernel) return DNAB_ERR_INVALID;
  DNAB_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  double tot = 0.0;
  uint64_t cnt = 0;
  for (auto& t : ctx->timers) {
    for (auto& pr : t.pending) {
      float ms = 0.0f;
      DNAB_HIP_CHECK(ctx, hipEventElapsedTime(&ms, pr.first, pr.second));
      t.total_ms += ms;
      t.launches += 1;
      ctx->event_pool.push_back(pr.first);
      ctx->event_pool.push_back(pr.second);
    }
    t.pending.clear();
    if (t.name == kernel) {
      tot = t.total_ms;
      cnt = t.launches;
    }
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = cnt;
  return DNAB_OK;
}

}  // extern "C"
