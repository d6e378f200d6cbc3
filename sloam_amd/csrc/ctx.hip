// Context, error reporting, timers and small utility kernels of the C ABI.
#include "common.hpp"

#include <cstdio>
#include <cstring>
#include <new>

int sloam_record_hip_error(sloam_ctx* ctx, hipError_t e, const char* what) {
  if (ctx) {
    ctx->last_error = std::string(what) + ": " + hipGetErrorString(e);
  }
  return SLOAM_ERR_HIP;
}

static hipEvent_t pool_get(sloam_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

SloamKernelScope::SloamKernelScope(sloam_ctx* c, const char* n) : ctx(c), name(n) {
  if (ctx && ctx->timing) {
    start = pool_get(ctx);
    if (start) hipEventRecord(start, ctx->stream);
  }
}

SloamKernelScope::~SloamKernelScope() {
  if (!start) return;
  hipEvent_t stop = pool_get(ctx);
  if (!stop) return;
  hipEventRecord(stop, ctx->stream);
  ctx->timers[name].pending.emplace_back(start, stop);
}

int sloam_dev_alloc(sloam_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return SLOAM_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    sloam_record_hip_error(ctx, e, "hipMalloc");
    return SLOAM_ERR_ALLOC;
  }
  return SLOAM_OK;
}

void sloam_dev_free(void* p) {
  if (p) (void)hipFree(p);
}

int sloam_ctx_scratch(sloam_ctx* ctx, int slot, size_t bytes, void** out) {
  *out = nullptr;
  if (slot < 0 || slot > 31) return SLOAM_ERR_INVALID_ARG;
  if ((int)ctx->scratch.size() <= slot) {
    ctx->scratch.resize(slot + 1, {nullptr, 0});
    ctx->scratch_pins.resize(slot + 1, 0);
  }
  auto& s = ctx->scratch[slot];
  if (ctx->capturing) ctx->capture_slots |= 1u << slot;
  if (s.second < bytes && ctx->capturing) {
    ctx->last_error = "scratch buffer too small during graph capture: run the sequence once "
                      "before capturing it";
    return SLOAM_ERR_INVALID_ARG;
  }
  if (s.second < bytes && ctx->scratch_pins[slot] > 0) {
    ctx->last_error = "scratch slot " + std::to_string(slot) +
                      " is held by a captured graph and is too small for this call: destroy "
                      "the graph, or use a context of its own for the larger inputs";
    return SLOAM_ERR_INVALID_ARG;
  }
  if (s.second < bytes) {
    // the stream may still use the old buffer
    SLOAM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    sloam_dev_free(s.first);
    s.first = nullptr;
    s.second = 0;
    const size_t sz = bytes < 4096 ? 4096 : bytes + bytes / 4;
    int st = sloam_dev_alloc(ctx, &s.first, sz);
    if (st) return st;
    s.second = sz;
  }
  *out = s.first;
  return SLOAM_OK;
}

void sloam_ctx_pin_slots(sloam_ctx* ctx, uint32_t slots, int delta) {
  if (!ctx) return;
  for (int i = 0; i < 32; ++i)
    if ((slots >> i) & 1u) {
      if ((int)ctx->scratch_pins.size() <= i) ctx->scratch_pins.resize(i + 1, 0);
      ctx->scratch_pins[i] += delta;
    }
}

__global__ void sl_fill_u32_kernel(uint32_t* __restrict__ d, size_t n, uint32_t value) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = value;
}

int sloam_fill_u32(sloam_ctx* ctx, uint32_t* d, size_t n, uint32_t value) {
  if (n == 0) return SLOAM_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(sl_fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d, n,
                     value);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

// xyz[i] = the first three floats of xyzi[i] (one point per thread)
__global__ __launch_bounds__(256) void sl_xyz_kernel(const float4* __restrict__ xyzi, int64_t n,
                                                      float* __restrict__ xyz) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 v = xyzi[i];
  xyz[3 * i] = v.x;
  xyz[3 * i + 1] = v.y;
  xyz[3 * i + 2] = v.z;
}

extern "C" int sloam_xyz_from_xyzi(sloam_ctx* ctx, const float* d_xyzi, int64_t n, float* d_xyz) {
  if (!ctx || n < 0 || (n > 0 && (!d_xyzi || !d_xyz))) return SLOAM_ERR_INVALID_ARG;
  if (((uintptr_t)d_xyzi & 15) != 0) return SLOAM_ERR_INVALID_ARG;  // read as float4
  if (n == 0) return SLOAM_OK;
  if ((n + 255) / 256 > 0x7fffffff) return SLOAM_ERR_TOO_LARGE;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(sl_xyz_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const float4*)d_xyzi, n, d_xyz);
  SLOAM_HIP(ctx, hipGetLastError());
  return SLOAM_OK;
}

extern "C" {

int sloam_ctx_create(int device, sloam_ctx** out) {
  if (!out) return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SLOAM_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = new (std::nothrow) sloam_ctx();
  if (!ctx) return SLOAM_ERR_ALLOC;
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return SLOAM_ERR_HIP;
  }
  // a failure here only disables sloam_project_partition_noalloc, which
  // then reports it
  (void)sloam_project_noalloc_init(ctx);
  *out = ctx;
  return SLOAM_OK;
}

static void graph_release(sloam_graph* g);

void sloam_ctx_destroy(sloam_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  // graphs still alive lose their HIP objects and their context
  for (sloam_graph* g : ctx->graphs) graph_release(g);
  ctx->graphs.clear();
  for (auto& kv : ctx->timers)
    for (auto& pr : kv.second.pending) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto& s : ctx->scratch) sloam_dev_free(s.first);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int sloam_sync(sloam_ctx* ctx) {
  if (!ctx) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SLOAM_OK;
}

const char* sloam_strerror(int status) {
  switch (status) {
    case SLOAM_OK: return "ok";
    case SLOAM_ERR_INVALID_ARG: return "invalid argument";
    case SLOAM_ERR_HIP: return "HIP runtime error";
    case SLOAM_ERR_NO_DEVICE: return "no HIP device";
    case SLOAM_ERR_TOO_LARGE: return "input exceeds a supported maximum";
    case SLOAM_ERR_ALLOC: return "device allocation failed";
    case SLOAM_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

const char* sloam_ctx_last_error(const sloam_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "";
}

void* sloam_ctx_stream(sloam_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

struct sloam_graph {
  sloam_ctx* ctx = nullptr;  // null once the context is destroyed (orphaned)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t slots = 0;  // scratch slots the graph's kernels use (held until destroyed)
};

// Releases g's HIP objects and detaches it from its context (whose stream
// has completed the graph's work).
static void graph_release(sloam_graph* g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
  g->ctx = nullptr;
}

int sloam_graph_begin(sloam_ctx* ctx) {
  if (!ctx || ctx->capturing) return SLOAM_ERR_INVALID_ARG;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  // the stream's earlier work is not part of the graph
  SLOAM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->timing_before_capture = ctx->timing;
  ctx->timing = false;
  const hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    ctx->timing = ctx->timing_before_capture;
    return sloam_record_hip_error(ctx, e, "hipStreamBeginCapture");
  }
  ctx->capturing = true;
  ctx->capture_slots = 0;
  return SLOAM_OK;
}

int sloam_graph_end(sloam_ctx* ctx, sloam_graph** out) {
  if (!ctx || !out || !ctx->capturing) return SLOAM_ERR_INVALID_ARG;
  *out = nullptr;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->capturing = false;
  ctx->timing = ctx->timing_before_capture;
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return sloam_record_hip_error(ctx, e, "hipStreamEndCapture");
  }
  sloam_graph* r = new (std::nothrow) sloam_graph();
  if (!r) {
    (void)hipGraphDestroy(g);
    return SLOAM_ERR_ALLOC;
  }
  r->ctx = ctx;
  r->graph = g;
  const hipError_t ei = hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    delete r;
    return sloam_record_hip_error(ctx, ei, "hipGraphInstantiate");
  }
  r->slots = ctx->capture_slots;
  sloam_ctx_pin_slots(ctx, r->slots, +1);
  ctx->graphs.push_back(r);
  *out = r;
  return SLOAM_OK;
}

int sloam_graph_launch(sloam_graph* g) {
  if (!g || !g->ctx) return SLOAM_ERR_INVALID_ARG;
  sloam_ctx* ctx = g->ctx;
  SLOAM_HIP(ctx, hipSetDevice(ctx->device));
  SloamKernelScope ks(ctx, "graph");
  SLOAM_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
  return SLOAM_OK;
}

void sloam_graph_destroy(sloam_graph* g) {
  if (!g) return;
  sloam_ctx* ctx = g->ctx;
  if (ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    sloam_ctx_pin_slots(ctx, g->slots, -1);
    auto& v = ctx->graphs;
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i] == g) {
        v.erase(v.begin() + (long)i);
        break;
      }
    graph_release(g);
  }
  delete g;
}

int sloam_memcpy_async(sloam_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return SLOAM_ERR_INVALID_ARG;
  if (!bytes) return SLOAM_OK;
  SLOAM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
  return SLOAM_OK;
}

int sloam_ctx_enable_timing(sloam_ctx* ctx, int enable) {
  if (!ctx) return SLOAM_ERR_INVALID_ARG;
  ctx->timing = enable != 0;
  return SLOAM_OK;
}

static int drain_timer(sloam_ctx* ctx, SloamTimer& t) {
  for (auto& pr : t.pending) {
    SLOAM_HIP(ctx, hipEventSynchronize(pr.second));
    float ms = 0.f;
    SLOAM_HIP(ctx, hipEventElapsedTime(&ms, pr.first, pr.second));
    t.ms += ms;
    t.launches += 1;
    ctx->event_pool.push_back(pr.first);
    ctx->event_pool.push_back(pr.second);
  }
  t.pending.clear();
  return SLOAM_OK;
}

int sloam_ctx_kernel_time(sloam_ctx* ctx, const char* kernel, double* ms_total, int64_t* launches) {
  if (!ctx || !kernel) return SLOAM_ERR_INVALID_ARG;
  auto it = ctx->timers.find(kernel);
  if (it == ctx->timers.end()) {
    if (ms_total) *ms_total = 0.0;
    if (launches) *launches = 0;
    return SLOAM_OK;
  }
  int st = drain_timer(ctx, it->second);
  if (st) return st;
  if (ms_total) *ms_total = it->second.ms;
  if (launches) *launches = it->second.launches;
  return SLOAM_OK;
}

int sloam_ctx_reset_timing(sloam_ctx* ctx) {
  if (!ctx) return SLOAM_ERR_INVALID_ARG;
  for (auto& kv : ctx->timers) {
    int st = drain_timer(ctx, kv.second);
    if (st) return st;
    kv.second.ms = 0.0;
    kv.second.launches = 0;
  }
  return SLOAM_OK;
}

}  // extern "C"
