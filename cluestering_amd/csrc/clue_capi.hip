// C ABI of the MI355X-native CLUE library: contexts, pooled workspaces, batch
// descriptors, profiling and graph plans. Declarations: include/clue_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "clue_amd.h"
#include "clue_assoc.h"
#include "clue_internal.h"
#include "clue_runtime.h"

using namespace clue_dev;

namespace {
using namespace clue_rt;
}  // namespace

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free entry point (clue_make_clusters_noalloc):
// one array per GPU the library runs on, held until the process ends. It holds
// the pipeline workspace of one event; kNoAllocScratchBytes covers about a
// million points in 2-D or 3-D. In host code `scratch` names a host-side stand-in:
// its device address comes from hipGetSymbolAddress, taken once per device when
// the first context on that device is created, and only that address is used.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(160) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
constexpr size_t kScratchAlign = 256;

struct ScratchTable {
  std::mutex mu;
  bool done[kScratchDevices] = {};
  hipError_t err[kScratchDevices] = {};
  void* ptr[kScratchDevices] = {};
};
ScratchTable& scratch_table() {
  static ScratchTable t;
  return t;
}

// Takes the device address of `scratch` on `device` (current) once; later calls
// return what the first one found.
hipError_t init_scratch(int device) {
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  ScratchTable& t = scratch_table();
  std::lock_guard<std::mutex> g(t.mu);
  if (!t.done[device]) {
    void* ptr = nullptr;
    t.err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    t.ptr[device] = t.err[device] == hipSuccess ? ptr : nullptr;
    t.done[device] = true;
  }
  return t.err[device];
}

void* scratch_ptr(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  ScratchTable& t = scratch_table();
  std::lock_guard<std::mutex> g(t.mu);
  return t.done[device] ? t.ptr[device] : nullptr;
}

// The descriptors of one event of n points at p_off 0 (what the host plan holds),
// written on the device: chunk c covers [c * kChunk, min(n, (c + 1) * kChunk)).
__global__ void k_single_event_descriptors(EventDesc ev, EventDesc* __restrict__ events,
                                           int32_t* __restrict__ order, Chunk* __restrict__ chunks,
                                           int n_chunks) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) {
    events[0] = ev;
    if (order) order[0] = 0;
  }
  if (c < n_chunks) {
    Chunk ch{};
    ch.event = 0;
    ch.begin = c * kChunk;
    ch.end = min(ev.n, (c + 1) * kChunk);
    chunks[c] = ch;
  }
}

}  // namespace

struct clue_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  Workspace ws;
  bool has_run = false;
  int64_t last_n = 0;
  int32_t* last_seeds = nullptr;
  int32_t* last_nseeds = nullptr;
  bool pending = false;  // an async run may still read the host descriptors below
  BatchPlan plan;        // host descriptors of the last run (kept alive for its copies)
  bool profiling = false;
  hipEvent_t ev[kNumStages + 1] = {};
  bool ev_ready = false;
  bool ev_recorded[kNumStages + 1] = {};
  Workspace aux;  // association maps (its own buffers: a pending run keeps ws)
  bool tuned = false;  // clue_set_tuning fixed `tuning`
  clue_tuning tuning{};
};

struct clue_plan {
  clue_ctx* ctx = nullptr;
  Workspace ws;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void mark_stage(void* self, int stage, hipStream_t s) {
  clue_ctx* c = static_cast<clue_ctx*>(self);
  if (hipEventRecord(c->ev[stage], s) == hipSuccess) c->ev_recorded[stage] = true;
}

clue_status read_stage_times(hipEvent_t* ev, const bool* recorded, const char** names, float* us,
                             int32_t capacity, int32_t* count) {
  int k = 0;
  for (int s = 0; s < kNumStages && k < capacity; ++s) {
    if (!recorded[s] || !recorded[s + 1]) continue;
    float ms = 0.0f;
    CLUE_HIP(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
    if (names) names[k] = stage_name(s);
    if (us) us[k] = ms * 1000.0f;
    ++k;
  }
  *count = k;
  return CLUE_OK;
}

template <class P, class Pts, class Res>
clue_status run_batch(clue_ctx* ctx, const P* p, const Pts* pts, const uint32_t* sizes,
                      int32_t n_events, Res* res, int32_t* event_of_cluster) {
  using T = value_t<P>;
  clue_status st = validate(p);
  if (st) return st;
  if (!res || !res->cluster_index) return fail(CLUE_ERR_INVALID_ARGUMENT, "cluster_index is required");
  if (!sizes || n_events < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "no events");
  ScopedDevice dev(ctx->device);
  if (ctx->pending) {
    CLUE_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pending = false;
  }
  BatchPlan& bp = ctx->plan;
  st = make_batch_plan(p, sizes, n_events, &bp, ctx->tuned ? &ctx->tuning : nullptr);
  if (st) return st;
  st = check_points(p, pts, bp.n_total);
  if (st) return st;
  BatchBuffersT<T> bb;
  ResultArraysT<T> ra;
  // One event: its descriptors are written by a kernel (one launch where the
  // host-to-device copies of the plan take two or three).
  const bool single = bp.events.size() == 1 && bp.events[0].p_off == 0;
  st = prepare_batch(ctx->ws, bp, p, pts, res, ctx->stream, &bb, &ra, !single);
  if (st) return st;
  if (single) {
    const int n_chunks = static_cast<int>(bp.chunks.size());
    hipLaunchKernelGGL(k_single_event_descriptors, dim3(static_cast<unsigned>((n_chunks + 255) / 256)),
                       dim3(256), 0, ctx->stream, bp.events[0],
                       static_cast<EventDesc*>(ctx->ws.buf[S_EVENTS].p),
                       bb.event_order ? static_cast<int32_t*>(ctx->ws.buf[S_EORDER].p) : nullptr,
                       static_cast<Chunk*>(ctx->ws.buf[S_CHUNKS].p), n_chunks);
  }
  ra.event_of_cluster = event_of_cluster;
  const RunParamsT<T> rp = make_run_params(p, pts, bp);
  StageHook hook{mark_stage, ctx};
  if (ctx->profiling && !ctx->ev_ready) {
    for (auto& e : ctx->ev) CLUE_HIP(hipEventCreate(&e));
    ctx->ev_ready = true;
  }
  for (bool& r : ctx->ev_recorded) r = false;
  launch_pipeline(rp, point_arrays(p, pts), bb, ra, ctx->stream,
                  ctx->profiling ? &hook : nullptr);
  CLUE_HIP(hipGetLastError());
  ctx->pending = true;
  ctx->has_run = true;
  ctx->last_n = bp.n_total;
  ctx->last_seeds = bb.seeds;
  ctx->last_nseeds = bb.n_seeds;
  return CLUE_OK;
}

}  // namespace

extern "C" {

const char* clue_last_error(void) { return clue_rt::last_error(); }

const char* clue_version(void) { return "cluestering_amd 0.1.0 (gfx950)"; }

clue_status clue_device_count(int32_t* count) {
  if (!count) return fail(CLUE_ERR_INVALID_ARGUMENT, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return fail(CLUE_ERR_NO_DEVICE, hipGetErrorString(e));
  }
  *count = c;
  return CLUE_OK;
}

clue_status clue_device_name(int32_t device, char* buf, int32_t buflen) {
  if (!buf || buflen < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "null buffer");
  hipDeviceProp_t prop;
  CLUE_HIP(hipGetDeviceProperties(&prop, device));
  std::snprintf(buf, static_cast<size_t>(buflen), "%s (%s)", prop.name, prop.gcnArchName);
  return CLUE_OK;
}

clue_status clue_context_create(int32_t device, void* stream, clue_ctx** out) {
  if (!out) return fail(CLUE_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(CLUE_ERR_NO_DEVICE, "no HIP device found");
  if (device < 0 || device >= count) return fail(CLUE_ERR_NO_DEVICE, "bad device id");
  std::unique_ptr<clue_ctx> c(new clue_ctx());
  c->device = device;
  ScopedDevice dev(device);
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
  } else {
    CLUE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
  }
  // Set-up of the allocation-free entry point: the scratch address on this device
  // (its failure is reported by that entry point only), and host descriptors with
  // room for any event the scratch can hold, so that planning it allocates nothing.
  (void)init_scratch(device);
  c->plan.events.reserve(1);
  c->plan.event_order.reserve(1);
  c->plan.chunks.reserve(kNoAllocScratchBytes / kChunk / 64 + 1);
  *out = c.release();
  return CLUE_OK;
}

clue_status clue_context_destroy(clue_ctx* ctx) {
  if (!ctx) return CLUE_OK;
  {
    ScopedDevice dev(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev_ready)
      for (auto& e : ctx->ev) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return CLUE_OK;
}

clue_status clue_context_stream(clue_ctx* ctx, void** stream) {
  if (!ctx || !stream) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  *stream = ctx->stream;
  return CLUE_OK;
}

clue_status clue_synchronize(clue_ctx* ctx) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  ctx->pending = false;
  return CLUE_OK;
}

clue_status clue_validate_params(const clue_params* params) { return validate(params); }

clue_status clue_make_clusters_async(clue_ctx* ctx, const clue_params* params,
                                     const clue_points* points, clue_results* results) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  if (!points) return fail(CLUE_ERR_INVALID_ARGUMENT, "null points");
  if (points->n < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "the number of points must be positive");
  const uint32_t n = static_cast<uint32_t>(points->n);
  return run_batch(ctx, params, points, &n, 1, results, nullptr);
}

clue_status clue_make_clusters(clue_ctx* ctx, const clue_params* params,
                               const clue_points* points, clue_results* results) {
  clue_status st = clue_make_clusters_async(ctx, params, points, results);
  if (st) return st;
  return clue_synchronize(ctx);
}

clue_status clue_make_clusters_batched(clue_ctx* ctx, const clue_params* params,
                                       const clue_points* points, const uint32_t* sizes,
                                       int32_t n_events, clue_results* results,
                                       int32_t* event_of_cluster, int64_t* n_clusters) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  clue_status st = run_batch(ctx, params, points, sizes, n_events, results, event_of_cluster);
  if (st) return st;
  if (n_clusters) return clue_get_n_clusters(ctx, n_clusters);
  return CLUE_OK;
}

// Allocation-free variant of clue_make_clusters: the pipeline workspace lives in
// the static scratch of the context's device instead of the context's pooled
// buffers. No hipMalloc, hipFree, memset or host-to-device copy: the event's
// descriptors are written by a kernel, and every stage initialises what it reads.
clue_status clue_make_clusters_noalloc(clue_ctx* ctx, const clue_params* params,
                                       const clue_points* points, clue_results* results) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  clue_status st = validate(params);
  if (st) return st;
  if (!points) return fail(CLUE_ERR_INVALID_ARGUMENT, "null points");
  if (points->n < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "the number of points must be positive");
  if (!results || !results->cluster_index)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "cluster_index is required");
  const size_t n_chunks = (static_cast<size_t>(points->n) + kChunk - 1) / kChunk;
  if (n_chunks > ctx->plan.chunks.capacity())
    return fail(CLUE_ERR_CAPACITY, "event larger than the allocation-free scratch holds (" +
                                       std::to_string(kNoAllocScratchBytes >> 20) + " MiB)");
  ScopedDevice dev(ctx->device);
  const hipError_t se = init_scratch(ctx->device);
  if (se != hipSuccess)
    return fail(CLUE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
  char* base = static_cast<char*>(scratch_ptr(ctx->device));
  if (ctx->pending) {
    CLUE_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pending = false;
  }
  BatchPlan& bp = ctx->plan;
  const uint32_t n = static_cast<uint32_t>(points->n);
  st = make_batch_plan(params, &n, 1, &bp, ctx->tuned ? &ctx->tuning : nullptr);
  if (st) return st;
  st = check_points(params, points, bp.n_total);
  if (st) return st;
  size_t need[S_NUM];
  workspace_requests(workspace_shape(bp, params, points, results), need);
  void* slot[S_NUM];
  size_t off = 0;
  for (int k = 0; k < S_NUM; ++k) {
    slot[k] = base + off;
    off += (need[k] + kScratchAlign - 1) / kScratchAlign * kScratchAlign;
  }
  if (off > kNoAllocScratchBytes)
    return fail(CLUE_ERR_CAPACITY, "event needs " + std::to_string(off) +
                                       " bytes of scratch; the allocation-free scratch holds " +
                                       std::to_string(kNoAllocScratchBytes));
  BatchBuffers bb;
  ResultArrays ra;
  st = bind_batch(slot, bp, params, points, results, ctx->stream, &bb, &ra, false);
  if (st) return st;
  hipLaunchKernelGGL(k_single_event_descriptors, dim3(static_cast<unsigned>((n_chunks + 255) / 256)),
                     dim3(256), 0, ctx->stream, bp.events[0], at<EventDesc>(slot, S_EVENTS),
                     bp.event_order.empty() ? nullptr : at<int32_t>(slot, S_EORDER),
                     at<Chunk>(slot, S_CHUNKS), static_cast<int>(n_chunks));
  const RunParams rp = make_run_params(params, points, bp);
  for (bool& r : ctx->ev_recorded) r = false;
  launch_pipeline(rp, point_arrays(params, points), bb, ra, ctx->stream, nullptr);
  CLUE_HIP(hipGetLastError());
  ctx->has_run = true;
  ctx->last_n = bp.n_total;
  ctx->last_seeds = bb.seeds;
  ctx->last_nseeds = bb.n_seeds;
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

}  // extern "C"

namespace {

// Host in, host out, in the precision of P.
template <class P, class Pts, class Res>
clue_status run_host(clue_ctx* ctx, const P* params, int32_t n, const value_t<P>* host_data,
                     int32_t* host_cluster_index, const uint32_t* sizes, int32_t n_events,
                     int32_t* host_is_seed, value_t<P>* host_rho, int32_t* host_nearest_higher) {
  using T = value_t<P>;
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  clue_status st = validate(params);
  if (st) return st;
  if (n < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "the number of points must be positive");
  if (!host_data || !host_cluster_index) return fail(CLUE_ERR_INVALID_ARGUMENT, "null host buffer");
  const int D = params->ndim;
  ScopedDevice dev(ctx->device);
  Workspace& ws = ctx->ws;
  const size_t col = sizeof(T) * static_cast<size_t>(n);
  CLUE_HIP(ws.buf[S_HIN].reserve(col * (D + 1)));
  CLUE_HIP(ws.buf[S_HCL].reserve(sizeof(int32_t) * static_cast<size_t>(n)));
  T* din = ws.buf[S_HIN].as<T>();
  CLUE_HIP(hipMemcpyAsync(din, host_data, col * (D + 1), hipMemcpyHostToDevice, ctx->stream));
  Pts pts{};
  pts.n = n;
  for (int d = 0; d < D; ++d) pts.coords[d] = din + static_cast<size_t>(n) * d;
  pts.weight = din + static_cast<size_t>(n) * D;
  Res res{};
  res.cluster_index = ws.buf[S_HCL].as<int32_t>();
  const uint32_t one = static_cast<uint32_t>(n);
  if (!sizes) {
    sizes = &one;
    n_events = 1;
  }
  st = run_batch(ctx, params, &pts, sizes, n_events, &res, nullptr);
  if (st) return st;
  CLUE_HIP(hipMemcpyAsync(host_cluster_index, res.cluster_index, sizeof(int32_t) * n,
                          hipMemcpyDeviceToHost, ctx->stream));
  if (host_is_seed)
    CLUE_HIP(hipMemcpyAsync(host_is_seed, ws.buf[S_ISSEED].p, sizeof(int32_t) * n,
                            hipMemcpyDeviceToHost, ctx->stream));
  if (host_rho)
    CLUE_HIP(hipMemcpyAsync(host_rho, ws.buf[S_RHO].p, sizeof(T) * n, hipMemcpyDeviceToHost,
                            ctx->stream));
  if (host_nearest_higher)
    CLUE_HIP(hipMemcpyAsync(host_nearest_higher, ws.buf[S_NH].p, sizeof(int32_t) * n,
                            hipMemcpyDeviceToHost, ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

}  // namespace

extern "C" {

clue_status clue_make_clusters_host(clue_ctx* ctx, const clue_params* params, int32_t n,
                                    const float* host_data, int32_t* host_cluster_index,
                                    const uint32_t* sizes, int32_t n_events,
                                    int32_t* host_is_seed, float* host_rho,
                                    int32_t* host_nearest_higher) {
  return run_host<clue_params, clue_points, clue_results>(ctx, params, n, host_data,
                                                          host_cluster_index, sizes, n_events,
                                                          host_is_seed, host_rho,
                                                          host_nearest_higher);
}

clue_status clue_make_clusters_host_f64(clue_ctx* ctx, const clue_params_f64* params, int32_t n,
                                        const double* host_data, int32_t* host_cluster_index,
                                        const uint32_t* sizes, int32_t n_events,
                                        int32_t* host_is_seed, double* host_rho,
                                        int32_t* host_nearest_higher) {
  return run_host<clue_params_f64, clue_points_f64, clue_results_f64>(
      ctx, params, n, host_data, host_cluster_index, sizes, n_events, host_is_seed, host_rho,
      host_nearest_higher);
}

clue_status clue_validate_params_f64(const clue_params_f64* params) { return validate(params); }

clue_status clue_make_clusters_f64(clue_ctx* ctx, const clue_params_f64* params,
                                   const clue_points_f64* points, clue_results_f64* results) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  if (!points) return fail(CLUE_ERR_INVALID_ARGUMENT, "null points");
  if (points->n < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "the number of points must be positive");
  const uint32_t n = static_cast<uint32_t>(points->n);
  clue_status st = run_batch(ctx, params, points, &n, 1, results, nullptr);
  if (st) return st;
  return clue_synchronize(ctx);
}

clue_status clue_make_clusters_batched_f64(clue_ctx* ctx, const clue_params_f64* params,
                                           const clue_points_f64* points, const uint32_t* sizes,
                                           int32_t n_events, clue_results_f64* results,
                                           int32_t* event_of_cluster, int64_t* n_clusters) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  clue_status st = run_batch(ctx, params, points, sizes, n_events, results, event_of_cluster);
  if (st) return st;
  if (n_clusters) return clue_get_n_clusters(ctx, n_clusters);
  return CLUE_OK;
}

clue_status clue_get_n_clusters(clue_ctx* ctx, int64_t* n_clusters) {
  if (!ctx || !n_clusters) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  if (!ctx->has_run)
    return fail(CLUE_ERR_RUNTIME, "Seeds are not available. Please run make_clusters first.");
  int32_t k = 0;
  CLUE_HIP(hipMemcpyAsync(&k, ctx->last_nseeds, sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  *n_clusters = k;
  return CLUE_OK;
}

clue_status clue_get_seeds(clue_ctx* ctx, int32_t* seeds, int64_t capacity, int64_t* count) {
  if (!ctx || !count) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  int64_t k = 0;
  clue_status st = clue_get_n_clusters(ctx, &k);
  if (st) return st;
  *count = k;
  if (k == 0) return CLUE_OK;
  if (!seeds || capacity < k) return fail(CLUE_ERR_CAPACITY, "seed buffer too small");
  CLUE_HIP(hipMemcpyAsync(seeds, ctx->last_seeds, sizeof(int32_t) * k, hipMemcpyDeviceToHost,
                          ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

clue_status clue_make_association(clue_ctx* ctx, const int32_t* keys, int32_t n, int32_t* offsets,
                                  int32_t offsets_capacity, int32_t* members, int32_t* n_bins,
                                  int32_t* n_members) {
  if (!ctx || !offsets || !n_bins || !n_members || (n > 0 && !keys))
    return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  if (n < 0) return fail(CLUE_ERR_INVALID_ARGUMENT, "negative size");
  ScopedDevice dev(ctx->device);
  Workspace& w = ctx->aux;
  int64_t hist = 0, scan = 0;
  assoc_scratch_elems(n, &hist, &scan);
  CLUE_HIP(w.buf[S_K0].reserve(sizeof(uint32_t) * n + 4));
  CLUE_HIP(w.buf[S_V0].reserve(sizeof(int32_t) * n + 4));
  CLUE_HIP(w.buf[S_K1].reserve(sizeof(uint32_t) * n + 4));
  CLUE_HIP(w.buf[S_V1].reserve(sizeof(int32_t) * n + 4));
  CLUE_HIP(w.buf[S_HIST].reserve(sizeof(int32_t) * hist));
  CLUE_HIP(w.buf[S_SCAN].reserve(sizeof(int32_t) * scan));
  CLUE_HIP(w.buf[S_NSEEDS].reserve(sizeof(int32_t) * 4));
  int32_t* stats = w.buf[S_NSEEDS].as<int32_t>();
  CLUE_HIP(assoc_stats(keys, n, stats, ctx->stream));
  int32_t hs[2] = {0, 0};
  CLUE_HIP(hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  // bins = max + 1 (make_associator), never below 0.
  const int64_t bins = std::max<int64_t>(0, static_cast<int64_t>(hs[0]) + 1);
  if (bins >= INT32_MAX) return fail(CLUE_ERR_CAPACITY, "key too large for an offsets array");
  *n_bins = static_cast<int32_t>(bins);
  *n_members = hs[1];
  if (bins + 1 > offsets_capacity)
    return fail(CLUE_ERR_CAPACITY, "offsets buffer too small: need max key + 2 entries");
  if (hs[1] > 0 && !members) return fail(CLUE_ERR_INVALID_ARGUMENT, "null members");
  AssocScratch sc{w.buf[S_K0].as<uint32_t>(), w.buf[S_V0].as<int32_t>(),
                  w.buf[S_K1].as<uint32_t>(), w.buf[S_V1].as<int32_t>(),
                  w.buf[S_HIST].as<int32_t>(), w.buf[S_SCAN].as<int32_t>()};
  CLUE_HIP(assoc_build(keys, hs[1] > 0 ? n : 0, static_cast<int>(bins), sc, offsets, members,
                       hs[1], ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

clue_status clue_debug_poison(clue_ctx* ctx) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  ScopedDevice dev(ctx->device);
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  ctx->pending = false;
  for (DevBuf& b : ctx->ws.buf)
    if (b.p) CLUE_HIP(hipMemsetAsync(b.p, 0xFF, b.cap, ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

clue_status clue_set_tuning(clue_ctx* ctx, const clue_tuning* tuning) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  if (tuning && (tuning->xcd_chunk_log2 < -1 || tuning->xcd_chunk_log2 > 12))
    return fail(CLUE_ERR_INVALID_ARGUMENT, "xcd_chunk_log2 must lie in [-1, 12]");
  if (tuning && tuning->block_waves != 0 && tuning->block_waves != 1 &&
      tuning->block_waves != 2 && tuning->block_waves != 4 && tuning->block_waves != 8 &&
      tuning->block_waves != 16)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "block_waves must be 1, 2, 4, 8 or 16");
  if (tuning && (tuning->persistent < 0 || tuning->persistent > 64))
    return fail(CLUE_ERR_INVALID_ARGUMENT, "persistent must lie in [0, 64] (waves per CU)");
  ctx->tuned = tuning != nullptr;
  if (tuning) {
    ctx->tuning = clue_tuning{};
    ctx->tuning.group_queries = tuning->group_queries ? 1 : 0;
    ctx->tuning.xcd_chunk_log2 = tuning->xcd_chunk_log2;
    ctx->tuning.block_waves = tuning->block_waves ? tuning->block_waves : 4;
    ctx->tuning.nh_plain_scan = tuning->nh_plain_scan ? 1 : 0;
    ctx->tuning.persistent = tuning->persistent;
  }
  return CLUE_OK;
}

clue_status clue_get_tuning(clue_ctx* ctx, const clue_params* params, int32_t n,
                            const uint32_t* sizes, int32_t n_events, clue_tuning* out) {
  if (!ctx || !params || !out) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->tuned) {
    *out = ctx->tuning;
    return CLUE_OK;
  }
  int64_t total = 0, events = 0;
  if (sizes) {
    for (int32_t e = 0; e < n_events; ++e) {
      total += sizes[e];
      events += sizes[e] > 0 ? 1 : 0;
    }
  } else {
    total = n;
    events = 1;
  }
  *out = table_tuning(params->ndim, total, events);
  return CLUE_OK;
}

clue_status clue_set_profiling(clue_ctx* ctx, int32_t enabled) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  ctx->profiling = enabled != 0;
  return CLUE_OK;
}

clue_status clue_stage_times(clue_ctx* ctx, const char** names, float* us, int32_t capacity,
                             int32_t* count) {
  if (!ctx || !count) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  *count = 0;
  if (!ctx->ev_ready) return fail(CLUE_ERR_RUNTIME, "profiling was not enabled for the last run");
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return read_stage_times(ctx->ev, ctx->ev_recorded, names, us, capacity, count);
}

clue_status clue_plan_create(clue_ctx* ctx, const clue_params* params, const clue_points* points,
                             const uint32_t* sizes, int32_t n_events, clue_results* results,
                             int32_t* event_of_cluster, clue_plan** out) {
  if (!ctx || !out) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  clue_status st = validate(params);
  if (st) return st;
  if (!results || !results->cluster_index)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "cluster_index is required");
  if (!sizes || n_events < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "no events");
  BatchPlan bp;
  st = make_batch_plan(params, sizes, n_events, &bp, ctx->tuned ? &ctx->tuning : nullptr);
  if (st) return st;
  st = check_points(params, points, bp.n_total);
  if (st) return st;
  ScopedDevice dev(ctx->device);
  std::unique_ptr<clue_plan> plan(new clue_plan());
  plan->ctx = ctx;
  BatchBuffers bb;
  ResultArrays ra;
  st = prepare_batch(plan->ws, bp, params, points, results, ctx->stream, &bb, &ra);
  if (st) return st;
  CLUE_HIP(hipStreamSynchronize(ctx->stream));  // descriptors uploaded before capture
  ra.event_of_cluster = event_of_cluster;
  const RunParams rp = make_run_params(params, points, bp);
  CLUE_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  launch_pipeline(rp, point_arrays(params, points), bb, ra, ctx->stream, nullptr);
  hipError_t le = hipGetLastError();
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
  if (le != hipSuccess) return fail(CLUE_ERR_HIP, std::string("capture: ") + hipGetErrorString(le));
  if (ce != hipSuccess) return fail(CLUE_ERR_HIP, std::string("end capture: ") + hipGetErrorString(ce));
  plan->graph = g;
  CLUE_HIP(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
  ctx->has_run = true;
  ctx->last_n = bp.n_total;
  ctx->last_seeds = bb.seeds;
  ctx->last_nseeds = bb.n_seeds;
  *out = plan.release();
  return CLUE_OK;
}

clue_status clue_plan_launch(clue_plan* plan) {
  if (!plan || !plan->exec) return fail(CLUE_ERR_INVALID_ARGUMENT, "null plan");
  ScopedDevice dev(plan->ctx->device);
  CLUE_HIP(hipGraphLaunch(plan->exec, plan->ctx->stream));
  return CLUE_OK;
}

clue_status clue_plan_destroy(clue_plan* plan) {
  if (!plan) return CLUE_OK;
  {
    ScopedDevice dev(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->ctx->last_seeds == plan->ws.buf[S_SEEDS].p) {
      plan->ctx->has_run = false;
      plan->ctx->last_seeds = nullptr;
      plan->ctx->last_nseeds = nullptr;
    }
  }
  delete plan;
  return CLUE_OK;
}

clue_status clue_malloc(clue_ctx* ctx, size_t bytes, void** ptr) {
  if (!ctx || !ptr) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  ScopedDevice dev(ctx->device);
  CLUE_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return CLUE_OK;
}

clue_status clue_free(clue_ctx* ctx, void* ptr) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  if (!ptr) return CLUE_OK;
  ScopedDevice dev(ctx->device);
  CLUE_HIP(hipFree(ptr));
  return CLUE_OK;
}

clue_status clue_memcpy_h2d(clue_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  if (!bytes) return CLUE_OK;
  ScopedDevice dev(ctx->device);
  CLUE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

clue_status clue_memcpy_d2h(clue_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  if (!bytes) return CLUE_OK;
  ScopedDevice dev(ctx->device);
  CLUE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

clue_status clue_memset_d8(clue_ctx* ctx, void* dst, int value, size_t bytes) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  if (!bytes) return CLUE_OK;
  ScopedDevice dev(ctx->device);
  CLUE_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
  return CLUE_OK;
}

}  // extern "C"

extern "C" clue_status clue_debug_math_impl(void* stream, int32_t n, const float* a,
                                            const float* b, float* out_sqrt, float* out_exp,
                                            float* out_div);

extern "C" clue_status clue_debug_math(clue_ctx* ctx, int32_t n, const float* a, const float* b,
                                       float* out_sqrt, float* out_exp, float* out_div) {
  if (!ctx || !a || (out_div && !b)) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  ScopedDevice dev(ctx->device);
  clue_status st = clue_debug_math_impl(ctx->stream, n, a, b, out_sqrt, out_exp, out_div);
  if (st) return fail(st, "debug math kernel failed");
  return CLUE_OK;
}

extern "C" clue_status clue_debug_math_f64_impl(void* stream, int32_t n, const double* a,
                                                const double* b, double* out_sqrt,
                                                double* out_exp, double* out_div);

extern "C" clue_status clue_debug_math_f64(clue_ctx* ctx, int32_t n, const double* a,
                                           const double* b, double* out_sqrt, double* out_exp,
                                           double* out_div) {
  if (!ctx || !a || (out_div && !b)) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  ScopedDevice dev(ctx->device);
  clue_status st = clue_debug_math_f64_impl(ctx->stream, n, a, b, out_sqrt, out_exp, out_div);
  if (st) return fail(st, "debug math kernel failed");
  return CLUE_OK;
}

extern "C" clue_status clue_debug_copy_impl(void* stream, void* dst, const void* src,
                                            size_t bytes);

extern "C" clue_status clue_debug_copy(clue_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || !dst || !src) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  if (bytes % 16) return fail(CLUE_ERR_INVALID_ARGUMENT, "bytes must be a multiple of 16");
  ScopedDevice dev(ctx->device);
  clue_status st = clue_debug_copy_impl(ctx->stream, dst, src, bytes);
  if (st) return fail(st, "debug copy kernel failed");
  return CLUE_OK;
}
