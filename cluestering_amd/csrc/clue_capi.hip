// C ABI of the MI355X-native CLUE library: contexts, pooled workspaces, batch
// descriptors, profiling and graph plans. Declarations: include/clue_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "clue_amd.h"
#include "clue_internal.h"

using namespace clue_dev;

namespace {

thread_local std::string g_last_error;

clue_status fail(clue_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define CLUE_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(CLUE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

constexpr int kChunk = 2048;  // points per chunk of the per-point grids
constexpr int kDefaultPointsPerTile = 128;

// Device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t reserve(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t want = bytes + bytes / 4 + 256;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Workspace slots of one batch run.
enum Slot {
  S_EVENTS, S_CHUNKS, S_EXT, S_GEO, S_K0, S_V0, S_K1, S_V1, S_HIST, S_SCAN, S_TOFF, S_ICOUNT,
  S_NITEMS, S_ITEMS, S_SRHO, S_SKEY, S_SNH, S_SDELTA, S_FLAG, S_RANK, S_NSEEDS, S_SEEDS, S_SX, S_SW,
  S_SSIG, S_SU, S_STAG, S_SFW, S_PERM, S_RHO, S_NH, S_ISSEED, S_DELTA, S_HIN, S_HCL,
  S_NUM
};

struct Workspace {
  DevBuf buf[S_NUM];
};

// Host-side description of one batch (all derived from the event sizes).
struct BatchPlan {
  std::vector<EventDesc> events;
  std::vector<Chunk> chunks;
  int64_t n_total = 0;
  int64_t n_tiles_total = 0;
  int64_t max_items = 0;
  int key_bits = 0;
};

clue_status make_batch_plan(const clue_params* p, const uint32_t* sizes, int32_t n_events,
                            BatchPlan* bp) {
  const int D = p->ndim;
  const int ppt = p->points_per_tile > 0 ? p->points_per_tile : kDefaultPointsPerTile;
  bp->events.clear();
  bp->chunks.clear();
  int64_t off = 0, tiles = 0;
  for (int32_t e = 0; e < n_events; ++e) {
    const int64_t n = sizes[e];
    if (n == 0) continue;
    // SetupTiles.hpp:30-34: ntiles = ceil(n / ppt); nper = min{k : k^D >= ntiles}.
    const int64_t nt0 = (n + ppt - 1) / ppt;
    int64_t nper = 1;
    auto pw = [D](int64_t b) {
      int64_t r = 1;
      for (int k = 0; k < D; ++k) r *= b;
      return r;
    };
    while (pw(nper) < nt0) ++nper;
    const int64_t nt = pw(nper);
    EventDesc ev{};
    ev.p_off = static_cast<int32_t>(off);
    ev.n = static_cast<int32_t>(n);
    ev.nper = static_cast<int32_t>(nper);
    ev.ntiles = static_cast<int32_t>(nt);
    ev.t_base = static_cast<int32_t>(tiles);
    const int32_t eidx = static_cast<int32_t>(bp->events.size());
    bp->events.push_back(ev);
    for (int64_t b = 0; b < n; b += kChunk) {
      Chunk c{};
      c.event = eidx;
      c.begin = static_cast<int32_t>(off + b);
      c.end = static_cast<int32_t>(off + std::min<int64_t>(n, b + kChunk));
      bp->chunks.push_back(c);
    }
    off += n;
    tiles += nt;
    if (off > INT32_MAX / 2 || tiles > INT32_MAX / 2)
      return fail(CLUE_ERR_CAPACITY, "batch too large for 32-bit indexing");
  }
  bp->n_total = off;
  bp->n_tiles_total = tiles;
  bp->max_items = (off + 63) / 64 + tiles;
  int bits = 0;
  while ((int64_t(1) << bits) < tiles) ++bits;
  bp->key_bits = bits;
  return CLUE_OK;
}

// Largest float t with sqrtf(t) <= r (sqrtf is correctly rounded on the host and on
// gfx950), so that sqrt(t) <= r  <=>  t <= sq_max(r) for every float t.
float sq_max(float r) {
  if (!(r > 0.0f)) return -1.0f;
  if (std::isinf(r)) return INFINITY;
  float t = r * r;
  if (std::isinf(t)) t = FLT_MAX;
  while (t > 0.0f && std::sqrt(t) > r) t = std::nextafter(t, 0.0f);
  while (true) {
    const float nx = std::nextafter(t, INFINITY);
    if (std::isinf(nx) || std::sqrt(nx) > r) break;
    t = nx;
  }
  return t;
}

clue_status validate(const clue_params* p) {
  if (!p) return fail(CLUE_ERR_INVALID_ARGUMENT, "null params");
  if (p->ndim < 1 || p->ndim > kMaxDim)
    return fail(CLUE_ERR_OUT_OF_RANGE, "This library only works up to 10 dimensions");
  if (!(p->dc > 0.0f) || !(p->rhoc >= 0.0f) || !(p->dm > 0.0f) || !(p->seed_dc > 0.0f))
    return fail(CLUE_ERR_INVALID_ARGUMENT,
                "Invalid clustering parameters. The parameters must be positive.");
  switch (p->kernel) {
    case CLUE_KERNEL_FLAT:
      if (!(p->kparam[0] > 0.0f))
        return fail(CLUE_ERR_INVALID_ARGUMENT, "Flat kernel parameter must be positive.");
      break;
    case CLUE_KERNEL_EXP:
      if (!(p->kparam[0] > 0.0f) || !(p->kparam[1] > 0.0f))
        return fail(CLUE_ERR_INVALID_ARGUMENT,
                    "Exponential kernel parameters must be positive.");
      break;
    case CLUE_KERNEL_GAUSS:
      if (!(p->kparam[0] > 0.0f) || !(p->kparam[1] > 0.0f))
        return fail(CLUE_ERR_INVALID_ARGUMENT, "Gaussian kernel parameters must be positive.");
      break;
    default:
      return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown convolutional kernel");
  }
  if (p->metric < CLUE_METRIC_EUCLIDEAN || p->metric > CLUE_METRIC_MAHALANOBIS)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown distance metric");
  if (p->points_per_tile < 0) return fail(CLUE_ERR_INVALID_ARGUMENT, "points_per_tile < 0");
  return CLUE_OK;
}

RunParams make_run_params(const clue_params* p, const clue_points* pts, int n_events,
                          int64_t n_total) {
  RunParams rp{};
  rp.D = p->ndim;
  rp.kernel = p->kernel;
  rp.metric = p->metric;
  rp.has_sigma = pts->sigma[0] != nullptr;
  rp.has_u = pts->density_uncertainty != nullptr;
  rp.has_tags = pts->tags != nullptr;
  rp.n_events = n_events;
  rp.n_total = static_cast<int32_t>(n_total);
  rp.dc = p->dc;
  rp.rhoc = p->rhoc;
  rp.dm = p->dm;
  rp.seed_dc = p->seed_dc;
  rp.dc_sq_max = sq_max(p->dc);
  rp.dm_sq_max = sq_max(p->dm);
  rp.seed_sq_max = sq_max(p->seed_dc);
  rp.kp0 = p->kparam[0];
  rp.kp1 = p->kparam[1];
  // 2 * sigma * sigma, evaluated left to right (ConvolutionalKernel.hpp:40).
  rp.gauss_den = (2.0f * p->kparam[0]) * p->kparam[0];
  {
    int ex = 0;
    const float m = std::frexp(rp.gauss_den, &ex);
    // Power of two with a normal, finite reciprocal.
    rp.gauss_pow2 = (m == 0.5f && rp.gauss_den >= 0x1p-120f && rp.gauss_den <= 0x1p120f) ? 1 : 0;
    rp.gauss_inv_den = rp.gauss_pow2 ? std::ldexp(1.0f, 1 - ex) : 0.0f;
  }
  // Within the radius the kernel's exp argument is -lambda*d (exp) or -d*d/(2s^2)
  // (gauss), with 0 <= d <= dc; bound its magnitude with a 1% margin for rounding.
  {
    const double dc = p->dc;
    double mag = 0.0;
    if (p->kernel == CLUE_KERNEL_EXP) mag = static_cast<double>(p->kparam[0]) * dc;
    if (p->kernel == CLUE_KERNEL_GAUSS) mag = dc * dc / static_cast<double>(rp.gauss_den);
    rp.exp_in_range = (std::isfinite(mag) && mag * 1.01 < 100.0) ? 1 : 0;
  }
  for (int d = 0; d < kMaxDim; ++d) {
    rp.wrapped[d] = d < p->ndim ? (p->wrapped[d] ? 1 : 0) : 0;
    rp.mparam[d] = d < p->ndim ? p->mparam[d] : 0.0f;
  }
  // Weighted Euclidean with unit weights evaluates (1 * dx) * dx == dx * dx.
  if (rp.metric == CLUE_METRIC_WEIGHTED_EUCLIDEAN) {
    bool unit = true;
    for (int d = 0; d < p->ndim; ++d) unit = unit && p->mparam[d] == 1.0f;
    if (unit) rp.metric = CLUE_METRIC_EUCLIDEAN;
  }
  return rp;
}

clue_status check_points(const clue_params* p, const clue_points* pts, int64_t n_expected) {
  if (!pts) return fail(CLUE_ERR_INVALID_ARGUMENT, "null points");
  if (pts->n < 1 || (n_expected >= 0 && pts->n != n_expected))
    return fail(CLUE_ERR_INVALID_ARGUMENT, "the number of points must be positive");
  for (int d = 0; d < p->ndim; ++d)
    if (!pts->coords[d]) return fail(CLUE_ERR_INVALID_ARGUMENT, "missing coordinate column");
  if (!pts->weight) return fail(CLUE_ERR_INVALID_ARGUMENT, "missing weights");
  const bool any_sigma = pts->sigma[0] != nullptr;
  for (int d = 0; d < p->ndim; ++d)
    if ((pts->sigma[d] != nullptr) != any_sigma)
      return fail(CLUE_ERR_INVALID_ARGUMENT, "sigma must be set for every dimension or none");
  if (p->metric == CLUE_METRIC_MAHALANOBIS && !any_sigma)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "the Mahalanobis metric needs sigma in every dimension");
  return CLUE_OK;
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
};

struct clue_plan {
  clue_ctx* ctx = nullptr;
  Workspace ws;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

struct ScopedDevice {
  int prev = 0;
  explicit ScopedDevice(int d) {
    (void)hipGetDevice(&prev);
    if (prev != d) (void)hipSetDevice(d);
  }
  ~ScopedDevice() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != prev) (void)hipSetDevice(prev);
  }
};

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

// Reserves the workspace of one batch and fills the BatchBuffers. Descriptor
// uploads are enqueued on the stream (they precede any kernel that reads them).
clue_status prepare_batch(Workspace& ws, const BatchPlan& bp, const clue_params* p,
                          const clue_points* pts, clue_results* res, hipStream_t s,
                          BatchBuffers* bb, ResultArrays* ra) {
  const int D = p->ndim;
  const int64_t n = bp.n_total;
  const int64_t E = static_cast<int64_t>(bp.events.size());
  const int64_t C = static_cast<int64_t>(bp.chunks.size());
  const int64_t T = bp.n_tiles_total;
  const int64_t hist = radix_hist_elems(n);
  const int64_t scan_elems =
      std::max({scan_temp_elems(n), scan_temp_elems(hist), scan_temp_elems(T)}) + 16;
  const int64_t stride = (n + 63) / 64 * 64;
  const bool has_sigma = pts->sigma[0] != nullptr;
  struct Req {
    Slot s;
    size_t bytes;
  };
  const Req reqs[] = {
      {S_EVENTS, sizeof(EventDesc) * E},
      {S_CHUNKS, sizeof(Chunk) * C},
      {S_EXT, sizeof(int32_t) * E * D * 2},
      {S_GEO, sizeof(float) * E * 3 * D},
      {S_K0, sizeof(uint32_t) * n},
      {S_V0, sizeof(int32_t) * n},
      {S_K1, sizeof(uint32_t) * n},
      {S_V1, sizeof(int32_t) * n},
      {S_HIST, sizeof(int32_t) * hist},
      {S_SCAN, sizeof(int32_t) * scan_elems},
      {S_TOFF, sizeof(int32_t) * (T + 1)},
      {S_ICOUNT, sizeof(int32_t) * T},
      {S_NITEMS, sizeof(int32_t) * 4},
      {S_ITEMS, sizeof(Item) * bp.max_items},
      {S_SRHO, sizeof(float) * n},
      {S_SKEY, sizeof(uint64_t) * n},
      {S_SFW, p->kernel == CLUE_KERNEL_FLAT ? sizeof(float) * n : 0},
      {S_SNH, sizeof(int32_t) * n},
      {S_SDELTA, sizeof(float) * n},
      {S_FLAG, sizeof(int32_t) * n},
      {S_RANK, sizeof(int32_t) * n},
      {S_NSEEDS, sizeof(int32_t) * 4},
      {S_SEEDS, sizeof(int32_t) * n},
      {S_SX, sizeof(float) * stride * D},
      {S_SW, sizeof(float) * n},
      {S_SSIG, has_sigma ? sizeof(float) * stride * D : 0},
      {S_SU, pts->density_uncertainty ? sizeof(float) * n : 0},
      {S_STAG, sizeof(uint32_t) * n},
      {S_PERM, sizeof(int32_t) * n},
      {S_RHO, res->rho ? 0 : sizeof(float) * n},
      {S_NH, res->nearest_higher ? 0 : sizeof(int32_t) * n},
      {S_ISSEED, res->is_seed ? 0 : sizeof(int32_t) * n},
  };
  for (const Req& r : reqs) CLUE_HIP(ws.buf[r.s].reserve(r.bytes));
  CLUE_HIP(hipMemcpyAsync(ws.buf[S_EVENTS].p, bp.events.data(), sizeof(EventDesc) * E,
                          hipMemcpyHostToDevice, s));
  CLUE_HIP(hipMemcpyAsync(ws.buf[S_CHUNKS].p, bp.chunks.data(), sizeof(Chunk) * C,
                          hipMemcpyHostToDevice, s));
  BatchBuffers& b = *bb;
  b = BatchBuffers{};
  b.n_events = static_cast<int32_t>(E);
  b.n_total = static_cast<int32_t>(n);
  b.n_chunks = static_cast<int32_t>(C);
  b.n_tiles_total = static_cast<int32_t>(T);
  b.max_items = static_cast<int32_t>(bp.max_items);
  b.key_bits = bp.key_bits;
  b.events = ws.buf[S_EVENTS].as<EventDesc>();
  b.chunks = ws.buf[S_CHUNKS].as<Chunk>();
  b.ext = ws.buf[S_EXT].as<int32_t>();
  b.geo = ws.buf[S_GEO].as<float>();
  b.keys0 = ws.buf[S_K0].as<uint32_t>();
  b.vals0 = ws.buf[S_V0].as<int32_t>();
  b.keys1 = ws.buf[S_K1].as<uint32_t>();
  b.vals1 = ws.buf[S_V1].as<int32_t>();
  b.rs_hist = ws.buf[S_HIST].as<int32_t>();
  b.scan_tmp = ws.buf[S_SCAN].as<int32_t>();
  b.toff = ws.buf[S_TOFF].as<int32_t>();
  b.icount = ws.buf[S_ICOUNT].as<int32_t>();
  b.n_items = ws.buf[S_NITEMS].as<int32_t>();
  b.items = ws.buf[S_ITEMS].as<Item>();
  b.srho = ws.buf[S_SRHO].as<float>();
  b.skey = ws.buf[S_SKEY].as<uint64_t>();
  b.snh = ws.buf[S_SNH].as<int32_t>();
  b.sdelta = ws.buf[S_SDELTA].as<float>();
  b.flag = ws.buf[S_FLAG].as<int32_t>();
  b.rank = ws.buf[S_RANK].as<int32_t>();
  b.n_seeds = ws.buf[S_NSEEDS].as<int32_t>();
  b.seeds = ws.buf[S_SEEDS].as<int32_t>();
  b.sorted_stride = stride;
  for (int d = 0; d < kMaxDim; ++d) {
    b.sorted.x[d] = d < D ? ws.buf[S_SX].as<float>() + stride * d : nullptr;
    b.sorted.sigma[d] = (has_sigma && d < D) ? ws.buf[S_SSIG].as<float>() + stride * d : nullptr;
  }
  b.sorted.w = ws.buf[S_SW].as<float>();
  b.sorted.fw = p->kernel == CLUE_KERNEL_FLAT ? ws.buf[S_SFW].as<float>() : nullptr;
  b.sorted.u = pts->density_uncertainty ? ws.buf[S_SU].as<float>() : nullptr;
  b.sorted.tag = ws.buf[S_STAG].as<uint32_t>();
  b.sorted.perm = ws.buf[S_PERM].as<int32_t>();
  ra->cluster = res->cluster_index;
  ra->is_seed = res->is_seed ? res->is_seed : ws.buf[S_ISSEED].as<int32_t>();
  ra->rho = res->rho ? res->rho : ws.buf[S_RHO].as<float>();
  ra->nh = res->nearest_higher ? res->nearest_higher : ws.buf[S_NH].as<int32_t>();
  ra->delta = res->delta;
  ra->event_of_cluster = nullptr;
  return CLUE_OK;
}

PointArrays point_arrays(const clue_params* p, const clue_points* pts) {
  PointArrays a{};
  for (int d = 0; d < kMaxDim; ++d) {
    a.x[d] = d < p->ndim ? pts->coords[d] : nullptr;
    a.sigma[d] = d < p->ndim ? pts->sigma[d] : nullptr;
  }
  a.w = pts->weight;
  a.u = pts->density_uncertainty;
  a.tags = pts->tags;
  return a;
}

clue_status run_batch(clue_ctx* ctx, const clue_params* p, const clue_points* pts,
                      const uint32_t* sizes, int32_t n_events, clue_results* res,
                      int32_t* event_of_cluster) {
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
  st = make_batch_plan(p, sizes, n_events, &bp);
  if (st) return st;
  st = check_points(p, pts, bp.n_total);
  if (st) return st;
  BatchBuffers bb;
  ResultArrays ra;
  st = prepare_batch(ctx->ws, bp, p, pts, res, ctx->stream, &bb, &ra);
  if (st) return st;
  ra.event_of_cluster = event_of_cluster;
  const RunParams rp = make_run_params(p, pts, static_cast<int>(bp.events.size()), bp.n_total);
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

const char* clue_last_error(void) { return g_last_error.c_str(); }

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

clue_status clue_make_clusters_host(clue_ctx* ctx, const clue_params* params, int32_t n,
                                    const float* host_data, int32_t* host_cluster_index,
                                    const uint32_t* sizes, int32_t n_events,
                                    int32_t* host_is_seed, float* host_rho,
                                    int32_t* host_nearest_higher) {
  if (!ctx) return fail(CLUE_ERR_INVALID_ARGUMENT, "null context");
  clue_status st = validate(params);
  if (st) return st;
  if (n < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "the number of points must be positive");
  if (!host_data || !host_cluster_index) return fail(CLUE_ERR_INVALID_ARGUMENT, "null host buffer");
  const int D = params->ndim;
  ScopedDevice dev(ctx->device);
  Workspace& ws = ctx->ws;
  const size_t col = sizeof(float) * static_cast<size_t>(n);
  CLUE_HIP(ws.buf[S_HIN].reserve(col * (D + 1)));
  CLUE_HIP(ws.buf[S_HCL].reserve(sizeof(int32_t) * static_cast<size_t>(n)));
  float* din = ws.buf[S_HIN].as<float>();
  CLUE_HIP(hipMemcpyAsync(din, host_data, col * (D + 1), hipMemcpyHostToDevice, ctx->stream));
  clue_points pts{};
  pts.n = n;
  for (int d = 0; d < D; ++d) pts.coords[d] = din + static_cast<size_t>(n) * d;
  pts.weight = din + static_cast<size_t>(n) * D;
  clue_results res{};
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
    CLUE_HIP(hipMemcpyAsync(host_rho, ws.buf[S_RHO].p, sizeof(float) * n, hipMemcpyDeviceToHost,
                            ctx->stream));
  if (host_nearest_higher)
    CLUE_HIP(hipMemcpyAsync(host_nearest_higher, ws.buf[S_NH].p, sizeof(int32_t) * n,
                            hipMemcpyDeviceToHost, ctx->stream));
  CLUE_HIP(hipStreamSynchronize(ctx->stream));
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
  st = make_batch_plan(params, sizes, n_events, &bp);
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
  const RunParams rp =
      make_run_params(params, points, static_cast<int>(bp.events.size()), bp.n_total);
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
