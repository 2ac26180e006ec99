// Continuous-batching server for CLUE events on one GPU.
//
// Clients submit events (host SoA, any size up to the per-launch cap) from any
// thread. The submitted points are copied into a pinned host staging ring. One
// worker thread owns the device. Whenever one of its lanes (a stream with its
// own pooled device buffers) is free, it folds every waiting event of the same
// parameter class into one batch, uploads the batch, runs the batched pipeline
// with per-event cluster numbering, and copies the labels back. Lanes run
// concurrently, so batch k+1 is uploaded and computed while batch k finishes.
// Admission is bounded by the HBM budget: a batch grows only while the device
// bytes its lane would then hold (the pipeline workspace sized by the same table
// as every run, plus the lane's staging, columns and labels) stay within the
// lane's share of the budget. The scheduling state machine is host-only code
// (clue_serve_core.h); this file supplies the HIP lanes behind it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "clue_amd.h"
#include "clue_runtime.h"
#include "clue_serve_core.h"

using namespace clue_rt;

namespace {

// Where one event of a batch sits in the device staging area.
struct ReqDesc {
  int64_t src;  // float offset of its (ndim+1) x n block
  int32_t n;
  int32_t dst;  // first point in the batch
};

// Staging block of every event -> batch SoA columns (coordinates, then weights).
__global__ void k_repack(const ReqDesc* __restrict__ reqs, const clue_dev::Chunk* __restrict__ chunks,
                         const float* __restrict__ staging, float* __restrict__ cols, int D,
                         int64_t n_total) {
  const clue_dev::Chunk ch = chunks[blockIdx.x];
  const ReqDesc r = reqs[ch.event];
  for (int i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
    const int k = i - r.dst;
    for (int d = 0; d <= D; ++d) cols[d * n_total + i] = staging[r.src + static_cast<int64_t>(d) * r.n + k];
  }
}

// Points per launch: 32-bit indexing of the batched pipeline, and the labels
// buffer each lane pins on the host.
constexpr int64_t kMaxLanePoints = int64_t(16) << 20;

// Device staging bytes of a batch whose columns take col_bytes: each event may
// bring up to 255 bytes of its ring block's padding.
inline size_t staging_bytes_for(size_t col_bytes, size_t events) { return col_bytes + 256 * events; }

// Device bytes of one lane for a batch of this shape: the pipeline workspace (the
// request table prepare_batch reserves from) plus the lane's own staging,
// columns, labels and descriptor buffers, each as its grow-only buffer would hold
// it. `caps` (NULL: nothing held yet) are the lane's current capacities.
int64_t lane_bytes(const clue_params& p, const clue_srv::BatchShape& shape, const size_t* ws_caps,
                   const size_t* own_caps) {
  BatchPlan bp;
  bp.n_total = shape.points;
  bp.n_tiles_total = shape.tiles;
  bp.max_items = (shape.points + 63) / 64 + shape.tiles;
  bp.tuning = table_tuning(p.ndim, shape.points, shape.events);
  bp.events.resize(static_cast<size_t>(shape.events));
  bp.chunks.resize(static_cast<size_t>(shape.chunks));
  clue_points pts{};
  clue_results res{};
  int32_t dummy = 0;
  res.cluster_index = &dummy;  // the lane's labels buffer; rho, nh, is_seed in the workspace
  size_t need[S_NUM];
  workspace_requests(workspace_shape(bp, &p, &pts, &res), need);
  int64_t total = 0;
  for (int k = 0; k < S_NUM; ++k) total += grown_capacity(ws_caps ? ws_caps[k] : 0, need[k]);
  const size_t col_bytes = sizeof(float) * static_cast<size_t>(shape.points) * (p.ndim + 1);
  const size_t own[4] = {staging_bytes_for(col_bytes, static_cast<size_t>(shape.events)), col_bytes,
                         sizeof(int32_t) * static_cast<size_t>(shape.points),
                         sizeof(ReqDesc) * static_cast<size_t>(shape.events)};
  for (int k = 0; k < 4; ++k) total += grown_capacity(own_caps ? own_caps[k] : 0, own[k]);
  return total;
}

// The lanes on HIP streams. Completion of a launch is signalled by a host
// function queued after its label download; the worker then reads the lane's
// event to tell success from an asynchronous failure.
class HipLanes : public clue_srv::ServeBackend {
 public:
  struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    BatchPlan plan;
    DevBuf staging, cols, labels, reqdesc;
    int32_t* labels_host = nullptr;   // pinned, lane point cap
    ReqDesc* reqdesc_host = nullptr;  // pinned
    size_t reqdesc_cap = 0;
    std::vector<int32_t> offsets;
    HipLanes* owner = nullptr;
  };

  explicit HipLanes(int device) : device_(device) {}
  ~HipLanes() override {
    ScopedDevice dev(device_);
    for (auto& L : lanes_) {
      if (L->stream) {
        release_buffers(*L);
        (void)hipStreamSynchronize(L->stream);
      }
      if (L->labels_host) (void)hipHostFree(L->labels_host);
      if (L->reqdesc_host) (void)hipHostFree(L->reqdesc_host);
      if (L->done) (void)hipEventDestroy(L->done);
      if (L->stream) (void)hipStreamDestroy(L->stream);
    }
    if (pool_) (void)hipMemPoolDestroy(pool_);
  }
  // The lanes' device buffers come from one pool of this device that keeps what
  // is freed (release threshold: no limit), so a lane that grows reuses blocks
  // instead of mapping new memory, and no growth waits for the device.
  clue_status init_pool() {
    ScopedDevice dev(device_);
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device_;
    CLUE_HIP(hipMemPoolCreate(&pool_, &props));
    uint64_t keep_all = UINT64_MAX;
    CLUE_HIP(hipMemPoolSetAttribute(pool_, hipMemPoolAttrReleaseThreshold, &keep_all));
    return CLUE_OK;
  }
  clue_status add_lane(int64_t lane_points) {
    auto L = std::make_unique<Lane>();
    L->owner = this;
    CLUE_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
    CLUE_HIP(hipEventCreateWithFlags(&L->done, hipEventDisableTiming));
    L->ws.set_stream(L->stream, pool_);
    for (DevBuf* b : {&L->staging, &L->cols, &L->labels, &L->reqdesc}) {
      b->stream = L->stream;
      b->pool = pool_;
    }
    CLUE_HIP(hipHostMalloc(reinterpret_cast<void**>(&L->labels_host),
                           sizeof(int32_t) * static_cast<size_t>(lane_points), hipHostMallocDefault));
    lanes_.push_back(std::move(L));
    return CLUE_OK;
  }
  void set_core(clue_srv::ServeCore* core) { core_ = core; }

  int lanes() const override { return static_cast<int>(lanes_.size()); }

  clue_status launch(int lane, const std::vector<clue_srv::RequestP>& batch, const char* ring,
                     std::string* err) override {
    const clue_status st = launch_impl(*lanes_[lane], batch, ring);
    if (st) *err = last_error();
    return st;
  }

  clue_srv::LanePoll poll(int lane, std::string* err) override {
    const hipError_t e = hipEventQuery(lanes_[lane]->done);
    if (e == hipSuccess) return clue_srv::LanePoll::kDone;
    if (e == hipErrorNotReady) return clue_srv::LanePoll::kPending;
    *err = hipGetErrorString(e);
    return clue_srv::LanePoll::kFailed;
  }

  void collect(int lane, const std::vector<clue_srv::RequestP>& batch) override {
    Lane& L = *lanes_[lane];
    for (size_t k = 0; k < batch.size(); ++k) {
      const int32_t* src = L.labels_host + L.offsets[k];
      batch[k]->labels.assign(src, src + batch[k]->n);
    }
  }

  void drain(int lane) override { (void)hipStreamSynchronize(lanes_[lane]->stream); }

  int64_t bytes_after(int lane, const clue_params& p,
                      const clue_srv::BatchShape& shape) const override {
    const Lane& L = *lanes_[lane];
    size_t ws_caps[S_NUM];
    for (int k = 0; k < S_NUM; ++k) ws_caps[k] = L.ws.buf[k].cap;
    const size_t own_caps[4] = {L.staging.cap, L.cols.cap, L.labels.cap, L.reqdesc.cap};
    return lane_bytes(p, shape, ws_caps, own_caps);
  }
  int64_t bytes_fresh(const clue_params& p, const clue_srv::BatchShape& shape) const override {
    return lane_bytes(p, shape, nullptr, nullptr);
  }
  // Gives the lane's device memory back to the driver: the frees are ordered on
  // the lane's stream, which is idle here, and the pool then drops every block no
  // lane holds.
  void trim(int lane) override {
    Lane& L = *lanes_[lane];
    ScopedDevice dev(device_);
    release_buffers(L);
    (void)hipStreamSynchronize(L.stream);
    (void)hipMemPoolTrimTo(pool_, 0);
  }

 private:
  static void release_buffers(Lane& L) {
    for (DevBuf& b : L.ws.buf) b.release();
    L.staging.release();
    L.cols.release();
    L.labels.release();
    L.reqdesc.release();
  }
  static void on_done(void* lane) { static_cast<Lane*>(lane)->owner->core_->notify(); }

  clue_status launch_impl(Lane& L, const std::vector<clue_srv::RequestP>& batch, const char* ring) {
    ScopedDevice dev(device_);
    const clue_params& p = batch.front()->params;
    const int D = p.ndim;
    std::vector<uint32_t> sizes(batch.size());
    int64_t n_total = 0;
    L.offsets.resize(batch.size());
    for (size_t k = 0; k < batch.size(); ++k) {
      sizes[k] = static_cast<uint32_t>(batch[k]->n);
      L.offsets[k] = static_cast<int32_t>(n_total);
      n_total += batch[k]->n;
    }
    clue_status st = make_batch_plan(&p, sizes.data(), static_cast<int32_t>(sizes.size()), &L.plan);
    if (st) return st;
    const size_t col_bytes = sizeof(float) * static_cast<size_t>(n_total) * (D + 1);
    CLUE_HIP(L.staging.reserve(staging_bytes_for(col_bytes, batch.size())));
    CLUE_HIP(L.cols.reserve(col_bytes));
    CLUE_HIP(L.labels.reserve(sizeof(int32_t) * n_total));
    CLUE_HIP(L.reqdesc.reserve(sizeof(ReqDesc) * batch.size()));
    if (L.reqdesc_cap < batch.size()) {
      if (L.reqdesc_host) (void)hipHostFree(L.reqdesc_host);
      L.reqdesc_host = nullptr;
      L.reqdesc_cap = std::max<size_t>(batch.size() * 2, 256);
      CLUE_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.reqdesc_host),
                             sizeof(ReqDesc) * L.reqdesc_cap, hipHostMallocDefault));
    }
    // Every submitted event has n >= 1, so the plan describes the batch's events
    // in order.
    // Requests that sit next to each other in the ring (admitted one after the
    // other) go up in one copy: within such a run staging holds the ring's bytes,
    // 256-byte block padding included, and runs follow each other.
    char* staging_bytes = L.staging.as<char>();
    size_t run_ring = 0, run_dst = 0, run_bytes = 0;
    for (size_t k = 0; k < batch.size(); ++k) {
      const clue_srv::Request& r = *batch[k];
      const size_t bytes = sizeof(float) * static_cast<size_t>(r.n) * (D + 1);
      if (run_bytes > 0 && r.ring_off >= run_ring + run_bytes &&
          r.ring_off - (run_ring + run_bytes) < 256) {
        run_bytes = r.ring_off + bytes - run_ring;
      } else {
        if (run_bytes > 0) {
          CLUE_HIP(hipMemcpyAsync(staging_bytes + run_dst, ring + run_ring, run_bytes,
                                  hipMemcpyHostToDevice, L.stream));
          run_dst += (run_bytes + 255) & ~size_t(255);
        }
        run_ring = r.ring_off;
        run_bytes = bytes;
      }
      const size_t at = run_dst + (r.ring_off - run_ring);  // bytes into staging
      L.reqdesc_host[k] = ReqDesc{static_cast<int64_t>(at / sizeof(float)), r.n, L.offsets[k]};
    }
    if (run_bytes > 0)
      CLUE_HIP(hipMemcpyAsync(staging_bytes + run_dst, ring + run_ring, run_bytes,
                              hipMemcpyHostToDevice, L.stream));
    float* staging = L.staging.as<float>();
    CLUE_HIP(hipMemcpyAsync(L.reqdesc.p, L.reqdesc_host, sizeof(ReqDesc) * batch.size(),
                            hipMemcpyHostToDevice, L.stream));
    clue_points pts{};
    pts.n = static_cast<int32_t>(n_total);
    float* cols = L.cols.as<float>();
    for (int d = 0; d < D; ++d) pts.coords[d] = cols + static_cast<int64_t>(d) * n_total;
    pts.weight = cols + static_cast<int64_t>(D) * n_total;
    clue_results res{};
    res.cluster_index = L.labels.as<int32_t>();
    BatchBuffers bb;
    ResultArrays ra;
    st = prepare_batch(L.ws, L.plan, &p, &pts, &res, L.stream, &bb, &ra);
    if (st) return st;
    ra.local_ids = 1;
    hipLaunchKernelGGL(k_repack, dim3(bb.n_chunks), dim3(256), 0, L.stream,
                       L.reqdesc.as<ReqDesc>(), bb.chunks, staging, cols, D, n_total);
    const RunParams rp = make_run_params(&p, &pts, L.plan);
    launch_pipeline(rp, point_arrays(&p, &pts), bb, ra, L.stream, nullptr);
    CLUE_HIP(hipGetLastError());
    CLUE_HIP(hipMemcpyAsync(L.labels_host, L.labels.p, sizeof(int32_t) * n_total,
                            hipMemcpyDeviceToHost, L.stream));
    CLUE_HIP(hipEventRecord(L.done, L.stream));
    CLUE_HIP(hipLaunchHostFunc(L.stream, &HipLanes::on_done, &L));
    return CLUE_OK;
  }

  int device_;
  hipMemPool_t pool_ = nullptr;
  std::vector<std::unique_ptr<Lane>> lanes_;
  clue_srv::ServeCore* core_ = nullptr;
};

}  // namespace

struct clue_server {
  clue_server_config cfg{};
  HipLanes lanes;
  char* ring = nullptr;
  std::unique_ptr<clue_srv::ServeCore> core;
  explicit clue_server(int device) : lanes(device) {}
  ~clue_server() {
    core.reset();  // finishes queued work and joins the worker
    if (ring) {
      ScopedDevice dev(cfg.device);
      (void)hipHostFree(ring);
    }
  }
};

extern "C" {

clue_status clue_server_create(const clue_server_config* cfg, clue_server** out) {
  if (!cfg || !out) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || cfg->device < 0 || cfg->device >= count)
    return fail(CLUE_ERR_NO_DEVICE, "no such HIP device");
  std::unique_ptr<clue_server> s(new clue_server(cfg->device));
  s->cfg = *cfg;
  if (s->cfg.n_lanes <= 0) s->cfg.n_lanes = 2;
  if (s->cfg.hbm_budget_bytes <= 0) s->cfg.hbm_budget_bytes = int64_t(64) << 30;
  if (s->cfg.staging_bytes <= 0) s->cfg.staging_bytes = int64_t(1) << 30;
  if (s->cfg.max_batch_events <= 0) s->cfg.max_batch_events = 1 << 20;
  clue_srv::ServeCore::Limits lim;
  lim.lane_budget_bytes = s->cfg.hbm_budget_bytes / s->cfg.n_lanes;
  lim.lane_points = s->cfg.max_batch_points > 0 ? std::min<int64_t>(s->cfg.max_batch_points, kMaxLanePoints)
                                                : kMaxLanePoints;
  lim.max_batch_events = s->cfg.max_batch_events;
  lim.staging_bytes = static_cast<size_t>(s->cfg.staging_bytes);
  ScopedDevice dev(cfg->device);
  {
    const clue_status st = s->lanes.init_pool();
    if (st) return st;
  }
  for (int k = 0; k < s->cfg.n_lanes; ++k) {
    const clue_status st = s->lanes.add_lane(lim.lane_points);
    if (st) return st;
  }
  CLUE_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->ring), lim.staging_bytes, hipHostMallocDefault));
  s->core = std::make_unique<clue_srv::ServeCore>(&s->lanes, s->ring, lim);
  s->lanes.set_core(s->core.get());
  s->core->start();
  *out = s.release();
  return CLUE_OK;
}

clue_status clue_server_destroy(clue_server* s) {
  delete s;  // finishes queued work first
  return CLUE_OK;
}

clue_status clue_server_submit(clue_server* s, const clue_params* params, int32_t n,
                               const float* host_data, uint64_t* id) {
  if (!s || !params || !host_data || !id) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  clue_status st = validate(params);
  if (st) return st;
  if (params->metric == CLUE_METRIC_MAHALANOBIS)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "served events carry no sigma (Mahalanobis)");
  if (n < 1) return fail(CLUE_ERR_INVALID_ARGUMENT, "the number of points must be positive");
  std::string err;
  st = s->core->submit(*params, n, host_data, id, &err);
  return st ? fail(st, err) : CLUE_OK;
}

clue_status clue_server_wait(clue_server* s, uint64_t id, int32_t* host_labels,
                             int64_t timeout_us, clue_server_timing* timing) {
  if (!s) return fail(CLUE_ERR_INVALID_ARGUMENT, "null server");
  std::string err;
  const clue_status st = s->core->wait(id, host_labels, timeout_us, timing, &err);
  return st ? fail(st, err) : CLUE_OK;
}

clue_status clue_server_get_stats(clue_server* s, clue_server_stats* out) {
  if (!s || !out) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  s->core->get_stats(out);
  return CLUE_OK;
}

int64_t clue_now_ns(void) { return clue_srv::now_ns(); }

}  // extern "C"
