// Continuous-batching server for CLUE events on one GPU.
//
// Clients submit events (host SoA, any size up to the per-launch cap) from any
// thread. The submitted points are copied into a pinned host staging ring. One
// worker thread owns the device. Whenever one of its lanes (a stream with its
// own pooled device buffers) is free, it folds every waiting event of the same
// parameter class into one batch, uploads the batch, runs the batched pipeline
// with per-event cluster numbering, and copies the labels back. Lanes run
// concurrently, so batch k+1 is uploaded and computed while batch k finishes.
// Admission is bounded by the HBM budget through the per-lane point cap.
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

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Device bytes per point of one lane (staging, columns, labels and the batch
// workspace), an upper estimate used for the HBM budget.
constexpr int64_t kLaneBytesPerPoint = 192;

bool same_class(const clue_params& a, const clue_params& b) {
  return std::memcmp(&a, &b, sizeof(clue_params)) == 0;
}

}  // namespace

struct clue_server {
  struct Request {
    uint64_t id = 0;
    clue_params params{};
    int32_t n = 0;
    size_t ring_off = 0;
    size_t ring_bytes = 0;
    int64_t submit_ns = 0, launch_ns = 0, done_ns = 0;
    int32_t batch_events = 0;
    int64_t batch_points = 0;
    std::vector<int32_t> labels;
    bool done = false;
    bool ring_released = false;
    clue_status status = CLUE_OK;
    std::string error;
  };
  using RequestP = std::shared_ptr<Request>;

  struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    BatchPlan plan;
    DevBuf staging, cols, labels, reqdesc;
    int32_t* labels_host = nullptr;  // pinned, lane capacity
    ReqDesc* reqdesc_host = nullptr;  // pinned
    size_t reqdesc_cap = 0;
    std::vector<RequestP> reqs;
    std::vector<int32_t> offsets;
    bool busy = false;
  };

  clue_server_config cfg{};
  int64_t lane_points = 0;
  std::vector<std::unique_ptr<Lane>> lanes;

  // Pinned staging ring (FIFO allocation, released in order of submission).
  char* ring = nullptr;
  size_t ring_cap = 0;
  std::deque<RequestP> ring_live;

  std::deque<RequestP> waiting;
  std::unordered_map<uint64_t, RequestP> table;
  std::mutex mu;
  std::condition_variable cv_done, cv_space, cv_work;
  std::thread worker;
  bool stop = false;
  uint64_t next_id = 1;
  clue_server_stats stats{};
  std::string fatal;

  // Worker-side failure: recorded and reported to every later call.
  void set_fatal(const std::string& msg) {
    std::lock_guard<std::mutex> g(mu);
    if (fatal.empty()) fatal = msg;
    cv_done.notify_all();
    cv_space.notify_all();
  }

  // Ring allocation; caller holds mu. Returns false when there is no room now.
  // Live blocks are contiguous in ring order from the oldest (tail) to the newest;
  // the live region has wrapped when the oldest block sits above the newest.
  bool ring_alloc(size_t bytes, size_t* off) {
    if (bytes > ring_cap) return false;
    if (ring_live.empty()) {
      *off = 0;
      return true;
    }
    const size_t tail = ring_live.front()->ring_off;
    const size_t head = ring_live.back()->ring_off + ring_live.back()->ring_bytes;
    const bool wrapped = ring_live.front()->ring_off > ring_live.back()->ring_off;
    if (!wrapped) {
      if (ring_cap - head >= bytes) {
        *off = head;
        return true;
      }
      if (tail >= bytes) {
        *off = 0;
        return true;
      }
      return false;
    }
    if (tail - head >= bytes) {
      *off = head;
      return true;
    }
    return false;
  }

  // Caller holds mu: releases the ring blocks of finished requests in FIFO order.
  void ring_release() {
    while (!ring_live.empty() && ring_live.front()->ring_released) ring_live.pop_front();
    cv_space.notify_all();
  }

  clue_status launch(Lane& L, std::vector<RequestP>& batch) {
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
    CLUE_HIP(L.staging.reserve(col_bytes));
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
    // Events with n > 0 are the batch's events, in order (make_batch_plan skips none
    // here: every submitted event has n >= 1).
    float* staging = L.staging.as<float>();
    int64_t src = 0;
    for (size_t k = 0; k < batch.size(); ++k) {
      const Request& r = *batch[k];
      const size_t bytes = sizeof(float) * static_cast<size_t>(r.n) * (D + 1);
      CLUE_HIP(hipMemcpyAsync(staging + src, ring + r.ring_off, bytes, hipMemcpyHostToDevice,
                              L.stream));
      L.reqdesc_host[k] = ReqDesc{src, r.n, L.offsets[k]};
      src += static_cast<int64_t>(r.n) * (D + 1);
    }
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
    const int64_t t = now_ns();
    for (auto& r : batch) {
      r->launch_ns = t;
      r->batch_events = static_cast<int32_t>(batch.size());
      r->batch_points = n_total;
    }
    L.reqs = std::move(batch);
    L.busy = true;
    return CLUE_OK;
  }

  void finish(Lane& L) {
    const int64_t t = now_ns();
    for (size_t k = 0; k < L.reqs.size(); ++k) {
      Request& r = *L.reqs[k];
      r.labels.assign(L.labels_host + L.offsets[k], L.labels_host + L.offsets[k] + r.n);
    }
    {
      std::lock_guard<std::mutex> g(mu);
      for (auto& r : L.reqs) {
        r->done_ns = t;
        r->done = true;
        r->ring_released = true;
        stats.events_done += 1;
        stats.points_done += r->n;
      }
      stats.batches += 1;
      ring_release();
    }
    cv_done.notify_all();
    L.reqs.clear();
    L.busy = false;
  }

  // Takes the waiting events of the oldest event's class, up to the lane cap.
  std::vector<RequestP> form_batch() {
    std::vector<RequestP> out;
    if (waiting.empty()) return out;
    const clue_params cls = waiting.front()->params;
    int64_t pts = 0;
    for (auto it = waiting.begin(); it != waiting.end();) {
      RequestP r = *it;
      if (same_class(r->params, cls) && pts + r->n <= lane_points &&
          static_cast<int64_t>(out.size()) < cfg.max_batch_events) {
        pts += r->n;
        out.push_back(r);
        it = waiting.erase(it);
      } else {
        ++it;
        if (pts >= lane_points) break;
      }
    }
    return out;
  }

  void run() {
    (void)hipSetDevice(cfg.device);
    while (true) {
      bool progressed = false;
      for (auto& L : lanes) {
        if (L->busy && hipEventQuery(L->done) == hipSuccess) {
          finish(*L);
          progressed = true;
        }
      }
      for (auto& L : lanes) {
        if (L->busy) continue;
        std::vector<RequestP> batch;
        {
          std::lock_guard<std::mutex> g(mu);
          batch = form_batch();
        }
        if (batch.empty()) break;
        std::vector<RequestP> keep = batch;
        const clue_status st = launch(*L, batch);
        if (st) {
          std::lock_guard<std::mutex> g(mu);
          for (auto& r : keep) {
            r->status = st;
            r->error = clue_rt::last_error();
            r->done = true;
            r->done_ns = now_ns();
            r->ring_released = true;
          }
          ring_release();
          cv_done.notify_all();
          L->busy = false;
          L->reqs.clear();
        }
        progressed = true;
      }
      std::unique_lock<std::mutex> lk(mu);
      bool any_busy = false;
      for (auto& L : lanes) any_busy = any_busy || L->busy;
      if (stop && waiting.empty() && !any_busy) break;
      if (!progressed) {
        if (any_busy)
          cv_work.wait_for(lk, std::chrono::microseconds(20));
        else
          cv_work.wait_for(lk, std::chrono::milliseconds(5),
                           [&] { return stop || !waiting.empty(); });
      }
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
  std::unique_ptr<clue_server> s(new clue_server());
  s->cfg = *cfg;
  if (s->cfg.n_lanes <= 0) s->cfg.n_lanes = 2;
  if (s->cfg.hbm_budget_bytes <= 0) s->cfg.hbm_budget_bytes = int64_t(64) << 30;
  if (s->cfg.staging_bytes <= 0) s->cfg.staging_bytes = int64_t(1) << 30;
  if (s->cfg.max_batch_events <= 0) s->cfg.max_batch_events = 1 << 20;
  const int64_t by_budget = s->cfg.hbm_budget_bytes / (s->cfg.n_lanes * kLaneBytesPerPoint);
  s->lane_points = s->cfg.max_batch_points > 0 ? std::min(s->cfg.max_batch_points, by_budget)
                                               : std::min<int64_t>(by_budget, 16 << 20);
  if (s->lane_points < 1) return fail(CLUE_ERR_CAPACITY, "HBM budget too small for one point");
  ScopedDevice dev(cfg->device);
  for (int k = 0; k < s->cfg.n_lanes; ++k) {
    auto L = std::make_unique<clue_server::Lane>();
    CLUE_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
    CLUE_HIP(hipEventCreateWithFlags(&L->done, hipEventDisableTiming));
    CLUE_HIP(hipHostMalloc(reinterpret_cast<void**>(&L->labels_host),
                           sizeof(int32_t) * static_cast<size_t>(s->lane_points),
                           hipHostMallocDefault));
    s->lanes.push_back(std::move(L));
  }
  s->ring_cap = static_cast<size_t>(s->cfg.staging_bytes);
  CLUE_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->ring), s->ring_cap, hipHostMallocDefault));
  s->worker = std::thread([srv = s.get()] { srv->run(); });
  *out = s.release();
  return CLUE_OK;
}

clue_status clue_server_destroy(clue_server* s) {
  if (!s) return CLUE_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    ScopedDevice dev(s->cfg.device);
    for (auto& L : s->lanes) {
      (void)hipStreamSynchronize(L->stream);
      if (L->labels_host) (void)hipHostFree(L->labels_host);
      if (L->reqdesc_host) (void)hipHostFree(L->reqdesc_host);
      (void)hipEventDestroy(L->done);
      (void)hipStreamDestroy(L->stream);
    }
    s->lanes.clear();
    if (s->ring) (void)hipHostFree(s->ring);
  }
  delete s;
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
  if (n > s->lane_points)
    return fail(CLUE_ERR_CAPACITY, "event larger than the per-launch point cap of the HBM budget");
  const size_t bytes = (sizeof(float) * static_cast<size_t>(n) * (params->ndim + 1) + 255) & ~size_t(255);
  auto r = std::make_shared<clue_server::Request>();
  r->params = *params;
  r->n = n;
  r->ring_bytes = bytes;
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->fatal.empty()) return fail(CLUE_ERR_RUNTIME, s->fatal);
  if (bytes > s->ring_cap) return fail(CLUE_ERR_CAPACITY, "event larger than the staging ring");
  size_t off = 0;
  s->cv_space.wait(lk, [&] { return s->ring_alloc(bytes, &off) || !s->fatal.empty(); });
  if (!s->fatal.empty()) return fail(CLUE_ERR_RUNTIME, s->fatal);
  r->ring_off = off;
  r->id = s->next_id++;
  s->ring_live.push_back(r);
  lk.unlock();
  // Copy outside the lock: the block is reserved for this request.
  std::memcpy(s->ring + off, host_data, sizeof(float) * static_cast<size_t>(n) * (params->ndim + 1));
  r->submit_ns = now_ns();
  lk.lock();
  s->table[r->id] = r;
  s->waiting.push_back(r);
  s->stats.events_submitted += 1;
  lk.unlock();
  s->cv_work.notify_one();
  *id = r->id;
  return CLUE_OK;
}

clue_status clue_server_wait(clue_server* s, uint64_t id, int32_t* host_labels,
                             int64_t timeout_us, clue_server_timing* timing) {
  if (!s) return fail(CLUE_ERR_INVALID_ARGUMENT, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->table.find(id);
  if (it == s->table.end()) return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown request id");
  clue_server::RequestP r = it->second;
  auto ready = [&] { return r->done || !s->fatal.empty(); };
  if (timeout_us < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
    return fail(CLUE_ERR_RUNTIME, "timeout");
  }
  if (!r->done) return fail(CLUE_ERR_RUNTIME, s->fatal);
  s->table.erase(id);
  lk.unlock();
  if (r->status) return fail(r->status, r->error);
  if (host_labels) std::memcpy(host_labels, r->labels.data(), sizeof(int32_t) * r->labels.size());
  if (timing) {
    timing->submit_ns = r->submit_ns;
    timing->launch_ns = r->launch_ns;
    timing->done_ns = r->done_ns;
    timing->batch_events = r->batch_events;
    timing->batch_points = r->batch_points;
  }
  return CLUE_OK;
}

clue_status clue_server_get_stats(clue_server* s, clue_server_stats* out) {
  if (!s || !out) return fail(CLUE_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;
  out->lane_points = s->lane_points;
  return CLUE_OK;
}

int64_t clue_now_ns(void) { return now_ns(); }

}  // extern "C"
