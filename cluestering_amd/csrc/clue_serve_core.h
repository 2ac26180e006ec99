// Host side of the continuous-batching server: requests, the pinned staging ring,
// admission under the HBM budget, and the worker's state machine. No HIP here:
// the device work sits behind ServeBackend, which clue_serve.hip implements with
// HIP streams and the CPU tests implement with a scripted fake (so the state
// machine, including its failure paths, runs under the sanitizers without a GPU).
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "clue_amd.h"
#include "clue_host.h"

namespace clue_srv {

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
  clue_status status = CLUE_OK;
  std::string error;
};
using RequestP = std::shared_ptr<Request>;

inline int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Byte ring of the pinned staging buffer. Blocks are taken in submission order
// and freed in that order: a block whose request finished early stays reserved
// until every older block is free.
class StagingRing {
 public:
  void reset(size_t cap) {
    cap_ = cap;
    live_.clear();
  }
  size_t capacity() const { return cap_; }
  size_t live_blocks() const { return live_.size(); }
  // Reserves `bytes` for request `id`; false when there is no room now.
  bool alloc(uint64_t id, size_t bytes, size_t* off) {
    if (bytes > cap_ || bytes == 0) return false;
    size_t at = 0;
    if (!live_.empty()) {
      const size_t tail = live_.front().off;
      const size_t head = live_.back().off + live_.back().bytes;
      if (tail < head) {  // live region [tail, head)
        if (cap_ - head >= bytes)
          at = head;
        else if (tail >= bytes)
          at = 0;
        else
          return false;
      } else {  // wrapped: live region [tail, cap) + [0, head)
        if (tail - head >= bytes)
          at = head;
        else
          return false;
      }
    }
    live_.push_back(Block{id, at, bytes, false});
    *off = at;
    return true;
  }
  // Frees the block of `id` once every older block is free too.
  void release(uint64_t id) {
    for (Block& b : live_)
      if (b.id == id) b.released = true;
    while (!live_.empty() && live_.front().released) live_.pop_front();
  }

 private:
  struct Block {
    uint64_t id;
    size_t off, bytes;
    bool released;
  };
  std::deque<Block> live_;
  size_t cap_ = 0;
};

enum class LanePoll { kPending, kDone, kFailed };

// Points, events and tiles of a batch being formed.
struct BatchShape {
  int64_t points = 0, events = 0, chunks = 0, tiles = 0;
  void add(const clue_params& p, int32_t n) {
    int64_t nper = 0, nt = 0;
    clue_rt::event_tiles(n, p.ndim, p.points_per_tile > 0 ? p.points_per_tile
                                                          : clue_rt::kDefaultPointsPerTile,
                         &nper, &nt);
    points += n;
    events += 1;
    chunks += (n + clue_rt::kChunk - 1) / clue_rt::kChunk;
    tiles += nt;
  }
};

// The device side of the lanes.
class ServeBackend {
 public:
  virtual ~ServeBackend() = default;
  virtual int lanes() const = 0;
  // Enqueues upload, pipeline and label download of `batch` on `lane`; the points
  // of request r are at ring + r->ring_off. On failure *err holds the reason.
  virtual clue_status launch(int lane, const std::vector<RequestP>& batch, const char* ring,
                             std::string* err) = 0;
  // Completion of the lane's last launch; kFailed sets *err.
  virtual LanePoll poll(int lane, std::string* err) = 0;
  // Copies the labels of the lane's completed launch into its requests.
  virtual void collect(int lane, const std::vector<RequestP>& batch) = 0;
  // Returns once every transfer and kernel queued on the lane has ended (their
  // outcome ignored); after it the batch's staging blocks may be reused.
  virtual void drain(int lane) = 0;
  // Device bytes the lane would hold after running a batch of this shape: with
  // the buffers it holds now (they only grow), and from none.
  virtual int64_t bytes_after(int lane, const clue_params& p, const BatchShape& shape) const = 0;
  virtual int64_t bytes_fresh(const clue_params& p, const BatchShape& shape) const = 0;
  // Frees the lane's device buffers (the lane is idle).
  virtual void trim(int lane) = 0;
};

// Admission, batching and completion. Clients call submit / wait / stats from any
// thread; one worker thread drives the backend. A lane failure that the device
// reports asynchronously is fatal: the lane's requests and every waiting request
// complete with that error, and later submits are refused.
class ServeCore {
 public:
  struct Limits {
    int64_t lane_budget_bytes = 0;  // device bytes one lane may hold
    int64_t lane_points = 0;        // points per launch (32-bit indexing and latency cap)
    int64_t max_batch_events = 1 << 20;
    size_t staging_bytes = 0;
    // Longest wait of the worker between polls of a busy lane; completions
    // normally wake it at once through notify().
    std::chrono::microseconds busy_poll{2000};
  };

  ServeCore(ServeBackend* be, char* ring, const Limits& lim) : be_(be), ring_base_(ring), lim_(lim) {
    ring_.reset(lim.staging_bytes);
    busy_.assign(static_cast<size_t>(be->lanes()), false);
    lane_reqs_.resize(static_cast<size_t>(be->lanes()));
  }
  ~ServeCore() { stop(); }
  ServeCore(const ServeCore&) = delete;
  ServeCore& operator=(const ServeCore&) = delete;

  void start() { worker_ = std::thread([this] { run(); }); }
  // Finishes queued work, then joins the worker.
  void stop() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      signalled_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
  }
  // A lane may have completed (called by the backend from any thread).
  void notify() {
    {
      std::lock_guard<std::mutex> g(mu_);
      signalled_ = true;
    }
    cv_work_.notify_one();
  }

  // Bytes of one event in the staging ring.
  static size_t staged_bytes(const clue_params& p, int32_t n) {
    return (sizeof(float) * static_cast<size_t>(n) * (p.ndim + 1) + 255) & ~size_t(255);
  }

  // Copies the event into the ring (waiting while it is full) and queues it.
  clue_status submit(const clue_params& p, int32_t n, const float* data, uint64_t* id,
                     std::string* err) {
    BatchShape one;
    one.add(p, n);
    if (n > lim_.lane_points || be_->bytes_fresh(p, one) > lim_.lane_budget_bytes) {
      *err = "event larger than the per-launch cap of the HBM budget";
      return CLUE_ERR_CAPACITY;
    }
    const size_t bytes = staged_bytes(p, n);
    if (bytes > ring_.capacity()) {
      *err = "event larger than the staging ring";
      return CLUE_ERR_CAPACITY;
    }
    auto r = std::make_shared<Request>();
    r->params = p;
    r->n = n;
    r->ring_bytes = bytes;
    std::unique_lock<std::mutex> lk(mu_);
    if (!fatal_.empty()) {
      *err = fatal_;
      return CLUE_ERR_RUNTIME;
    }
    r->id = next_id_++;
    size_t off = 0;
    cv_space_.wait(lk, [&] { return !fatal_.empty() || ring_.alloc(r->id, bytes, &off); });
    if (!fatal_.empty()) {
      *err = fatal_;
      return CLUE_ERR_RUNTIME;
    }
    r->ring_off = off;
    lk.unlock();
    // The block is this request's alone until it is released: copy outside the lock.
    std::memcpy(ring_base_ + off, data, sizeof(float) * static_cast<size_t>(n) * (p.ndim + 1));
    r->submit_ns = now_ns();
    lk.lock();
    table_[r->id] = r;
    waiting_.push_back(r);
    stats_.events_submitted += 1;
    signalled_ = true;
    lk.unlock();
    cv_work_.notify_one();
    *id = r->id;
    return CLUE_OK;
  }

  // Waits for an event (timeout_us < 0: no limit), copies its labels and forgets it.
  clue_status wait(uint64_t id, int32_t* labels, int64_t timeout_us, clue_server_timing* timing,
                   std::string* err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = table_.find(id);
    if (it == table_.end()) {
      *err = "unknown request id";
      return CLUE_ERR_INVALID_ARGUMENT;
    }
    RequestP r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_us < 0) {
      cv_done_.wait(lk, ready);
    } else if (!cv_done_.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
      *err = "timeout";
      return CLUE_ERR_RUNTIME;
    }
    table_.erase(id);
    lk.unlock();
    if (r->status != CLUE_OK) {
      *err = r->error;
      return r->status;
    }
    if (labels) std::memcpy(labels, r->labels.data(), sizeof(int32_t) * r->labels.size());
    if (timing) {
      timing->submit_ns = r->submit_ns;
      timing->launch_ns = r->launch_ns;
      timing->done_ns = r->done_ns;
      timing->batch_events = r->batch_events;
      timing->reserved = 0;
      timing->batch_points = r->batch_points;
    }
    return CLUE_OK;
  }

  void get_stats(clue_server_stats* out) {
    std::lock_guard<std::mutex> g(mu_);
    *out = stats_;
    out->lane_points = lim_.lane_points;
  }
  std::string fatal() {
    std::lock_guard<std::mutex> g(mu_);
    return fatal_;
  }

 private:
  // Caller holds mu_. Completes requests with `st`, frees their staging blocks.
  void complete_locked(const std::vector<RequestP>& reqs, clue_status st, const std::string& msg) {
    const int64_t t = now_ns();
    for (const RequestP& r : reqs) {
      r->status = st;
      if (st != CLUE_OK) r->error = msg;
      r->done_ns = t;
      r->done = true;
      ring_.release(r->id);
      if (st == CLUE_OK) {
        stats_.events_done += 1;
        stats_.points_done += r->n;
      }
    }
    cv_done_.notify_all();
    cv_space_.notify_all();
  }

  // Caller holds mu_: the waiting events of the oldest event's class that fit
  // the lane's budget, the point cap and the event cap, in arrival order. When the
  // buffers the lane already holds leave no room for the first event, *trim says
  // to free them first (the batch is then sized from none).
  std::vector<RequestP> form_batch_locked(int lane, bool* trim) {
    std::vector<RequestP> out;
    *trim = false;
    if (waiting_.empty()) return out;
    const clue_params cls = waiting_.front()->params;
    BatchShape shape;
    auto fits = [&](const BatchShape& s) {
      return (*trim ? be_->bytes_fresh(cls, s) : be_->bytes_after(lane, cls, s)) <=
             lim_.lane_budget_bytes;
    };
    for (auto it = waiting_.begin(); it != waiting_.end();) {
      const RequestP& r = *it;
      if (std::memcmp(&r->params, &cls, sizeof(clue_params)) != 0) {
        ++it;
        continue;
      }
      if (static_cast<int64_t>(out.size()) >= lim_.max_batch_events) break;
      BatchShape next = shape;
      next.add(r->params, r->n);
      if (next.points > lim_.lane_points && !out.empty()) break;
      if (!fits(next)) {
        if (!out.empty()) break;
        *trim = true;  // submit admitted the event: it fits a lane that holds nothing
      }
      shape = next;
      out.push_back(r);
      it = waiting_.erase(it);
    }
    return out;
  }

  void run() {
    const int L = be_->lanes();
    while (true) {
      bool progressed = false;
      // Completions.
      for (int l = 0; l < L; ++l) {
        if (!busy_[l]) continue;
        std::string err;
        const LanePoll st = be_->poll(l, &err);
        if (st == LanePoll::kPending) continue;
        progressed = true;
        if (st == LanePoll::kDone) be_->collect(l, lane_reqs_[l]);
        else be_->drain(l);
        std::lock_guard<std::mutex> g(mu_);
        if (st == LanePoll::kDone) {
          complete_locked(lane_reqs_[l], CLUE_OK, std::string());
          stats_.batches += 1;
        } else {
          if (fatal_.empty()) fatal_ = "device failure: " + err;
          complete_locked(lane_reqs_[l], CLUE_ERR_HIP, fatal_);
        }
        lane_reqs_[l].clear();
        busy_[l] = false;
      }
      // After a failure nothing more runs: every waiting event completes with it.
      {
        std::lock_guard<std::mutex> g(mu_);
        if (!fatal_.empty() && !waiting_.empty()) {
          std::vector<RequestP> rest(waiting_.begin(), waiting_.end());
          waiting_.clear();
          complete_locked(rest, CLUE_ERR_HIP, fatal_);
        }
      }
      // Launches on free lanes.
      for (int l = 0; l < L; ++l) {
        if (busy_[l]) continue;
        std::vector<RequestP> batch;
        bool trim = false;
        {
          std::lock_guard<std::mutex> g(mu_);
          if (!fatal_.empty()) break;
          batch = form_batch_locked(l, &trim);
        }
        if (batch.empty()) break;
        if (trim) be_->trim(l);
        progressed = true;
        int64_t pts = 0;
        for (const RequestP& r : batch) pts += r->n;
        const int64_t t = now_ns();
        for (const RequestP& r : batch) {
          r->launch_ns = t;
          r->batch_events = static_cast<int32_t>(batch.size());
          r->batch_points = pts;
        }
        std::string err;
        const clue_status st = be_->launch(l, batch, ring_base_, &err);
        if (st != CLUE_OK) {
          // Copies out of the ring may already be queued: wait for them before
          // the blocks go back to the ring.
          be_->drain(l);
          std::lock_guard<std::mutex> g(mu_);
          complete_locked(batch, st, err);
          continue;
        }
        lane_reqs_[l] = std::move(batch);
        busy_[l] = true;
      }
      std::unique_lock<std::mutex> lk(mu_);
      bool any_busy = false;
      for (int l = 0; l < L; ++l) any_busy = any_busy || busy_[l];
      if (stop_ && waiting_.empty() && !any_busy) break;
      if (!progressed && !signalled_) {
        if (any_busy)
          cv_work_.wait_for(lk, lim_.busy_poll, [&] { return signalled_; });
        else
          cv_work_.wait(lk, [&] { return signalled_; });
      }
      signalled_ = false;
    }
  }

  ServeBackend* be_;
  char* ring_base_;
  Limits lim_;
  StagingRing ring_;
  std::vector<bool> busy_;                       // worker thread only
  std::vector<std::vector<RequestP>> lane_reqs_;  // worker thread only
  std::deque<RequestP> waiting_;
  std::unordered_map<uint64_t, RequestP> table_;
  std::mutex mu_;
  std::condition_variable cv_done_, cv_space_, cv_work_;
  std::thread worker_;
  bool stop_ = false;
  bool signalled_ = false;
  uint64_t next_id_ = 1;
  clue_server_stats stats_{};
  std::string fatal_;
};

}  // namespace clue_srv
