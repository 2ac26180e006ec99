// Host-only parts of the library under the address and undefined-behaviour
// sanitizers (built with g++, no GPU): batch plans, workspace sizing, the staging
// ring, and the server's state machine driven by a scripted fake device,
// including asynchronous device failures and failed launches.
#include <atomic>
#include <cstdio>
#include <numeric>
#include <thread>
#include <vector>

#include "clue_serve_core.h"

static int g_fail = 0, g_pass = 0;
#define CHECK(cond)                                              \
  do {                                                           \
    if (cond) {                                                  \
      ++g_pass;                                                  \
    } else {                                                     \
      ++g_fail;                                                  \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
    }                                                            \
  } while (0)

using namespace clue_rt;
using clue_srv::BatchShape;
using clue_srv::LanePoll;
using clue_srv::RequestP;
using clue_srv::ServeCore;

static void test_plan() {
  // SetupTiles.hpp:30-34: 129 points -> 2 tiles -> 2 per dimension in 2-D -> 4 tiles.
  std::vector<uint32_t> sizes{129, 0, 4096, 1};
  BatchPlan bp;
  std::string err;
  CHECK(plan_batch(2, 0, sizes.data(), 4, nullptr, &bp, &err) == CLUE_OK);
  CHECK(bp.events.size() == 3);  // the empty event is not described
  CHECK(bp.events[0].nper == 2 && bp.events[0].ntiles == 4);
  CHECK(bp.events[1].nper == 6 && bp.events[1].ntiles == 36 && bp.events[1].index == 2);
  CHECK(bp.events[2].nper == 1 && bp.events[2].ntiles == 1 && bp.events[2].p_off == 129 + 4096);
  CHECK(bp.n_total == 129 + 4096 + 1 && bp.n_tiles_total == 41);
  CHECK(bp.chunks.size() == 1 + 2 + 1);
  CHECK(bp.key_bits == 6);
  CHECK(bp.max_items == (bp.n_total + 63) / 64 + 41);
  // Chunks cover every point once, in order, within their event.
  int64_t covered = 0;
  bool ordered = true;
  for (const Chunk& c : bp.chunks) {
    ordered = ordered && c.begin == covered && c.end > c.begin &&
              c.begin >= bp.events[c.event].p_off &&
              c.end <= bp.events[c.event].p_off + bp.events[c.event].n;
    covered = c.end;
  }
  CHECK(ordered && covered == bp.n_total);
  // 3-D: 4096 points -> 32 tiles -> 4 per dimension (4^3 = 64 >= 32 > 3^3).
  uint32_t one = 4096;
  CHECK(plan_batch(3, 0, &one, 1, nullptr, &bp, &err) == CLUE_OK);
  CHECK(bp.events[0].nper == 4 && bp.events[0].ntiles == 64);
  // points_per_tile honoured; 32-bit overflow refused.
  CHECK(plan_batch(2, 1024, &one, 1, nullptr, &bp, &err) == CLUE_OK && bp.events[0].ntiles == 4);
  std::vector<uint32_t> huge(3, 0x30000000u);
  CHECK(plan_batch(2, 0, huge.data(), 3, nullptr, &bp, &err) == CLUE_ERR_CAPACITY);
  CHECK(!err.empty());
  clue_tuning fixed{};
  fixed.block_waves = 8;
  CHECK(plan_batch(2, 0, &one, 1, &fixed, &bp, &err) == CLUE_OK && bp.tuning.block_waves == 8);
}

static void test_workspace() {
  CHECK(grown_capacity(100, 50) == 100);
  CHECK(grown_capacity(100, 400) == 400 + 100 + 256);
  // Bytes per point of a served-style run (labels given, the rest in the
  // workspace) for every ndim, at 2^20 points in one event: bounded by a few
  // hundred bytes (it is not monotone in D: the tile count n^D jumps).
  for (int D = 1; D <= 10; ++D) {
    uint32_t n = 1u << 20;
    BatchPlan bp;
    std::string err;
    CHECK(plan_batch(D, 0, &n, 1, nullptr, &bp, &err) == CLUE_OK);
    WorkspaceShape w;
    w.D = D;
    w.n = bp.n_total;
    w.events = 1;
    w.chunks = static_cast<int64_t>(bp.chunks.size());
    w.tiles = bp.n_tiles_total;
    w.max_items = bp.max_items;
    w.group_queries = true;
    w.nh_blocks = true;
    w.own_rho = w.own_nh = w.own_is_seed = true;
    size_t b[S_NUM];
    workspace_requests(w, b);
    const double total = std::accumulate(b, b + S_NUM, 0.0);
    const double per_point = total / n;
    std::printf("workspace bytes per point, D=%d: %.1f (tiles %lld)\n", D, per_point,
                static_cast<long long>(bp.n_tiles_total));
    CHECK(per_point > 90.0 && per_point < 96.0 + 32.0 * D);
    // Every per-point column the stages index is there.
    CHECK(b[S_SX] >= sizeof(float) * n * D && b[S_K0] >= 4 * n);
    CHECK(b[S_HX] >= sizeof(float) * D * n && b[S_HIDX] >= 4 * n && b[S_HKEY] >= 8 * n);
    // With the block layout the keys live there only.
    CHECK(b[S_SKEY] == 0);
  }
  // Optional columns only when used.
  WorkspaceShape w;
  w.D = 2;
  w.n = 1000;
  w.events = 1;
  w.chunks = 1;
  w.tiles = 9;
  w.max_items = 25;
  size_t b[S_NUM];
  workspace_requests(w, b);
  CHECK(b[S_SSIG] == 0 && b[S_SU] == 0 && b[S_SFW] == 0 && b[S_HX] == 0 && b[S_QPERM] == 0);
  CHECK(b[S_RHO] == 0 && b[S_NH] == 0 && b[S_ISSEED] == 0);
  w.has_sigma = w.has_u = w.flat_kernel = w.group_queries = w.separate_dm_order = true;
  w.value_bytes = w.key_bytes = 8;
  workspace_requests(w, b);
  CHECK(b[S_SSIG] == 8 * 1024 * 2 && b[S_SU] == 8000 && b[S_SFW] == 8000 && b[S_QPERM_NH] == 4000);
  CHECK(b[S_QPERM] == 4000 && b[S_SKEY] == 8000);
  // Density in Morton order (candidate boxes): no class order for dc.
  w.sorted_boxes = w.density_morton = true;
  workspace_requests(w, b);
  CHECK(b[S_QPERM] == 0 && b[S_QFINE] == 4000 && b[S_QPERM_NH] == 4000 && b[S_SBOX] > 0);
  CHECK(b[S_EXT] == 8 * 1 * 2 * 2 && b[S_GEO] == 8 * 3 * 2);
}

static void test_ring() {
  clue_srv::StagingRing r;
  r.reset(1000);
  size_t off = 0;
  CHECK(r.alloc(1, 400, &off) && off == 0);
  CHECK(r.alloc(2, 400, &off) && off == 400);
  CHECK(!r.alloc(3, 400, &off));  // 200 left at the end, nothing at the front
  CHECK(!r.alloc(3, 0, &off) && !r.alloc(3, 1001, &off));
  r.release(2);  // younger block first: stays reserved until block 1 goes
  CHECK(r.live_blocks() == 2 && !r.alloc(3, 400, &off));
  r.release(1);
  CHECK(r.live_blocks() == 0);
  CHECK(r.alloc(3, 700, &off) && off == 0);
  CHECK(r.alloc(4, 300, &off) && off == 700);  // exact fit at the end
  r.release(3);
  CHECK(r.alloc(5, 600, &off) && off == 0);    // wraps to the front
  CHECK(!r.alloc(6, 200, &off));               // [600, 700) free but too small
  CHECK(r.alloc(6, 100, &off) && off == 600);  // exactly fills the gap
  CHECK(!r.alloc(7, 1, &off));
  r.release(4);
  r.release(5);
  r.release(6);
  CHECK(r.live_blocks() == 0);
}

// A scripted device: launches are recorded, completions are released by the
// test (or immediately), and poll can report a failure.
class FakeLanes : public clue_srv::ServeBackend {
 public:
  explicit FakeLanes(int n) : lanes_(n) {}
  int lanes() const override { return static_cast<int>(lanes_.size()); }
  clue_status launch(int lane, const std::vector<RequestP>& batch, const char* ring,
                     std::string* err) override {
    Lane& L = lanes_[lane];
    launches++;
    if (fail_launch_at == launches.load()) {
      *err = "scripted launch failure";
      return CLUE_ERR_HIP;
    }
    L.labels.clear();
    for (const RequestP& r : batch) {
      // The staged points must be there: the first float of every event is its id.
      float first;
      std::memcpy(&first, ring + r->ring_off, sizeof(float));
      L.labels.push_back(static_cast<int32_t>(first));
    }
    L.sizes.clear();
    for (const RequestP& r : batch) L.sizes.push_back(r->n);
    L.pending = true;
    max_batch = std::max<int64_t>(max_batch, static_cast<int64_t>(batch.size()));
    return CLUE_OK;
  }
  LanePoll poll(int lane, std::string* err) override {
    Lane& L = lanes_[lane];
    polls++;
    if (fail_poll_at > 0 && polls.load() >= fail_poll_at && L.pending) {
      *err = "scripted event-query failure";
      return LanePoll::kFailed;
    }
    if (hold.load()) return LanePoll::kPending;
    L.pending = false;
    return LanePoll::kDone;
  }
  void collect(int lane, const std::vector<RequestP>& batch) override {
    Lane& L = lanes_[lane];
    for (size_t k = 0; k < batch.size(); ++k) batch[k]->labels.assign(L.sizes[k], L.labels[k]);
  }
  void drain(int) override { drains++; }
  int64_t bytes_after(int lane, const clue_params&, const BatchShape& s) const override {
    return std::max<int64_t>(lanes_[lane].cap, 10 * s.points);
  }
  int64_t bytes_fresh(const clue_params&, const BatchShape& s) const override {
    return 10 * s.points;
  }
  void trim(int lane) override {
    lanes_[lane].cap = 0;
    trims++;
  }
  struct Lane {
    std::vector<int32_t> labels, sizes;
    bool pending = false;
    int64_t cap = 0;
  };
  std::vector<Lane> lanes_;
  std::atomic<int> launches{0}, polls{0}, drains{0}, trims{0};
  std::atomic<bool> hold{false};
  int fail_launch_at = -1;
  int fail_poll_at = -1;
  int64_t max_batch = 0;
};

static clue_params params2d() {
  clue_params p{};
  p.ndim = 2;
  p.dc = 1.f;
  p.rhoc = 1.f;
  p.dm = 1.f;
  p.seed_dc = 1.f;
  p.kparam[0] = .5f;
  return p;
}

static std::vector<float> event_data(int id, int n) {
  std::vector<float> d(3 * static_cast<size_t>(n), 0.f);
  d[0] = static_cast<float>(id);
  return d;
}

static ServeCore::Limits limits(int64_t budget, size_t ring) {
  ServeCore::Limits lim;
  lim.lane_budget_bytes = budget;
  lim.lane_points = 1 << 20;
  lim.staging_bytes = ring;
  lim.busy_poll = std::chrono::microseconds(200);
  return lim;
}

static void test_serve_normal_and_budget() {
  FakeLanes fake(2);
  std::vector<char> ring(1 << 20);
  ServeCore core(&fake, ring.data(), limits(10 * 1000, ring.size()));
  fake.hold = true;  // launched batches stay pending until released
  const clue_params p = params2d();
  std::vector<uint64_t> ids;
  std::string err;
  for (int k = 0; k < 40; ++k) {
    auto d = event_data(k, 100);
    uint64_t id = 0;
    CHECK(core.submit(p, 100, d.data(), &id, &err) == CLUE_OK);
    ids.push_back(id);
  }
  // An event whose footprint exceeds one lane's budget is refused at once.
  auto big = event_data(0, 2000);
  uint64_t bid = 0;
  CHECK(core.submit(p, 2000, big.data(), &bid, &err) == CLUE_ERR_CAPACITY);
  core.start();  // every event queued: batches form under the lane budget
  // Not done yet: a bounded wait times out.
  std::vector<int32_t> labels(100);
  CHECK(core.wait(ids[0], labels.data(), 1000, nullptr, &err) == CLUE_ERR_RUNTIME && err == "timeout");
  fake.hold = false;
  core.notify();
  for (int k = 0; k < 40; ++k) {
    clue_server_timing t{};
    CHECK(core.wait(ids[k], labels.data(), -1, &t, &err) == CLUE_OK);
    CHECK(labels[0] == k && labels[99] == k);  // each event got its own labels back
    CHECK(t.done_ns >= t.launch_ns && t.launch_ns >= t.submit_ns && t.batch_events >= 1);
    CHECK(t.batch_points <= 1000);  // the lane budget (10 bytes a point here)
  }
  CHECK(fake.max_batch > 1 && fake.max_batch <= 10);
  CHECK(core.wait(ids[0], labels.data(), -1, nullptr, &err) == CLUE_ERR_INVALID_ARGUMENT);
  clue_server_stats st{};
  core.get_stats(&st);
  CHECK(st.events_done == 40 && st.events_submitted == 40 && st.batches >= 4);
  // A lane holding buffers too large to add the next event is trimmed first.
  fake.lanes_[0].cap = fake.lanes_[1].cap = 20000;
  auto d = event_data(7, 100);
  uint64_t id = 0;
  CHECK(core.submit(p, 100, d.data(), &id, &err) == CLUE_OK);
  CHECK(core.wait(id, labels.data(), -1, nullptr, &err) == CLUE_OK && labels[0] == 7);
  CHECK(fake.trims.load() >= 1);
  core.stop();
}

static void test_serve_device_failure() {
  FakeLanes fake(2);
  std::vector<char> ring(1 << 16);
  ServeCore core(&fake, ring.data(), limits(1 << 30, ring.size()));
  fake.fail_poll_at = 1;  // the first completion check reports a device error
  fake.hold = true;
  const clue_params p = params2d();
  std::string err;
  std::vector<uint64_t> ids;
  for (int k = 0; k < 6; ++k) {
    auto d = event_data(k, 50);
    uint64_t id = 0;
    CHECK(core.submit(p, 50, d.data(), &id, &err) == CLUE_OK);
    ids.push_back(id);
  }
  core.start();
  std::vector<int32_t> labels(50);
  for (uint64_t id : ids) {
    // Every event ends, with the device's error: none is left waiting forever.
    CHECK(core.wait(id, labels.data(), 5000000, nullptr, &err) == CLUE_ERR_HIP);
    CHECK(err.find("scripted event-query failure") != std::string::npos);
  }
  CHECK(fake.drains.load() >= 1);  // staging blocks reused only after the lane drained
  auto d = event_data(9, 50);
  uint64_t id = 0;
  CHECK(core.submit(p, 50, d.data(), &id, &err) == CLUE_ERR_RUNTIME);
  CHECK(!core.fatal().empty());
  core.stop();  // returns although the device failed
}

static void test_serve_launch_failure() {
  FakeLanes fake(1);
  std::vector<char> ring(1 << 16);
  ServeCore core(&fake, ring.data(), limits(1 << 30, ring.size()));
  fake.fail_launch_at = 1;
  core.start();
  const clue_params p = params2d();
  std::string err;
  auto d0 = event_data(0, 50);
  uint64_t id0 = 0, id1 = 0;
  CHECK(core.submit(p, 50, d0.data(), &id0, &err) == CLUE_OK);
  std::vector<int32_t> labels(50);
  CHECK(core.wait(id0, labels.data(), -1, nullptr, &err) == CLUE_ERR_HIP);
  CHECK(err == "scripted launch failure" && fake.drains.load() == 1);
  // The server goes on: the next event runs.
  auto d1 = event_data(1, 50);
  CHECK(core.submit(p, 50, d1.data(), &id1, &err) == CLUE_OK);
  CHECK(core.wait(id1, labels.data(), -1, nullptr, &err) == CLUE_OK && labels[0] == 1);
  CHECK(core.fatal().empty());
  core.stop();
}

static void test_serve_backpressure() {
  // A ring of 4 blocks of 256 bytes: submitters wait for completions to free room.
  FakeLanes fake(2);
  std::vector<char> ring(4 * 256);
  ServeCore core(&fake, ring.data(), limits(1 << 30, ring.size()));
  core.start();
  const clue_params p = params2d();
  std::atomic<int> ok{0};
  std::vector<std::thread> clients;
  for (int c = 0; c < 4; ++c)
    clients.emplace_back([&, c] {
      for (int k = 0; k < 50; ++k) {
        const int tag = c * 1000 + k;
        auto d = event_data(tag, 20);  // 240 bytes -> one 256-byte block
        uint64_t id = 0;
        std::string err;
        if (core.submit(p, 20, d.data(), &id, &err) != CLUE_OK) continue;
        std::vector<int32_t> labels(20);
        if (core.wait(id, labels.data(), -1, nullptr, &err) == CLUE_OK && labels[19] == tag) ++ok;
      }
    });
  for (auto& t : clients) t.join();
  CHECK(ok.load() == 200);
  core.stop();
}

int main() {
  test_plan();
  test_workspace();
  test_ring();
  test_serve_normal_and_budget();
  test_serve_device_failure();
  test_serve_launch_failure();
  test_serve_backpressure();
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
