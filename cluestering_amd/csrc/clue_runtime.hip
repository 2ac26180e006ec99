// Host runtime shared by the C ABI and the server (declarations: clue_runtime.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "clue_runtime.h"
#include "clue_tuning_table.h"

namespace clue_rt {

namespace {
thread_local std::string g_last_error;
}  // namespace

clue_status fail(clue_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

const char* last_error() { return g_last_error.c_str(); }

clue_tuning table_tuning(int ndim, int64_t n_total, int64_t n_events) {
  const int64_t mean = n_events > 0 ? n_total / n_events : n_total;
  for (const TuningEntry& t : kTuningTable)
    if ((t.ndim == 0 || t.ndim == ndim) && mean <= t.max_mean_event) {
      clue_tuning out{};
      out.group_queries = t.group_queries;
      out.xcd_chunk_log2 = t.xcd_chunk_log2;
      out.block_waves = t.block_waves;
      return out;
    }
  clue_tuning out{};
  out.group_queries = 1;
  out.xcd_chunk_log2 = 4;
  out.block_waves = 4;
  return out;
}

clue_status make_batch_plan(const clue_params* p, const uint32_t* sizes, int32_t n_events,
                            BatchPlan* bp, const clue_tuning* fixed) {
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
    ev.index = e;
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
  bp->tuning = fixed ? *fixed
                     : table_tuning(D, off, static_cast<int64_t>(bp->events.size()));
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

RunParams make_run_params(const clue_params* p, const clue_points* pts, const BatchPlan& bp) {
  const int n_events = static_cast<int>(bp.events.size());
  const int64_t n_total = bp.n_total;
  RunParams rp{};
  rp.xcd_chunk_log2 = bp.tuning.xcd_chunk_log2;
  rp.block_waves = bp.tuning.block_waves > 0 ? bp.tuning.block_waves : 4;
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
      {S_QPERM, bp.tuning.group_queries ? sizeof(int32_t) * n : 0},
      {S_QPERM_NH, (bp.tuning.group_queries && p->dm != p->dc) ? sizeof(int32_t) * n : 0},
      {S_QFINE, bp.tuning.group_queries ? sizeof(int32_t) * n : 0},
      {S_QKEY, bp.tuning.group_queries ? sizeof(uint32_t) * n : 0},
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
  b.qperm = bp.tuning.group_queries ? ws.buf[S_QPERM].as<int32_t>() : nullptr;
  b.qperm_nh = (bp.tuning.group_queries && p->dm != p->dc) ? ws.buf[S_QPERM_NH].as<int32_t>()
                                                          : b.qperm;
  b.qfine = bp.tuning.group_queries ? ws.buf[S_QFINE].as<int32_t>() : nullptr;
  b.qkey = bp.tuning.group_queries ? ws.buf[S_QKEY].as<uint32_t>() : nullptr;
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
  ra->local_ids = 0;
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


}  // namespace clue_rt
