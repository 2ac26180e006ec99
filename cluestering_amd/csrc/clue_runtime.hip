// Host runtime shared by the C ABI and the server (declarations: clue_runtime.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "clue_runtime.h"

namespace clue_rt {

namespace {
thread_local std::string g_last_error;
}  // namespace

clue_status fail(clue_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

const char* last_error() { return g_last_error.c_str(); }

template <class P>
clue_status make_batch_plan(const P* p, const uint32_t* sizes, int32_t n_events, BatchPlan* bp,
                            const clue_tuning* fixed) {
  std::string err;
  const clue_status st = plan_batch(p->ndim, p->points_per_tile, sizes, n_events, fixed, bp, &err);
  return st == CLUE_OK ? CLUE_OK : fail(st, err);
}

// Largest t with sqrt(t) <= r (sqrt is correctly rounded on the host and on
// gfx950 in both precisions), so that sqrt(t) <= r  <=>  t <= sq_max(r) for every t.
template <class T>
T sq_max(T r) {
  if (!(r > T(0))) return T(-1);
  if (std::isinf(r)) return std::numeric_limits<T>::infinity();
  T t = r * r;
  if (std::isinf(t)) t = std::numeric_limits<T>::max();
  while (t > T(0) && std::sqrt(t) > r) t = std::nextafter(t, T(0));
  while (true) {
    const T nx = std::nextafter(t, std::numeric_limits<T>::infinity());
    if (std::isinf(nx) || std::sqrt(nx) > r) break;
    t = nx;
  }
  return t;
}

template <class P>
clue_status validate(const P* p) {
  if (!p) return fail(CLUE_ERR_INVALID_ARGUMENT, "null params");
  if (p->ndim < 1 || p->ndim > kMaxDim)
    return fail(CLUE_ERR_OUT_OF_RANGE, "This library only works up to 10 dimensions");
  if (!(p->dc > 0) || !(p->rhoc >= 0) || !(p->dm > 0) || !(p->seed_dc > 0))
    return fail(CLUE_ERR_INVALID_ARGUMENT,
                "Invalid clustering parameters. The parameters must be positive.");
  switch (p->kernel) {
    case CLUE_KERNEL_FLAT:
      if (!(p->kparam[0] > 0))
        return fail(CLUE_ERR_INVALID_ARGUMENT, "Flat kernel parameter must be positive.");
      break;
    case CLUE_KERNEL_EXP:
      if (!(p->kparam[0] > 0) || !(p->kparam[1] > 0))
        return fail(CLUE_ERR_INVALID_ARGUMENT,
                    "Exponential kernel parameters must be positive.");
      break;
    case CLUE_KERNEL_GAUSS:
      if (!(p->kparam[0] > 0) || !(p->kparam[1] > 0))
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

// Whether the nearest-higher stage runs over the Morton-ordered block layout:
// it needs the Morton order of the query grouping, candidates visited in any
// order (unique tags: the default ones), and a metric with a lower bound over a
// bounding box (not Mahalanobis; weighted metrics with weights >= 0).
// Bounding boxes of each kSortedBox tile-sorted positions, for the density stage's
// block culling: every metric with a lower bound over a box (not Mahalanobis),
// weighted metrics with weights >= 0.
template <class P>
bool sorted_boxes_enabled(const P* p) {
  if (p->metric == CLUE_METRIC_MAHALANOBIS) return false;
  if (p->metric == CLUE_METRIC_WEIGHTED_EUCLIDEAN || p->metric == CLUE_METRIC_WEIGHTED_CHEBYSHEV)
    for (int d = 0; d < p->ndim; ++d)
      if (!(p->mparam[d] >= 0)) return false;
  return true;
}

template <class P, class Pts>
bool nh_blocks_enabled(const P* p, const Pts* pts, const BatchPlan& bp) {
  if (bp.tuning.nh_plain_scan || !bp.tuning.group_queries || pts->tags != nullptr) return false;
  if (p->metric == CLUE_METRIC_MAHALANOBIS) return false;
  if (p->metric == CLUE_METRIC_WEIGHTED_EUCLIDEAN || p->metric == CLUE_METRIC_WEIGHTED_CHEBYSHEV)
    for (int d = 0; d < p->ndim; ++d)
      if (!(p->mparam[d] >= 0)) return false;
  return true;
}

// Density's queries in Morton order: with candidate boxes the class order buys
// nothing (measured on MI355X on the 4096-event 2-D batch: 24.9 ms either way), and
// the block scan's layout then receives the keys from the density stage itself.
template <class P>
bool density_morton(const P* p, const BatchPlan& bp) {
  return bp.tuning.group_queries && sorted_boxes_enabled(p);
}

// The plain nearest-higher scan's own (class, Morton) order for dm, when the
// density stage's order is not that order.
template <class P>
bool separate_dm_order(const P* p, const BatchPlan& bp, bool nhb) {
  return bp.tuning.group_queries && !nhb && (p->dm != p->dc || density_morton(p, bp));
}

template <class P, class Pts>
RunParamsT<value_t<P>> make_run_params(const P* p, const Pts* pts, const BatchPlan& bp) {
  using T = value_t<P>;
  const int n_events = static_cast<int>(bp.events.size());
  const int64_t n_total = bp.n_total;
  RunParamsT<T> rp{};
  rp.xcd_chunk_log2 = bp.tuning.xcd_chunk_log2;
  rp.block_waves = bp.tuning.block_waves > 0 ? bp.tuning.block_waves : 4;
  rp.persistent = bp.tuning.persistent > 0 ? bp.tuning.persistent : 0;
  rp.nh_blocks = nh_blocks_enabled(p, pts, bp) ? 1 : 0;
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
  rp.gauss_den = (T(2) * p->kparam[0]) * p->kparam[0];
  {
    int ex = 0;
    const T m = std::frexp(rp.gauss_den, &ex);
    // Power of two with a normal, finite reciprocal.
    const T lim = sizeof(T) == 4 ? T(0x1p120) : T(0x1p1000);
    rp.gauss_pow2 = (m == T(0.5) && rp.gauss_den >= T(1) / lim && rp.gauss_den <= lim) ? 1 : 0;
    rp.gauss_inv_den = rp.gauss_pow2 ? std::ldexp(T(1), 1 - ex) : T(0);
    // Float: fold gauss_den = 2^s into the exp's constants when |s| <= 8.
    rp.gauss_fold = 0;
    if (sizeof(T) == 4 && rp.gauss_pow2 && ex - 1 >= -8 && ex - 1 <= 8) {
      rp.gauss_fold = 1;
      rp.es = exp_scaled(ex - 1);
    }
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
    rp.mparam[d] = d < p->ndim ? p->mparam[d] : T(0);
  }
  // Weighted Euclidean with unit weights evaluates (1 * dx) * dx == dx * dx.
  if (rp.metric == CLUE_METRIC_WEIGHTED_EUCLIDEAN) {
    bool unit = true;
    for (int d = 0; d < p->ndim; ++d) unit = unit && p->mparam[d] == 1;
    if (unit) rp.metric = CLUE_METRIC_EUCLIDEAN;
  }
  return rp;
}

template <class P, class Pts>
clue_status check_points(const P* p, const Pts* pts, int64_t n_expected) {
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


template <class P, class Pts, class Res>
WorkspaceShape workspace_shape(const BatchPlan& bp, const P* p, const Pts* pts, const Res* res) {
  using T = value_t<P>;
  WorkspaceShape w;
  w.D = p->ndim;
  w.value_bytes = sizeof(T);
  w.key_bytes = sizeof(typename FloatBits<T>::Int);
  w.n = bp.n_total;
  w.events = static_cast<int64_t>(bp.events.size());
  w.chunks = static_cast<int64_t>(bp.chunks.size());
  w.tiles = bp.n_tiles_total;
  w.max_items = bp.max_items;
  w.event_tiles = !bp.event_order.empty();
  w.group_queries = bp.tuning.group_queries != 0;
  w.nh_blocks = nh_blocks_enabled(p, pts, bp);
  w.separate_dm_order = separate_dm_order(p, bp, w.nh_blocks);
  w.flat_kernel = p->kernel == CLUE_KERNEL_FLAT;
  w.has_sigma = pts->sigma[0] != nullptr;
  w.has_u = pts->density_uncertainty != nullptr;
  w.own_rho = res->rho == nullptr;
  w.own_nh = res->nearest_higher == nullptr;
  w.own_is_seed = res->is_seed == nullptr;
  w.sorted_boxes = sorted_boxes_enabled(p);
  w.density_morton = density_morton(p, bp);
  return w;
}

// Fills the BatchBuffers and result pointers of one batch from workspace slots
// (each at least workspace_requests' bytes) and, with `upload`, enqueues the
// descriptor uploads on the stream (they precede any kernel that reads them).
template <class P, class Pts, class Res>
clue_status bind_batch(void* const (&slot)[S_NUM], const BatchPlan& bp, const P* p, const Pts* pts,
                       Res* res, hipStream_t s, BatchBuffersT<value_t<P>>* bb,
                       ResultArraysT<value_t<P>>* ra, bool upload) {
  using T = value_t<P>;
  using I = typename FloatBits<T>::Int;
  const int D = p->ndim;
  const int64_t n = bp.n_total;
  const int64_t E = static_cast<int64_t>(bp.events.size());
  const int64_t C = static_cast<int64_t>(bp.chunks.size());
  const int64_t NT = bp.n_tiles_total;
  const int64_t stride = (n + 63) / 64 * 64;
  const bool has_sigma = pts->sigma[0] != nullptr;
  const bool nhb = nh_blocks_enabled(p, pts, bp);
  const int64_t h_blocks = nhb ? (n + kNhBlock - 1) / kNhBlock + NT + 1 : 0;
  const int64_t h_slots = h_blocks * kNhBlock;
  if (upload) {
    CLUE_HIP(hipMemcpyAsync(slot[S_EVENTS], bp.events.data(), sizeof(EventDesc) * E,
                            hipMemcpyHostToDevice, s));
    CLUE_HIP(hipMemcpyAsync(slot[S_CHUNKS], bp.chunks.data(), sizeof(Chunk) * C,
                            hipMemcpyHostToDevice, s));
    if (!bp.event_order.empty())
      CLUE_HIP(hipMemcpyAsync(slot[S_EORDER], bp.event_order.data(), sizeof(int32_t) * E,
                              hipMemcpyHostToDevice, s));
  }
  auto& b = *bb;
  b = BatchBuffersT<T>{};
  b.n_events = static_cast<int32_t>(E);
  b.n_total = static_cast<int32_t>(n);
  b.n_chunks = static_cast<int32_t>(C);
  b.n_tiles_total = static_cast<int32_t>(NT);
  b.max_items = static_cast<int32_t>(bp.max_items);
  b.key_bits = bp.key_bits;
  b.events = at<EventDesc>(slot, S_EVENTS);
  b.event_order = bp.event_order.empty() ? nullptr : at<int32_t>(slot, S_EORDER);
  b.chunks = at<Chunk>(slot, S_CHUNKS);
  b.ext = at<I>(slot, S_EXT);
  b.geo = at<T>(slot, S_GEO);
  b.keys0 = at<uint32_t>(slot, S_K0);
  b.vals0 = at<int32_t>(slot, S_V0);
  b.keys1 = at<uint32_t>(slot, S_K1);
  b.vals1 = at<int32_t>(slot, S_V1);
  b.rs_hist = at<int32_t>(slot, S_HIST);
  b.scan_tmp = at<int32_t>(slot, S_SCAN);
  b.toff = at<int32_t>(slot, S_TOFF);
  b.icount = at<int32_t>(slot, S_ICOUNT);
  b.n_items = at<int32_t>(slot, S_NITEMS);
  b.items = at<Item>(slot, S_ITEMS);
  b.qfine = bp.tuning.group_queries ? at<int32_t>(slot, S_QFINE) : nullptr;
  b.qperm = !bp.tuning.group_queries ? nullptr
            : density_morton(p, bp)  ? b.qfine
                                     : at<int32_t>(slot, S_QPERM);
  // The block scan orders its queries by qfine: no separate order for dm.
  b.qperm_nh = separate_dm_order(p, bp, nhb) ? at<int32_t>(slot, S_QPERM_NH) : b.qperm;
  b.qkey = bp.tuning.group_queries ? at<uint32_t>(slot, S_QKEY) : nullptr;
  b.srho = at<T>(slot, S_SRHO);
  b.skey = nhb ? nullptr : at<uint64_t>(slot, S_SKEY);
  b.snh = at<int32_t>(slot, S_SNH);
  b.sdelta = at<T>(slot, S_SDELTA);
  b.flag = at<int32_t>(slot, S_FLAG);
  b.event_seeds = at<int32_t>(slot, S_ESEEDS);
  b.event_offsets = at<int32_t>(slot, S_EOFF);
  b.iperm = bp.event_order.empty() ? nullptr : at<int32_t>(slot, S_IPERM);
  b.rank = at<int32_t>(slot, S_RANK);
  b.n_seeds = at<int32_t>(slot, S_NSEEDS);
  b.seeds = at<int32_t>(slot, S_SEEDS);
  b.sorted_stride = stride;
  b.hx = nhb ? at<T>(slot, S_HX) : nullptr;
  b.hkey = nhb ? at<uint64_t>(slot, S_HKEY) : nullptr;
  b.hidx = nhb ? at<int32_t>(slot, S_HIDX) : nullptr;
  b.hbox = nhb ? at<T>(slot, S_HBOX) : nullptr;
  b.hbkey = nhb ? at<uint64_t>(slot, S_HBKEY) : nullptr;
  b.h_stride = h_slots;
  b.sbox = sorted_boxes_enabled(p) ? at<T>(slot, S_SBOX) : nullptr;
  for (int d = 0; d < kMaxDim; ++d) {
    b.sorted.x[d] = d < D ? at<T>(slot, S_SX) + stride * d : nullptr;
    b.sorted.sigma[d] = (has_sigma && d < D) ? at<T>(slot, S_SSIG) + stride * d : nullptr;
  }
  b.sorted.w = at<T>(slot, S_SW);
  b.sorted.fw = p->kernel == CLUE_KERNEL_FLAT ? at<T>(slot, S_SFW) : nullptr;
  b.sorted.u = pts->density_uncertainty ? at<T>(slot, S_SU) : nullptr;
  b.sorted.tag = at<uint32_t>(slot, S_STAG);
  b.sorted.perm = at<int32_t>(slot, S_PERM);
  ra->cluster = res->cluster_index;
  ra->is_seed = res->is_seed ? res->is_seed : at<int32_t>(slot, S_ISSEED);
  ra->rho = res->rho ? res->rho : at<T>(slot, S_RHO);
  ra->nh = res->nearest_higher ? res->nearest_higher : at<int32_t>(slot, S_NH);
  ra->delta = res->delta;
  ra->event_of_cluster = nullptr;
  ra->local_ids = 0;
  // The density stage fills the block layout's keys only from the Morton order
  // (the block scan implies it: no Mahalanobis, weights >= 0, query grouping).
  if (nhb && b.qperm != b.qfine) return CLUE_ERR_RUNTIME;
  return CLUE_OK;
}


// Reserves the workspace of one batch in grow-only buffers, then binds it.
template <class P, class Pts, class Res>
clue_status prepare_batch(Workspace& ws, const BatchPlan& bp, const P* p, const Pts* pts,
                          Res* res, hipStream_t s, BatchBuffersT<value_t<P>>* bb,
                          ResultArraysT<value_t<P>>* ra, bool upload) {
  size_t need[S_NUM];
  workspace_requests(workspace_shape(bp, p, pts, res), need);
  for (int k = 0; k < S_NUM; ++k) CLUE_HIP(ws.buf[k].reserve(need[k]));
  void* slot[S_NUM];
  for (int k = 0; k < S_NUM; ++k) slot[k] = ws.buf[k].p;
  return bind_batch(slot, bp, p, pts, res, s, bb, ra, upload);
}

template <class P, class Pts>
PointArraysT<value_t<P>> point_arrays(const P* p, const Pts* pts) {
  PointArraysT<value_t<P>> a{};
  for (int d = 0; d < kMaxDim; ++d) {
    a.x[d] = d < p->ndim ? pts->coords[d] : nullptr;
    a.sigma[d] = d < p->ndim ? pts->sigma[d] : nullptr;
  }
  a.w = pts->weight;
  a.u = pts->density_uncertainty;
  a.tags = pts->tags;
  return a;
}


template clue_status make_batch_plan<clue_params>(const clue_params*, const uint32_t*, int32_t,
                                                  BatchPlan*, const clue_tuning*);
template clue_status make_batch_plan<clue_params_f64>(const clue_params_f64*, const uint32_t*,
                                                      int32_t, BatchPlan*, const clue_tuning*);
template float sq_max<float>(float);
template double sq_max<double>(double);
template clue_status validate<clue_params>(const clue_params*);
template clue_status validate<clue_params_f64>(const clue_params_f64*);
template RunParamsT<float> make_run_params(const clue_params*, const clue_points*,
                                           const BatchPlan&);
template RunParamsT<double> make_run_params(const clue_params_f64*, const clue_points_f64*,
                                            const BatchPlan&);
template clue_status check_points(const clue_params*, const clue_points*, int64_t);
template clue_status check_points(const clue_params_f64*, const clue_points_f64*, int64_t);
template PointArraysT<float> point_arrays(const clue_params*, const clue_points*);
template PointArraysT<double> point_arrays(const clue_params_f64*, const clue_points_f64*);
template WorkspaceShape workspace_shape(const BatchPlan&, const clue_params*, const clue_points*,
                                       const clue_results*);
template WorkspaceShape workspace_shape(const BatchPlan&, const clue_params_f64*,
                                       const clue_points_f64*, const clue_results_f64*);
template clue_status bind_batch(void* const (&)[S_NUM], const BatchPlan&, const clue_params*,
                                const clue_points*, clue_results*, hipStream_t,
                                BatchBuffersT<float>*, ResultArraysT<float>*, bool);
template clue_status bind_batch(void* const (&)[S_NUM], const BatchPlan&, const clue_params_f64*,
                                const clue_points_f64*, clue_results_f64*, hipStream_t,
                                BatchBuffersT<double>*, ResultArraysT<double>*, bool);
template clue_status prepare_batch(Workspace&, const BatchPlan&, const clue_params*,
                                   const clue_points*, clue_results*, hipStream_t,
                                   BatchBuffersT<float>*, ResultArraysT<float>*, bool);
template clue_status prepare_batch(Workspace&, const BatchPlan&, const clue_params_f64*,
                                   const clue_points_f64*, clue_results_f64*, hipStream_t,
                                   BatchBuffersT<double>*, ResultArraysT<double>*, bool);

}  // namespace clue_rt
