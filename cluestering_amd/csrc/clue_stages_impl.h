// The five CLUE stages on gfx950:
//   (a) tile construction: per-event extremes, tile keys, stable radix sort into a
//       tile-sorted SoA copy, tile offsets, work items;
//   (b) local density and (c) nearest-higher: one wave per work item (up to 64
//       points of one tile, one lane each) walks the union of the lanes' search
//       boxes in the reference's lexicographic tile order; candidates are
//       wave-uniform and come through the scalar cache. With the default tags,
//       (c) instead scans each tile's candidates in Morton-ordered blocks and
//       skips the blocks that cannot hold the answer (its result does not
//       depend on the visiting order);
//   (d) seeds, numbered by an exclusive scan in input order;
//   (e) follower labels: pointer jumping, then a chase to the seed.
//
// Semantics follow the reference's serial backend:
//   geometry  core/detail/SetupTiles.hpp:23-57, core/detail/ComputeTiles.hpp:17-37
//   bins/box  data_structures/internal/TilesView.hpp:38-104
//   density   core/detail/ClusteringKernels.hpp:29-140
//   NH        core/detail/ClusteringKernels.hpp:142-291
//   seeds     core/detail/ClusteringKernels.hpp:293-343
//
// Included by one translation unit per point type (clue_stages_f32.hip,
// clue_stages_f64.hip), which compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <limits>
#include <type_traits>

#include "clue_device.h"
#include "clue_internal.h"

namespace clue_dev {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
// The neighbour stages take 1 to 16 waves per workgroup (a tuning choice): a
// workgroup runs on one CU, so its consecutive work items (neighbouring tiles)
// share that CU's scalar cache, while smaller workgroups spread more evenly.
constexpr int kMaxNeighbourBlock = 1024;

enum { M_EUCLID = 0, M_WEUCLID = 1, M_PERIODIC = 2, M_MANHATTAN = 3, M_CHEBYSHEV = 4,
       M_WCHEBYSHEV = 5, M_MAHALANOBIS = 6 };
enum { K_FLAT = 0, K_EXP = 1, K_GAUSS = 2 };

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b mod 8), each with its own L2. The neighbour stages deal chunks of G = 2^k
// consecutive work-item blocks to the XCDs in turn: an XCD's L2 then sees runs
// of neighbouring tiles, while the work stays spread evenly over the XCDs (a
// contiguous eighth per XCD leaves them unequal, since events differ in cost).
// Blocks past the last whole round of chunks keep their index.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb, int log2_chunk) {
  if (log2_chunk < 0) return b;
  const int round = kXcds << log2_chunk;
  if (b >= nb / round * round) return b;
  const int xcd = b % kXcds;
  const int k = b / kXcds;  // b is the k-th block on its XCD
  return ((((k >> log2_chunk) * kXcds) + xcd) << log2_chunk) + (k & ((1 << log2_chunk) - 1));
}

// Work items of a neighbour stage. In the grid form each wave takes one item (its
// workgroup mapped by xcd_block). In the persistent form (rp.persistent waves per
// CU) waves pull items from device queues: the item space is dealt to the XCDs in
// chunks of 2^k items, as xcd_block deals workgroups, each XCD counts off its own
// chunks (ctl[x], zeroed by k_item_fill), and a wave whose XCD has no items left
// takes them from the next XCDs in turn. A queue's items come in increasing order,
// so a wave stops once every one of the eight counters has passed the last item.
constexpr int kQueueDensity = 8;  // counters in the n_items block: density
constexpr int kQueueNh = 16;      // nearest-higher
constexpr int kQueueWords = 24;   // words of the block k_item_fill zeroes
struct ItemCursor {
  int32_t* ctl;
  int n;
  int log2c;
  int home;

  template <class P>
  __device__ ItemCursor(int32_t* n_items, int queue, const P& rp)
      : ctl(n_items + queue), n(*n_items), log2c(rp.xcd_chunk_log2), home(blockIdx.x % kXcds) {}

  __device__ int pull() {
    for (int t = 0; t < kXcds; ++t) {
      const int x = (home + t) & (kXcds - 1);
      int j = 0;
      if (lane_id() == 0) j = atomicAdd(&ctl[x], 1);
      j = uniform(j);
      const int item = log2c < 0 ? j * kXcds + x
                                 : ((x + kXcds * (j >> log2c)) << log2c) + (j & ((1 << log2c) - 1));
      if (item < n) return item;
    }
    return -1;
  }
};

template <int M>
struct MetricTraits {
  // Metrics ending in a square root are compared through the squared sum against
  // the largest float whose correctly rounded square root is within the radius.
  static constexpr bool kSqrt =
      (M == M_EUCLID || M == M_WEUCLID || M == M_PERIODIC || M == M_MAHALANOBIS);
};

// ---------------------------------------------------------------------------
// Utility kernels.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// (a1) Per-event coordinate extremes (ComputeTiles.hpp:17-36). NaN coordinates
// are skipped, as the reference's std::max/std::min fold skips them.
// ---------------------------------------------------------------------------
template <class T>
struct CoordPtrs {
  const T* x[kMaxDim];
};

template <class T>
__global__ __launch_bounds__(kBlock) void k_extremes(CoordPtrs<T> c, int D,
                                                     const Chunk* __restrict__ chunks,
                                                     typename FloatBits<T>::Int* __restrict__ ext) {
  using I = typename FloatBits<T>::Int;
  constexpr I kHi = std::numeric_limits<I>::max(), kLo = std::numeric_limits<I>::min();
  const Chunk ch = chunks[blockIdx.x];
  __shared__ I smin[kWaves][kMaxDim], smax[kWaves][kMaxDim];
  const int w = threadIdx.x / kWave;
  for (int d = 0; d < D; ++d) {
    I lo = kHi, hi = kLo;
    for (int i = ch.begin + threadIdx.x; i < ch.end; i += kBlock) {
      const T v = c.x[d][i];
      if (v != v) continue;
      const I o = float_to_ordered(v);
      lo = o < lo ? o : lo;
      hi = o > hi ? o : hi;
    }
    lo = wave_min_i(lo);
    hi = wave_max_i(hi);
    if (lane_id() == 0) {
      smin[w][d] = lo;
      smax[w][d] = hi;
    }
  }
  __syncthreads();
  if (threadIdx.x < D) {
    const int d = threadIdx.x;
    I lo = kHi, hi = kLo;
    for (int k = 0; k < kWaves; ++k) {
      lo = smin[k][d] < lo ? smin[k][d] : lo;
      hi = smax[k][d] > hi ? smax[k][d] : hi;
    }
    if (lo != kHi) atomicMin(&ext[(ch.event * D + d) * 2 + 0], lo);
    if (hi != kLo) atomicMax(&ext[(ch.event * D + d) * 2 + 1], hi);
  }
}

// Extremes start from the identities of min and max.
template <class I>
__global__ void k_init_extremes(I* __restrict__ ext, int64_t pairs) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < pairs) {
    ext[2 * k] = std::numeric_limits<I>::max();
    ext[2 * k + 1] = std::numeric_limits<I>::min();
  }
}

// geo[e] = {mn[D], mx[D], side[D]}, side = (max - min) / nper (ComputeTiles.hpp:33-34).
template <class T>
__global__ void k_geometry(const typename FloatBits<T>::Int* __restrict__ ext,
                           const EventDesc* __restrict__ events, int n_events, int D,
                           T* __restrict__ geo) {
  using I = typename FloatBits<T>::Int;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_events) return;
  const int nper = events[e].nper;
  for (int d = 0; d < D; ++d) {
    const I lo = ext[(e * D + d) * 2 + 0];
    const I hi = ext[(e * D + d) * 2 + 1];
    const T mn = lo == std::numeric_limits<I>::max() ? std::numeric_limits<T>::max()
                                                     : ordered_to_float(lo);
    const T mx = hi == std::numeric_limits<I>::min() ? std::numeric_limits<T>::lowest()
                                                     : ordered_to_float(hi);
    T* g = geo + static_cast<int64_t>(e) * 3 * D;
    g[d] = mn;
    g[D + d] = mx;
    g[2 * D + d] = (mx - mn) / static_cast<T>(nper);
  }
}

template <class T>
__device__ __forceinline__ GeoT<T> event_geo(const T* geo, const RunParamsT<T>& rp, int e,
                                             int nper) {
  GeoT<T> g;
  g.D = rp.D;
  g.nper = nper;
  g.mn = geo + static_cast<int64_t>(e) * 3 * rp.D;
  g.mx = g.mn + rp.D;
  g.side = g.mn + 2 * rp.D;
  g.wrapped = rp.wrapped;
  return g;
}

// ---------------------------------------------------------------------------
// (a2) Tile key of every point: t_base + sum_d nper^(D-1-d) * bin_d
// (TilesView.hpp:54-64), and the identity permutation.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_tile_keys(CoordPtrs<T> c, const RunParamsT<T> rp,
                                                      const Chunk* __restrict__ chunks,
                                                      const EventDesc* __restrict__ events,
                                                      const T* __restrict__ geo,
                                                      uint32_t* __restrict__ keys,
                                                      int32_t* __restrict__ vals) {
  const Chunk ch = chunks[blockIdx.x];
  const EventDesc ev = events[ch.event];
  const GeoT<T> g = event_geo(geo, rp, ch.event, ev.nper);
  for (int i = ch.begin + threadIdx.x; i < ch.end; i += kBlock) {
    int t = 0;
    for (int d = 0; d < rp.D; ++d) t = t * ev.nper + get_bin(g, c.x[d][i], d);
    keys[i] = static_cast<uint32_t>(ev.t_base + t);
    vals[i] = i;
  }
}

// ---------------------------------------------------------------------------
// (a3) Gather the tile-sorted SoA copy and the tile offsets. toff[k] is the first
// sorted position whose key is >= k, so tile k spans [toff[k], toff[k+1]).
// ---------------------------------------------------------------------------
template <class T>
struct GatherArgs {
  PointArraysT<T> in;
  SortedArraysT<T> out;
  int64_t cs;  // column stride of the sorted coordinate block
};

template <class T>
__global__ __launch_bounds__(kBlock) void k_gather(const GatherArgs<T> a, const RunParamsT<T> rp,
                                                   const uint32_t* __restrict__ keys,
                                                   const int32_t* __restrict__ perm, int n,
                                                   int n_tiles_total, int32_t* __restrict__ toff) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  for (int d = 0; d < rp.D; ++d) a.out.x[d][s] = a.in.x[d][i];
  const T wi = a.in.w[i];
  a.out.w[s] = wi;
  if (a.out.fw) a.out.fw[s] = rp.kp0 * wi;  // flat kernel: K * w for j != i
  if (rp.has_sigma)
    for (int d = 0; d < rp.D; ++d) a.out.sigma[d][s] = a.in.sigma[d][i];
  if (rp.has_u) a.out.u[s] = a.in.u[i];
  a.out.tag[s] = rp.has_tags ? a.in.tags[i] : static_cast<uint32_t>(i);
  a.out.perm[s] = i;
  const int key = static_cast<int>(keys[s]);
  const int prev = s == 0 ? -1 : static_cast<int>(keys[s - 1]);
  for (int k = prev + 1; k <= key; ++k) toff[k] = s;
  if (s == n - 1)
    for (int k = key + 1; k <= n_tiles_total; ++k) toff[k] = n;
}

// ---------------------------------------------------------------------------
// (a1-a3 fused) One workgroup per event when every event of the batch has at
// most kEvtMaxTiles tiles: extremes and geometry, tile ids, a stable counting
// sort by tile into the sorted SoA copy and the tile offsets, all in one
// kernel. Same results as k_extremes + k_geometry + k_tile_keys + the radix
// sort + k_gather: the extremes are the same ordered-integer min/max (NaN
// skipped), the bins the same get_bin, and the order (tile, input index).
// Each wave owns a contiguous sixteenth of the event, walked in
// rounds of 64; lanes with the same tile are ranked by ballot, per-wave tile
// counters live in LDS, and waves, rounds and lanes go in index order, so the
// sort is stable. Events are taken largest first (order[]).
// ---------------------------------------------------------------------------
constexpr int kEvtWaves = 16;
constexpr int kEvtBlock = kEvtWaves * kWave;
static_assert(kEvtMaxTiles <= kEvtBlock, "one tile per thread in the offset scan");

// Rounds of 64 points whose loads a wave issues together (memory-level
// parallelism; the rounds are still ranked in order).
constexpr int kEvtRounds = 4;

template <int D, class T>
__global__ __launch_bounds__(kEvtBlock) void k_event_tiles(
    const RunParamsT<T> rp, const EventDesc* __restrict__ events, const int32_t* __restrict__ order,
    const PointArraysT<T> in, const SortedArraysT<T> out, T* __restrict__ geo,
    int32_t* __restrict__ toff, int32_t* __restrict__ iperm, int n_tiles_total, int n_total) {
  using I = typename FloatBits<T>::Int;
  constexpr I kHi = std::numeric_limits<I>::max(), kLo = std::numeric_limits<I>::min();
  __shared__ int32_t cnt[kEvtWaves][kEvtMaxTiles];
  __shared__ I red[kEvtWaves][2 * D];
  __shared__ T sgeo[3 * D];
  __shared__ int32_t wsum[kEvtWaves];
  const int e = order[blockIdx.x];
  const EventDesc ev = events[e];
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int b0 = ev.p_off, n = ev.n;

  // Extremes (ComputeTiles.hpp:17-36) and geometry (:33-34).
#pragma unroll
  for (int d = 0; d < D; ++d) {
    I lo = kHi, hi = kLo;
#pragma unroll 4
    for (int k = threadIdx.x; k < n; k += kEvtBlock) {
      const T v = in.x[d][b0 + k];
      const I o = float_to_ordered(v);
      const bool ok = v == v;  // NaN coordinates are skipped
      lo = (ok && o < lo) ? o : lo;
      hi = (ok && o > hi) ? o : hi;
    }
    lo = wave_min_i(lo);
    hi = wave_max_i(hi);
    if (lane == 0) {
      red[w][2 * d] = lo;
      red[w][2 * d + 1] = hi;
    }
  }
  __syncthreads();
  if (threadIdx.x < D) {
    const int d = threadIdx.x;
    I lo = kHi, hi = kLo;
    for (int k = 0; k < kEvtWaves; ++k) {
      lo = red[k][2 * d] < lo ? red[k][2 * d] : lo;
      hi = red[k][2 * d + 1] > hi ? red[k][2 * d + 1] : hi;
    }
    const T mn = lo == kHi ? std::numeric_limits<T>::max() : ordered_to_float(lo);
    const T mx = hi == kLo ? std::numeric_limits<T>::lowest() : ordered_to_float(hi);
    const T side = (mx - mn) / static_cast<T>(ev.nper);
    sgeo[d] = mn;
    sgeo[D + d] = mx;
    sgeo[2 * D + d] = side;
    T* gq = geo + static_cast<int64_t>(e) * 3 * D;
    gq[d] = mn;
    gq[D + d] = mx;
    gq[2 * D + d] = side;
  }
  for (int k = threadIdx.x; k < kEvtWaves * kEvtMaxTiles; k += kEvtBlock) (&cnt[0][0])[k] = 0;
  __syncthreads();
  GeoT<T> g;
  g.D = D;
  g.nper = ev.nper;
  g.mn = sgeo;
  g.mx = sgeo + D;
  g.side = sgeo + 2 * D;
  g.wrapped = rp.wrapped;
  // Tile id (TilesView.hpp:54-64) of the point with coordinates x.
  auto tile_of = [&](const T (&x)[D]) {
    int t = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) t = t * g.nper + get_bin(g, x[d], d);
    return static_cast<uint32_t>(t);
  };

  int bits = 0;
  while ((1 << bits) < ev.ntiles) ++bits;
  const int per_wave = (n + kEvtWaves - 1) / kEvtWaves;
  const int wb = w * per_wave;
  const int we = min(n, wb + per_wave);
  const uint64_t lt = lanemask_lt();
  // Per-wave tile counts.
  for (int k0 = wb; k0 < we; k0 += kWave * kEvtRounds) {
    uint32_t t[kEvtRounds];
    bool valid[kEvtRounds];
#pragma unroll
    for (int u = 0; u < kEvtRounds; ++u) {
      const int k = k0 + u * kWave + lane;
      valid[u] = k < we;
      T x[D];
#pragma unroll
      for (int d = 0; d < D; ++d) x[d] = valid[u] ? in.x[d][b0 + k] : T(0);
      t[u] = valid[u] ? tile_of(x) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kEvtRounds; ++u) {
      const uint64_t peers = match_peers(t[u], bits, valid[u]);
      if (valid[u] && (peers & lt) == 0) cnt[w][t[u]] += __popcll(peers);
    }
  }
  __syncthreads();
  // Tile offsets: exclusive scan over tiles (one per thread), then each wave's
  // first position within each tile.
  {
    const int t = threadIdx.x;
    int tot = 0;
    if (t < ev.ntiles)
      for (int k = 0; k < kEvtWaves; ++k) tot += cnt[k][t];
    int incl = tot;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(incl, off, 64);
      if (lane >= off) incl += y;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < w; ++k) before += wsum[k];
    const int base = before + incl - tot;
    if (t < ev.ntiles) {
      toff[ev.t_base + t] = b0 + base;
      int run = base;
      for (int k = 0; k < kEvtWaves; ++k) {
        const int c = cnt[k][t];
        cnt[k][t] = run;
        run += c;
      }
    }
    if (t == 0 && ev.t_base + ev.ntiles == n_tiles_total) toff[n_tiles_total] = n_total;
  }
  __syncthreads();
  // Stable placement: rounds in order, lanes ranked by ballot. Only the
  // permutation is scattered (perm[s] = i, and iperm[i] = s in input order);
  // the sorted columns are then gathered in sorted order, so their writes are
  // whole lines instead of one scattered word per column.
  for (int k0 = wb; k0 < we; k0 += kWave * kEvtRounds) {
    uint32_t t[kEvtRounds];
    bool valid[kEvtRounds];
#pragma unroll
    for (int u = 0; u < kEvtRounds; ++u) {
      const int k = k0 + u * kWave + lane;
      valid[u] = k < we;
      T x[D];
#pragma unroll
      for (int d = 0; d < D; ++d) x[d] = valid[u] ? in.x[d][b0 + k] : T(0);
      t[u] = valid[u] ? tile_of(x) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kEvtRounds; ++u) {
      const uint64_t peers = match_peers(t[u], bits, valid[u]);
      int pos = 0;
      if (valid[u]) pos = cnt[w][t[u]] + __popcll(peers & lt);
      __builtin_amdgcn_wave_barrier();
      if (valid[u] && (peers & lt) == 0) cnt[w][t[u]] += __popcll(peers);
      if (valid[u]) {
        const int i = b0 + k0 + u * kWave + lane;
        out.perm[b0 + pos] = i;
        iperm[i] = b0 + pos;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // perm of the whole event
  __syncthreads();
#pragma unroll 4
  for (int k = threadIdx.x; k < n; k += kEvtBlock) {
    const int s = b0 + k;
    const int i = out.perm[s];
#pragma unroll
    for (int d = 0; d < D; ++d) out.x[d][s] = in.x[d][i];
    const T wt = in.w[i];
    out.w[s] = wt;
    if (out.fw) out.fw[s] = rp.kp0 * wt;  // flat kernel: K * w for j != i
    if (rp.has_sigma)
#pragma unroll
      for (int d = 0; d < D; ++d) out.sigma[d][s] = in.sigma[d][i];
    if (rp.has_u) out.u[s] = in.u[i];
    out.tag[s] = rp.has_tags ? in.tags[i] : static_cast<uint32_t>(i);
  }
}

template <int D, class T>
static void launch_event_tiles_d(const RunParamsT<T>& rp, const PointArraysT<T>& in,
                                 const BatchBuffersT<T>& bb, hipStream_t s) {
  hipLaunchKernelGGL((k_event_tiles<D, T>), dim3(static_cast<unsigned>(bb.n_events)),
                     dim3(kEvtBlock), 0, s, rp, bb.events, bb.event_order, in, bb.sorted, bb.geo,
                     bb.toff, bb.iperm, bb.n_tiles_total, bb.n_total);
}

template <class T>
static void launch_event_tiles(const RunParamsT<T>& rp, const PointArraysT<T>& in,
                               const BatchBuffersT<T>& bb, hipStream_t s) {
  switch (rp.D) {
    case 1: launch_event_tiles_d<1>(rp, in, bb, s); break;
    case 2: launch_event_tiles_d<2>(rp, in, bb, s); break;
    case 3: launch_event_tiles_d<3>(rp, in, bb, s); break;
    case 4: launch_event_tiles_d<4>(rp, in, bb, s); break;
    case 5: launch_event_tiles_d<5>(rp, in, bb, s); break;
    case 6: launch_event_tiles_d<6>(rp, in, bb, s); break;
    case 7: launch_event_tiles_d<7>(rp, in, bb, s); break;
    case 8: launch_event_tiles_d<8>(rp, in, bb, s); break;
    case 9: launch_event_tiles_d<9>(rp, in, bb, s); break;
    default: launch_event_tiles_d<10>(rp, in, bb, s); break;
  }
}

// Work items: ceil(count / 64) per tile.
static __global__ void k_item_count(const int32_t* __restrict__ toff, int n_tiles,
                             int32_t* __restrict__ icount) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_tiles) return;
  icount[k] = (toff[k + 1] - toff[k] + kWave - 1) / kWave;
}

static __global__ void k_item_fill(const int32_t* __restrict__ toff, const int32_t* __restrict__ ioff,
                            int n_tiles, const EventDesc* __restrict__ events, int n_events,
                            Item* __restrict__ items, int32_t* __restrict__ queues) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kQueueDensity && k < kQueueWords) queues[k] = 0;  // work-item queue counters
  if (k >= n_tiles) return;
  // Event of tile k: last event with t_base <= k.
  int lo = 0, hi = n_events - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (events[mid].t_base <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int b = toff[k], e = toff[k + 1];
  int o = ioff[k];
  for (int s = b; s < e; s += kWave) {
    Item it;
    it.tile = k;
    it.begin = s;
    it.end = min(e, s + kWave);
    it.event = lo;
    items[o++] = it;
  }
}

// The three item kernels above in one workgroup, for batches of at most
// kItemsOneBlockTiles tiles (one launch instead of five): the tiles' item counts
// and their exclusive scan (kBlock tiles at a time, with a carry), the total and
// the queue counters, then the items. Each thread fills the tiles whose offsets it
// wrote, so the fill reads only its own stores.
constexpr int kItemsOneBlockTiles = 4096;
static __global__ __launch_bounds__(kBlock) void k_items_one_block(
    const int32_t* __restrict__ toff, int n_tiles, const EventDesc* __restrict__ events,
    int n_events, int32_t* __restrict__ ioff, Item* __restrict__ items, int32_t* __restrict__ ctl) {
  int carry = 0;
  for (int base = 0; base < n_tiles; base += kBlock) {
    const int t = base + threadIdx.x;
    const int c = t < n_tiles ? (toff[t + 1] - toff[t] + kWave - 1) / kWave : 0;
    int tot;
    const int ex = block_excl_scan<kBlock>(c, &tot);
    if (t < n_tiles) ioff[t] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) ctl[0] = carry;
  for (int k = kQueueDensity + threadIdx.x; k < kQueueWords; k += kBlock) ctl[k] = 0;
  for (int t = threadIdx.x; t < n_tiles; t += kBlock) {
    int lo = 0, hi = n_events - 1;  // event of tile t: last event with t_base <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (events[mid].t_base <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int b = toff[t], e = toff[t + 1];
    int o = ioff[t];
    for (int s = b; s < e; s += kWave) {
      Item it;
      it.tile = t;
      it.begin = s;
      it.end = min(e, s + kWave);
      it.event = lo;
      items[o++] = it;
    }
  }
}

// ---------------------------------------------------------------------------
// Distances (DistanceMetrics.hpp), right fold over dimensions
// (internal/meta/accumulate.hpp): t0 + (t1 + (... + (t_{D-1} + 0))).
// For square-root metrics the value returned is the sum under the root.
// ---------------------------------------------------------------------------
template <int D, int M, class T>
__device__ __forceinline__ T metric_pair(const T (&xi)[D], const T (&si)[D], const T (&xj)[D],
                                         const T (&sj)[D], const RunParamsT<T>& rp) {
  if constexpr (M == M_EUCLID) {
    T s;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const T dx = xi[d] - xj[d];
      s = (d == D - 1) ? dx * dx : dx * dx + s;  // t + 0 == t for t = dx*dx >= +0
    }
    return s;
  } else if constexpr (M == M_WEUCLID) {
    T s = T(0);
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const T dx = xi[d] - xj[d];
      s = rp.mparam[d] * dx * dx + s;
    }
    return s;
  } else if constexpr (M == M_PERIODIC) {
    T s;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const T a = __builtin_fabs(xi[d] - xj[d]);
      const T pa = rp.mparam[d] - a;
      const T m = (pa < a) ? pa : a;  // std::min(a, P - a)
      s = (d == D - 1) ? m * m : m * m + s;
    }
    return s;
  } else if constexpr (M == M_MANHATTAN) {
    T s;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const T a = __builtin_fabs(xi[d] - xj[d]);
      s = (d == D - 1) ? a : a + s;
    }
    return s;
  } else if constexpr (M == M_CHEBYSHEV) {
    T m = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T a = __builtin_fabs(xi[d] - xj[d]);
      m = (m < a) ? a : m;  // std::max(m, a)
    }
    return m;
  } else if constexpr (M == M_WCHEBYSHEV) {
    T m = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T a = rp.mparam[d] * __builtin_fabs(xi[d] - xj[d]);
      m = (m < a) ? a : m;
    }
    return m;
  } else {  // M_MAHALANOBIS: (dx*dx) / (si*si + sj*sj)
    T s = T(0);
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const T dx = xi[d] - xj[d];
      s = dx * dx / (si[d] * si[d] + sj[d] * sj[d]) + s;
    }
    return s;
  }
}

// The same for candidate j read from the sorted columns.
template <int D, int M, class T>
__device__ __forceinline__ T metric_value(const T (&xi)[D], const T (&si)[D],
                                          const T* __restrict__ sx, int64_t cs, int j,
                                          const T* __restrict__ ssig, const RunParamsT<T>& rp) {
  T xj[D], sj[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    xj[d] = sx[d * cs + j];
    sj[d] = (M == M_MAHALANOBIS) ? ssig[d * cs + j] : T(0);
  }
  return metric_pair<D, M>(xi, si, xj, sj, rp);
}

// Normal: the caller has shown v >= 2^-96 or v == 0 (float sqrt without its
// tiny-input branch).
template <int M, bool Normal = false, class T>
__device__ __forceinline__ T finish_distance(T v) {
  if constexpr (MetricTraits<M>::kSqrt) {
    if constexpr (Normal && sizeof(T) == 4)
      return sqrt_rn_normal(v);
    else
      return sqrt_rn(v);
  } else {
    return v;
  }
}

// Search box of one point (ClusteringKernels.hpp:111-121, TilesView.hpp:77-87).
// sqrt(2) correctly rounded in each precision (std::sqrt of 2 in the reference).
template <class T>
struct Sqrt2;
template <>
struct Sqrt2<float> {
  static constexpr float v = 0x1.6a09e6p+0f;
};
template <>
struct Sqrt2<double> {
  static constexpr double v = 0x1.6a09e667f3bcdp+0;
};

template <int D, class T>
__device__ __forceinline__ void search_box(const GeoT<T>& g, const T (&xi)[D], const T (&si)[D],
                                           T R, int (&lo)[D], int (&hi)[D]) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T rs = R * si[d] * Sqrt2<T>::v;
    const T r = (R < rs) ? rs : R;  // std::max(R, R * sigma * sqrt(2))
    int a = get_bin(g, xi[d] - r, d);
    int b = get_bin(g, xi[d] + r, d);
    if (g.wrapped[d] && a > b) b += g.nper;
    lo[d] = a;
    hi[d] = b;
  }
}

// Walks the union of the wave's search boxes in lexicographic order (dim 0
// outermost). For every tile that at least one lane's own box contains, calls
// visit(tile_key, in_box) with in_box true on the lanes whose box contains it.
// Each lane thus sees exactly its own boxes' tiles, in the reference's order.
template <int D, class F>
__device__ __forceinline__ void walk_tiles(const int (&lo)[D], const int (&hi)[D], int nper,
                                           const int32_t* wrapped, int t_base, F&& visit) {
  int ulo[D], uhi[D], b[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    ulo[d] = wave_min_uniform(lo[d]);
    uhi[d] = wave_max_uniform(hi[d]);
    b[d] = ulo[d];
  }
  if (ulo[0] > uhi[0]) return;  // no active lane
  while (true) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) in = in && (lo[d] <= b[d]) && (b[d] <= hi[d]);
    if (ballot(in) != 0ull) {
      int t = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) t = t * nper + (wrapped[d] ? b[d] % nper : b[d]);
      visit(uniform(t_base + t), in);
    }
    int d = D - 1;
    for (; d >= 0; --d) {
      if (b[d] < uhi[d]) {
        ++b[d];
        break;
      }
      b[d] = ulo[d];
    }
    if (d < 0) break;
  }
}

// ---------------------------------------------------------------------------
// Query order within a tile. A wave walks the union of its lanes' search boxes
// and runs a candidate's kernel when any lane has it within the radius, so waves
// of lanes with alike boxes and nearby points waste fewer slots. The points of
// every tile with more than one wave's worth are ordered by
//   (box class, Morton index of the point's sub-cell of the tile),
// where the class says, per dimension (first three), whether the search box
// reaches below and above the tile's bin; items then take 64 consecutive entries
// of that order. Results do not depend on the order: a point's traversal is its
// own, whichever lane runs it. k_query_order builds both stages' orders, one
// wave per tile (below).
// ---------------------------------------------------------------------------
constexpr int kQueryDims = 3;  // dimensions that enter the keys

template <int D, class T>
__device__ __forceinline__ uint32_t query_class(const GeoT<T>& g, const T (&xi)[D],
                                                const T (&si)[D], T R, const int (&tb)[D]) {
  int lo[D], hi[D];
  search_box<D>(g, xi, si, R, lo, hi);
  uint32_t c = 0;
#pragma unroll
  for (int d = 0; d < (D < kQueryDims ? D : kQueryDims); ++d)
    c |= ((lo[d] < tb[d] ? 2u : 0u) | (hi[d] > tb[d] ? 1u : 0u)) << (2 * d);
  return c;
}

// Morton index of the point's sub-cell: 8 bits over the first (up to) three
// dimensions of the tile, bits interleaved from the most significant.
template <int D, class T>
__device__ __forceinline__ uint32_t query_fine(const GeoT<T>& g, const T (&xi)[D],
                                               const int (&tb)[D]) {
  constexpr int kD = D < kQueryDims ? D : kQueryDims;
  constexpr int kB = 8 / kD;  // bits per dimension
  constexpr int kS = 1 << kB;
  uint32_t q[kD];
#pragma unroll
  for (int d = 0; d < kD; ++d) {
    const T c = g.wrapped[d] ? normalize(g, xi[d], d) : xi[d];
    const T f = ((c - g.mn[d]) / g.side[d] - static_cast<T>(tb[d])) * static_cast<T>(kS);
    q[d] = (f >= T(0)) ? (f < static_cast<T>(kS) ? static_cast<uint32_t>(f) : kS - 1u) : 0u;
  }
  uint32_t key = 0;
#pragma unroll
  for (int b = kB - 1; b >= 0; --b)
#pragma unroll
    for (int d = 0; d < kD; ++d) key = (key << 1) | ((q[d] >> b) & 1u);
  return key;
}

// Makes the wave's global stores visible to the wave's later loads (lanes read
// entries other lanes wrote): the stores complete before any later access.
__device__ __forceinline__ void wave_memory_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Stable counting sort of the points at positions [b, e) of `order` (identity
// when NULL) by a Bits-bit key, into out[b, e). One wave; cnt is its LDS
// counter array of 2^Bits entries. The counts need no order (one LDS atomic per
// point; with Counted the caller has made them already, while it computed the
// keys); the placement visits rounds and lanes in position order and ranks equal
// keys by ballot, so equal keys keep their input order. Placement takes four
// rounds' positions and keys with their loads in flight together, then ranks
// the rounds one after the other.
template <int Bits, bool Counted = false, class KeyOf>
__device__ __forceinline__ void wave_counting_sort(int b, int e, const int32_t* order,
                                                   int32_t* cnt, KeyOf key_of, int32_t* out) {
  constexpr int kKeys = 1 << Bits;
  constexpr int kRounds = 4;
  const int lane = lane_id();
  if constexpr (!Counted) {
    for (int c = lane; c < kKeys; c += kWave) cnt[c] = 0;
    __builtin_amdgcn_wave_barrier();
    for (int i = b + lane; i < e; i += kWave) atomicAdd(&cnt[key_of(order ? order[i] : i)], 1);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  int carry = b;
  for (int c0 = 0; c0 < kKeys; c0 += kWave) {
    const int c = c0 + lane;
    const int v = c < kKeys ? cnt[c] : 0;
    int x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (c < kKeys) cnt[c] = carry + x - v;
    carry += __shfl(x, kWave - 1, 64);
  }
  __builtin_amdgcn_wave_barrier();
  for (int i0 = b; i0 < e; i0 += kRounds * kWave) {
    int sr[kRounds];
    uint32_t kr[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int i = i0 + r * kWave + lane;
      sr[r] = i < e ? (order ? order[i] : i) : 0;
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) kr[r] = (i0 + r * kWave + lane < e) ? key_of(sr[r]) : 0u;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      if (i0 + r * kWave >= e) break;  // wave-uniform
      const bool valid = i0 + r * kWave + lane < e;
      const uint32_t k = kr[r];
      const uint64_t peers = match_peers(k, Bits, valid);
      int pos = 0;
      if (valid) pos = cnt[k] + __popcll(peers & lanemask_lt());
      __builtin_amdgcn_wave_barrier();
      if (valid && (peers & lanemask_lt()) == 0) cnt[k] += __popcll(peers);
      __builtin_amdgcn_wave_barrier();
      if (valid) out[pos] = sr[r];
    }
  }
}

__device__ __forceinline__ float fmin_t(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double fmin_t(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float fmax_t(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double fmax_t(double a, double b) { return __builtin_fmax(a, b); }

// Minimum and maximum over each group of N = 8 or 16 lanes (a half row or a row),
// in every lane of the group (float through DPP lane moves: pair swap, quad swap,
// half-row mirror, then for 16 the row mirror; lanes a move does not write take
// the identity, so each move fuses into its operation).
template <int N, class F>
__device__ __forceinline__ float row_reduce_dpp(float v, float id, F op) {
  static_assert(N == 8 || N == 16, "a half row or a row");
  const int i = __float_as_int(id);
  int x = __float_as_int(v);
  x = __float_as_int(op(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(i, x, 0xB1, 0xF, 0xF, false))));
  x = __float_as_int(op(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(i, x, 0x4E, 0xF, 0xF, false))));
  x = __float_as_int(op(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(i, x, 0x141, 0xF, 0xF, false))));
  if constexpr (N == 16)
    x = __float_as_int(op(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(i, x, 0x140, 0xF, 0xF, false))));
  return __int_as_float(x);
}
template <int N = 16>
__device__ __forceinline__ float row_min(float v) {
  return row_reduce_dpp<N>(v, __builtin_inff(), [](float a, float b) { return __builtin_fminf(a, b); });
}
template <int N = 16>
__device__ __forceinline__ float row_max(float v) {
  return row_reduce_dpp<N>(v, -__builtin_inff(), [](float a, float b) { return __builtin_fmaxf(a, b); });
}
template <int N = 16>
__device__ __forceinline__ double row_min(double v) {
#pragma unroll
  for (int off = 1; off < N; off <<= 1) v = fmin_t(v, __shfl_xor(v, off, 64));
  return v;
}
template <int N = 16>
__device__ __forceinline__ double row_max(double v) {
#pragma unroll
  for (int off = 1; off < N; off <<= 1) v = fmax_t(v, __shfl_xor(v, off, 64));
  return v;
}
// Largest 64-bit key over each group of N = 8 or 16 lanes (both halves moved by DPP).
template <int Ctrl>
__device__ __forceinline__ uint64_t max_dpp_u64(uint64_t v) {
  const int lo = static_cast<int>(static_cast<uint32_t>(v));
  const int hi = static_cast<int>(static_cast<uint32_t>(v >> 32));
  const uint32_t olo = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, lo, Ctrl, 0xF, 0xF, false));
  const uint32_t ohi = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, hi, Ctrl, 0xF, 0xF, false));
  const uint64_t o = (static_cast<uint64_t>(ohi) << 32) | olo;
  return o > v ? o : v;
}
template <int N = 16>
__device__ __forceinline__ uint64_t row_max_u64(uint64_t v) {
  static_assert(N == 8 || N == 16, "a half row or a row");
  v = max_dpp_u64<0xB1>(v);   // quad_perm 1,0,3,2
  v = max_dpp_u64<0x4E>(v);   // quad_perm 2,3,0,1
  v = max_dpp_u64<0x141>(v);  // row_half_mirror
  if constexpr (N == 16) v = max_dpp_u64<0x140>(v);  // row_mirror
  return v;
}

// First layout block of tile t whose candidates begin at sorted position b:
// tile starts rounded up to whole blocks, one spare block per tile, so that the
// tiles' block ranges never overlap.
__device__ __forceinline__ int64_t nh_first_block(int b, int t) {
  return static_cast<int64_t>((b + kNhBlock - 1) / kNhBlock) + t;
}

// The nearest-higher layout of tile t (one wave), written by the query-order
// pass: the tile's candidates in Morton order (order[k], k < cnt) from block
// nh_first_block on, padded to whole blocks (index -1), with each block's
// bounding box. The keys, known after the density stage, are written by the
// density stage itself (k_density, from the same Morton order).
template <int D, class T>
__device__ __forceinline__ void write_nh_layout(int t, int b, int cnt, const int32_t* order,
                                                const T* __restrict__ sx, int64_t cs,
                                                T* __restrict__ hx, int64_t hs,
                                                int32_t* __restrict__ hidx, T* __restrict__ hbox) {
  static_assert(kWave % kNhBlock == 0, "blocks must not straddle waves");
  static_assert(kNhBlock == 8 || kNhBlock == 16, "block boxes are reduced over DPP half rows or rows");
  const int lane = lane_id();
  const int64_t blk0 = nh_first_block(b, t);
  const int span = (cnt + kNhBlock - 1) / kNhBlock * kNhBlock;
  for (int k0 = 0; k0 < span; k0 += kWave) {
    const int k = k0 + lane;
    const bool real = k < cnt;
    T x[D];
    int s = -1;
    if (real) s = order ? order[b + k] : b + k;
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = real ? sx[d * cs + s] : T(0);
    const int64_t p = blk0 * kNhBlock + k;
    if (k < span) {
#pragma unroll
      for (int d = 0; d < D; ++d) hx[d * hs + p] = x[d];
      hidx[p] = s;
    }
    T lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lo[d] = real ? x[d] : std::numeric_limits<T>::infinity();
      hi[d] = real ? x[d] : -std::numeric_limits<T>::infinity();
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lo[d] = row_min<kNhBlock>(lo[d]);
      hi[d] = row_max<kNhBlock>(hi[d]);
    }
    if (k < span && (k % kNhBlock) == 0) {
      const int64_t q = blk0 + k / kNhBlock;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        hbox[q * 2 * D + d] = lo[d];
        hbox[q * 2 * D + D + d] = hi[d];
      }
    }
  }
}

// Bounding box of each kSortedBox consecutive tile-sorted positions (positions
// past n hold nothing). A box may cover the end of one tile and the start of the
// next, or of two events: it then bounds more points than a walk takes from it,
// which only makes it skipped less often. Float: a thread takes four positions
// with one 16-byte load per column (the columns start 64-aligned) and four lanes
// make a box; double: sixteen lanes, one position each.
template <class F>
__device__ __forceinline__ float quad_reduce_dpp(float v, float id, F op) {
  const int i = __float_as_int(id);
  int x = __float_as_int(v);
  x = __float_as_int(op(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(i, x, 0xB1, 0xF, 0xF, false))));
  x = __float_as_int(op(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(i, x, 0x4E, 0xF, 0xF, false))));
  return __int_as_float(x);
}

template <int D, class T>
__global__ __launch_bounds__(kBlock) void k_sorted_boxes(const T* __restrict__ sx, int64_t cs,
                                                         int n, T* __restrict__ sbox) {
  static_assert(kSortedBox == 16, "a box is four lanes of four floats, or one row of 16 lanes");
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if constexpr (sizeof(T) == 4) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    const int64_t p = 4 * t;  // first of the thread's four positions
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float lo = __builtin_inff(), hi = -__builtin_inff();
      if (p + 3 < n) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(sx + d * cs + p);
        lo = __builtin_fminf(__builtin_fminf(v.x, v.y), __builtin_fminf(v.z, v.w));
        hi = __builtin_fmaxf(__builtin_fmaxf(v.x, v.y), __builtin_fmaxf(v.z, v.w));
      } else {
        for (int k = 0; k < 4; ++k)
          if (p + k < n) {
            const float x = sx[d * cs + p + k];
            lo = __builtin_fminf(lo, x);
            hi = __builtin_fmaxf(hi, x);
          }
      }
      lo = quad_reduce_dpp(lo, __builtin_inff(), [](float a, float b) { return __builtin_fminf(a, b); });
      hi = quad_reduce_dpp(hi, -__builtin_inff(), [](float a, float b) { return __builtin_fmaxf(a, b); });
      if (p < n && (t % 4) == 0) {
        const int64_t q = p / kSortedBox;
        sbox[q * 2 * D + d] = lo;
        sbox[q * 2 * D + D + d] = hi;
      }
    }
  } else {
    const int64_t p = t;
    const bool real = p < n;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T x = real ? sx[d * cs + p] : T(0);
      const T lo = row_min(real ? x : std::numeric_limits<T>::infinity());
      const T hi = row_max(real ? x : -std::numeric_limits<T>::infinity());
      if (real && (p % kSortedBox) == 0) {
        const int64_t q = p / kSortedBox;
        sbox[q * 2 * D + d] = lo;
        sbox[q * 2 * D + D + d] = hi;
      }
    }
  }
}

// The query orders of one tile per wave. Pass 1 packs each point's keys into
// qkey (Morton index in bits 0-7, class for dc in 8-13, class for dm in 14-19);
// then up to three stable counting sorts: sorted order -> Morton order (qfine),
// and Morton order -> (class, Morton) for dc (qperm) and for dm (qperm_nh). An
// order that aliases qfine, or qperm_nh aliasing qperm, is not sorted again.
template <int D, class T>
__global__ __launch_bounds__(kBlock) void k_query_order(
    const RunParamsT<T> rp, const EventDesc* __restrict__ events, const Item* __restrict__ items,
    const int32_t* __restrict__ ioff, const T* __restrict__ geo, const int32_t* __restrict__ toff,
    int n_tiles, const T* __restrict__ sx, int64_t cs, const T* __restrict__ ssig,
    uint32_t* __restrict__ qkey, int32_t* __restrict__ qfine, int32_t* __restrict__ qperm,
    int32_t* __restrict__ qperm_nh, T* __restrict__ hx, int64_t hs, int32_t* __restrict__ hidx,
    T* __restrict__ hbox) {
  constexpr int kD = D < kQueryDims ? D : kQueryDims;
  constexpr int kFineBits = (8 / kD) * kD;
  constexpr int kClassBits = 2 * kD;
  __shared__ int32_t cnt[kWaves][1 << kFineBits];
  const int w = threadIdx.x / kWave;
  const int t = blockIdx.x * kWaves + w;
  if (t >= n_tiles) return;
  const int lane = lane_id();
  const int b = toff[t], e = toff[t + 1];
  const bool class_dc = qperm != qfine;
  const bool class_dm = qperm_nh != qperm && qperm_nh != qfine;
  if (e - b <= kWave) {  // one item: the query order is irrelevant
    if (b + lane < e) {
      qfine[b + lane] = b + lane;
      if (class_dc) qperm[b + lane] = b + lane;
      if (class_dm) qperm_nh[b + lane] = b + lane;
    }
    if (hx) write_nh_layout<D>(t, b, e - b, nullptr, sx, cs, hx, hs, hidx, hbox);
    return;
  }
  // Event of tile t: that of its first work item (ioff: the items' exclusive
  // scan), two dependent loads where a search over the events takes a dozen.
  const int ev_i = items[ioff[t]].event;
  const EventDesc ev = events[ev_i];
  const GeoT<T> g = event_geo(geo, rp, ev_i, ev.nper);
  int tb[D];
  {
    int r = t - ev.t_base;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      tb[d] = r % ev.nper;
      r /= ev.nper;
    }
  }
  // The keys, and the Morton sort's counts as they are made.
  int32_t* c = cnt[w];
  for (int k = lane; k < (1 << kFineBits); k += kWave) c[k] = 0;
  __builtin_amdgcn_wave_barrier();
#pragma unroll 4
  for (int s = b + lane; s < e; s += kWave) {
    T xi[D], si[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      xi[d] = sx[d * cs + s];
      si[d] = rp.has_sigma ? ssig[d * cs + s] : T(0);
    }
    const uint32_t fine = query_fine<D>(g, xi, tb);
    const uint32_t c_dc = class_dc ? query_class<D>(g, xi, si, rp.dc, tb) : 0u;
    const uint32_t c_dm = class_dm ? query_class<D>(g, xi, si, rp.dm, tb) : 0u;
    qkey[s] = fine | (c_dc << 8) | (c_dm << 14);
    atomicAdd(&c[fine], 1);
  }
  wave_memory_sync();  // qkey of every point of the tile
  wave_counting_sort<kFineBits, true>(b, e, nullptr, c, [&](int s) { return qkey[s] & 0xFFu; },
                                      qfine);
  wave_memory_sync();  // qfine, the input order of the class sorts
  if (class_dc)
    wave_counting_sort<kClassBits>(b, e, qfine, c, [&](int s) { return (qkey[s] >> 8) & 63u; },
                                   qperm);
  if (class_dm)
    wave_counting_sort<kClassBits>(b, e, qfine, c, [&](int s) { return (qkey[s] >> 14) & 63u; },
                                   qperm_nh);
  if (hx) write_nh_layout<D>(t, b, e - b, qfine, sx, cs, hx, hs, hidx, hbox);
}

// ---------------------------------------------------------------------------
// Kernel values (ConvolutionalKernel.hpp:21-57) for a candidate within the radius.
// K_GAUSS2 is the gaussian whose 2*sigma*sigma is a power of two: there division
// by it and multiplication by its reciprocal are the same correctly rounded
// operation on the same exact value, so the cheaper one is used (in float, the
// power of two is folded into the exp's constants instead, ExpScaled). The _UNIT
// forms are those of amplitude 1, whose product A * e is e itself (denormals are
// kept, so even a subnormal e is unchanged).
// ---------------------------------------------------------------------------
enum { K_GAUSS2 = 3, K_EXP_UNIT = 4, K_GAUSS2_UNIT = 5 };

// exp of an argument the host has shown to lie in [LO, HI] for every candidate
// within the radius (rp.exp_in_range, passed as InRange) skips the range
// handling; both forms give the same value there.
template <bool InRange>
__device__ __forceinline__ float kernel_exp(float x, const RunParamsT<float>&) {
  return InRange ? expf_cw_core(x) : expf_cw(x);
}
template <bool InRange>
__device__ __forceinline__ double kernel_exp(double x, const RunParamsT<double>&) {
  return InRange ? expd_cw_core(x) : expd_cw(x);
}

// c0: rp.es.c[0] (float gaussian with a folded scale), possibly in a vector register.
template <int M, int K, bool InRange, bool Normal, class T>
__device__ __forceinline__ T kernel_value(T v, const RunParamsT<T>& rp, T c0) {
  if constexpr (K == K_FLAT) {
    return rp.kp0;
  } else if constexpr (K == K_EXP || K == K_EXP_UNIT) {
    const T e = kernel_exp<InRange>(-rp.kp0 * finish_distance<M, Normal>(v), rp);
    return K == K_EXP_UNIT ? e : rp.kp1 * e;
  } else if constexpr (K == K_GAUSS) {
    const T dd = finish_distance<M, Normal>(v);
    return rp.kp1 * kernel_exp<InRange>(-dd * dd / rp.gauss_den, rp);
  } else if constexpr (sizeof(T) == 4) {
    // exp((-dd*dd) * 2^-s) with 2^-s = 1 / gauss_den folded into rp.es.
    const T dd = finish_distance<M, Normal>(v);
    const T y = -dd * dd;
    const T e = InRange ? expf_cw_core_scaled(y, rp.es, c0) : expf_cw_scaled(y, rp.es, c0);
    return K == K_GAUSS2_UNIT ? e : rp.kp1 * e;
  } else {
    const T dd = finish_distance<M, Normal>(v);
    const T e = kernel_exp<InRange>((-dd * dd) * rp.gauss_inv_den, rp);
    return K == K_GAUSS2_UNIT ? e : rp.kp1 * e;
  }
}

// Ordering key stored with each density: float packs (rho, tag) into 64 bits;
// double keeps rho's order alone and compares tags separately.
__device__ __forceinline__ uint64_t density_key(float rho, uint32_t tag) {
  return rho_tag_key(rho, tag, false);
}
__device__ __forceinline__ uint64_t density_key(double rho, uint32_t) {
  return rho_key_f64(rho, false);
}

// Loads U consecutive wave-uniform column values (a block of candidates) with
// one wide scalar load and holds them in scalar registers at this point: left to
// itself the compiler may sink each value's load into the candidate's branch,
// where every use then waits on its own load.
template <int U, class T>
__device__ __forceinline__ void load_block(const T* __restrict__ p, T (&out)[U]) {
  typedef T Vec __attribute__((ext_vector_type(U), aligned(sizeof(T))));
  Vec v = *reinterpret_cast<const Vec*>(p);
  asm volatile("" : "+s"(v));
#pragma unroll
  for (int k = 0; k < U; ++k) out[k] = v[k];
}

// A lower bound, over every query point in [qlo, qhi] and every candidate inside
// the box [bl, bh], of the value metric_value computes (the same operations on
// the per-dimension gaps, which never exceed the pairs' own differences after
// rounding). With qlo = qhi = x it bounds one query point. Weighted metrics
// need weights >= 0 (checked on the host).
template <int D, int M, class T>
__device__ __forceinline__ T metric_lb(const T (&qlo)[D], const T (&qhi)[D], const T (&bl)[D],
                                       const T (&bh)[D], const RunParamsT<T>& rp) {
  T g[D];
#pragma unroll
  for (int d = 0; d < D; ++d) g[d] = fmax_t(fmax_t(bl[d] - qhi[d], qlo[d] - bh[d]), T(0));
  if constexpr (M == M_EUCLID) {
    T s;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) s = (d == D - 1) ? g[d] * g[d] : g[d] * g[d] + s;
    return s;
  } else if constexpr (M == M_WEUCLID) {
    T s = T(0);
#pragma unroll
    for (int d = D - 1; d >= 0; --d) s = rp.mparam[d] * g[d] * g[d] + s;
    return s;
  } else if constexpr (M == M_PERIODIC) {
    // a = |dx| lies in [g, G]. With period P > 0, min(a, P - a) >= min(g, P - G)
    // when P - G >= 0 (else 0 bounds its square). With P <= 0 (an unwrapped
    // dimension), min(a, P - a) = P - a and |P - a| >= a >= g.
    T s;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const T P = rp.mparam[d];
      const T G = fmax_t(qhi[d] - bl[d], bh[d] - qlo[d]);
      const T pg = P - G;
      const T m = (P <= T(0)) ? g[d] : ((pg >= T(0)) ? fmin_t(g[d], pg) : T(0));
      s = (d == D - 1) ? m * m : m * m + s;
    }
    return s;
  } else if constexpr (M == M_MANHATTAN) {
    T s;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) s = (d == D - 1) ? g[d] : g[d] + s;
    return s;
  } else if constexpr (M == M_CHEBYSHEV) {
    T m = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) m = (m < g[d]) ? g[d] : m;
    return m;
  } else {  // M_WCHEBYSHEV
    T m = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T a = rp.mparam[d] * g[d];
      m = (m < a) ? a : m;
    }
    return m;
  }
}

// Compile-time loop modes of the density stage: kernel exp arguments known to be
// in range, and square roots known to need no tiny-input branch.
template <bool InRange, bool Normal>
struct ModeC {
  static constexpr bool kInRange = InRange;
  static constexpr bool kNormal = Normal;
};

// Candidates are walked in blocks of U so that their wave-uniform data arrives
// in a few wide scalar loads (measured on MI355X: 8 for density, 16 for
// nearest-higher, whose blocks carry less work per candidate).
template <int D>
struct Unroll {
  static constexpr int U = D <= 3 ? 8 : 4;
  static constexpr int U_NH = D <= 3 ? 16 : 4;
};

// ---------------------------------------------------------------------------
// (b) Local density: rho_i = sum over visited j of ((float)(d <= dc) * K) * w_j
// (ClusteringKernels.hpp:53-69). Candidates beyond dc add +0 and are skipped;
// the additions keep the candidate order.
// ---------------------------------------------------------------------------
template <int D, int M, int K, bool Persistent, class T>
__global__ __launch_bounds__(kMaxNeighbourBlock) void k_density(
    const RunParamsT<T> rp, const EventDesc* __restrict__ events, const T* __restrict__ geo,
    const int32_t* __restrict__ toff, const Item* __restrict__ items,
    int32_t* __restrict__ n_items, const T* __restrict__ sx, int64_t cs,
    const T* __restrict__ sw, const T* __restrict__ sfw, const T* __restrict__ ssig,
    const uint32_t* __restrict__ stag, const int32_t* __restrict__ qperm,
    T* __restrict__ srho, uint64_t* __restrict__ skey, const T* __restrict__ sbox,
    uint64_t* __restrict__ hkey, uint64_t* __restrict__ hbkey) {
  constexpr int U = Unroll<D>::U;
  auto process = [&](const int item) {
    const Item it = items[item];
    const EventDesc ev = events[it.event];
    const GeoT<T> g = event_geo(geo, rp, it.event, ev.nper);
    const int q = it.begin + lane_id();
    const bool active = q < it.end;
    const int s = active ? (qperm ? qperm[q] : q) : -1;  // sorted position of this lane's point

    T xi[D], si[D];
  #pragma unroll
    for (int d = 0; d < D; ++d) {
      xi[d] = active ? sx[d * cs + s] : T(0);
      si[d] = (active && rp.has_sigma) ? ssig[d * cs + s] : T(0);
    }
    int lo[D], hi[D];
    search_box<D>(g, xi, si, rp.dc, lo, hi);
    if (!active) {
  #pragma unroll
      for (int d = 0; d < D; ++d) {
        lo[d] = INT32_MAX;
        hi[d] = INT32_MIN;
      }
    }
    const T thr_in = MetricTraits<M>::kSqrt ? rp.dc_sq_max : rp.dc;
    T rho = T(0);
    // The diagonal term K(i,i) * w_i = w_i; the flat kernel's off-diagonal weights
    // come pre-multiplied (fw = f * w).
    const T wself = active ? sw[s] : T(0);
    const T* __restrict__ wcol = (K == K_FLAT) ? sfw : sw;
    // The folded gaussian's leading coefficient, held in a vector register for the
    // whole walk (an fma takes one scalar operand; the next coefficient is one).
    T c0 = T(0);
    if constexpr (sizeof(T) == 4) {
      c0 = rp.es.c[0];
      asm volatile("" : "+v"(c0));
    }

    // term(j) = K(d_ij) * w_j, with K = 1 on the diagonal (self only in its own tile).
    // wj is the candidate's column value, loaded with its block of candidates.
    // own_c: std::true_type for the point's own tile, where the diagonal term lies;
    // mode_c: ModeC<in range, normal> (see below).
    auto add = [&](T v, int j, T wj, T thr, auto own_c, auto mode_c) {
      constexpr bool own = decltype(own_c)::value;
      constexpr bool inr = decltype(mode_c)::kInRange;
      constexpr bool nrm = decltype(mode_c)::kNormal;
      if constexpr (K == K_FLAT) {
        // rho + 0 == rho here (rho is never -0), so selecting the unchanged sum is
        // the reference's addition of a zero term.
        const T next = rho + ((own && j == s) ? wself : wj);
        rho = (v <= thr) ? next : rho;
      } else if constexpr (K == K_EXP_UNIT || K == K_GAUSS2_UNIT) {
        // Amplitude 1: the diagonal's distance is 0 and exp(-0) is exactly 1, the
        // reference's K(i, i), so the diagonal needs no case of its own.
        if (v <= thr) rho += kernel_value<M, K, inr, nrm>(v, rp, c0) * wj;
      } else {
        if (v <= thr) {
          const T kv = (own && j == s) ? T(1) : kernel_value<M, K, inr, nrm>(v, rp, c0);
          rho += kv * wj;
        }
      }
    };
    auto walk = [&](int j0, int j1, T thr, auto own_c, auto mode_c) {
      int j = j0;
      for (; j + U <= j1; j += U) {
        T v[U], w[U];
  #pragma unroll
        for (int k = 0; k < U; ++k) v[k] = metric_value<D, M>(xi, si, sx, cs, j + k, ssig, rp);
        load_block<U>(wcol + j, w);
  #pragma unroll
        for (int k = 0; k < U; ++k) add(v[k], j + k, w[k], thr, own_c, mode_c);
      }
      for (; j < j1; ++j)
        add(metric_value<D, M>(xi, si, sx, cs, j, ssig, rp), j, wcol[j], thr, own_c, mode_c);
    };
    // A neighbour tile's candidates go by boxes of kSortedBox sorted positions: a
    // box that no lane can reach (a lower bound of the metric over it above every
    // lane's threshold) holds only candidates beyond every lane's radius, whose
    // terms the reference adds as +0, so it is skipped without changing the sums.
    // (The own tile is the dense one: there nearly every box is reached.)
    auto walk_boxes = [&](int j0, int j1, T thr, auto mode_c) {
      if constexpr (M != M_MAHALANOBIS) {
        for (int j = j0; j < j1;) {
          const int qb = j / kSortedBox;
          const int je = min(j1, (qb + 1) * kSortedBox);
          T bl[D], bh[D];
  #pragma unroll
          for (int d = 0; d < D; ++d) {
            bl[d] = sbox[static_cast<int64_t>(qb) * 2 * D + d];
            bh[d] = sbox[static_cast<int64_t>(qb) * 2 * D + D + d];
          }
          if (ballot(metric_lb<D, M>(xi, xi, bl, bh, rp) <= thr) != 0ull)
            walk(j, je, thr, std::false_type{}, mode_c);
          j = je;
        }
      } else {
        walk(j0, j1, thr, std::false_type{}, mode_c);
      }
    };
    auto run = [&](auto mode_c) {
      walk_tiles<D>(lo, hi, ev.nper, rp.wrapped, ev.t_base, [&](int t, bool in) {
        const int j0 = toff[t];
        const int j1 = toff[t + 1];
        const T thr = in ? thr_in : T(-1);
        if (t == it.tile)
          walk(j0, j1, thr, std::true_type{}, mode_c);
        else if (sbox)
          walk_boxes(j0, j1, thr, mode_c);
        else
          walk(j0, j1, thr, std::false_type{}, mode_c);
      });
    };
    // Both decisions are made once, outside the candidate loop:
    //  - in range: launch-wide, every exp argument within the radius is in range;
    //  - normal: Euclidean in float, and no lane of the wave has a coordinate of
    //    magnitude below 2^-24. Then every pair with v > 0 has some |dx| >= 2^-48
    //    (floats of magnitude >= 2^-24 are multiples of 2^-47, and the largest float
    //    below 2^-24 is 2^-24 - 2^-48), so v >= 2^-96 and the square root needs no
    //    tiny-input branch.
    if constexpr (K == K_FLAT) {
      run(ModeC<true, false>{});
    } else {
      bool tiny = false;
  #pragma unroll
      for (int d = 0; d < D; ++d) tiny = tiny || (__builtin_fabs(xi[d]) < T(0x1p-24));
      const bool normal = (M == M_EUCLID) && sizeof(T) == 4 && ballot(active && tiny) == 0ull;
      if (rp.exp_in_range && normal)
        run(ModeC<true, true>{});
      else if (rp.exp_in_range)
        run(ModeC<true, false>{});
      else
        run(ModeC<false, false>{});
    }
    if (hkey) {
      // The nearest-higher block layout takes the keys here: the lanes are the
      // tile's points in Morton order (qperm is qfine) from a multiple of 64, so
      // lane k's point fills layout slot k of the tile and each row of 16 lanes
      // is one block, whose largest key a row reduction gives. Lanes past the
      // tile's points write the padding's key 0 up to the last block.
      static_assert(kWave % kNhBlock == 0, "a block is a group of lanes of one wave");
      const int tb = toff[it.tile];
      const int k = q - tb;
      const int span = (toff[it.tile + 1] - tb + kNhBlock - 1) / kNhBlock * kNhBlock;
      const uint64_t key = active ? density_key(rho, stag[s]) : 0ull;
      const uint64_t m = row_max_u64<kNhBlock>(key);
      if (k < span) {
        const int64_t p = nh_first_block(tb, it.tile) * kNhBlock + k;
        hkey[p] = key;
        if ((k % kNhBlock) == 0) hbkey[p / kNhBlock] = m;
      }
      if (active) srho[s] = rho;
    } else if (active) {
      srho[s] = rho;
      skey[s] = density_key(rho, stag[s]);
    }
  };
  if constexpr (Persistent) {
    ItemCursor cursor(n_items, kQueueDensity, rp);
    for (int item = cursor.pull(); item >= 0; item = cursor.pull()) process(item);
  } else {
    const int item = xcd_block(blockIdx.x, gridDim.x, rp.xcd_chunk_log2) * (blockDim.x / kWave) +
                     threadIdx.x / kWave;
    if (item < *n_items) process(item);
  }
}

// ---------------------------------------------------------------------------
// (c) Nearest-higher (ClusteringKernels.hpp:164-206, 247-289). The search box uses
// dm; the acceptance radius is seed_dc for points at or above rhoc*u, else dm.
// A candidate is "higher" when its (rho, tag) key exceeds the point's. Only
// candidates at or below the current best distance can change the result; the
// running threshold below over-approximates that bound, and every candidate that
// passes it goes through the reference's exact comparison.
// ---------------------------------------------------------------------------
template <int M>
__device__ __forceinline__ float prune_threshold(float eff_thr, float delta) {
  if constexpr (MetricTraits<M>::kSqrt) {
    // sqrt_rn(t) <= delta implies t <= delta^2 (1 + 2^-22); delta*delta*(1 + 2^-20)
    // rounded bounds that from above for delta >= 2^-60 (no underflow).
    if (!(delta >= 0x1p-60f)) return eff_thr;
    const float b = (delta * delta) * 0x1.00001p+0f;
    return b < eff_thr ? b : eff_thr;
  } else {
    return delta < eff_thr ? delta : eff_thr;
  }
}
template <int M>
__device__ __forceinline__ double prune_threshold(double eff_thr, double delta) {
  if constexpr (MetricTraits<M>::kSqrt) {
    // Double: t <= delta^2 (1 + 2^-51.9) follows; delta*delta*(1 + 2^-48) rounded
    // bounds it from above for delta >= 2^-500.
    if (!(delta >= 0x1p-500)) return eff_thr;
    const double b = (delta * delta) * 0x1.000000000001p+0;
    return b < eff_thr ? b : eff_thr;
  } else {
    return delta < eff_thr ? delta : eff_thr;
  }
}

// "j is higher than i" (ClusteringKernels.hpp:180-182) from the stored keys. Float
// keys hold (rho, tag); double keys hold rho's order only, so equal densities
// compare the tags here (rho > 0 is the key above that of +0).
__device__ __forceinline__ bool key_higher(uint64_t kj, uint32_t, uint64_t ki, uint32_t, float) {
  return kj > ki;
}
__device__ __forceinline__ bool key_higher(uint64_t kj, uint32_t tj, uint64_t ki, uint32_t ti,
                                           double) {
  return (kj > ki) | ((kj == ki) & (kj > 0x8000000000000000ull) & (tj > ti));
}
__device__ __forceinline__ uint64_t query_key(float rho, uint32_t tag) {
  return rho_tag_key(rho, tag, true);
}
__device__ __forceinline__ uint64_t query_key(double rho, uint32_t) {
  return rho_key_f64(rho, true);
}

template <int D, int M, class T>
__global__ __launch_bounds__(kMaxNeighbourBlock) void k_nearest_higher(
    const RunParamsT<T> rp, const EventDesc* __restrict__ events, const T* __restrict__ geo,
    const int32_t* __restrict__ toff, const Item* __restrict__ items,
    const int32_t* __restrict__ n_items, const T* __restrict__ sx, int64_t cs,
    const T* __restrict__ ssig, const T* __restrict__ su,
    const uint32_t* __restrict__ stag, const T* __restrict__ srho,
    const uint64_t* __restrict__ skey, const int32_t* __restrict__ qperm,
    int32_t* __restrict__ snh, T* __restrict__ sdelta) {
  constexpr bool kF64 = sizeof(T) == 8;
  constexpr int U = Unroll<D>::U_NH;
  const int item = xcd_block(blockIdx.x, gridDim.x, rp.xcd_chunk_log2) * (blockDim.x / kWave) +
                   threadIdx.x / kWave;
  if (item >= *n_items) return;
  const Item it = items[item];
  const EventDesc ev = events[it.event];
  const GeoT<T> g = event_geo(geo, rp, it.event, ev.nper);
  const int q = it.begin + lane_id();
  const bool active = q < it.end;
  const int s = active ? (qperm ? qperm[q] : q) : -1;  // sorted position of this lane's point

  T xi[D], si[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    xi[d] = active ? sx[d * cs + s] : T(0);
    si[d] = (active && rp.has_sigma) ? ssig[d * cs + s] : T(0);
  }
  const T rho_i = active ? srho[s] : T(0);
  const uint32_t tag_i = active ? stag[s] : 0u;
  const uint64_t key_i = active ? query_key(rho_i, tag_i) : ~0ull;
  const T ui = (active && rp.has_u) ? su[s] : T(1);
  const T eff_min = rp.rhoc * ui;
  const bool use_seed = rho_i >= eff_min;
  const T eff = use_seed ? rp.seed_dc : rp.dm;
  const T eff_thr =
      MetricTraits<M>::kSqrt ? (use_seed ? rp.seed_sq_max : rp.dm_sq_max) : eff;

  int lo[D], hi[D];
  search_box<D>(g, xi, si, rp.dm, lo, hi);
  if (!active) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lo[d] = INT32_MAX;
      hi[d] = INT32_MIN;
    }
  }
  T delta = std::numeric_limits<T>::max();
  int nh = -1;
  T rho_nh = T(0);
  uint32_t tag_nh = 0u;
  T bound = eff_thr;  // prune_threshold(eff_thr, delta)

  // The reference's acceptance test, in candidate order. Its radius test, dist <=
  // eff, holds for every candidate that reaches it (callers pass only v <= bound
  // <= eff_thr, the largest value whose distance is <= eff), so it is left out.
  auto consider = [&](T v, T rj, uint32_t tj, int j) {
    const T dist = finish_distance<M>(v);
    if ((dist < delta) ||
        ((dist == delta) && (nh >= 0) && ((rj > rho_nh) || ((rj == rho_nh) && (tj > tag_nh))))) {
      delta = dist;
      nh = j;
      rho_nh = rj;
      tag_nh = tj;
      bound = prune_threshold<M>(eff_thr, delta);
    }
  };

  walk_tiles<D>(lo, hi, ev.nper, rp.wrapped, ev.t_base, [&](int t, bool in) {
    const int j0 = toff[t];
    const int j1 = toff[t + 1];
    if (!in) return;  // lanes whose box does not hold this tile idle through it
    int j = j0;
    for (; j + U <= j1; j += U) {
      // m: the smallest distance among the block's higher candidates (vector
      // selects and minima instead of per-candidate scalar mask logic).
      T v[U];
      uint32_t tj[U];
      T m = std::numeric_limits<T>::infinity();
#pragma unroll
      for (int k = 0; k < U; ++k) tj[k] = kF64 ? stag[j + k] : 0u;
#pragma unroll
      for (int k = 0; k < U; ++k) {
        v[k] = metric_value<D, M>(xi, si, sx, cs, j + k, ssig, rp);
        const bool h = key_higher(skey[j + k], tj[k], key_i, tag_i, T(0));
        m = fmin_t(m, h ? v[k] : std::numeric_limits<T>::infinity());
      }
      // Most blocks hold no candidate for any lane: one wave-wide test skips them.
      if (ballot(m <= bound) != 0ull) {
        T rj[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
          rj[k] = srho[j + k];
          if (!kF64) tj[k] = stag[j + k];
        }
#pragma unroll
        for (int k = 0; k < U; ++k)
          if (key_higher(skey[j + k], tj[k], key_i, tag_i, T(0)) && v[k] <= bound)
            consider(v[k], rj[k], tj[k], j + k);
      }
    }
    for (; j < j1; ++j) {
      const T v = metric_value<D, M>(xi, si, sx, cs, j, ssig, rp);
      if (key_higher(skey[j], stag[j], key_i, tag_i, T(0)) & (v <= bound))
        consider(v, srho[j], stag[j], j);
    }
  });
  if (active) {
    snh[s] = nh;
    sdelta[s] = delta;
  }
}

// ---------------------------------------------------------------------------
// (c') Nearest-higher over Morton-ordered candidate blocks. The accepted
// candidate of the reference's scan is the one of smallest distance, equal
// distances going to the larger (rho, tag); with unique tags (the default ones:
// the input index) no two candidates tie on all three, so the result does not
// depend on the order in which candidates are visited. Each tile's candidates
// are laid out in Morton order in blocks of kNhBlock with the block's bounding
// box and largest key; a wave skips a block when for each lane either no
// candidate of the block is higher or the box lies beyond the lane's current
// bound. A skipped block holds no candidate that the scan would accept (the box
// bound never exceeds a candidate's computed distance value: rounding is
// monotone), so the results are those of the plain scan.
// ---------------------------------------------------------------------------

// Wave-wide minimum and maximum of floating-point values, and a lower bound of
// the wave's smallest ordering key. (LDS permutes here: the DPP form costs more
// vector instructions, and this scan is short of them; measured on MI355X.)
template <class T>
__device__ __forceinline__ T wave_min_f(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin_t(v, __shfl_xor(v, off, 64));
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max_f(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax_t(v, __shfl_xor(v, off, 64));
  return v;
}
// The smallest high word with a zero low word: never above the smallest key, so
// "a block key above it" keeps every block that could hold a higher candidate.
__device__ __forceinline__ uint64_t wave_min_key_bound(uint64_t v) {
  uint32_t h = static_cast<uint32_t>(v >> 32);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(h), off, 64));
    h = o < h ? o : h;
  }
  return static_cast<uint64_t>(h) << 32;
}

// Lane i's value in every lane (i wave-uniform).
__device__ __forceinline__ float read_lane(float v, int i) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
}
__device__ __forceinline__ double read_lane(double v, int i) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), i);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), i);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<uint32_t>(lo)));
}
__device__ __forceinline__ uint64_t read_lane(uint64_t v, int i) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), i));
  const uint32_t hi =
      static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), i));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Candidates per register block of the layout scan (a divisor of kNhBlock).
template <int D>
constexpr int nh_sub_block() {
  return D <= 3 ? 8 : 4;
}

// May a block whose largest key is bk hold a candidate higher than the query?
// Float keys decide it exactly; double keys leave equal densities to the tags.
__device__ __forceinline__ bool block_may_be_higher(uint64_t bk, uint64_t ki, float) {
  return bk > ki;
}
__device__ __forceinline__ bool block_may_be_higher(uint64_t bk, uint64_t ki, double) {
  return bk >= ki;
}
__device__ __forceinline__ bool key_maybe_higher(uint64_t kj, uint64_t ki, float) {
  return kj > ki;
}
__device__ __forceinline__ bool key_maybe_higher(uint64_t kj, uint64_t ki, double) {
  return (kj > ki) | ((kj == ki) & (kj > 0x8000000000000000ull));
}

// Eight waves per SIMD in 1-D and 2-D: the scan waits on chains of scalar loads,
// so resident waves matter more than registers (measured on MI355X on the
// 4096-event 2-D batch: 12.3 ms with the compiler's own choice, 11.0 ms with
// this bound, 10.2 ms with it and register blocks of 8 candidates instead of
// 16). From 3-D on the bound would spill to scratch; there the compiler chooses
// (equal time on the 2^20-point 3-D event).
template <int D, int M, bool Persistent, class T>
__global__ __launch_bounds__(kMaxNeighbourBlock)
__attribute__((amdgpu_waves_per_eu(D <= 2 ? 8 : 1))) void k_nh_blocks(
    const RunParamsT<T> rp, const EventDesc* __restrict__ events, const T* __restrict__ geo,
    const int32_t* __restrict__ toff, const Item* __restrict__ items,
    int32_t* __restrict__ n_items, const T* __restrict__ sx, int64_t cs,
    const T* __restrict__ ssig, const T* __restrict__ su, const uint32_t* __restrict__ stag,
    const T* __restrict__ srho, const int32_t* __restrict__ qperm, const T* __restrict__ hx,
    int64_t hs,
    const uint64_t* __restrict__ hkey, const int32_t* __restrict__ hidx,
    const T* __restrict__ hbox, const uint64_t* __restrict__ hbkey, int32_t* __restrict__ snh,
    T* __restrict__ sdelta) {
  constexpr int V = nh_sub_block<D>();
  auto process = [&](const int item) {
    const Item it = items[item];
    const EventDesc ev = events[it.event];
    const GeoT<T> g = event_geo(geo, rp, it.event, ev.nper);
    const int q = it.begin + lane_id();
    const bool active = q < it.end;
    const int s = active ? (qperm ? qperm[q] : q) : -1;  // sorted position of this lane's point

    T xi[D], si[D];
  #pragma unroll
    for (int d = 0; d < D; ++d) {
      xi[d] = active ? sx[d * cs + s] : T(0);
      si[d] = (active && rp.has_sigma) ? ssig[d * cs + s] : T(0);
    }
    const T rho_i = active ? srho[s] : T(0);
    const uint32_t tag_i = active ? stag[s] : 0u;
    const uint64_t key_i = active ? query_key(rho_i, tag_i) : ~0ull;
    const T ui = (active && rp.has_u) ? su[s] : T(1);
    const T eff_min = rp.rhoc * ui;
    const bool use_seed = rho_i >= eff_min;
    const T eff = use_seed ? rp.seed_dc : rp.dm;
    const T eff_thr =
        MetricTraits<M>::kSqrt ? (use_seed ? rp.seed_sq_max : rp.dm_sq_max) : eff;

    int lo[D], hi[D];
    search_box<D>(g, xi, si, rp.dm, lo, hi);
    if (!active) {
  #pragma unroll
      for (int d = 0; d < D; ++d) {
        lo[d] = INT32_MAX;
        hi[d] = INT32_MIN;
      }
    }
    T delta = std::numeric_limits<T>::max();
    int64_t nh_slot = -1;  // layout slot of the nearest-higher found so far
    T rho_nh = T(0);
    uint32_t tag_nh = 0u;
    T bound = eff_thr;  // prune_threshold(eff_thr, delta)

    // The reference's acceptance test (ClusteringKernels.hpp:180-195). Its radius
    // test, dist <= eff, holds for every candidate that reaches it: callers pass
    // only v <= bound <= eff_thr, the largest value whose distance is <= eff.
    auto consider = [&](T v, T rj, uint32_t tj, int64_t slot) {
      const T dist = finish_distance<M>(v);
      if ((dist < delta) ||
          ((dist == delta) && (nh_slot >= 0) &&
           ((rj > rho_nh) || ((rj == rho_nh) && (tj > tag_nh))))) {
        delta = dist;
        nh_slot = slot;
        rho_nh = rj;
        tag_nh = tj;
        bound = prune_threshold<M>(eff_thr, delta);
      }
    };

    // One block's candidates, V at a time.
    auto scan_block = [&](int64_t p0, bool in) {
      for (int k0 = 0; k0 < kNhBlock; k0 += V) {
        const int64_t p = p0 + k0;
        T v[V];
        T m = std::numeric_limits<T>::infinity();
  #pragma unroll
        for (int k = 0; k < V; ++k) {
          T xj[D], sj[D];
  #pragma unroll
          for (int d = 0; d < D; ++d) {
            xj[d] = hx[d * hs + p + k];
            sj[d] = T(0);
          }
          v[k] = metric_pair<D, M>(xi, si, xj, sj, rp);
          const bool h = in & key_maybe_higher(hkey[p + k], key_i, T(0));
          m = fmin_t(m, h ? v[k] : std::numeric_limits<T>::infinity());
        }
        if (ballot(m <= bound) != 0ull) {
  #pragma unroll
          for (int k = 0; k < V; ++k) {
            const uint64_t kj = hkey[p + k];
            if (in & key_maybe_higher(kj, key_i, T(0)) & (v[k] <= bound)) {
              if constexpr (sizeof(T) == 4) {
                // Float keys are exact and carry rho, and the tag when rho > 0: no
                // dependent loads on this path.
                const float rj = key_rho(kj);
                const uint32_t tj = rj > 0.0f ? static_cast<uint32_t>(kj) : stag[hidx[p + k]];
                consider(v[k], rj, tj, p + k);
              } else {
                const int j = hidx[p + k];
                const uint32_t tj = stag[j];
                if (key_higher(kj, tj, key_i, tag_i, T(0))) consider(v[k], srho[j], tj, p + k);
              }
            }
          }
        }
      }
    };

    // in: the lane's search box holds tile t. The boxes and keys of up to 64 blocks
    // arrive in one vector load (one block per lane); a wave-wide test against the
    // box, smallest key and largest bound of the lanes that hold the tile drops
    // most blocks at once, and each block left is tested per lane before its scan.
    // The tile's blocks are visited from block sb on, wrapping around: in the own
    // tile sb is the block that holds the wave's first point, whose neighbours
    // settle most bounds before the rest of the tile is tested. The result does
    // not depend on the order (the acceptance test is a strict total order).
    auto scan_tile = [&](int t, bool in, int sb) {
      const int b = toff[t];
      const int nblk = (toff[t + 1] - b + kNhBlock - 1) / kNhBlock;
      const int64_t blk0 = nh_first_block(b, t);
      T qlo[D], qhi[D];
  #pragma unroll
      for (int d = 0; d < D; ++d) {
        qlo[d] = wave_min_f(in ? xi[d] : std::numeric_limits<T>::infinity());
        qhi[d] = wave_max_f(in ? xi[d] : -std::numeric_limits<T>::infinity());
      }
      const uint64_t kmin = wave_min_key_bound(in ? key_i : ~0ull);
      for (int c0 = 0; c0 < nblk; c0 += kWave) {
        const T bmax = wave_max_f(in ? bound : -std::numeric_limits<T>::infinity());
        const int l = lane_id();
        const bool valid = c0 + l < nblk;
        int rl = c0 + (valid ? l : 0) + sb;
        rl -= rl >= nblk ? nblk : 0;
        const int64_t qb = blk0 + rl;
        T bl[D], bh[D];
  #pragma unroll
        for (int d = 0; d < D; ++d) {
          bl[d] = hbox[qb * 2 * D + d];
          bh[d] = hbox[qb * 2 * D + D + d];
        }
        const uint64_t bk = hbkey[qb];
        uint64_t cand = ballot(valid & block_may_be_higher(bk, kmin, T(0)) &
                               (metric_lb<D, M>(qlo, qhi, bl, bh, rp) <= bmax));
        while (cand) {
          const int i = __builtin_ctzll(cand);
          cand &= cand - 1ull;
          T sl[D], sh[D];
  #pragma unroll
          for (int d = 0; d < D; ++d) {
            sl[d] = read_lane(bl[d], i);
            sh[d] = read_lane(bh[d], i);
          }
          const bool need = in & block_may_be_higher(read_lane(bk, i), key_i, T(0)) &
                            (metric_lb<D, M>(xi, xi, sl, sh, rp) <= bound);
          int ri = c0 + i + sb;
          ri -= ri >= nblk ? nblk : 0;
          if (ballot(need) != 0ull) scan_block((blk0 + ri) * kNhBlock, in);
        }
      }
    };

    // The lanes' own tile first: the nearest candidates tighten the bounds early.
    // A lane scans it only if its box holds it: on a wrapped dimension a box wider
    // than half the data's extent can leave it out (the walk below reaches the
    // own tile at bin b or, wrapped, at b + nper).
    const int own = uniform(it.tile);  // wave-uniform: keeps the scan's loads scalar
    {
      bool in_own = active;
      int r = own - ev.t_base;
  #pragma unroll
      for (int d = D - 1; d >= 0; --d) {
        const int b = r % ev.nper;
        r /= ev.nper;
        in_own = in_own && (((lo[d] <= b) && (b <= hi[d])) ||
                            (rp.wrapped[d] && (lo[d] <= b + ev.nper) && (b + ev.nper <= hi[d])));
      }
      if (ballot(in_own) != 0ull) scan_tile(own, in_own, uniform(it.begin - toff[own]) / kNhBlock);
    }
    walk_tiles<D>(lo, hi, ev.nper, rp.wrapped, ev.t_base, [&](int t, bool in) {
      if (t != own) scan_tile(t, in, 0);
    });
    if (active) {
      snh[s] = nh_slot >= 0 ? hidx[nh_slot] : -1;
      sdelta[s] = delta;
    }
  };
  if constexpr (Persistent) {
    ItemCursor cursor(n_items, kQueueNh, rp);
    for (int item = cursor.pull(); item >= 0; item = cursor.pull()) process(item);
  } else {
    const int item = xcd_block(blockIdx.x, gridDim.x, rp.xcd_chunk_log2) * (blockDim.x / kWave) +
                     threadIdx.x / kWave;
    if (item < *n_items) process(item);
  }
}

// ---------------------------------------------------------------------------
// (d) Seed flags (ClusteringKernels.hpp:300-316) and the results in input order.
// snh is rewritten in place as the follower link of the label stage:
// -2 for a seed, -1 for a point without nearest-higher, else the sorted
// position of its nearest-higher.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_unpermute(
    const RunParamsT<T> rp, int n, const int32_t* __restrict__ perm, const T* __restrict__ srho,
    int32_t* __restrict__ snh, const T* __restrict__ sdelta, const T* __restrict__ su,
    T* __restrict__ rho, int32_t* __restrict__ nh, T* __restrict__ delta,
    int32_t* __restrict__ is_seed, int32_t* __restrict__ flag) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const T r = srho[s];
  const int h = snh[s];
  const T u = rp.has_u ? su[s] : T(1);
  const int seed = (h == -1) && (r >= rp.rhoc * u) ? 1 : 0;
  rho[i] = r;
  nh[i] = h >= 0 ? perm[h] : -1;
  if (delta) delta[i] = sdelta[s];
  is_seed[i] = seed;
  flag[i] = seed;
  snh[s] = seed ? -2 : h;
}

// ---------------------------------------------------------------------------
// (d fused) With the per-event tile build, one workgroup per event does the seed
// stage of k_unpermute, numbers the event's seeds in input order through a bit
// per point in LDS (one popcount scan over the bits: the rank of a seed among
// the event's seeds, stored in rank[i]), counts them (event_seeds[e]) and
// shortens the follower links with the pointer-jumping passes of k_link_jump.
// The seeds of earlier events are added by the label stage from an exclusive
// scan of event_seeds. The seed flags and links are settled in sorted order;
// the results are then written in input order, each point's sorted values
// gathered through iperm, so that every store is coalesced. An event of at
// most kEvtLdsLinks points keeps its links in LDS as 16-bit event positions:
// the results pass reads them there, the jumping passes run there (more of
// them, as they cost no memory traffic) and the links go back to memory once.
// ---------------------------------------------------------------------------
constexpr int kLinkJumpPasses = 4;     // pointer-jumping passes in memory before the label chase
constexpr int kLdsJumpPasses = 6;      // in LDS: chains of up to 64 links end at their root
constexpr int kEvtLdsLinks = 1 << 15;  // events whose links fit LDS (16-bit positions)
constexpr int kEvtBitWords = kEvtMaxPoints / 32;
constexpr int kLdsBitWords = kEvtLdsLinks / 32;
// One LDS pool: the seed bits of any event, or the bits and links of a small one
// (two workgroups per CU, as the 16 waves of each allow).
constexpr int kSeedPoolWords =
    kEvtBitWords > kLdsBitWords + kEvtLdsLinks / 2 ? kEvtBitWords : kLdsBitWords + kEvtLdsLinks / 2;
static_assert(kEvtBitWords % kEvtBlock == 0, "bit words per thread");
static_assert(kEvtLdsLinks <= 32768, "links are 16-bit event positions");

template <class T>
__global__ __launch_bounds__(kEvtBlock) void k_event_seeds(
    const RunParamsT<T> rp, const EventDesc* __restrict__ events, const int32_t* __restrict__ order,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
    const T* __restrict__ srho, int32_t* __restrict__ snh, const T* __restrict__ sdelta,
    const T* __restrict__ su, T* __restrict__ rho, int32_t* __restrict__ nh, T* __restrict__ delta,
    int32_t* __restrict__ is_seed, int32_t* __restrict__ rank, int32_t* __restrict__ event_seeds,
    int32_t* __restrict__ lone_offsets, int32_t* __restrict__ lone_total,
    int32_t* __restrict__ cluster) {
  constexpr int kWordsPerThread = kEvtBitWords / kEvtBlock;
  __shared__ uint32_t pool[kSeedPoolWords];
  __shared__ int32_t wsum[kEvtWaves];
  uint32_t* bits = pool;
  int16_t* links = reinterpret_cast<int16_t*>(pool + kLdsBitWords);
  const int e = order[blockIdx.x];
  const EventDesc ev = events[e];
  const int b0 = ev.p_off, n = ev.n;
  const bool in_lds = n <= kEvtLdsLinks;
  const int n_words = in_lds ? kLdsBitWords : kEvtBitWords;  // bit words in use
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  for (int k = threadIdx.x; k < n_words; k += kEvtBlock) bits[k] = 0u;
  __syncthreads();
  // Seed flags (ClusteringKernels.hpp:300-316) in sorted order: a seed's link
  // becomes -2, and its bit (input order) is set.
#pragma unroll 4
  for (int k = threadIdx.x; k < n; k += kEvtBlock) {
    const int s = b0 + k;
    const int h = snh[s];
    const T u = rp.has_u ? su[s] : T(1);
    const bool seed = (h == -1) && (srho[s] >= rp.rhoc * u);
    if (in_lds) links[k] = static_cast<int16_t>(seed ? -2 : (h >= 0 ? h - b0 : h));
    if (seed) {
      if (!in_lds) snh[s] = -2;
      const int i = perm[s] - b0;
      atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // the -2 links
  __syncthreads();
  // Results in input order.
#pragma unroll 4
  for (int k = threadIdx.x; k < n; k += kEvtBlock) {
    const int i = b0 + k;
    const int s = iperm[i];
    int h;
    if (in_lds) {
      h = links[s - b0];
      h = h >= 0 ? h + b0 : h;
    } else {
      h = snh[s];
    }
    rho[i] = srho[s];
    nh[i] = h >= 0 ? perm[h] : -1;
    if (delta) delta[i] = sdelta[s];
    is_seed[i] = h == -2 ? 1 : 0;
  }
  // Ranks in input order: each thread owns kWordsPerThread consecutive words
  // (words past the bits in use count nothing).
  const int w0 = threadIdx.x * kWordsPerThread;
  int mine = 0;
#pragma unroll
  for (int k = 0; k < kWordsPerThread; ++k) mine += w0 + k < n_words ? __popc(bits[w0 + k]) : 0;
  int incl = mine;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == kWave - 1) wsum[w] = incl;
  __syncthreads();
  int before = 0, total = 0;
  for (int k = 0; k < kEvtWaves; ++k) {
    before += k < w ? wsum[k] : 0;
    total += wsum[k];
  }
  int run = before + incl - mine;
#pragma unroll
  for (int k = 0; k < kWordsPerThread; ++k) {
    uint32_t word = w0 + k < n_words ? bits[w0 + k] : 0u;
    while (word) {
      const int bit = __builtin_ctz(word);
      word &= word - 1u;
      const int ii = (w0 + k) * 32 + bit;  // the seed's input position in the event
      if (in_lds) {
        // The event's seed list in rank order, and the seed's link: its rank as
        // a code (-3 - rank) that the jumping passes carry to its followers.
        rank[b0 + run] = b0 + ii;
        links[iperm[b0 + ii] - b0] = static_cast<int16_t>(-3 - run);
        ++run;
      } else {
        rank[b0 + ii] = run++;
      }
    }
  }
  if (threadIdx.x == 0) {
    event_seeds[e] = total;
    if (lone_total) {  // a batch of one event: the scan over events is the identity
      lone_offsets[0] = 0;
      *lone_total = total;
    }
  }
  // Pointer jumping on the event's links (as k_link_jump; any interleaving of the
  // updates leaves every link on its own chain).
  if (in_lds) {
    // A link is a position, a seed's code or -1 (the end of an outlier chain), and
    // every point shares the label of what its link names; replacing a position
    // by that point's link keeps this, in any interleaving. After the passes the
    // few chains longer than 2^kLdsJumpPasses are finished by a chase in LDS, and
    // each point's label within the event goes to cluster[] in input order (the
    // label stage adds the event's offset).
    for (int pass = 0; pass < kLdsJumpPasses; ++pass) {
      __syncthreads();
      for (int k = threadIdx.x; k < n; k += kEvtBlock) {
        const int l = links[k];
        if (l >= 0) links[k] = links[l];
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int k = threadIdx.x; k < n; k += kEvtBlock) {
      const int i = b0 + k;
      int l = links[iperm[i] - b0];
      while (l >= 0) l = links[l];
      cluster[i] = l <= -3 ? -3 - l : -1;
    }
  } else {
    for (int pass = 0; pass < kLinkJumpPasses; ++pass) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __syncthreads();
      for (int k = threadIdx.x; k < n; k += kEvtBlock) {
        const int s = b0 + k;
        const int l = snh[s];
        if (l < 0) continue;
        const int l2 = snh[l];
        if (l2 >= 0) snh[s] = l2;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// (e) Labels: seeds take their rank in input order (the serial push_back order,
// ClusteringKernels.hpp:310, 318-327); followers chase nearest_higher to a seed
// or to a point without one (:329-343). The chase runs in tile-sorted space,
// where a chain stays among neighbouring tiles of one event. An event occupies
// the same range of positions in input and sorted order.
// ---------------------------------------------------------------------------
// Pointer jumping on the follower links before the chase: each pass replaces a
// link by its target's link when that target is itself a follower, so a pass
// halves every chain (at least) and the chase's dependent loads shrink with it.
// Every link still points to an ancestor on the same chain, whatever order the
// threads run in, so the chase ends at the same seed or outlier.
static __global__ __launch_bounds__(kBlock) void k_link_jump(int32_t* __restrict__ link, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int l = link[s];
  if (l < 0) return;
  const int l2 = link[l];
  if (l2 >= 0) link[s] = l2;
}

// rank[i] of a seed i is its global rank (seeds of the whole batch before it in
// input order), or, with event_offsets, its rank within its event, to which the
// seeds of the earlier events (event_offsets[e]) are added here. Chains are
// chased from sorted positions, so their first links are read in order and the
// label stores scatter within each event (chasing from input order instead
// stores in order but reads twice the bytes, for the same time on MI355X).
// An event the fused seed stage resolved in LDS (at most kEvtLdsLinks points)
// has its labels within the event in cluster[] already and its seed list in
// rank[p_off, p_off + event_seeds[e]): only the offset is added.
static __global__ __launch_bounds__(kBlock) void k_labels(const Chunk* __restrict__ chunks,
                                                   const EventDesc* __restrict__ events,
                                                   const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ link,
                                                   const int32_t* __restrict__ rank,
                                                   const int32_t* __restrict__ event_offsets,
                                                   const int32_t* __restrict__ event_seeds,
                                                   int32_t* __restrict__ cluster,
                                                   int32_t* __restrict__ seeds,
                                                   int32_t* __restrict__ event_of_cluster,
                                                   int local_ids) {
  const Chunk ch = chunks[blockIdx.x];
  const EventDesc ev = events[ch.event];
  // Seeds of the earlier events: added to local ranks, subtracted from global
  // ranks with per-event numbering.
  const int first = event_offsets ? event_offsets[ch.event] : rank[ev.p_off];
  if (event_offsets && ev.n <= kEvtLdsLinks) {
    const int count = event_seeds[ch.event];
    for (int p = ch.begin + threadIdx.x; p < ch.end; p += kBlock) {
      const int c = cluster[p];
      if (c >= 0 && !local_ids) cluster[p] = c + first;
      const int r = p - ev.p_off;
      if (r < count) {
        seeds[first + r] = rank[p];
        if (event_of_cluster) event_of_cluster[first + r] = ev.index;
      }
    }
    return;
  }
  const int add = event_offsets ? first : 0;
  const int sub = local_ids ? first : 0;
  for (int s = ch.begin + threadIdx.x; s < ch.end; s += kBlock) {
    int cur = s;
    int nx = link[cur];
    while (nx >= 0) {
      cur = nx;
      nx = link[cur];
    }
    const int i = perm[s];
    const int c = (nx == -2) ? rank[perm[cur]] + add : -1;
    cluster[i] = c >= 0 ? c - sub : -1;
    if (cur == s && nx == -2) {
      seeds[c] = i;
      if (event_of_cluster) event_of_cluster[c] = ev.index;
    }
  }
}

// ---------------------------------------------------------------------------
// Dispatch over (D, metric, kernel).
// ---------------------------------------------------------------------------
// Compute units of the current device (cached per device).
inline int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

template <class T>
struct NeighbourLaunch {
  const RunParamsT<T>* rp;
  const BatchBuffersT<T>* bb;
  int64_t cs;
  unsigned grid;             // one wave per work item (the grid form)
  unsigned grid_persistent;  // rp.persistent waves per CU (kernels with persistent variants)
  unsigned block;  // threads per workgroup of the neighbour stages
  hipStream_t s;
};

// Persistent waves (the device work-item queues) are built for 1-3 dimensions,
// the ones the tuning table covers; other dimensions take the grid form.
template <int D>
constexpr bool kPersistentBuilt = D <= 3;

template <int D, int M, int K, class T>
static void launch_density_dmk(const NeighbourLaunch<T>& L) {
  const BatchBuffersT<T>& b = *L.bb;
  auto go = [&](auto kern, unsigned grid) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(L.block), 0, L.s, *L.rp, b.events, b.geo, b.toff,
                       b.items, b.n_items, b.sorted.x[0], L.cs, b.sorted.w, b.sorted.fw,
                       b.sorted.sigma[0], b.sorted.tag, b.qperm, b.srho, b.skey, b.sbox,
                       b.hkey, b.hbkey);
  };
  if constexpr (kPersistentBuilt<D>) {
    if (L.rp->persistent > 0) {
      go(k_density<D, M, K, true, T>, L.grid_persistent);
      return;
    }
  }
  go(k_density<D, M, K, false, T>, L.grid);
}

template <int D, int M, class T>
static void launch_density_dm(const NeighbourLaunch<T>& L) {
  switch (L.rp->kernel) {
    case K_FLAT: launch_density_dmk<D, M, K_FLAT>(L); break;
    case K_EXP:
      if (L.rp->kp1 == T(1))
        launch_density_dmk<D, M, K_EXP_UNIT>(L);
      else
        launch_density_dmk<D, M, K_EXP>(L);
      break;
    default: {
      // Float folds the power of two into the exp's constants (|s| <= 8 only).
      const bool pow2 = sizeof(T) == 4 ? L.rp->gauss_fold != 0 : L.rp->gauss_pow2 != 0;
      if (pow2 && L.rp->kp1 == T(1))
        launch_density_dmk<D, M, K_GAUSS2_UNIT>(L);
      else if (pow2)
        launch_density_dmk<D, M, K_GAUSS2>(L);
      else
        launch_density_dmk<D, M, K_GAUSS>(L);
      break;
    }
  }
}

template <int D, int M, class T>
static void launch_nh_dm(const NeighbourLaunch<T>& L) {
  const BatchBuffersT<T>& b = *L.bb;
  hipLaunchKernelGGL((k_nearest_higher<D, M, T>), dim3(L.grid), dim3(L.block), 0, L.s, *L.rp,
                     b.events, b.geo, b.toff, b.items, b.n_items, b.sorted.x[0], L.cs,
                     b.sorted.sigma[0], b.sorted.u, b.sorted.tag, b.srho, b.skey, b.qperm_nh,
                     b.snh, b.sdelta);
}

template <int D, int M, class T>
static void launch_nh_blocks_dm(const NeighbourLaunch<T>& L) {
  const BatchBuffersT<T>& b = *L.bb;
  auto go = [&](auto kern, unsigned grid) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(L.block), 0, L.s, *L.rp, b.events, b.geo, b.toff,
                       b.items, b.n_items, b.sorted.x[0], L.cs, b.sorted.sigma[0], b.sorted.u,
                       b.sorted.tag, b.srho, b.qfine, b.hx, b.h_stride, b.hkey, b.hidx,
                       b.hbox, b.hbkey, b.snh, b.sdelta);
  };
  if constexpr (kPersistentBuilt<D>) {
    if (L.rp->persistent > 0) {
      go(k_nh_blocks<D, M, true, T>, L.grid_persistent);
      return;
    }
  }
  go(k_nh_blocks<D, M, false, T>, L.grid);
}

// Nearest-higher over the block layout: the layout pass, then the scan. Its
// waves take each tile's points in Morton order (qfine): the scan visits only the
// blocks its lanes want, so grouping lanes by search-box class buys nothing here.
template <int D, class T>
static void launch_nh_blocks_d(const NeighbourLaunch<T>& L) {
  switch (L.rp->metric) {
    case M_EUCLID: launch_nh_blocks_dm<D, M_EUCLID>(L); break;
    case M_WEUCLID: launch_nh_blocks_dm<D, M_WEUCLID>(L); break;
    case M_PERIODIC: launch_nh_blocks_dm<D, M_PERIODIC>(L); break;
    case M_MANHATTAN: launch_nh_blocks_dm<D, M_MANHATTAN>(L); break;
    case M_CHEBYSHEV: launch_nh_blocks_dm<D, M_CHEBYSHEV>(L); break;
    default: launch_nh_blocks_dm<D, M_WCHEBYSHEV>(L); break;
  }
}

// Defined here, instantiated per dimension in clue_neighbour.hip (one object per
// dimension, compiled in parallel); the pipeline's translation units only call it.
template <int D, class T>
void launch_neighbour_d(const NeighbourLaunch<T>& L, bool density) {
  if (!density && L.rp->nh_blocks && L.rp->metric != M_MAHALANOBIS) {
    launch_nh_blocks_d<D>(L);
    return;
  }
#define CLUE_METRIC_CASE(MID)               \
  case MID:                                 \
    if (density)                            \
      launch_density_dm<D, MID>(L);         \
    else                                    \
      launch_nh_dm<D, MID>(L);              \
    break;
  switch (L.rp->metric) {
    CLUE_METRIC_CASE(M_EUCLID)
    CLUE_METRIC_CASE(M_WEUCLID)
    CLUE_METRIC_CASE(M_PERIODIC)
    CLUE_METRIC_CASE(M_MANHATTAN)
    CLUE_METRIC_CASE(M_CHEBYSHEV)
    CLUE_METRIC_CASE(M_WCHEBYSHEV)
    default:
      if (density)
        launch_density_dm<D, M_MAHALANOBIS>(L);
      else
        launch_nh_dm<D, M_MAHALANOBIS>(L);
  }
#undef CLUE_METRIC_CASE
}

#define CLUE_NEIGHBOUR_EXTERN(D)                                                          \
  extern template void launch_neighbour_d<D, float>(const NeighbourLaunch<float>&, bool); \
  extern template void launch_neighbour_d<D, double>(const NeighbourLaunch<double>&, bool);
CLUE_NEIGHBOUR_EXTERN(1)
CLUE_NEIGHBOUR_EXTERN(2)
CLUE_NEIGHBOUR_EXTERN(3)
CLUE_NEIGHBOUR_EXTERN(4)
CLUE_NEIGHBOUR_EXTERN(5)
CLUE_NEIGHBOUR_EXTERN(6)
CLUE_NEIGHBOUR_EXTERN(7)
CLUE_NEIGHBOUR_EXTERN(8)
CLUE_NEIGHBOUR_EXTERN(9)
CLUE_NEIGHBOUR_EXTERN(10)
#undef CLUE_NEIGHBOUR_EXTERN

template <int D, class T>
static void launch_query_order_d(const NeighbourLaunch<T>& L) {
  const BatchBuffersT<T>& b = *L.bb;
  const unsigned grid = static_cast<unsigned>((b.n_tiles_total + kWaves - 1) / kWaves);
  hipLaunchKernelGGL((k_query_order<D, T>), dim3(grid), dim3(kBlock), 0, L.s, *L.rp, b.events,
                     b.items, b.icount, b.geo, b.toff, b.n_tiles_total, b.sorted.x[0], L.cs,
                     b.sorted.sigma[0], b.qkey, b.qfine, b.qperm, b.qperm_nh, b.hx, b.h_stride,
                     b.hidx, b.hbox);
}

template <int D, class T>
static void launch_sorted_boxes_d(const NeighbourLaunch<T>& L) {
  const BatchBuffersT<T>& b = *L.bb;
  // Float: four positions per thread; double: one.
  const int64_t per = sizeof(T) == 4 ? 4 : 1;
  const int64_t n16 = (static_cast<int64_t>(b.n_total) + kSortedBox - 1) / kSortedBox * kSortedBox;
  const int64_t threads = n16 / per;
  hipLaunchKernelGGL((k_sorted_boxes<D, T>), dim3(static_cast<unsigned>((threads + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, b.sorted.x[0], L.cs, b.n_total, b.sbox);
}

template <class T>
static void launch_sorted_boxes(const NeighbourLaunch<T>& L) {
  switch (L.rp->D) {
    case 1: launch_sorted_boxes_d<1>(L); break;
    case 2: launch_sorted_boxes_d<2>(L); break;
    case 3: launch_sorted_boxes_d<3>(L); break;
    case 4: launch_sorted_boxes_d<4>(L); break;
    case 5: launch_sorted_boxes_d<5>(L); break;
    case 6: launch_sorted_boxes_d<6>(L); break;
    case 7: launch_sorted_boxes_d<7>(L); break;
    case 8: launch_sorted_boxes_d<8>(L); break;
    case 9: launch_sorted_boxes_d<9>(L); break;
    default: launch_sorted_boxes_d<10>(L); break;
  }
}

// The query orders of both neighbour stages (qperm for dc, qperm_nh for dm).
template <class T>
static void launch_query_order(const NeighbourLaunch<T>& L) {
  switch (L.rp->D) {
    case 1: launch_query_order_d<1>(L); break;
    case 2: launch_query_order_d<2>(L); break;
    case 3: launch_query_order_d<3>(L); break;
    case 4: launch_query_order_d<4>(L); break;
    case 5: launch_query_order_d<5>(L); break;
    case 6: launch_query_order_d<6>(L); break;
    case 7: launch_query_order_d<7>(L); break;
    case 8: launch_query_order_d<8>(L); break;
    case 9: launch_query_order_d<9>(L); break;
    default: launch_query_order_d<10>(L); break;
  }
}

template <class T>
static void launch_neighbour(const NeighbourLaunch<T>& L, bool density) {
  switch (L.rp->D) {
    case 1: launch_neighbour_d<1>(L, density); break;
    case 2: launch_neighbour_d<2>(L, density); break;
    case 3: launch_neighbour_d<3>(L, density); break;
    case 4: launch_neighbour_d<4>(L, density); break;
    case 5: launch_neighbour_d<5>(L, density); break;
    case 6: launch_neighbour_d<6>(L, density); break;
    case 7: launch_neighbour_d<7>(L, density); break;
    case 8: launch_neighbour_d<8>(L, density); break;
    case 9: launch_neighbour_d<9>(L, density); break;
    default: launch_neighbour_d<10>(L, density); break;
  }
}


template <class T>
void launch_pipeline(const RunParamsT<T>& rp, const PointArraysT<T>& in, const BatchBuffersT<T>& bb,
                     const ResultArraysT<T>& out, hipStream_t s, const StageHook* hook) {
  using I = typename FloatBits<T>::Int;
  auto mark = [&](int stage) {
    if (hook && hook->mark) hook->mark(hook->self, stage, s);
  };
  const int D = rp.D;
  const int n = bb.n_total;
  CoordPtrs<T> cp{};
  for (int d = 0; d < D; ++d) cp.x[d] = in.x[d];

  // (a) tiles
  mark(kStageExtremes);
  if (bb.event_order) {
    // Every event fits one workgroup's tile counters: the fused per-event build.
    mark(kStageTiles);
    launch_event_tiles(rp, in, bb, s);
  } else {
  hipLaunchKernelGGL(k_init_extremes<I>,
                     dim3(static_cast<unsigned>((static_cast<int64_t>(bb.n_events) * D + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, bb.ext, static_cast<int64_t>(bb.n_events) * D);
  hipLaunchKernelGGL(k_extremes<T>, dim3(bb.n_chunks), dim3(kBlock), 0, s, cp, D, bb.chunks, bb.ext);
  hipLaunchKernelGGL(k_geometry<T>, dim3((bb.n_events + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     bb.ext, bb.events, bb.n_events, D, bb.geo);
  mark(kStageTiles);
  hipLaunchKernelGGL(k_tile_keys<T>, dim3(bb.n_chunks), dim3(kBlock), 0, s, cp, rp, bb.chunks,
                     bb.events, bb.geo, bb.keys0, bb.vals0);
  bool in_first = true;
  radix_sort_pairs(bb.keys0, bb.vals0, bb.keys1, bb.vals1, n, bb.key_bits, bb.rs_hist, bb.scan_tmp,
                   &in_first, s);
  const uint32_t* skeys = in_first ? bb.keys0 : bb.keys1;
  const int32_t* sperm = in_first ? bb.vals0 : bb.vals1;
  GatherArgs<T> ga;
  ga.in = in;
  ga.out = bb.sorted;
  ga.cs = 0;
  hipLaunchKernelGGL(k_gather<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, ga, rp, skeys,
                     sperm, n, bb.n_tiles_total, bb.toff);
  }
  if (bb.n_tiles_total <= kItemsOneBlockTiles) {
    hipLaunchKernelGGL(k_items_one_block, dim3(1), dim3(kBlock), 0, s, bb.toff, bb.n_tiles_total,
                       bb.events, bb.n_events, bb.icount, bb.items, bb.n_items);
  } else {
    hipLaunchKernelGGL(k_item_count, dim3((bb.n_tiles_total + kBlock - 1) / kBlock), dim3(kBlock),
                       0, s, bb.toff, bb.n_tiles_total, bb.icount);
    scan_exclusive(bb.icount, bb.icount, bb.n_tiles_total, bb.scan_tmp, bb.n_items, s);
    hipLaunchKernelGGL(k_item_fill,
                       dim3(static_cast<unsigned>(
                           (std::max<int64_t>(bb.n_tiles_total, kQueueWords) + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0,
                       s, bb.toff, bb.icount, bb.n_tiles_total, bb.events, bb.n_events, bb.items,
                       bb.n_items);
  }

  NeighbourLaunch<T> L;
  L.rp = &rp;
  L.bb = &bb;
  L.cs = bb.sorted_stride;
  const int waves = rp.block_waves;
  L.block = static_cast<unsigned>(waves * kWave);
  const int64_t blocks = (bb.max_items + waves - 1) / waves;
  L.grid = static_cast<unsigned>((blocks + kXcds - 1) / kXcds * kXcds);
  // Persistent waves: rp.persistent per CU, never more than the grid form's.
  const int64_t pblocks = std::min<int64_t>(
      blocks, (static_cast<int64_t>(device_cus()) * std::max(rp.persistent, 0) + waves - 1) / waves);
  L.grid_persistent = static_cast<unsigned>((pblocks + kXcds - 1) / kXcds * kXcds);
  L.s = s;
  if (bb.sbox) launch_sorted_boxes(L);  // candidate boxes of the density stage
  if (bb.qperm) launch_query_order(L);  // query orders of the two neighbour stages

  // (b), (c)
  mark(kStageDensity);
  launch_neighbour(L, true);
  mark(kStageNearestHigher);
  launch_neighbour(L, false);

  // (d), (e)
  mark(kStageSeeds);
  if (bb.event_order) {
    const bool lone = bb.n_events == 1;
    hipLaunchKernelGGL(k_event_seeds<T>, dim3(static_cast<unsigned>(bb.n_events)), dim3(kEvtBlock),
                       0, s, rp, bb.events, bb.event_order, bb.sorted.perm, bb.iperm, bb.srho, bb.snh,
                       bb.sdelta, bb.sorted.u, out.rho, out.nh, out.delta, out.is_seed, bb.rank,
                       bb.event_seeds, lone ? bb.event_offsets : nullptr, lone ? bb.n_seeds : nullptr,
                       out.cluster);
    if (!lone) scan_exclusive(bb.event_seeds, bb.event_offsets, bb.n_events, bb.scan_tmp, bb.n_seeds, s);
    mark(kStageLabels);
  } else {
    hipLaunchKernelGGL(k_unpermute<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rp, n,
                       bb.sorted.perm, bb.srho, bb.snh, bb.sdelta, bb.sorted.u, out.rho, out.nh,
                       out.delta, out.is_seed, bb.flag);
    scan_exclusive(bb.flag, bb.rank, n, bb.scan_tmp, bb.n_seeds, s);
    mark(kStageLabels);
    for (int pass = 0; pass < kLinkJumpPasses; ++pass)
      hipLaunchKernelGGL(k_link_jump, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, bb.snh, n);
  }
  hipLaunchKernelGGL(k_labels, dim3(bb.n_chunks), dim3(kBlock), 0, s, bb.chunks, bb.events,
                     bb.sorted.perm, bb.snh, bb.rank,
                     bb.event_order ? bb.event_offsets : nullptr, bb.event_seeds,
                     out.cluster, bb.seeds, out.event_of_cluster, out.local_ids);
  mark(kNumStages);
}

}  // namespace clue_dev
