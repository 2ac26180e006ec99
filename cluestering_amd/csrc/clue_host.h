// Host-only sizing shared by the C ABI, the server and the CPU tests: the batch
// descriptor built from the event sizes, the launch tuning it selects, and the
// device bytes each workspace buffer needs. No HIP types here, so the same code
// builds with g++ under the address and undefined-behaviour sanitizers.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "clue_amd.h"
#include "clue_tuning_table.h"
#include "clue_types.h"

namespace clue_rt {

using clue_dev::Chunk;
using clue_dev::EventDesc;
using clue_dev::Item;
using clue_dev::radix_hist_elems;
using clue_dev::scan_temp_elems;

constexpr int kChunk = 2048;  // points per chunk of the per-point grids
constexpr int kEvtSoloMaxPoints = 8192;  // largest lone event for the fused per-event stages
constexpr int kDefaultPointsPerTile = 128;

// Constants of exp(y * 2^-s) (clue_dev::ExpScaled): each constant of the float
// exp scaled by a power of two, exact for |s| <= 8.
inline clue_dev::ExpScaled exp_scaled(int s) {
  clue_dev::ExpScaled e{};
  e.log2e = std::ldexp(clue_dev::kExpLog2e, -s);
  e.ln2_hi = std::ldexp(clue_dev::kExpLn2Hi, s);
  e.ln2_lo = std::ldexp(clue_dev::kExpLn2Lo, s);
  for (int i = 0; i < 5; ++i) e.c[i] = std::ldexp(clue_dev::kExpPoly[i], -s * (6 - i));
  e.c[5] = std::ldexp(1.0f, -s);
  e.lo = std::ldexp(clue_dev::kExpCoreLo, s);
  e.hi = std::ldexp(clue_dev::kExpCoreHi, s);
  return e;
}

// Capacity a grow-only device buffer holds after a request of `bytes`: unchanged
// when it already fits, else the request plus a quarter and 256 bytes.
inline size_t grown_capacity(size_t cap, size_t bytes) {
  return bytes <= cap ? cap : bytes + bytes / 4 + 256;
}

// Workspace slots of one batch run.
enum Slot {
  S_EVENTS, S_CHUNKS, S_EXT, S_GEO, S_K0, S_V0, S_K1, S_V1, S_HIST, S_SCAN, S_TOFF, S_ICOUNT,
  S_NITEMS, S_ITEMS, S_SRHO, S_SKEY, S_SNH, S_SDELTA, S_FLAG, S_RANK, S_NSEEDS, S_SEEDS, S_SX,
  S_SW, S_SSIG, S_SU, S_STAG, S_SFW, S_PERM, S_RHO, S_NH, S_ISSEED, S_DELTA, S_HIN, S_HCL,
  S_QPERM, S_QPERM_NH, S_QFINE, S_QKEY, S_HX, S_HKEY, S_HIDX, S_HBOX, S_HBKEY, S_EORDER, S_ESEEDS, S_EOFF, S_IPERM, S_SBOX, S_NUM
};

// Host-side description of one batch (all derived from the event sizes).
struct BatchPlan {
  std::vector<EventDesc> events;
  std::vector<int32_t> event_order;  // fused per-event tile build: events largest first; empty: radix path
  std::vector<Chunk> chunks;
  int64_t n_total = 0;
  int64_t n_tiles_total = 0;
  int64_t max_items = 0;
  int key_bits = 0;
  clue_tuning tuning{};
};

// Tiles per dimension of an event of n points (SetupTiles.hpp:30-34):
// ntiles = ceil(n / ppt), nper = min{k : k^D >= ntiles}; tiles = nper^D.
inline void event_tiles(int64_t n, int D, int ppt, int64_t* nper_out, int64_t* tiles_out) {
  const int64_t nt0 = (n + ppt - 1) / ppt;
  auto pw = [D](int64_t b) {
    int64_t r = 1;
    for (int k = 0; k < D; ++k) r *= b;
    return r;
  };
  int64_t nper = 1;
  while (pw(nper) < nt0) ++nper;
  *nper_out = nper;
  *tiles_out = pw(nper);
}

// Tuning of a batch: the built-in table's entry for its ndim and mean event size.
inline clue_tuning table_tuning(int ndim, int64_t n_total, int64_t n_events) {
  const int64_t mean = n_events > 0 ? n_total / n_events : n_total;
  clue_tuning out{};
  for (const TuningEntry& t : kTuningTable)
    if ((t.ndim == 0 || t.ndim == ndim) && mean <= t.max_mean_event) {
      out.group_queries = t.group_queries;
      out.xcd_chunk_log2 = t.xcd_chunk_log2;
      out.block_waves = t.block_waves;
      out.persistent = t.persistent;
      return out;
    }
  out.group_queries = 1;
  out.xcd_chunk_log2 = 4;
  out.block_waves = 4;
  return out;
}

// Builds the batch descriptor: one EventDesc per non-empty event, its chunks,
// tile bases, work-item bound and radix key width. `fixed` (may be NULL) replaces
// the table's tuning. Returns CLUE_OK, or CLUE_ERR_CAPACITY with *err set when the
// batch exceeds 32-bit indexing.
inline clue_status plan_batch(int D, int points_per_tile, const uint32_t* sizes, int32_t n_events,
                              const clue_tuning* fixed, BatchPlan* bp, std::string* err) {
  const int ppt = points_per_tile > 0 ? points_per_tile : kDefaultPointsPerTile;
  bp->events.clear();
  bp->chunks.clear();
  int64_t off = 0, tiles = 0;
  for (int32_t e = 0; e < n_events; ++e) {
    const int64_t n = sizes[e];
    if (n == 0) continue;
    int64_t nper = 0, nt = 0;
    event_tiles(n, D, ppt, &nper, &nt);
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
    if (off > INT32_MAX / 2 || tiles > INT32_MAX / 2) {
      *err = "batch too large for 32-bit indexing";
      return CLUE_ERR_CAPACITY;
    }
  }
  bp->n_total = off;
  bp->n_tiles_total = tiles;
  // The fused per-event stages when every event's tiles and points fit their LDS,
  // except for one event alone above kEvtSoloMaxPoints: there the many-workgroup
  // stages win over one workgroup's serial passes (per-event calls on MI355X over
  // 512 events of 256-65,536 points: 3.3k -> 3.75k events/s, p99 1.19 -> 1.00 ms,
  // median 0.18 ms either way; a 16k-point event 0.31 -> 0.29 ms).
  bool local = !bp->events.empty() &&
               (bp->events.size() >= 2 || bp->events[0].n <= kEvtSoloMaxPoints);
  for (const EventDesc& ev : bp->events)
    local = local && ev.ntiles <= clue_dev::kEvtMaxTiles && ev.n <= clue_dev::kEvtMaxPoints;
  bp->event_order.clear();
  if (local) {
    for (int32_t k = 0; k < static_cast<int32_t>(bp->events.size()); ++k) bp->event_order.push_back(k);
    if (bp->event_order.size() > 1)
      std::stable_sort(bp->event_order.begin(), bp->event_order.end(),
                       [&](int32_t a, int32_t b) { return bp->events[a].n > bp->events[b].n; });
  }
  bp->tuning = fixed ? *fixed : table_tuning(D, off, static_cast<int64_t>(bp->events.size()));
  bp->max_items = (off + 63) / 64 + tiles;
  int bits = 0;
  while ((int64_t(1) << bits) < tiles) ++bits;
  bp->key_bits = bits;
  return CLUE_OK;
}

// What the workspace of one batch depends on.
struct WorkspaceShape {
  int D = 1;
  int value_bytes = 4;      // sizeof the point type
  int key_bytes = 4;        // sizeof the ordered-integer extremes (the point type's width)
  int64_t n = 0;            // points
  int64_t events = 0;       // non-empty events
  int64_t chunks = 0;
  int64_t tiles = 0;        // sum of nper^D
  int64_t max_items = 0;
  bool event_tiles = false;        // the fused per-event tile build (no radix-sort buffers)
  bool group_queries = false;
  bool separate_dm_order = false;  // dm != dc and the plain nearest-higher scan
  bool nh_blocks = false;          // nearest-higher over the Morton block layout
  bool flat_kernel = false;
  bool has_sigma = false;
  bool has_u = false;
  bool own_rho = false;            // results the caller does not provide live in the workspace
  bool own_nh = false;
  bool own_is_seed = false;
  bool sorted_boxes = false;       // density culls candidate blocks by their bounding boxes
  bool density_morton = false;     // density takes the Morton order (no class order for dc)
};

// Bytes of every workspace slot for a batch of this shape (0: unused).
inline void workspace_requests(const WorkspaceShape& w, size_t (&bytes)[S_NUM]) {
  std::memset(bytes, 0, sizeof(bytes));
  const int64_t n = w.n, NT = w.tiles;
  const int64_t hist = radix_hist_elems(n);
  const int64_t scan_elems =
      std::max({scan_temp_elems(n), scan_temp_elems(hist), scan_temp_elems(NT)}) + 16;
  const int64_t stride = (n + 63) / 64 * 64;
  const int64_t h_blocks = w.nh_blocks ? (n + clue_dev::kNhBlock - 1) / clue_dev::kNhBlock + NT + 1 : 0;
  const int64_t h_slots = h_blocks * clue_dev::kNhBlock;
  const size_t T = static_cast<size_t>(w.value_bytes);
  const size_t D = static_cast<size_t>(w.D);
  bytes[S_EVENTS] = sizeof(EventDesc) * w.events;
  bytes[S_CHUNKS] = sizeof(Chunk) * w.chunks;
  bytes[S_EXT] = w.event_tiles ? 0 : static_cast<size_t>(w.key_bytes) * w.events * D * 2;
  bytes[S_GEO] = T * w.events * 3 * D;
  bytes[S_K0] = bytes[S_K1] = w.event_tiles ? 0 : sizeof(uint32_t) * n;
  bytes[S_V0] = bytes[S_V1] = w.event_tiles ? 0 : sizeof(int32_t) * n;
  bytes[S_HIST] = w.event_tiles ? 0 : sizeof(int32_t) * hist;
  bytes[S_EORDER] = w.event_tiles ? sizeof(int32_t) * w.events : 0;
  bytes[S_ESEEDS] = bytes[S_EOFF] = w.event_tiles ? sizeof(int32_t) * w.events : 0;
  bytes[S_IPERM] = w.event_tiles ? sizeof(int32_t) * n : 0;
  bytes[S_SCAN] = sizeof(int32_t) * scan_elems;
  bytes[S_TOFF] = sizeof(int32_t) * (NT + 1);
  bytes[S_ICOUNT] = sizeof(int32_t) * NT;
  bytes[S_NITEMS] = sizeof(int32_t) * 32;  // item count, then the work-item queue counters
  bytes[S_ITEMS] = sizeof(Item) * w.max_items;
  bytes[S_QPERM] = (w.group_queries && !w.density_morton) ? sizeof(int32_t) * n : 0;
  bytes[S_QPERM_NH] = (w.group_queries && w.separate_dm_order) ? sizeof(int32_t) * n : 0;
  bytes[S_QFINE] = w.group_queries ? sizeof(int32_t) * n : 0;
  bytes[S_QKEY] = w.group_queries ? sizeof(uint32_t) * n : 0;
  bytes[S_HX] = T * h_slots * D;
  bytes[S_HKEY] = sizeof(uint64_t) * h_slots;
  bytes[S_HIDX] = sizeof(int32_t) * h_slots;
  bytes[S_HBOX] = T * h_blocks * 2 * D;
  bytes[S_HBKEY] = sizeof(uint64_t) * h_blocks;
  bytes[S_SBOX] = w.sorted_boxes ? T * ((n + clue_dev::kSortedBox - 1) / clue_dev::kSortedBox) * 2 * D : 0;
  bytes[S_SRHO] = T * n;
  bytes[S_SKEY] = w.nh_blocks ? 0 : sizeof(uint64_t) * n;  // the block scan's keys: S_HKEY
  bytes[S_SFW] = w.flat_kernel ? T * n : 0;
  bytes[S_SNH] = sizeof(int32_t) * n;
  bytes[S_SDELTA] = T * n;
  bytes[S_FLAG] = w.event_tiles ? 0 : sizeof(int32_t) * n;
  bytes[S_RANK] = sizeof(int32_t) * n;
  bytes[S_NSEEDS] = sizeof(int32_t) * 4;
  bytes[S_SEEDS] = sizeof(int32_t) * n;
  bytes[S_SX] = T * stride * D;
  bytes[S_SW] = T * n;
  bytes[S_SSIG] = w.has_sigma ? T * stride * D : 0;
  bytes[S_SU] = w.has_u ? T * n : 0;
  bytes[S_STAG] = sizeof(uint32_t) * n;
  bytes[S_PERM] = sizeof(int32_t) * n;
  bytes[S_RHO] = w.own_rho ? T * n : 0;
  bytes[S_NH] = w.own_nh ? sizeof(int32_t) * n : 0;
  bytes[S_ISSEED] = w.own_is_seed ? sizeof(int32_t) * n : 0;
}

}  // namespace clue_rt
