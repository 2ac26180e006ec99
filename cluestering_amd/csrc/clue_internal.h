// Host-side declarations of the device stages and primitives.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clue_types.h"

namespace clue_dev {

// ---- primitives (clue_primitives.hip) ----
// Exclusive scan of n int32 (in may equal out). temp: scan_temp_elems(n) ints.
// total (device, may be NULL) receives the sum.
void scan_exclusive(const int32_t* in, int32_t* out, int64_t n, int32_t* temp, int32_t* total,
                    hipStream_t s);
void radix_sort_pairs(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int64_t n, int key_bits,
                      int32_t* hist, int32_t* scan_temp, bool* in_first, hipStream_t s);

// ---- stages (clue_stages_impl.h, instantiated in clue_stages_f32.hip / _f64.hip) ----

// Integer types of the same width as the point type: ordered encodings of
// coordinates (extremes) and of densities (nearest-higher keys).
template <class T>
struct FloatBits;
template <>
struct FloatBits<float> {
  using Int = int32_t;
  using UInt = uint32_t;
};
template <>
struct FloatBits<double> {
  using Int = long long;  // the 64-bit type of the HIP atomics
  using UInt = unsigned long long;
};

// Input columns and the sorted copies, all device pointers.
template <class T>
struct PointArraysT {
  const T* x[kMaxDim];
  const T* w;
  const T* sigma[kMaxDim];
  const T* u;
  const uint32_t* tags;
};
using PointArrays = PointArraysT<float>;

template <class T>
struct SortedArraysT {
  T* x[kMaxDim];
  T* w;
  T* fw;          // flat kernel only: flat value times weight
  T* sigma[kMaxDim];
  T* u;
  uint32_t* tag;  // user tags, or the original index when no tags are given
  int32_t* perm;  // sorted position -> original index
};

// Everything one batch run needs on the device.
template <class T>
struct BatchBuffersT {
  int32_t n_events;
  int32_t n_total;
  int32_t n_chunks;
  int32_t n_tiles_total;  // sum of ntiles over events
  int32_t max_items;      // upper bound of work items
  int32_t key_bits;
  const EventDesc* events;
  const int32_t* event_order;  // events largest first for the fused tile build; NULL: radix path
  const Chunk* chunks;
  typename FloatBits<T>::Int* ext;  // [E][D][2] ordered-integer extremes
  T* geo;                           // [E][3][D] mn, mx, side
  uint32_t* keys0;
  int32_t* vals0;
  uint32_t* keys1;
  int32_t* vals1;
  int32_t* rs_hist;
  int32_t* scan_tmp;
  int32_t* toff;       // [n_tiles_total + 1]
  int32_t* icount;     // [n_tiles_total]
  int32_t* n_items;    // [1]
  Item* items;         // [max_items]
  int32_t* qperm;      // [n_total] density query order within tiles (qfine with candidate
                       // boxes); NULL = sorted order
  int32_t* qperm_nh;   // [n_total] nearest-higher query order (may alias qperm)
  int32_t* qfine;      // [n_total] Morton order within tiles (input of both)
  uint32_t* qkey;      // [n_total] packed query keys (Morton, class for dc, class for dm)
  T* srho;
  uint64_t* skey;      // ordering key of every sorted point: float (rho, tag); double rho
                       // (the plain nearest-higher scan; NULL with the block layout)
  int32_t* snh;
  T* sdelta;
  int32_t* flag;       // [n_total]
  int32_t* event_seeds;    // [n_events] seeds per event (fused per-event seed stage)
  int32_t* event_offsets;  // [n_events] exclusive scan of event_seeds
  int32_t* iperm;          // [n_total] sorted position of every input point (fused per-event path)
  int32_t* rank;       // [n_total]
  int32_t* n_seeds;    // [1]
  int32_t* seeds;      // [n_total]
  SortedArraysT<T> sorted;
  int64_t sorted_stride;  // column stride of sorted.x[] and sorted.sigma[] blocks
  // Nearest-higher block layout (NULL when not used): each tile's candidates in
  // Morton order from block ceil(toff[t] / kNhBlock) + t on, padded to whole
  // blocks with key 0 (never higher) and index -1.
  T* hx;             // [D][h_stride] coordinates
  uint64_t* hkey;    // [h_stride] ordering keys
  int32_t* hidx;     // [h_stride] sorted position of the candidate
  T* hbox;           // [h_blocks][2][D] bounding box (lower corner, upper corner)
  uint64_t* hbkey;   // [h_blocks] largest key of the block
  int64_t h_stride;  // slots = h_blocks * kNhBlock
  // Bounding box of each kSortedBox consecutive tile-sorted positions (NULL when
  // the metric has no box bound): [ceil(n / kSortedBox)][2][D].
  T* sbox;
};
using BatchBuffers = BatchBuffersT<float>;

template <class T>
struct ResultArraysT {
  int32_t* cluster;
  int32_t* is_seed;
  T* rho;
  int32_t* nh;
  T* delta;
  int32_t* event_of_cluster;  // may be NULL
  int32_t local_ids;          // number clusters per event from 0 (served events)
};
using ResultArrays = ResultArraysT<float>;

// Stage boundaries reported by the profiler.
enum Stage {
  kStageExtremes = 0,
  kStageTiles,
  kStageDensity,
  kStageNearestHigher,
  kStageSeeds,
  kStageLabels,
  kNumStages
};
const char* stage_name(int s);

// Timing hook called between stages (may be a no-op).
struct StageHook {
  void (*mark)(void* self, int stage, hipStream_t s);
  void* self;
};

// Launches the whole pipeline for one batch on stream s. No host synchronisation,
// no allocation, no memset: safe to capture in a hipGraph.
// Instantiated for float and double.
template <class T>
void launch_pipeline(const RunParamsT<T>& rp, const PointArraysT<T>& in, const BatchBuffersT<T>& bb,
                     const ResultArraysT<T>& out, hipStream_t s, const StageHook* hook);

}  // namespace clue_dev
