// Device association maps (key -> members CSR); kernels in clue_assoc.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clue_internal.h"

namespace clue_dev {

// Scratch of one build: two key/value ping-pong pairs of n entries, the radix
// histogram and the scan temporary (sizes from assoc_scratch_elems).
struct AssocScratch {
  uint32_t* k0;
  int32_t* v0;
  uint32_t* k1;
  int32_t* v1;
  int32_t* hist;
  int32_t* scan;
};

void assoc_scratch_elems(int64_t n, int64_t* hist_elems, int64_t* scan_elems);
// stats_dev[0] = max key (INT32_MIN when n == 0), stats_dev[1] = number of keys >= 0.
hipError_t assoc_stats(const int32_t* keys, int n, int32_t* stats_dev, hipStream_t s);
// offsets: bins + 1 entries; members: n_members entries (the keys >= 0).
hipError_t assoc_build(const int32_t* keys, int n, int bins, const AssocScratch& w,
                       int32_t* offsets, int32_t* members, int n_members, hipStream_t s);

}  // namespace clue_dev
