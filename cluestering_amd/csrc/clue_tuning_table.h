// Built-in launch tuning, by ndim (0 = any) and mean event size (first match).
// Regenerated from MI355X measurements by cluestering_amd/autotune.py.
#pragma once

#include <cstdint>

namespace clue_rt {

struct TuningEntry {
  int32_t ndim;
  int64_t max_mean_event;
  int32_t group_queries;
  int32_t xcd_chunk_log2;
  int32_t block_waves;
  int32_t persistent;  // persistent waves per CU (0: hardware dispatch)
};

inline constexpr TuningEntry kTuningTable[] = {
    {2, 2048, 1, 6, 1, 0},
    {2, 16384, 1, 4, 1, 0},
    {2, 131072, 1, 6, 1, 0},
    {2, INT64_MAX, 1, 4, 2, 0},
    {3, 2048, 1, 8, 1, 0},
    {3, 16384, 1, 8, 1, 0},
    {3, 131072, 1, 8, 1, 0},
    {3, INT64_MAX, 1, 8, 2, 0},
    {4, 2048, 1, 8, 1, 0},
    {4, 16384, 1, 4, 1, 0},
    {4, INT64_MAX, 1, 6, 1, 0},
    {6, 2048, 0, 6, 1, 0},
    {6, 16384, 0, 8, 1, 0},
    {6, INT64_MAX, 0, 6, 1, 0},
    {0, INT64_MAX, 0, 6, 1, 0},  // other ndim
};

}  // namespace clue_rt
