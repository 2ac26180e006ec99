// The pipeline stages for float points (Clusterer<Ndim, float>).
#include "clue_stages_impl.h"

namespace clue_dev {

const char* stage_name(int s) {
  static const char* names[kNumStages] = {"extremes", "tiles", "density", "nearest_higher",
                                          "seeds", "labels"};
  return (s >= 0 && s < kNumStages) ? names[s] : "?";
}

template void launch_pipeline<float>(const RunParamsT<float>&, const PointArraysT<float>&,
                                     const BatchBuffersT<float>&, const ResultArraysT<float>&,
                                     hipStream_t, const StageHook*);

}  // namespace clue_dev
