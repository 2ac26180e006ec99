// The pipeline stages for double points (Clusterer<Ndim, double>).
#include "clue_stages_impl.h"

namespace clue_dev {

template void launch_pipeline<double>(const RunParamsT<double>&, const PointArraysT<double>&,
                                      const BatchBuffersT<double>&, const ResultArraysT<double>&,
                                      hipStream_t, const StageHook*);

}  // namespace clue_dev
