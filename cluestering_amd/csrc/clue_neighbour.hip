// The neighbour stages (density and nearest-higher, every metric and kernel) of
// one dimension, CLUE_NEIGHBOUR_D, in both precisions. The Makefile compiles this
// file once per dimension so that the heaviest instantiations build in parallel.
#include "clue_stages_impl.h"

#ifndef CLUE_NEIGHBOUR_D
#error "CLUE_NEIGHBOUR_D (1..10) must be defined"
#endif

namespace clue_dev {

template void launch_neighbour_d<CLUE_NEIGHBOUR_D, float>(const NeighbourLaunch<float>&, bool);
template void launch_neighbour_d<CLUE_NEIGHBOUR_D, double>(const NeighbourLaunch<double>&, bool);

}  // namespace clue_dev
