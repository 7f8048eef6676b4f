// Launch interface of the forward pathwise Greeks kernel for GBM
// (ofg_pathwise.hip), shared with the host side (ofg_api.cpp). Arguments,
// per-item slots and outputs are the adjoint path's (AdjArgs, AdjPartial,
// AdjOut) with gradient slots 0 S0, 1 sigma, 2 r, 3 T.
#pragma once

#include "ofg_adjoint.hpp"

#include <hip/hip_runtime.h>

namespace ofg {

hipError_t launch_pathwise(const AdjArgs &a, hipStream_t stream);
int pathwise_grid(int block, int rng_mode, int antithetic, int device);

}  // namespace ofg
