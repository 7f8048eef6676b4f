// Launch interface of the forward pathwise Greeks kernel (ofg_pathwise.hip):
// GBM by running sums of the path, Heston by a forward tangent of the Euler
// step. Shared with the host side (ofg_api.cpp). Arguments, per-item slots
// and outputs are the adjoint path's (AdjArgs, AdjPartial, AdjOut) with its
// gradient slots (0 S0, 1 sigma | v0, 2 r, 3 T, 4 kappa, 5 theta, 6 xi, 7 rho).
#pragma once

#include "ofg_adjoint.hpp"

#include <hip/hip_runtime.h>

namespace ofg {

hipError_t launch_pathwise(const AdjArgs &a, hipStream_t stream);
int pathwise_grid(int block, int rng_mode, int model, int antithetic, int device);

}  // namespace ofg
