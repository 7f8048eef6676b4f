// Launch interface of the Monte Carlo Greeks estimator kernels
// (ofg_mcgreeks.hip), shared with the host side (ofg_api.cpp). The launch
// arguments, per-item slots and per-option outputs are the adjoint path's
// (AdjArgs, AdjPartial, AdjOut): g[0..4] hold the sums of pathwise delta,
// pathwise gamma, LR delta, LR gamma and LR vega, n the sample count.
#pragma once

#include "ofg_adjoint.hpp"

#include <hip/hip_runtime.h>

namespace ofg {

hipError_t launch_mc_greeks(const AdjArgs &a, hipStream_t stream);
hipError_t launch_mc_greeks_reduce(const int64_t *item_first, const int32_t *item_count,
                                   const AdjPartial *partials, AdjOut *out, int32_t n_opts,
                                   hipStream_t stream);
int mc_greeks_grid(int block, int rng_mode, int antithetic, int device);

}  // namespace ofg
