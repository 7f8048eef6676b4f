// Device-resident inputs and outputs of ofg_price_device: launch interface of
// the staging and scatter kernels (ofg_device_io.hip).
#pragma once

#include "../../include/openferric_amd.h"
#include "ofg_adjoint.hpp"
#include "ofg_internal.hpp"

namespace ofg {

// Error codes the staging kernel reports (the first offending option's).
enum DeviceInputError : int32_t {
    DIE_NONE = 0,
    DIE_STRIKE = 1,      // strike must be > 0
    DIE_EXPIRY = 2,      // maturity must be > 0 (device maturities)
    DIE_VOL = 3,         // market volatility must be > 0
    DIE_HESTON = 4,      // invalid Heston parameters
    DIE_FAST_RANGE = 5,  // fast mode: per-step log-return beyond the exp table
    DIE_NOT_FINITE = 6   // an input is NaN or infinite
};

struct StageArgs {
    OptDev *opts;              // the plan's descriptors, updated in place
    const int32_t *active;     // plan index -> caller index
    int32_t n_plan;
    int32_t rng_mode;
    const double *host;        // per plan option: dividend yield, maturity (host values)
    int32_t *err;              // [code, caller index], zeroed before the launch
};

// Recomputes every market- and model-dependent field of the plan's
// descriptors from the caller's device arrays (io, by value: its pointers
// are device addresses), as build_plan forms them on the host, and
// validates them.
hipError_t launch_stage(const StageArgs &a, const ofg_device_io &io, hipStream_t stream);

// Writes price, stderr and (with grads) the eight gradient slots of each
// plan option to the caller's arrays at its caller index, unless the staging
// kernel reported an error (err[0] != 0): then nothing is written.
hipError_t launch_scatter_price(const OptOut *outs, const int32_t *active, int32_t n_plan,
                                const ofg_device_io &io, const int32_t *err, hipStream_t stream);
hipError_t launch_scatter_adjoint(const AdjOut *outs, const int32_t *active, int32_t n_plan,
                                  const ofg_device_io &io, const int32_t *err, hipStream_t stream);

}  // namespace ofg
