// Exact-terminal European pricers on the reference's single streams:
// device layouts shared by ofg_exact.hip and ofg_exact_api.cpp.
#pragma once

#include "ofg_lsm.hpp"

namespace ofg {

// One xoshiro256++ stream of the reference (a whole pricer, or one chunk of
// mc_european_parallel): thread g of the job jumps g * 64 draws from the
// seeded state (LsmSeedJob: base, paths, polynomials) and prices the next 64
// paths, one draw each. thread_first is a multiple of 64, so each wave
// belongs to one job.
struct ExactJob {
    LsmSeedJob seed;
    double spot, strike, drift, diffusion;
    int64_t pay_offset;   // first path's slot in the optional payoff output
    int64_t wave_first;   // first wave of the job (thread_first / 64)
    int32_t wave_count;
    int32_t put;
};

struct ExactSums {
    double sum, sum_sq;
};

// Prices every job's paths; wave_sums receives one (sum, sum_sq) per wave,
// job_sums one per job (its waves summed in order). payoffs may be null.
hipError_t launch_exact(const ExactJob *jobs, int32_t n_jobs, int64_t total_threads,
                        const uint64_t *polys, ExactSums *wave_sums, ExactSums *job_sums,
                        double *payoffs, hipStream_t stream);

}  // namespace ofg
