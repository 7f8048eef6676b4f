// Two-asset spread options under correlated GBM: device layouts shared by
// ofg_spread.hip and ofg_spread_api.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ofg {

// Samples per thread of the spread kernel: a wave covers 64 * SPREAD_PER
// consecutive samples of one option.
constexpr int SPREAD_PER = 4;

// One option of ofg_spread_price with the per-option constants of
// SpreadMonteCarloEngine::price (spread_mc.rs:106-115).
struct SpreadJob {
    double s1, s2, k, rho, corr_tail, drift1, drift2, vol_term1, vol_term2;
    uint64_t seed;
    int64_t samples;      // antithetic ? ceil(paths / 2) : paths
    int64_t pay_offset;   // first sample's slot in the optional payoff output
    int64_t wave_first;   // first wave of the option
    int32_t wave_count;
    int32_t pad;
};

struct SpreadSums {
    double sum, sum_sq;
};

// Simulates every job's samples; wave_sums receives one (sum, sum_sq) per
// wave, job_sums one per job (its waves summed in wave order). rng_mode is
// OFG_RNG_REFERENCE, OFG_RNG_PCG64 or OFG_RNG_FAST; payoffs may be null.
hipError_t launch_spread(const SpreadJob *jobs, int32_t n_jobs, int64_t total_waves,
                         int rng_mode, bool antithetic, SpreadSums *wave_sums,
                         SpreadSums *job_sums, double *payoffs, hipStream_t stream);

}  // namespace ofg
