// Internal layouts shared by the HIP kernels and the host side of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/openferric_amd.h"

namespace ofg {

// Path-dependence class of a payoff: decides which running state a path
// carries in registers.
enum PayoffKind : int32_t {
    PK_VANILLA = 0,    // terminal spot only
    PK_ASIAN = 1,      // weighted running sum of observed spots
    PK_ASIAN_LOG = 2,  // + running sum of observed log-spots (geometric)
    PK_BARRIER = 3,    // running hit flag (inclusive, path[0] included), up
    PK_BARRIER_DOWN = 4  // the same, down barrier
};

__host__ __device__ constexpr bool is_barrier(int pk) {
    return pk == PK_BARRIER || pk == PK_BARRIER_DOWN;
}

enum OptFlags : int32_t {
    F_PUT = 1,            // option type put (else call)
    F_FLOATING = 2,       // Asian floating strike
    F_GEOMETRIC = 4,      // Asian geometric averaging payoff
    F_DOWN = 8,           // barrier direction down (else up)
    F_KNOCK_IN = 16,      // barrier style in (else out)
    F_CV = 32,            // control variate active
    F_CV_GEO = 64         // CV evaluator = geometric Asian (else terminal vanilla)
};

// Device-side contract descriptor. All fields are wave-uniform: a wave only
// ever simulates one option at a time, so these are read with scalar loads.
struct alignas(16) OptDev {
    int32_t model;        // 0 GBM, 1 Heston
    int32_t kind;         // PayoffKind
    int32_t flags;        // OptFlags
    int32_t steps;
    int64_t samples;      // ceil(paths/2) with antithetic, else paths
    int64_t out_offset;   // offset of sample 0 in per-sample output arrays
    int64_t normals_offset;   // offset of this option's [step][sample] normals
    int64_t obs_offset;   // offset of this option's steps+1 observation weights
    uint64_t seed;        // base seed of stream_seed(seed, sample)
    double s0, strike, mu, sigma;
    double dt, sqrt_dt, drift, diffusion;
    double kappa, theta, xi, rho, rho_c, v0;
    double barrier, rebate, nobs, log_s0;
    double df, cv_expected, rate;
    // Heston step constants of the fast mode's re-associated Euler step
    double kdt, kthdt, xsdt, mudt, hdt;
    // fast mode: exponent constants pre-scaled by 256/ln2 for exp2_256
    double drift2, diffusion2, sqrt_dt2, mudt2, hdt2;
    // Monte Carlo Greeks estimators: the bumped spots of the pathwise gamma
    double spot_up, spot_dn;
    // sum_j j w_j of the observation weights (forward pathwise Greeks)
    double obs_jsum;
    int32_t antithetic;
    int32_t has_cv;
};

// One work item: a contiguous range of samples of one option, simulated by
// one wavefront (lane l takes samples begin + l, begin + 64 + l, ...).
struct alignas(16) WorkItem {
    int32_t option;
    int32_t pad;
    int64_t begin;
    int64_t end;
    int64_t pad2;
};

// Per-item partial sums (simulation.rs:205-259 accumulator) at a fixed slot.
struct Partial {
    double sx, sx2, sy, sxy, sy2, n;
};

// Per-option reduced sums and closed statistics.
struct OptOut {
    double price, stderr_, sx, sx2, sy, sxy, sy2, n;
};

// Close of MonteCarloEngine::run (simulation.rs:305-347).
__host__ __device__ inline void close_stats(double sum_x, double sum_x2, double sum_y,
                                            double sum_xy, double sum_y2, double n,
                                            int has_cv, double cv_expected, double df,
                                            double *price, double *stderr_out) {
    if (has_cv) {
        const double denom = fmax(n - 1.0, 1.0);
        const double cov_xy = (sum_xy - sum_x * sum_y / n) / denom;
        const double var_y = (sum_y2 - sum_y * sum_y / n) / denom;
        const double beta = var_y > 1e-16 ? cov_xy / var_y : 0.0;
        const double e = cv_expected;
        const double sum_adj = sum_x + beta * (n * e - sum_y);
        const double sum_adj_sq = sum_x2 + 2.0 * beta * e * sum_x - 2.0 * beta * sum_xy +
                                  beta * beta * (n * e * e - 2.0 * e * sum_y + sum_y2);
        const double mean = sum_adj / n;
        const double var = n > 1.0 ? (sum_adj_sq - sum_adj * sum_adj / n) / (n - 1.0) : 0.0;
        *price = df * mean;
        *stderr_out = df * sqrt(var / n);
    } else {
        const double mean = sum_x / n;
        const double var = n > 1.0 ? (sum_x2 - sum_x * sum_x / n) / (n - 1.0) : 0.0;
        *price = df * mean;
        *stderr_out = df * sqrt(var / n);
    }
}

// Launch interface of the kernels (ofg_kernels.hip).
struct LaunchArgs {
    const OptDev *opts;
    const WorkItem *items;
    const int32_t *order;     // queue order (cost-descending) of item ids
    int32_t order_begin;      // this launch pops order[order_begin, order_end)
    int32_t order_end;
    uint32_t *queue_head;     // this launch's queue counter
    Partial *partials;
    const double *obs_w;      // observation weights, per option steps+1
    double *xs, *ys;          // optional per-sample outputs
    const double *normals1, *normals2;  // external normals
    int32_t rng_mode;
    int32_t model;            // every item of one launch has this model
    int32_t antithetic;       // and this antithetic flag
    int32_t block;            // threads per block
    int32_t grid;             // persistent blocks
};

// Zeroes n queue counters (a kernel, so it is safe inside captured graphs).
hipError_t launch_reset_queue(uint32_t *head, int32_t n, hipStream_t stream);
hipError_t launch_pricing(const LaunchArgs &a, hipStream_t stream);
hipError_t launch_reduce(const OptDev *opts, const int64_t *item_first, const int32_t *item_count,
                         const Partial *partials, OptOut *out, int32_t n_opts,
                         hipStream_t stream);
// Occupancy-derived persistent grid size for a block size.
int persistent_grid(int block, int rng_mode, int model, int antithetic, int device);

}  // namespace ofg
