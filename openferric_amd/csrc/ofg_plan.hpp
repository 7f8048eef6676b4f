// Host-side launch plans of the pricing kernels: instrument validation with
// the reference's messages (src/instruments/*.rs validate(),
// src/engines/monte_carlo/mc_engine.rs:876-912), the closed forms of the
// built-in control variates, the device contract descriptors, observation
// weights and the cost-ordered work queue of MonteCarloPricingEngine::price
// (mc_engine.rs:889-945). No HIP calls: this translation unit also builds
// with a host compiler, under the address and undefined-behaviour
// sanitizers, for the host tests.
#pragma once

#include "../../include/openferric_amd.h"
#include "ofg_host.hpp"
#include "ofg_internal.hpp"

#include <string>
#include <vector>

namespace ofg_plan {

// Zero observation weights after the last option's: the reference-draw
// kernels read a block's weights (up to 16 steps) together, past the end of
// an option's own weights at its last block (the values read there are
// never applied).
constexpr size_t OBS_PAD = 16;

// A call whose whole work would make fewer work items than this at the
// tuned item size gets smaller items (down to one sample per lane): about
// eight per resident wave of the 256 CUs, so the queue keeps every SIMD fed.
constexpr double MIN_ITEMS = 8192.0;

// Host-side plan of one pricing launch.
struct Plan {
    std::vector<ofg::OptDev> opts;
    std::vector<double> obs_w;
    std::vector<ofg::WorkItem> items;
    std::vector<int32_t> order;
    std::vector<int64_t> item_first;
    std::vector<int32_t> item_count;
    std::vector<int32_t> active;      // plan index -> caller option index
    std::vector<double> vols;         // per caller option
    int64_t total_samples = 0;
    int64_t normals_elems = 0;
    bool any_heston = false;
    int32_t antithetic = 0;
    int32_t model_begin[2] = {0, 0};   // order[] range of each model's items
    int32_t model_end[2] = {0, 0};
};

// Closed forms of the control variates: hart_norm_cdf (src/math/fast_norm.rs:
// 27-50), the bs_price generic branch (src/engines/analytic/black_scholes.rs:
// 62-107), geometric_asian_discrete_fixed_closed_form (src/pricing/asian.rs:
// 86-136).
double hart_norm_cdf(double x);
double vanilla_payoff(int32_t type, double spot, double strike);
double bs_price(int32_t type, double spot, double strike, double rate, double q, double vol,
                double expiry);
double geometric_asian_discrete(int32_t type, double s0, double k, double r, double q,
                                double sigma, const double *obs, int64_t n_obs);

// Heston::validate (src/models/stochastic.rs:47-54)
bool heston_valid(const ofg_option &o);
// payoff_from_path(&[spot]) of the maturity-zero shortcut (mc_engine.rs:896-904)
double payoff_at_spot(const ofg_option &o);

// Autotuner key of a plan: model mix, payoff mix, RNG mode and a step bucket.
std::string workload_class(const std::vector<ofg::OptDev> &opts, int32_t rng_mode);

// Validates the options (throws ofg_host::OfgError with the reference's
// message) and builds the device descriptors, observation weights and work
// items; maturity-zero options get no descriptor (Plan::active maps plan
// indices back to caller indices).
Plan build_plan(const ofg_option *in, size_t n, const ofg_engine_cfg &cfg,
                const ofg_host::Tuning &tune);

}  // namespace ofg_plan
