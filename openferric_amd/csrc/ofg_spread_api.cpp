// C ABI of the spread-option Monte Carlo engine (ofg_spread_price):
// SpreadMonteCarloEngine::price (src/engines/monte_carlo/spread_mc.rs:83-175)
// with the reference's validation (SpreadOption::validate,
// src/instruments/spread.rs:31-58), early returns and statistics.
#include "../../include/openferric_amd.h"
#include "ofg_host.hpp"
#include "ofg_spread.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

using namespace ofg_host;

namespace ofg_host {

struct SpreadBuffers {
    DevBuf jobs, waves, sums, pay;
};

void spread_release(ofg_ctx *ctx) {
    if (!ctx->spread) return;
    for (DevBuf *b : {&ctx->spread->jobs, &ctx->spread->waves, &ctx->spread->sums,
                      &ctx->spread->pay})
        b->release();
    delete ctx->spread;
    ctx->spread = nullptr;
}

}  // namespace ofg_host

namespace {

constexpr int64_t SAMPLES_PER_WAVE = 64 * ofg::SPREAD_PER;
constexpr int64_t MAX_WAVES = int64_t(1) << 30;

// SpreadOption::validate (spread.rs:31-58): the same comparisons in the same
// order, so NaN fields pass exactly where they pass there.
void validate_spread(const ofg_spread_option &o) {
    if (o.s1 <= 0.0 || o.s2 <= 0.0) fail(OFG_INVALID_INPUT, "spread spots s1 and s2 must be > 0");
    if (o.k < 0.0) fail(OFG_INVALID_INPUT, "spread strike k must be >= 0");
    if (o.vol1 < 0.0 || o.vol2 < 0.0)
        fail(OFG_INVALID_INPUT, "spread volatilities vol1 and vol2 must be >= 0");
    if (!(o.rho >= -1.0 && o.rho <= 1.0))
        fail(OFG_INVALID_INPUT, "spread correlation rho must be in [-1, 1]");
    if (o.t < 0.0) fail(OFG_INVALID_INPUT, "spread maturity t must be >= 0");
}

int32_t spread_impl(ofg_ctx *ctx, const ofg_spread_option *opts, size_t n, int32_t antithetic,
                    int32_t rng_mode, ofg_result *out, double *payoffs) {
    if (!ctx || (!opts && n) || (!out && n)) fail(OFG_INVALID_INPUT, "null argument");
    if (rng_mode != OFG_RNG_REFERENCE && rng_mode != OFG_RNG_PCG64 && rng_mode != OFG_RNG_FAST)
        fail(OFG_INVALID_INPUT,
             "ofg_spread_price takes OFG_RNG_REFERENCE, OFG_RNG_PCG64 or OFG_RNG_FAST");
    const bool anti = antithetic != 0;
    std::vector<ofg::SpreadJob> jobs;
    std::vector<int32_t> job_of(n, -1);
    std::vector<double> discount(n, 0.0);
    int64_t waves = 0, pay_offset = 0;
    for (size_t i = 0; i < n; ++i) {
        const ofg_spread_option &o = opts[i];
        ofg_result &r = out[i];
        std::memset(&r, 0, sizeof r);
        r.price = NAN;
        validate_spread(o);
        if (o.num_paths < 0) fail(OFG_INVALID_INPUT, "num_paths must be >= 0");
        if (o.num_paths == 0) fail(OFG_INVALID_INPUT, "spread MC num_paths must be > 0");
        const int64_t samples = anti ? (o.num_paths + 1) / 2 : o.num_paths;
        const int64_t my_pay = pay_offset;
        pay_offset += samples;
        if (o.t <= 0.0) {   // spread_mc.rs:94-101: intrinsic, stderr 0, no diagnostics
            r.price = std::fmax(o.s1 - o.s2 - o.k, 0.0);
            r.stderr_ = 0.0;
            r.has_stderr = 1;
            continue;
        }
        // spread_mc.rs:103-113
        const double sqrt_t = std::sqrt(o.t);
        ofg::SpreadJob j{};
        j.s1 = o.s1;
        j.s2 = o.s2;
        j.k = o.k;
        j.rho = o.rho;
        j.corr_tail = std::sqrt(std::fmax(1.0 - o.rho * o.rho, 0.0));
        j.drift1 = (o.r - o.q1 - 0.5 * o.vol1 * o.vol1) * o.t;
        j.drift2 = (o.r - o.q2 - 0.5 * o.vol2 * o.vol2) * o.t;
        j.vol_term1 = o.vol1 * sqrt_t;
        j.vol_term2 = o.vol2 * sqrt_t;
        j.seed = o.seed;
        j.samples = samples;
        j.pay_offset = my_pay;
        j.wave_first = waves;
        const int64_t wc = (samples + SAMPLES_PER_WAVE - 1) / SAMPLES_PER_WAVE;
        if (wc > MAX_WAVES - waves) fail(OFG_INVALID_INPUT, "too many spread samples in one call");
        j.wave_count = (int32_t)wc;
        waves += wc;
        discount[i] = std::exp(-o.r * o.t);
        job_of[i] = (int32_t)jobs.size();
        jobs.push_back(j);
        r.num_paths = (double)o.num_paths;
    }
    if (jobs.empty()) return OFG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->spread) ctx->spread = new SpreadBuffers();
    SpreadBuffers &b = *ctx->spread;
    hipStream_t s = ctx->stream;
    b.jobs.ensure(jobs.size() * sizeof(ofg::SpreadJob));
    hip_check(hipMemcpyAsync(b.jobs.p, jobs.data(), jobs.size() * sizeof(ofg::SpreadJob),
                             hipMemcpyHostToDevice, s),
              "H2D spread jobs");
    b.waves.ensure((size_t)waves * sizeof(ofg::SpreadSums));
    b.sums.ensure(jobs.size() * sizeof(ofg::SpreadSums));
    double *pay_dev = nullptr;
    if (payoffs) {
        b.pay.ensure(std::max<size_t>((size_t)pay_offset * sizeof(double), 16));
        pay_dev = b.pay.as<double>();
    }
    hip_check(hipEventRecord(ctx->ev0, s), "event");
    hip_check(ofg::launch_spread(b.jobs.as<ofg::SpreadJob>(), (int32_t)jobs.size(), waves,
                                 rng_mode, anti, b.waves.as<ofg::SpreadSums>(),
                                 b.sums.as<ofg::SpreadSums>(), pay_dev, s),
              "spread kernel launch");
    hip_check(hipEventRecord(ctx->ev2, s), "event");
    std::vector<ofg::SpreadSums> sums(jobs.size());
    hip_check(hipMemcpyAsync(sums.data(), b.sums.p, sums.size() * sizeof(ofg::SpreadSums),
                             hipMemcpyDeviceToHost, s),
              "D2H spread sums");
    hip_check(hipStreamSynchronize(s), "spread execution");
    if (payoffs) {   // simulated options' samples only
        for (const ofg::SpreadJob &j : jobs)
            hip_check(hipMemcpy(payoffs + j.pay_offset, pay_dev + j.pay_offset,
                                (size_t)j.samples * sizeof(double), hipMemcpyDeviceToHost),
                      "D2H spread payoffs");
    }
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev2), "event time");
    for (size_t i = 0; i < n; ++i) {
        out[i].kernel_ms = ms;
        if (job_of[i] < 0) continue;
        // spread_mc.rs:150-174: single-pass statistics over the samples
        const ofg::SpreadSums &t = sums[job_of[i]];
        const int64_t m = jobs[job_of[i]].samples;
        const double nn = (double)m;
        const double mean = t.sum / nn;
        const double variance = m > 1 ? (t.sum_sq - t.sum * t.sum / nn) / (nn - 1.0) : 0.0;
        out[i].price = discount[i] * mean;
        out[i].stderr_ = discount[i] * std::sqrt(variance / nn);
        out[i].has_stderr = 1;
    }
    return OFG_OK;
}

}  // namespace

extern "C" int32_t ofg_spread_price(ofg_ctx *ctx, const ofg_spread_option *opts, size_t n,
                                    int32_t antithetic, int32_t rng_mode, ofg_result *out,
                                    double *payoffs) {
    return guarded([&]() { return spread_impl(ctx, opts, n, antithetic, rng_mode, out, payoffs); });
}
