// C ABI of the exact-terminal European pricers (ofg_european_exact): the
// reference's mc_european_sequential / mc_european_parallel
// (src/engines/monte_carlo/mc_parallel.rs:264-434) and mc_european_with_arena
// (mc_engine.rs:186-343) in its default build (no "simd" feature), on the
// reference's own single streams split across GPU threads by jump-ahead.
#include "../../include/openferric_amd.h"
#include "ofg_exact.hpp"
#include "ofg_host.hpp"
#include "ofg_jump.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <vector>

using namespace ofg_host;

namespace ofg_host {

struct ExactBuffers {
    DevBuf jobs, polys, waves, sums, pay;
};

void exact_release(ofg_ctx *ctx) {
    if (!ctx->exact) return;
    for (DevBuf *b : {&ctx->exact->jobs, &ctx->exact->polys, &ctx->exact->waves,
                      &ctx->exact->sums, &ctx->exact->pay})
        b->release();
    delete ctx->exact;
    ctx->exact = nullptr;
}

}  // namespace ofg_host

namespace {

constexpr uint64_t SEQUENTIAL_SEED = 0xDEADBEEFCAFEBABEull;   // mc_parallel.rs:325, 404
constexpr uint64_t ARENA_SEED = 42;                             // ARENA_MC_SEED, mc_engine.rs:178
constexpr uint64_t CHUNK_SEED_STEP = 6364136223846793005ull;    // mc_parallel.rs:330

double vanilla(int32_t type, double spot, double strike) {
    return type == OFG_CALL ? std::fmax(spot - strike, 0.0) : std::fmax(strike - spot, 0.0);
}

template <class T>
void put(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    b.ensure(std::max<size_t>(v.size() * sizeof(T), 16));
    if (!v.empty())
        hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync H2D");
}

struct OptJobs {
    size_t first = 0, count = 0;   // jobs of this option, in chunk order
    double discount = 0.0;
    bool simulated = false;
};

int32_t exact_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n, int32_t kind,
                   int32_t n_chunks, ofg_result *out, double *payoffs) {
    if (!ctx || (!opts && n) || (!out && n)) fail(OFG_INVALID_INPUT, "null argument");
    if (kind != OFG_EXACT_SEQUENTIAL && kind != OFG_EXACT_PARALLEL && kind != OFG_EXACT_ARENA)
        fail(OFG_INVALID_INPUT, "unknown exact-terminal pricer");
    if (kind == OFG_EXACT_PARALLEL && n_chunks < 1) fail(OFG_INVALID_INPUT, "n_chunks must be >= 1");
    std::vector<ofg::ExactJob> jobs;
    std::vector<uint64_t> polys;
    std::map<uint64_t, std::vector<ofg_jump::Poly>> cache;
    std::vector<OptJobs> oj(n);
    int64_t thread_first = 0, pay_offset = 0;
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        ofg_result &r = out[i];
        std::memset(&r, 0, sizeof r);
        r.price = NAN;
        if (o.model != OFG_MODEL_GBM || o.payoff != OFG_PAYOFF_VANILLA)
            fail(OFG_INVALID_INPUT, "exact-terminal pricers take European vanillas under GBM");
        if (o.option_type != OFG_CALL && o.option_type != OFG_PUT)
            fail(OFG_INVALID_INPUT, "unknown option type");
        if (o.num_paths < 0) fail(OFG_INVALID_INPUT, "num_paths must be >= 0");
        const int64_t np = o.num_paths;
        const int64_t my_pay = pay_offset;
        pay_offset += np;
        // the reference's early returns, in its order
        if (np == 0) continue;                                    // price NaN, no stderr
        if (kind == OFG_EXACT_ARENA && (!(o.strike > 0.0) || o.expiry < 0.0))
            continue;                                             // validate() failed
        if (o.expiry <= 0.0) {                                    // intrinsic, stderr 0
            r.price = vanilla(o.option_type, o.spot, o.strike);
            r.stderr_ = 0.0;
            r.has_stderr = 1;
            continue;
        }
        if (!(o.vol > 0.0) || !std::isfinite(o.vol)) continue;   // price NaN
        const double t = o.expiry;
        // rate - q: the effective dividend yield (parallel, sequential) or
        // the dividend yield (arena), as the caller passes it in div_yield
        const double drift = (o.rate - o.div_yield - 0.5 * o.vol * o.vol) * t;
        const double diffusion = o.vol * std::sqrt(t);
        oj[i].discount = std::exp(-o.rate * t);
        oj[i].simulated = true;
        oj[i].first = jobs.size();
        // streams: (paths, seed)
        std::vector<std::pair<int64_t, uint64_t>> streams;
        if (kind == OFG_EXACT_SEQUENTIAL) {
            streams.push_back({np, SEQUENTIAL_SEED});
        } else if (kind == OFG_EXACT_ARENA) {
            streams.push_back({np, ARENA_SEED});
        } else {
            // split_paths (mc_parallel.rs:37-46): the first n % chunks chunks
            // take one path more; empty chunks are dropped before numbering
            const int64_t base = np / n_chunks, rem = np % n_chunks;
            uint64_t idx = 0;
            for (int64_t c = 0; c < n_chunks; ++c) {
                const int64_t len = c < rem ? base + 1 : base;
                if (len == 0) continue;
                streams.push_back({len, SEQUENTIAL_SEED + idx * CHUNK_SEED_STEP});
                ++idx;
            }
        }
        int64_t chunk_first = 0;
        for (const auto &st : streams) {
            ofg::ExactJob j{};
            j.seed = ofg_jump::make_seed_job(st.second, st.first, 1, 0, thread_first, polys, cache);
            if (j.seed.chunk != 64) fail(OFG_NUMERICAL_ERROR, "unexpected stream chunk");
            const int64_t threads = ofg_jump::seed_job_threads(j.seed);
            j.spot = o.spot;
            j.strike = o.strike;
            j.drift = drift;
            j.diffusion = diffusion;
            j.put = o.option_type == OFG_PUT ? 1 : 0;
            j.pay_offset = my_pay + chunk_first;
            j.wave_first = thread_first / 64;
            j.wave_count = (int32_t)((threads + 63) / 64);
            thread_first += (int64_t)j.wave_count * 64;
            chunk_first += st.first;
            jobs.push_back(j);
        }
        oj[i].count = jobs.size() - oj[i].first;
        r.num_paths = (double)np;
        r.num_steps = 1.0;
        r.vol = o.vol;
    }
    if (jobs.empty()) return OFG_OK;
    if (jobs.size() > (size_t)INT32_MAX) fail(OFG_INVALID_INPUT, "too many streams");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->exact) ctx->exact = new ExactBuffers();
    ExactBuffers &b = *ctx->exact;
    hipStream_t s = ctx->stream;
    const int64_t total_threads = thread_first;
    const int64_t n_waves = total_threads / 64;
    put(b.jobs, jobs, s);
    put(b.polys, polys, s);
    b.waves.ensure((size_t)n_waves * sizeof(ofg::ExactSums));
    b.sums.ensure(jobs.size() * sizeof(ofg::ExactSums));
    double *pay_dev = nullptr;
    if (payoffs) {
        b.pay.ensure(std::max<size_t>((size_t)pay_offset * sizeof(double), 16));
        pay_dev = b.pay.as<double>();
    }
    hip_check(hipEventRecord(ctx->ev0, s), "event");
    hip_check(ofg::launch_exact(b.jobs.as<ofg::ExactJob>(), (int32_t)jobs.size(), total_threads,
                                b.polys.as<uint64_t>(), b.waves.as<ofg::ExactSums>(),
                                b.sums.as<ofg::ExactSums>(), pay_dev, s),
              "exact-terminal kernel launch");
    hip_check(hipEventRecord(ctx->ev2, s), "event");
    std::vector<ofg::ExactSums> sums(jobs.size());
    hip_check(hipMemcpyAsync(sums.data(), b.sums.p, sums.size() * sizeof(ofg::ExactSums),
                             hipMemcpyDeviceToHost, s),
              "D2H sums");
    hip_check(hipStreamSynchronize(s), "exact-terminal execution");
    // payoffs of simulated options only (the others' slots stay untouched)
    if (payoffs) {
        for (size_t i = 0; i < n; ++i) {
            if (!oj[i].simulated) continue;
            const ofg::ExactJob &j0 = jobs[oj[i].first];
            hip_check(hipMemcpy(payoffs + j0.pay_offset, pay_dev + j0.pay_offset,
                                (size_t)opts[i].num_paths * sizeof(double), hipMemcpyDeviceToHost),
                      "D2H payoffs");
        }
    }
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev2), "event time");
    for (size_t i = 0; i < n; ++i) {
        out[i].kernel_ms = ms;
        if (!oj[i].simulated) continue;
        // chunk sums in chunk order (the reference reduces them with rayon)
        double sum = 0.0, sum_sq = 0.0;
        for (size_t k = 0; k < oj[i].count; ++k) {
            sum += sums[oj[i].first + k].sum;
            sum_sq += sums[oj[i].first + k].sum_sq;
        }
        const int64_t np = opts[i].num_paths;
        const double nn = (double)np;
        const double mean = sum / nn;
        double variance = np > 1 ? (sum_sq - sum * sum / nn) / (nn - 1.0) : 0.0;
        if (kind != OFG_EXACT_ARENA) variance = std::fmax(variance, 0.0);  // mc_parallel.rs:340
        out[i].price = oj[i].discount * mean;
        out[i].stderr_ = oj[i].discount * std::sqrt(variance / nn);
        out[i].has_stderr = 1;
    }
    return OFG_OK;
}

}  // namespace

extern "C" int32_t ofg_european_exact(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                                      int32_t kind, int32_t n_chunks, ofg_result *out,
                                      double *payoffs) {
    return guarded([&]() { return exact_impl(ctx, opts, n, kind, n_chunks, out, payoffs); });
}
