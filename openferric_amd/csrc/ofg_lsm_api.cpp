// Host side of the least-squares Monte Carlo entry points (ofg_lsm_price,
// ofg_stream_jump).
//
// Validation with the reference's messages and order
// (src/engines/lsm/longstaff_schwartz.rs:291-320, 448-480, 683-705 and the
// instruments' validate()), the exercise schedule of each instrument
// (can_exercise :567-584, effective_schedule instruments/bermudan.rs:85-103
// and step_schedule :314-320), the device layout of ofg_lsm.hpp, and the
// stream partition of the reference draws: the reference simulates path
// after path from ONE xoshiro256++ stream, so path p starts at draw p * d
// (d draws per path). The xoshiro256 transition T is linear over GF(2), so
// T^k = r(T) with r(x) = x^k mod P(x), P the characteristic polynomial of T
// (found once by Berlekamp-Massey on a state bit's sequence); the seed kernel
// applies such polynomials to reach each path's first draw.
#include "../../include/openferric_amd.h"
#include "ofg_host.hpp"
#include "ofg_internal.hpp"
#include "ofg_lsm.hpp"
#include "ofg_math.hpp"
#include "ofg_jump.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <vector>

using namespace ofg_host;
using namespace ofg_jump;

// the seed kernel's thread-start mode assumes 256-thread tiles
static_assert(ofg::LSM_BLOCK == 256, "tile width of ofg::thread_stream_start");

namespace ofg_host {

struct LsmBuffers {
    DevBuf opts, tile_opt, dates, date_out, opt_out, partials, spots, values, seeds, polys, jobs,
        n1, n2;
    // device time of the last call: seeds + path kernel, backward rounds
    double paths_ms = 0.0, rounds_ms = 0.0;
    int32_t rounds = 0;
};

void lsm_release(ofg_ctx *ctx) {
    if (!ctx || !ctx->lsm) return;
    LsmBuffers *b = ctx->lsm;
    for (DevBuf *d : {&b->opts, &b->tile_opt, &b->dates, &b->date_out, &b->opt_out, &b->partials,
                      &b->spots, &b->values, &b->seeds, &b->polys, &b->jobs, &b->n1, &b->n2})
        d->release();
    delete b;
    ctx->lsm = nullptr;
}

}  // namespace ofg_host

namespace {

// ---------------------------------------------------------------------------
// Planning
// ---------------------------------------------------------------------------
double intrinsic(int32_t type, double s, double k) {
    return type == OFG_CALL ? std::fmax(s - k, 0.0) : std::fmax(k - s, 0.0);
}

struct HostOpt {
    bool simulate = false;
    double shortcut_price = 0.0;   // expiry-0 vanilla
    int32_t dev = -1;
    bool bermudan = false;
    std::vector<double> times, strikes;   // per exercise date (boundary output)
    double terminal_strike = 0.0;
    double diag_vol = NAN;
};

struct LsmPlan {
    std::vector<ofg::LsmDev> opts;
    std::vector<ofg::LsmDate> dates;
    std::vector<int32_t> tile_opt;
    std::vector<HostOpt> host;
    std::vector<ofg::LsmSeedJob> jobs;
    std::vector<uint64_t> polys;
    int64_t spots = 0, values = 0, seeds = 0, normals = 0, seed_threads = 0;
    int32_t max_rounds = 0;
};

void validate_vanilla(const ofg_lsm_option &o) {
    // instruments/vanilla.rs:69-96
    if (o.strike <= 0.0) fail(OFG_INVALID_INPUT, "vanilla strike must be > 0");
    if (o.expiry < 0.0) fail(OFG_INVALID_INPUT, "vanilla expiry must be >= 0");
    if (o.exercise == OFG_BERMUDAN) {
        if (o.n_dates <= 0 || !o.dates)
            fail(OFG_INVALID_INPUT, "bermudan exercise dates cannot be empty");
        for (int32_t i = 0; i < o.n_dates; ++i)
            if (o.dates[i] <= 0.0 || o.dates[i] > o.expiry)
                fail(OFG_INVALID_INPUT, "bermudan exercise dates must lie in (0, expiry]");
    }
}

void validate_bermudan(const ofg_lsm_option &o) {
    // BermudanOption::validate (instruments/bermudan.rs:106-144)
    if (o.expiry <= 0.0 || !std::isfinite(o.expiry))
        fail(OFG_INVALID_INPUT, "bermudan expiry must be finite and > 0");
    if (o.n_dates <= 0 || !o.dates) fail(OFG_INVALID_INPUT, "bermudan exercise_dates cannot be empty");
    if (!o.strikes)
        fail(OFG_INVALID_INPUT, "bermudan strike_schedule must match exercise_dates length");
    for (int32_t i = 0; i < o.n_dates; ++i) {
        const double t = o.dates[i];
        if (!std::isfinite(t) || t <= 0.0 || t > o.expiry)
            fail(OFG_INVALID_INPUT, "bermudan exercise_dates must be finite and in (0, expiry]");
    }
    for (int32_t i = 1; i < o.n_dates; ++i)
        if (o.dates[i] <= o.dates[i - 1])
            fail(OFG_INVALID_INPUT, "bermudan exercise_dates must be strictly increasing");
    for (int32_t i = 0; i < o.n_dates; ++i)
        if (!std::isfinite(o.strikes[i]) || o.strikes[i] <= 0.0)
            fail(OFG_INVALID_INPUT, "bermudan strike_schedule entries must be finite and > 0");
}

void fill_path_constants(ofg::OptDev &d, const ofg_lsm_option &o, double vol, double mu,
                         int32_t rng_mode) {
    d.steps = (int32_t)o.num_steps;
    d.seed = o.seed;
    d.s0 = o.spot;
    d.strike = o.strike;
    d.mu = mu;
    d.sigma = vol;
    d.dt = o.expiry / (double)o.num_steps;
    d.sqrt_dt = std::sqrt(d.dt);
    d.drift = (mu - 0.5 * vol * vol) * d.dt;
    d.diffusion = vol * d.sqrt_dt;
    d.kappa = o.kappa;
    d.theta = o.theta;
    d.xi = o.xi;
    d.rho = o.rho;
    d.rho_c = std::sqrt(1.0 - o.rho * o.rho);
    d.v0 = o.v0;
    d.rate = o.rate;
    d.kdt = o.kappa * d.dt;
    d.kthdt = o.kappa * o.theta * d.dt;
    d.xsdt = o.xi * d.sqrt_dt;
    d.mudt = mu * d.dt;
    d.hdt = 0.5 * d.dt;
    d.drift2 = d.drift * OFG_INV_LN2_256;
    d.diffusion2 = d.diffusion * OFG_INV_LN2_256;
    d.sqrt_dt2 = d.sqrt_dt * OFG_INV_LN2_256;
    d.mudt2 = d.mudt * OFG_INV_LN2_256;
    d.hdt2 = d.hdt * OFG_INV_LN2_256;
    (void)rng_mode;
}

// Paths per thread of one option's tiles. A tile streams LSM_BLOCK * ppt
// paths per launch after a short dependent prologue (tile -> option -> date
// data), so tiles are made fat: about LSM_TILES_PER_OPTION per option (at
// most 64 paths per thread). The tile size depends on the option alone, so
// its summation order, and with it every bit of its price, does not depend
// on what else is in the batch.
constexpr int64_t LSM_TILES_PER_OPTION = 512;

int32_t tile_ppt(int64_t paths) {
    const int64_t blk = ofg::LSM_BLOCK;
    const int64_t ppt = (paths + blk * LSM_TILES_PER_OPTION - 1) / (blk * LSM_TILES_PER_OPTION);
    return (int32_t)std::min<int64_t>(64, std::max<int64_t>(1, ppt));
}

LsmPlan build_lsm_plan(const ofg_lsm_option *in, size_t n, int32_t rng_mode) {
    LsmPlan p;
    p.host.resize(n);
    std::map<uint64_t, std::vector<Poly>> poly_cache;
    for (size_t i = 0; i < n; ++i) {
        const ofg_lsm_option &o = in[i];
        HostOpt &h = p.host[i];
        if (o.option_type != OFG_CALL && o.option_type != OFG_PUT)
            fail(OFG_INVALID_INPUT, "unknown option type");
        const int64_t steps = o.num_steps, paths = o.num_paths;
        if (steps > 1000000) fail(OFG_INVALID_INPUT, "num_steps must be <= 1000000");
        std::vector<int64_t> ex_steps;
        std::vector<double> ex_strikes, ex_times;
        double vol = o.vol;
        int32_t dynamics = ofg::LSM_GBM;
        const double mu = o.rate - o.div_yield;
        ofg::LsmDev d{};

        if (o.instrument == OFG_LSM_BARRIER) {
            ofg_option b{};
            b.payoff = OFG_PAYOFF_BARRIER;
            b.option_type = o.option_type;
            b.barrier_dir = o.barrier_dir;
            b.barrier_style = o.barrier_style;
            b.strike = o.strike;
            b.expiry = o.expiry;
            b.barrier_level = o.barrier_level;
            b.rebate = o.rebate;
            validate_instrument(b);   // instruments/barrier.rs:36-58
            if (steps <= 0) fail(OFG_INVALID_INPUT, "num_steps must be > 0");
            if (paths <= 0) fail(OFG_INVALID_INPUT, "num_paths must be > 0");
            if (vol <= 0.0) fail(OFG_INVALID_INPUT, "market volatility must be > 0");
            d.barrier = 1;
            d.barrier_down = o.barrier_dir == OFG_DOWN;
            d.knock_in = o.barrier_style == OFG_IN;
            d.barrier_level = o.barrier_level;
            d.rebate = o.rebate;
            d.df_total = std::exp(-o.rate * o.expiry);
            h.diag_vol = vol;
        } else if (o.instrument == OFG_LSM_VANILLA) {
            if (o.exercise != OFG_EUROPEAN && o.exercise != OFG_AMERICAN &&
                o.exercise != OFG_BERMUDAN)
                fail(OFG_INVALID_INPUT, "unknown exercise style");
            validate_vanilla(o);
            if (steps < 2) fail(OFG_INVALID_INPUT, "num_steps must be >= 2 for Longstaff-Schwartz");
            if (paths < 3) fail(OFG_INVALID_INPUT, "num_paths must be >= 3 for Longstaff-Schwartz");
            if (o.expiry == 0.0) {
                h.shortcut_price = intrinsic(o.option_type, o.spot, o.strike);
                continue;
            }
            if (vol <= 0.0) fail(OFG_INVALID_INPUT, "market volatility must be > 0");
            h.diag_vol = vol;
            std::vector<char> can((size_t)steps + 1, 0);
            if (o.exercise == OFG_AMERICAN) {
                for (int64_t t = 1; t < steps; ++t) can[(size_t)t] = 1;
            } else if (o.exercise == OFG_BERMUDAN) {
                for (int32_t k = 0; k < o.n_dates; ++k) {
                    const double x = std::round((o.dates[k] / o.expiry) * (double)steps);
                    int64_t idx = x > 0.0 ? (int64_t)x : 0;
                    can[(size_t)std::min(idx, steps)] = 1;
                }
            }
            for (int64_t t = 1; t < steps; ++t)
                if (can[(size_t)t]) {
                    ex_steps.push_back(t);
                    ex_strikes.push_back(o.strike);
                    ex_times.push_back(o.expiry * (double)t / (double)steps);
                }
            h.terminal_strike = o.strike;
        } else if (o.instrument == OFG_LSM_BERMUDAN) {
            validate_bermudan(o);
            if (steps < 2) fail(OFG_INVALID_INPUT, "num_steps must be >= 2 for Longstaff-Schwartz");
            if (paths < 3) fail(OFG_INVALID_INPUT, "num_paths must be >= 3 for Longstaff-Schwartz");
            // effective_schedule: the expiry with the last strike is appended
            // when the last date is before it
            std::vector<std::pair<double, double>> sched;
            for (int32_t k = 0; k < o.n_dates; ++k) sched.emplace_back(o.dates[k], o.strikes[k]);
            if (o.expiry - sched.back().first > 1.0e-12)
                sched.emplace_back(o.expiry, sched.back().second);
            h.terminal_strike = sched.back().second;
            std::vector<int> has((size_t)steps + 1, 0);
            std::vector<std::pair<double, double>> at((size_t)steps + 1);
            for (const auto &tk : sched) {
                const double x = std::round((tk.first / o.expiry) * (double)steps);
                int64_t idx = x > 0.0 ? (int64_t)x : 0;
                idx = std::min(std::max<int64_t>(idx, 1), steps);
                has[(size_t)idx] = 1;
                at[(size_t)idx] = tk;
            }
            for (int64_t t = 1; t < steps; ++t)
                if (has[(size_t)t]) {
                    ex_steps.push_back(t);
                    ex_times.push_back(at[(size_t)t].first);
                    ex_strikes.push_back(at[(size_t)t].second);
                }
            dynamics = o.dynamics;
            if (dynamics == OFG_LSM_GBM) {
                if (!std::isfinite(vol) || vol <= 0.0)
                    fail(OFG_INVALID_INPUT, "market volatility must be finite and > 0");
                h.diag_vol = vol;
            } else if (dynamics == OFG_LSM_LOCAL_VOL) {
                if (!std::isfinite(vol) || vol <= 0.0)
                    fail(OFG_INVALID_INPUT, "local volatility surface returned non-positive value");
            } else if (dynamics == OFG_LSM_HESTON) {
                if (!(o.kappa > 0.0 && o.theta >= 0.0 && o.xi >= 0.0 && o.v0 >= 0.0 &&
                      o.rho > -1.0 && o.rho < 1.0))
                    fail(OFG_INVALID_INPUT, "invalid Heston parameters for Bermudan LSM dynamics");
                vol = 0.0;
            } else {
                fail(OFG_INVALID_INPUT, "unknown LSM dynamics");
            }
            h.bermudan = true;
            h.times = ex_times;
            h.strikes = ex_strikes;
        } else {
            fail(OFG_INVALID_INPUT, "unknown LSM instrument");
        }
        fill_path_constants(d.m, o, vol, mu, rng_mode);
        if (rng_mode == OFG_RNG_FAST && dynamics != ofg::LSM_HESTON &&
            !(std::fabs(d.m.drift2) + 8.6 * d.m.diffusion2 < OFG_EXP2_256_YMAX))
            fail(OFG_INVALID_INPUT,
                 "fast rng mode needs per-step log-returns below 2.9e6 in magnitude");
        d.dynamics = dynamics;
        d.put = o.option_type == OFG_PUT;
        d.n_ex = (int32_t)ex_steps.size();
        d.paths = paths;
        d.ppt = tile_ppt(paths);
        if ((paths + (int64_t)ofg::LSM_BLOCK * d.ppt - 1) / ((int64_t)ofg::LSM_BLOCK * d.ppt) >
            ofg::LSM_MAX_TILES)
            fail(OFG_INVALID_INPUT, "num_paths must be <= 16777216 per option");
        const int64_t tile_paths = (int64_t)ofg::LSM_BLOCK * d.ppt;
        d.n_tiles = (int32_t)((paths + tile_paths - 1) / tile_paths);
        d.tile_first = (int32_t)p.tile_opt.size();
        d.date_offset = (int32_t)p.dates.size();
        d.term_gap = (int32_t)(d.n_ex ? steps - ex_steps.back() : steps);
        d.s_offset = p.spots;
        d.v_offset = p.values;
        d.seed_offset = p.seeds;
        d.normals_offset = p.normals;
        d.disc = std::exp(-o.rate * d.m.dt);
        d.terminal_strike = h.terminal_strike;
        for (int32_t k = 0; k < d.n_ex; ++k) {
            ofg::LsmDate dt{};
            dt.step = (int32_t)ex_steps[(size_t)k];
            dt.gap = (int32_t)(k == 0 ? ex_steps[0] : ex_steps[(size_t)k] - ex_steps[(size_t)k - 1]);
            dt.strike = ex_strikes[(size_t)k];
            p.dates.push_back(dt);
        }
        const int32_t opt_index = (int32_t)p.opts.size();
        for (int32_t t = 0; t < d.n_tiles; ++t) p.tile_opt.push_back(opt_index);
        p.spots += (int64_t)d.n_ex * paths;
        p.values += d.barrier ? 0 : paths;
        p.normals += steps * paths;
        p.max_rounds = std::max(p.max_rounds, d.barrier ? 0 : d.n_ex);

        if (rng_mode == OFG_RNG_REFERENCE) {
            // one stream state per path thread, its paths consecutive in
            // the option's stream
            const ofg::LsmSeedJob j = make_thread_seed_job(
                o.seed, paths, steps * (dynamics == ofg::LSM_HESTON ? 2 : 1), d.ppt, p.seeds,
                p.seed_threads, p.polys, poly_cache);
            p.jobs.push_back(j);
            p.seed_threads += seed_job_threads(j);
            p.seeds += 4 * (int64_t)d.n_tiles * ofg::LSM_BLOCK;
        }
        h.simulate = true;
        h.dev = opt_index;
        p.opts.push_back(d);
    }
    return p;
}

template <class T>
T *upload(DevBuf &b, const std::vector<T> &v, hipStream_t s, const char *what) {
    b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty())
        hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s),
                  what);
    return b.as<T>();
}

int32_t lsm_price_impl(ofg_ctx *ctx, const ofg_lsm_option *opts, size_t n,
                       const ofg_engine_cfg *cfg, ofg_result *out, ofg_lsm_boundary *boundary,
                       int64_t *boundary_len) {
    if (!ctx || !cfg || (n && (!opts || !out))) fail(OFG_INVALID_INPUT, "null argument");
    if (boundary && !boundary_len) fail(OFG_INVALID_INPUT, "boundary needs boundary_len");
    const int32_t rng = cfg->rng_mode;
    if (rng != OFG_RNG_REFERENCE && rng != OFG_RNG_EXTERNAL && rng != OFG_RNG_FAST)
        fail(OFG_INVALID_INPUT, "unknown rng mode");
    for (size_t i = 0; i < n; ++i) {
        std::memset(&out[i], 0, sizeof out[i]);
        if (boundary_len) boundary_len[i] = 0;
    }
    LsmPlan p = build_lsm_plan(opts, n, rng);
    bool any_heston = false;
    for (const auto &d : p.opts) any_heston |= d.dynamics == ofg::LSM_HESTON && !d.barrier;
    if (rng == OFG_RNG_EXTERNAL && !p.opts.empty()) {
        if (!cfg->normals1) fail(OFG_INVALID_INPUT, "external rng mode needs normals1");
        if (any_heston && !cfg->normals2)
            fail(OFG_INVALID_INPUT, "external rng mode needs normals2 for Heston dynamics");
    }

    std::vector<ofg::LsmOptOut> opt_out(p.opts.size());
    std::vector<ofg::LsmDateOut> date_out(p.dates.size());
    float ms = 0.f;
    if (!p.opts.empty()) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!ctx->lsm) ctx->lsm = new LsmBuffers();
        LsmBuffers &b = *ctx->lsm;
        hipStream_t s = ctx->stream;
        trace("lsm: upload");
        ofg::LsmArgs a{};
        a.opts = upload(b.opts, p.opts, s, "H2D lsm options");
        a.tile_opt = upload(b.tile_opt, p.tile_opt, s, "H2D lsm tiles");
        a.dates = upload(b.dates, p.dates, s, "H2D lsm dates");
        b.date_out.ensure(std::max<size_t>(p.dates.size(), 1) * sizeof(ofg::LsmDateOut));
        b.opt_out.ensure(p.opts.size() * sizeof(ofg::LsmOptOut));
        b.partials.ensure(p.tile_opt.size() * sizeof(ofg::LsmPartial));
        b.spots.ensure((size_t)std::max<int64_t>(p.spots, 1) * sizeof(double));
        b.values.ensure((size_t)std::max<int64_t>(p.values, 1) * sizeof(double));
        a.date_out = b.date_out.as<ofg::LsmDateOut>();
        a.opt_out = b.opt_out.as<ofg::LsmOptOut>();
        a.partials = b.partials.as<ofg::LsmPartial>();
        a.n_opts = (int32_t)p.opts.size();
        a.spots = b.spots.as<double>();
        a.values = b.values.as<double>();
        a.rng_mode = rng;
        a.n_tiles = (int32_t)p.tile_opt.size();
        if (rng == OFG_RNG_REFERENCE) {
            b.seeds.ensure((size_t)p.seeds * sizeof(uint64_t));
            a.seeds = b.seeds.as<uint64_t>();
        } else if (rng == OFG_RNG_EXTERNAL) {
            const size_t bytes = (size_t)p.normals * sizeof(double);
            b.n1.ensure(bytes);
            hip_check(hipMemcpyAsync(b.n1.p, cfg->normals1, bytes, hipMemcpyHostToDevice, s),
                      "H2D normals1");
            a.normals1 = b.n1.as<double>();
            if (any_heston) {
                b.n2.ensure(bytes);
                hip_check(hipMemcpyAsync(b.n2.p, cfg->normals2, bytes, hipMemcpyHostToDevice, s),
                          "H2D normals2");
                a.normals2 = b.n2.as<double>();
            }
        }
        const ofg::LsmSeedJob *jobs = nullptr;
        const uint64_t *polys = nullptr;
        if (rng == OFG_RNG_REFERENCE) {
            jobs = upload(b.jobs, p.jobs, s, "H2D seed jobs");
            polys = upload(b.polys, p.polys, s, "H2D jump polynomials");
        }
        hip_check(hipEventRecord(ctx->ev0, s), "event");
        if (rng == OFG_RNG_REFERENCE)
            hip_check(ofg::launch_lsm_seeds(jobs, (int32_t)p.jobs.size(), p.seed_threads, polys,
                                            b.seeds.as<uint64_t>(), s),
                      "lsm seed kernel launch");
        trace("lsm: launches");
        a.launch = 0;
        hip_check(ofg::launch_lsm(a, s), "lsm path kernel launch");
        hip_check(hipEventRecord(ctx->ev1, s), "event");
        for (int32_t r = 1; r <= p.max_rounds; ++r) {
            a.launch = r;
            hip_check(ofg::launch_lsm(a, s), "lsm round launch");
        }
        hip_check(hipEventRecord(ctx->ev2, s), "event");
        hip_check(hipMemcpyAsync(opt_out.data(), a.opt_out, opt_out.size() * sizeof(opt_out[0]),
                                 hipMemcpyDeviceToHost, s),
                  "D2H lsm results");
        if (!date_out.empty())
            hip_check(hipMemcpyAsync(date_out.data(), a.date_out,
                                     date_out.size() * sizeof(date_out[0]), hipMemcpyDeviceToHost, s),
                      "D2H lsm dates");
        hip_check(hipStreamSynchronize(s), "lsm kernel execution");
        trace("lsm: synchronized");
        float ms_paths = 0.f;
        hip_check(hipEventElapsedTime(&ms_paths, ctx->ev0, ctx->ev1), "event time");
        hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev2), "event time");
        b.paths_ms = ms_paths;
        b.rounds_ms = ms - ms_paths;
        b.rounds = p.max_rounds;
    }

    int64_t slot = 0;
    for (size_t i = 0; i < n; ++i) {
        const HostOpt &h = p.host[i];
        ofg_result &r = out[i];
        r.has_stderr = 1;
        r.kernel_ms = ms;
        if (!h.simulate) {
            // expiry-0 shortcut (:469-476): empty diagnostics
            r.price = h.shortcut_price;
            r.stderr_ = 0.0;
            r.num_paths = r.num_steps = r.vol = NAN;
        } else {
            const ofg::LsmOptOut &oo = opt_out[(size_t)h.dev];
            const ofg::LsmDev &d = p.opts[(size_t)h.dev];
            r.price = oo.price;
            r.stderr_ = oo.stderr_;
            r.num_paths = (double)opts[i].num_paths;
            r.num_steps = (double)opts[i].num_steps;
            r.vol = h.diag_vol;
            if (boundary && h.bermudan) {
                ofg_lsm_boundary *bo = boundary + slot;
                int64_t m = 0;
                for (int32_t k = 0; k < d.n_ex; ++k, ++m) {
                    const ofg::LsmDateOut &dd = date_out[(size_t)(d.date_offset + k)];
                    ofg_lsm_boundary &pt = bo[m];
                    std::memset(&pt, 0, sizeof pt);
                    pt.time = h.times[(size_t)k];
                    pt.strike = h.strikes[(size_t)k];
                    pt.itm_paths = (int64_t)dd.itm;
                    pt.exercised_paths = dd.valid ? (int64_t)dd.exercised : 0;
                    pt.has_boundary = pt.exercised_paths > 0;
                    pt.boundary_spot = pt.has_boundary ? dd.extreme : 0.0;
                }
                ofg_lsm_boundary &term = bo[m++];
                std::memset(&term, 0, sizeof term);
                term.time = opts[i].expiry;
                term.strike = h.terminal_strike;
                term.boundary_spot = h.terminal_strike;
                term.has_boundary = 1;
                term.itm_paths = term.exercised_paths = (int64_t)oo.term_itm;
                boundary_len[i] = m;
            }
        }
        slot += std::max<int64_t>(opts[i].num_steps, 0);
    }
    return OFG_OK;
}

}  // namespace

extern "C" {

int32_t ofg_lsm_price(ofg_ctx *ctx, const ofg_lsm_option *opts, size_t n,
                      const ofg_engine_cfg *cfg, ofg_result *out, ofg_lsm_boundary *boundary,
                      int64_t *boundary_len) {
    return guarded([&]() { return lsm_price_impl(ctx, opts, n, cfg, out, boundary, boundary_len); });
}

int32_t ofg_lsm_last_timing(ofg_ctx *ctx, double *paths_ms, double *rounds_ms,
                            int32_t *rounds) {
    return guarded([&]() -> int32_t {
        if (!ctx) fail(OFG_INVALID_INPUT, "null argument");
        std::lock_guard<std::mutex> lock(ctx->mu);
        const LsmBuffers *b = ctx->lsm;
        if (paths_ms) *paths_ms = b ? b->paths_ms : 0.0;
        if (rounds_ms) *rounds_ms = b ? b->rounds_ms : 0.0;
        if (rounds) *rounds = b ? b->rounds : 0;
        return OFG_OK;
    });
}

int32_t ofg_stream_jump(const uint64_t state[4], uint64_t k, uint64_t out[4]) {
    return guarded([&]() -> int32_t {
        if (!state || !out) fail(OFG_INVALID_INPUT, "null argument");
        uint64_t s[4] = {state[0], state[1], state[2], state[3]};
        apply_poly(x_pow(k), s);
        std::memcpy(out, s, sizeof s);
        return OFG_OK;
    });
}

}  // extern "C"
