// Host side of the openferric_amd C ABI (include/openferric_amd.h).
//
// Responsibilities: input validation with the reference's error semantics
// (src/engines/monte_carlo/mc_engine.rs:875-912, src/instruments/*.rs
// validate()), the instrument -> path-generator wiring of
// MonteCarloPricingEngine::price (mc_engine.rs:889-945), the built-in control
// variates (mc_engine.rs:458-485, 546-583), building the device work queue,
// launching the kernels on the context's stream, hipGraph capture of batches,
// and the autotuner. No exception crosses the C ABI.
#include "../../include/openferric_amd.h"
#include "ofg_internal.hpp"
#include "ofg_math.hpp"
#include "ofg_adjoint.hpp"
#include "ofg_mcgreeks.hpp"
#include "ofg_pathwise.hpp"
#include "ofg_host.hpp"
#include "ofg_plan.hpp"
#include "ofg_normals.hpp"
#include "ofg_device_io.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace ofg_host;
using namespace ofg_plan;

struct ofg_batch {
    ofg_ctx *ctx = nullptr;
    Plan plan;
    ofg_engine_cfg cfg{};
    DevBuf opts, obs_w, items, order, item_first, item_count, partials, outs, head;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    ofg::LaunchArgs args{};
    Tuning tune;
    std::vector<ofg::OptOut> host_out;
    std::vector<ofg_result> results;
    double last_ms = 0.0, last_pricing_ms = 0.0;
    bool use_graph = false;
    bool ran = false;
};

namespace {

template <class T>
void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    b.ensure(std::max<size_t>(v.size() * sizeof(T), 16));
    if (!v.empty())
        hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync H2D");
}

void fill_result(ofg_result &r, const ofg_option &o, const ofg::OptOut *oo, double ms) {
    std::memset(&r, 0, sizeof r);
    r.num_paths = (double)o.num_paths;
    r.num_steps = (double)o.num_steps;
    r.vol = o.vol;
    r.kernel_ms = ms;
    if (oo == nullptr) {  // maturity-zero shortcut: no diagnostics, stderr 0
        r.price = payoff_at_spot(o);
        r.stderr_ = 0.0;
        r.has_stderr = 1;
        r.num_paths = r.num_steps = r.vol = 0.0;
        return;
    }
    r.price = oo->price;
    r.stderr_ = oo->stderr_;
    r.has_stderr = 1;
}

Tuning tuning_for(ofg_ctx *ctx, const std::vector<ofg::OptDev> &opts, const ofg_engine_cfg &cfg) {
    Tuning t;
    auto it = ctx->tuned.find(workload_class(opts, cfg.rng_mode));
    if (it != ctx->tuned.end()) t = it->second;
    if (cfg.tune_block > 0) t.block = cfg.tune_block;
    if (cfg.tune_item_steps > 0) {
        t.item_steps = cfg.tune_item_steps;
        t.auto_items = false;
    }
    return t;
}

// Enqueues the queue reset and one persistent pricing launch per model.
void enqueue_pricing(const ofg::LaunchArgs &base, const Plan &p, const Tuning &t, int device,
                     uint32_t *heads, hipStream_t s) {
    hip_check(ofg::launch_reset_queue(heads, 2, s), "reset queue");
    for (int m = 0; m < 2; ++m) {
        const int32_t n = p.model_end[m] - p.model_begin[m];
        if (n <= 0) continue;
        ofg::LaunchArgs a = base;
        a.model = m;
        a.antithetic = p.antithetic;
        a.order_begin = p.model_begin[m];
        a.order_end = p.model_end[m];
        a.queue_head = heads + m;
        a.block = t.block;
        const int full = ofg::persistent_grid(t.block, a.rng_mode, m, a.antithetic, device);
        const int wpb = t.block / 64;
        a.grid = std::max(1, (int)std::min<int64_t>(full, ((int64_t)n + wpb - 1) / wpb));
        hip_check(ofg::launch_pricing(a, s), "pricing kernel launch");
    }
}

// Runs a plan on the context stream; outs receives one OptOut per plan option.
double run_plan(ofg_ctx *ctx, const Plan &p, const ofg_engine_cfg &cfg, const Tuning &t,
                std::vector<ofg::OptOut> &outs, double *xs_host, double *ys_host,
                const double *n1_host, const double *n2_host, double *pricing_ms) {
    hipStream_t s = ctx->stream;
    const size_t nopt = p.opts.size();
    outs.assign(nopt, ofg::OptOut{});
    if (nopt == 0) return 0.0;
    upload(ctx->opts, p.opts, s);
    upload(ctx->obs_w, p.obs_w, s);
    upload(ctx->items, p.items, s);
    upload(ctx->order, p.order, s);
    upload(ctx->item_first, p.item_first, s);
    upload(ctx->item_count, p.item_count, s);
    ctx->partials.ensure(p.items.size() * sizeof(ofg::Partial));
    ctx->outs.ensure(nopt * sizeof(ofg::OptOut));
    ctx->head.ensure(256);
    double *xs_dev = nullptr, *ys_dev = nullptr;
    if (xs_host) {
        ctx->xs.ensure((size_t)p.total_samples * sizeof(double));
        xs_dev = ctx->xs.as<double>();
    }
    if (ys_host) {
        ctx->ys.ensure((size_t)p.total_samples * sizeof(double));
        ys_dev = ctx->ys.as<double>();
    }
    const double *n1_dev = nullptr, *n2_dev = nullptr;
    if (cfg.rng_mode == OFG_RNG_EXTERNAL) {
        if (!n1_host) fail(OFG_INVALID_INPUT, "external normals required for OFG_RNG_EXTERNAL");
        if (p.any_heston && !n2_host)
            fail(OFG_INVALID_INPUT, "Heston needs a second external normal stream");
        const size_t bytes = (size_t)p.normals_elems * sizeof(double);
        ctx->n1.ensure(bytes);
        hip_check(hipMemcpyAsync(ctx->n1.p, n1_host, bytes, hipMemcpyHostToDevice, s), "H2D normals");
        n1_dev = ctx->n1.as<double>();
        if (n2_host) {
            ctx->n2.ensure(bytes);
            hip_check(hipMemcpyAsync(ctx->n2.p, n2_host, bytes, hipMemcpyHostToDevice, s), "H2D normals");
            n2_dev = ctx->n2.as<double>();
        }
    }
    ofg::LaunchArgs a{};
    a.opts = ctx->opts.as<ofg::OptDev>();
    a.items = ctx->items.as<ofg::WorkItem>();
    a.order = ctx->order.as<int32_t>();
    a.partials = ctx->partials.as<ofg::Partial>();
    a.obs_w = ctx->obs_w.as<double>();
    a.xs = xs_dev;
    a.ys = ys_dev;
    a.normals1 = n1_dev;
    a.normals2 = n2_dev;
    a.rng_mode = cfg.rng_mode;

    trace("run_plan: uploads enqueued");
    hip_check(hipEventRecord(ctx->ev0, s), "event");
    enqueue_pricing(a, p, t, ctx->device, ctx->head.as<uint32_t>(), s);
    trace("run_plan: kernels enqueued");
    hip_check(hipEventRecord(ctx->ev1, s), "event");
    hip_check(ofg::launch_reduce(a.opts, ctx->item_first.as<int64_t>(), ctx->item_count.as<int32_t>(),
                                 a.partials, ctx->outs.as<ofg::OptOut>(), (int32_t)nopt, s),
              "reduce kernel launch");
    hip_check(hipEventRecord(ctx->ev2, s), "event");
    hip_check(hipMemcpyAsync(outs.data(), ctx->outs.p, nopt * sizeof(ofg::OptOut),
                             hipMemcpyDeviceToHost, s),
              "D2H results");
    if (xs_host)
        hip_check(hipMemcpyAsync(xs_host, xs_dev, (size_t)p.total_samples * sizeof(double),
                                 hipMemcpyDeviceToHost, s), "D2H xs");
    if (ys_host)
        hip_check(hipMemcpyAsync(ys_host, ys_dev, (size_t)p.total_samples * sizeof(double),
                                 hipMemcpyDeviceToHost, s), "D2H ys");
    hip_check(hipStreamSynchronize(s), "kernel execution");
    trace("run_plan: synchronized");
    float ms_all = 0.f, ms_price = 0.f;
    hip_check(hipEventElapsedTime(&ms_price, ctx->ev0, ctx->ev1), "event time");
    hip_check(hipEventElapsedTime(&ms_all, ctx->ev0, ctx->ev2), "event time");
    if (pricing_ms) *pricing_ms = ms_price;
    return ms_all;
}


int32_t price_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                   ofg_result *out);

// MonteCarloGreeksEngine::run_estimators (mc_greeks.rs:131-243): validation
// in the reference's order, one exact-terminal step per sample on the GPU,
// then discount * sum / n per estimate.
int32_t mc_greeks_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                       double spot_bump_rel, ofg_mc_greeks *out) {
    if (!ctx || !cfg || (!opts && n) || (!out && n)) fail(OFG_INVALID_INPUT, "null argument");
    if (cfg->rng_mode < 0 || cfg->rng_mode > 3) fail(OFG_INVALID_INPUT, "unknown rng mode");
    check_tuning(*cfg);
    std::vector<ofg_option> sim;
    std::vector<size_t> sim_of;
    std::vector<double> ups, dns;
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        std::memset(&out[i], 0, sizeof out[i]);
        validate_instrument(o);
        if (o.payoff != OFG_PAYOFF_VANILLA || o.model != OFG_MODEL_GBM)
            fail(OFG_INVALID_INPUT, "MonteCarloGreeksEngine supports European vanillas under GBM only");
        if (o.num_paths <= 0) fail(OFG_INVALID_INPUT, "num_paths must be > 0");
        if (!(spot_bump_rel > 0.0) || !std::isfinite(spot_bump_rel))
            fail(OFG_INVALID_INPUT, "spot_bump_rel must be finite and > 0");
        if (o.expiry <= 0.0) continue;  // default (all-zero) estimate
        if (o.spot <= 0.0) fail(OFG_INVALID_INPUT, "market spot must be > 0");
        if (o.vol <= 0.0 || !std::isfinite(o.vol))
            fail(OFG_INVALID_INPUT, "market volatility must be finite and > 0");
        ofg_option s = o;
        s.num_steps = 1;  // exact terminal value: one step over T
        sim.push_back(s);
        sim_of.push_back(i);
        const double h = std::max(o.spot * spot_bump_rel, 1.0e-6);
        ups.push_back(o.spot + h);
        dns.push_back(std::max(o.spot - h, 1.0e-8));
    }
    if (sim.empty()) return OFG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ofg_engine_cfg c = *cfg;
    c.control_variate = 0;
    c.greek_mode = OFG_GREEKS_NONE;
    Tuning t;
    t.block = cfg->tune_block > 0 ? std::min(cfg->tune_block, 256) : 256;
    t.item_steps = cfg->tune_item_steps > 0 ? cfg->tune_item_steps : 4096;
    t.auto_items = cfg->tune_item_steps <= 0;
    Plan p = build_plan(sim.data(), sim.size(), c, t);
    const size_t nopt = p.opts.size();
    for (size_t k = 0; k < nopt; ++k) {
        p.opts[k].spot_up = ups[(size_t)p.active[k]];
        p.opts[k].spot_dn = dns[(size_t)p.active[k]];
    }
    hipStream_t s = ctx->stream;
    upload(ctx->opts, p.opts, s);
    upload(ctx->items, p.items, s);
    upload(ctx->order, p.order, s);
    upload(ctx->item_first, p.item_first, s);
    upload(ctx->item_count, p.item_count, s);
    ctx->partials.ensure(p.items.size() * sizeof(ofg::AdjPartial));
    ctx->outs.ensure(nopt * sizeof(ofg::AdjOut));
    ctx->head.ensure(256);
    const double *n1_dev = nullptr;
    if (c.rng_mode == OFG_RNG_EXTERNAL) {
        if (!c.normals1) fail(OFG_INVALID_INPUT, "external normals required for OFG_RNG_EXTERNAL");
        const size_t bytes = (size_t)p.normals_elems * sizeof(double);
        ctx->n1.ensure(bytes);
        hip_check(hipMemcpyAsync(ctx->n1.p, c.normals1, bytes, hipMemcpyHostToDevice, s),
                  "H2D normals");
        n1_dev = ctx->n1.as<double>();
    }
    ofg::AdjArgs a{};
    a.opts = ctx->opts.as<ofg::OptDev>();
    a.items = ctx->items.as<ofg::WorkItem>();
    a.order = ctx->order.as<int32_t>();
    a.order_begin = p.model_begin[OFG_MODEL_GBM];
    a.order_end = p.model_end[OFG_MODEL_GBM];
    a.queue_head = ctx->head.as<uint32_t>();
    a.partials = ctx->partials.as<ofg::AdjPartial>();
    a.normals1 = n1_dev;
    a.rng_mode = c.rng_mode;
    a.model = OFG_MODEL_GBM;
    a.antithetic = p.antithetic;
    a.block = t.block;
    const int wpb = t.block / 64;
    const int32_t cnt = a.order_end - a.order_begin;
    const int full = ofg::mc_greeks_grid(t.block, c.rng_mode, p.antithetic, ctx->device);
    a.grid = std::max(1, (int)std::min<int64_t>(full, ((int64_t)cnt + wpb - 1) / wpb));
    hip_check(ofg::launch_reset_queue(ctx->head.as<uint32_t>(), 1, s), "reset queue");
    hip_check(hipEventRecord(ctx->ev0, s), "event");
    hip_check(ofg::launch_mc_greeks(a, s), "mc greeks kernel launch");
    hip_check(ofg::launch_mc_greeks_reduce(ctx->item_first.as<int64_t>(),
                                           ctx->item_count.as<int32_t>(),
                                           ctx->partials.as<ofg::AdjPartial>(),
                                           ctx->outs.as<ofg::AdjOut>(), (int32_t)nopt, s),
              "mc greeks reduce launch");
    hip_check(hipEventRecord(ctx->ev2, s), "event");
    std::vector<ofg::AdjOut> sums(nopt);
    hip_check(hipMemcpyAsync(sums.data(), ctx->outs.p, nopt * sizeof(ofg::AdjOut),
                             hipMemcpyDeviceToHost, s),
              "D2H mc greeks");
    hip_check(hipStreamSynchronize(s), "mc greeks execution");
    float f = 0.f;
    hip_check(hipEventElapsedTime(&f, ctx->ev0, ctx->ev2), "event time");
    for (size_t k = 0; k < nopt; ++k) {
        const ofg_option &o = sim[(size_t)p.active[k]];
        ofg_mc_greeks &r = out[sim_of[(size_t)p.active[k]]];
        const double discount = std::exp(-o.rate * o.expiry);
        const double nn = sums[k].n;
        r.pathwise_delta = discount * sums[k].g[0] / nn;
        r.pathwise_gamma = discount * sums[k].g[1] / nn;
        r.lr_delta = discount * sums[k].g[2] / nn;
        r.lr_gamma = discount * sums[k].g[3] / nn;
        r.lr_vega = discount * sums[k].g[4] / nn;
        r.num_samples = nn;
        r.kernel_ms = f;
    }
    return OFG_OK;
}

// Central finite differences with common random numbers: every bumped copy
// uses the same seed, so the same streams (greeks/bsm.rs:130-212 harness,
// priced in one batched launch).
int32_t price_fd(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                 ofg_result *out) {
    const double h = cfg->bump_rel > 0 ? cfg->bump_rel : 1e-4;
    std::vector<ofg_option> all;
    // Bumped expiries carry their Asian schedule scaled by T'/T, so the
    // observation index map round((t/T) n) is unchanged (the adjoint holds it
    // fixed as well) and every fixing stays inside [0, T'].
    std::vector<std::vector<double>> scaled;
    scaled.reserve(2 * n);
    struct Spec { size_t base; int nb; };
    std::vector<Spec> spec(n);
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        spec[i].base = all.size();
        all.push_back(o);
        auto add = [&](double ofg_option::*f, double rel_abs_floor) {
            const double b = std::max(std::fabs(o.*f) * h, rel_abs_floor);
            ofg_option up = o, dn = o;
            up.*f += b;
            dn.*f -= b;
            all.push_back(up);
            all.push_back(dn);
            return b;
        };
        (void)add;
        if (o.model == OFG_MODEL_GBM) {
            add(&ofg_option::spot, 1e-8);
            add(&ofg_option::vol, 1e-8);
            add(&ofg_option::rate, h);
            add(&ofg_option::expiry, 1e-8);
            if (o.payoff == OFG_PAYOFF_ASIAN && o.n_obs > 0 && o.obs_times && o.expiry > 0.0) {
                for (int side = 0; side < 2; ++side) {
                    ofg_option &b = all[all.size() - 2 + side];
                    scaled.emplace_back(o.obs_times, o.obs_times + o.n_obs);
                    for (double &t : scaled.back()) t = std::fmin(t * (b.expiry / o.expiry), b.expiry);
                    b.obs_times = scaled.back().data();
                }
            }
            spec[i].nb = 4;
        } else {
            add(&ofg_option::spot, 1e-8);
            add(&ofg_option::v0, 1e-8);
            add(&ofg_option::kappa, 1e-8);
            add(&ofg_option::theta, 1e-8);
            add(&ofg_option::xi, 1e-8);
            add(&ofg_option::rho, 1e-8);
            spec[i].nb = 6;
        }
    }
    ofg_engine_cfg c = *cfg;
    c.greek_mode = OFG_GREEKS_NONE;
    std::vector<ofg_result> res(all.size());
    const int32_t rc = price_impl(ctx, all.data(), all.size(), &c, res.data());
    if (rc != OFG_OK) return rc;
    double ms = 0;
    for (auto &r : res) ms = std::max(ms, r.kernel_ms);
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        ofg_result r = res[spec[i].base];
        r.has_greeks = 1;
        r.kernel_ms = ms;
        auto fd = [&](int k, double ofg_option::*f, double floor_) {
            const double b = std::max(std::fabs(o.*f) * h, floor_);
            const double up = res[spec[i].base + 1 + 2 * k].price;
            const double dn = res[spec[i].base + 2 + 2 * k].price;
            return std::make_pair((up - dn) / (2.0 * b), (up - 2.0 * r.price + dn) / (b * b));
        };
        if (o.model == OFG_MODEL_GBM) {
            auto d = fd(0, &ofg_option::spot, 1e-8);
            r.delta = d.first;
            r.gamma = d.second;
            r.vega = fd(1, &ofg_option::vol, 1e-8).first;
            r.rho = fd(2, &ofg_option::rate, h).first;
            r.theta = -fd(3, &ofg_option::expiry, 1e-8).first;
        } else {
            auto d = fd(0, &ofg_option::spot, 1e-8);
            r.delta = d.first;
            r.gamma = d.second;
            r.heston_grad[0] = d.first;
            r.heston_grad[1] = fd(1, &ofg_option::v0, 1e-8).first;
            r.heston_grad[2] = fd(2, &ofg_option::kappa, 1e-8).first;
            r.heston_grad[3] = fd(3, &ofg_option::theta, 1e-8).first;
            r.heston_grad[4] = fd(4, &ofg_option::xi, 1e-8).first;
            r.heston_grad[5] = fd(5, &ofg_option::rho, 1e-8).first;
        }
        out[i] = r;
    }
    return OFG_OK;
}

// The GenericPricer revaluations of finite_difference_greeks and
// bump_and_reprice (greeks/bsm.rs:130-212): `specs` holds, per option, the
// (s, r, sigma, t) of each revaluation; all are priced in one batched launch
// on the option's own seed. prices[i * m + j] receives revaluation j of
// option i. Returns the device time.
struct Reval {
    double s, r, v, t;
};

double price_revaluations(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                          const ofg_engine_cfg *cfg, const std::vector<Reval> &specs, size_t m,
                          std::vector<double> &prices) {
    std::vector<ofg_option> all;
    all.reserve(n * m);
    std::vector<std::vector<double>> scaled;
    scaled.reserve(n * m);
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        if (o.model != OFG_MODEL_GBM)
            fail(OFG_INVALID_INPUT,
                 "finite-difference Greeks take GBM options (the pricer's sigma is the vol)");
        for (size_t j = 0; j < m; ++j) {
            const Reval &rv = specs[i * m + j];
            ofg_option b = o;
            b.spot = rv.s;
            b.rate = rv.r;
            b.vol = rv.v;
            b.expiry = rv.t;
            if (o.payoff == OFG_PAYOFF_ASIAN && o.n_obs > 0 && o.obs_times && o.expiry > 0.0 &&
                rv.t != o.expiry) {
                scaled.emplace_back(o.obs_times, o.obs_times + o.n_obs);
                for (double &x : scaled.back()) x = std::fmin(x * (rv.t / o.expiry), rv.t);
                b.obs_times = scaled.back().data();
            }
            all.push_back(b);
        }
    }
    ofg_engine_cfg c = *cfg;
    c.greek_mode = OFG_GREEKS_NONE;
    std::vector<ofg_result> res(all.size());
    const int32_t rc = price_impl(ctx, all.data(), all.size(), &c, res.data());
    if (rc != OFG_OK) fail(rc, g_last_error);
    prices.resize(all.size());
    double ms = 0.0;
    for (size_t k = 0; k < all.size(); ++k) {
        prices[k] = res[k].price;
        ms = std::max(ms, res[k].kernel_ms);
    }
    return ms;
}

void check_bumps(const ofg_option *opts, const ofg_fd_bumps *bumps, size_t n) {
    if (n && (!opts || !bumps)) fail(OFG_INVALID_INPUT, "null argument");
    for (size_t i = 0; i < n; ++i) {
        const ofg_fd_bumps &b = bumps[i];
        if (!(std::isfinite(b.ds) && std::isfinite(b.dr) && std::isfinite(b.dv) &&
              std::isfinite(b.dt)))
            fail(OFG_INVALID_INPUT, "finite-difference bumps must be finite");
    }
}

// finite_difference_greeks (greeks/bsm.rs:165-212).
int32_t fd_greeks_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                       const ofg_fd_bumps *bumps, ofg_fd_greeks *out) {
    if (!ctx || !cfg || (!out && n)) fail(OFG_INVALID_INPUT, "null argument");
    check_bumps(opts, bumps, n);
    constexpr size_t M = 13;
    std::vector<Reval> specs(n * M);
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        const ofg_fd_bumps &b = bumps[i];
        const double s = o.spot, r = o.rate, v = o.vol, t = o.expiry;
        const double s_dn = std::fmax(s - b.ds, 1e-8);
        const double v_dn = std::fmax(v - b.dv, 1e-8);
        Reval *x = &specs[i * M];
        x[0] = {s, r, v, t};
        x[1] = {s + b.ds, r, v, t};
        x[2] = {s_dn, r, v, t};
        x[3] = {s, r + b.dr, v, t};
        x[4] = {s, r - b.dr, v, t};
        x[5] = {s, r, v + b.dv, t};
        x[6] = {s, r, v_dn, t};
        x[7] = {s, r, v, t + b.dt};
        x[8] = {s, r, v, std::fmax(t - b.dt, 1e-8)};
        x[9] = {s + b.ds, r, v + b.dv, t};
        x[10] = {s + b.ds, r, v_dn, t};
        x[11] = {s_dn, r, v + b.dv, t};
        x[12] = {s_dn, r, v_dn, t};
    }
    std::vector<double> p;
    const double ms = price_revaluations(ctx, opts, n, cfg, specs, M, p);
    for (size_t i = 0; i < n; ++i) {
        const double *q = &p[i * M];
        const ofg_fd_bumps &b = bumps[i];
        const double p0 = q[0], p_s_up = q[1], p_s_dn = q[2], p_r_up = q[3], p_r_dn = q[4];
        const double p_v_up = q[5], p_v_dn = q[6], p_t_up = q[7], p_t_dn = q[8];
        const double p_uu = q[9], p_ud = q[10], p_du = q[11], p_dd = q[12];
        ofg_fd_greeks &g = out[i];
        g.delta = (p_s_up - p_s_dn) / (2.0 * b.ds);
        g.gamma = (p_s_up - 2.0 * p0 + p_s_dn) / (b.ds * b.ds);
        g.rho = (p_r_up - p_r_dn) / (2.0 * b.dr);
        g.vega = (p_v_up - p_v_dn) / (2.0 * b.dv);
        // market convention: theta is dV/d(calendar time), -dV/d(maturity)
        g.theta = -(p_t_up - p_t_dn) / (2.0 * b.dt);
        g.vanna = (p_uu - p_ud - p_du + p_dd) / (4.0 * b.ds * b.dv);
        g.volga = (p_v_up - 2.0 * p0 + p_v_dn) / (b.dv * b.dv);
        g.kernel_ms = ms;
    }
    return OFG_OK;
}

// bump_and_reprice (greeks/bsm.rs:130-163).
int32_t bump_reprice_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                          const ofg_engine_cfg *cfg, const ofg_fd_bumps *bumps, double *out) {
    if (!ctx || !cfg || (!out && n)) fail(OFG_INVALID_INPUT, "null argument");
    check_bumps(opts, bumps, n);
    constexpr size_t M = 6;
    std::vector<Reval> specs(n * M);
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        const ofg_fd_bumps &b = bumps[i];
        const double s = o.spot, r = o.rate, v = o.vol, t = o.expiry;
        Reval *x = &specs[i * M];
        x[0] = {s, r, v, t};
        x[1] = {s + b.ds, r, v, t};
        x[2] = {s, r + b.dr, v, t};
        x[3] = {s, r, v + b.dv, t};
        x[4] = {s, r, v, std::fmax(t + b.dt, 1e-8)};
        x[5] = {s - b.ds, r, v, t};
    }
    std::vector<double> p;
    price_revaluations(ctx, opts, n, cfg, specs, M, p);
    std::copy(p.begin(), p.end(), out);
    return OFG_OK;
}

// Uploads a plan and enqueues the adjoint (or, with `forward`, the forward
// pathwise) kernels of both models and their segmented reduce on the context
// stream; ctx->outs then holds one AdjOut per plan option. `staged` runs on
// the stream after the uploads (device-side staging of the descriptors).
// Returns nothing until the caller synchronizes; ev0 / ev2 bracket the work.
void enqueue_adjoint(ofg_ctx *ctx, const Plan &p, const ofg_engine_cfg &c, bool forward,
                     const Tuning &t, const std::function<void(hipStream_t)> &staged) {
    hipStream_t s = ctx->stream;
    const size_t nopt = p.opts.size();
    upload(ctx->opts, p.opts, s);
    upload(ctx->obs_w, p.obs_w, s);
    upload(ctx->items, p.items, s);
    upload(ctx->order, p.order, s);
    upload(ctx->item_first, p.item_first, s);
    upload(ctx->item_count, p.item_count, s);
    ctx->partials.ensure(p.items.size() * sizeof(ofg::AdjPartial));
    ctx->outs.ensure(nopt * sizeof(ofg::AdjOut));
    ctx->head.ensure(256);
    const double *n1_dev = nullptr, *n2_dev = nullptr;
    if (c.rng_mode == OFG_RNG_EXTERNAL) {
        if (!c.normals1) fail(OFG_INVALID_INPUT, "external normals required for OFG_RNG_EXTERNAL");
        if (p.any_heston && !c.normals2)
            fail(OFG_INVALID_INPUT, "Heston needs a second external normal stream");
        const size_t bytes = (size_t)p.normals_elems * sizeof(double);
        ctx->n1.ensure(bytes);
        hip_check(hipMemcpyAsync(ctx->n1.p, c.normals1, bytes, hipMemcpyHostToDevice, s), "H2D normals");
        n1_dev = ctx->n1.as<double>();
        if (c.normals2) {
            ctx->n2.ensure(bytes);
            hip_check(hipMemcpyAsync(ctx->n2.p, c.normals2, bytes, hipMemcpyHostToDevice, s), "H2D normals");
            n2_dev = ctx->n2.as<double>();
        }
    }
    // per-model launch shapes and checkpoint scratch
    ofg::AdjArgs la[2];
    size_t scratch_doubles = 0;
    for (int m = 0; m < 2; ++m) {
        const int32_t cnt = p.model_end[m] - p.model_begin[m];
        if (cnt <= 0) continue;
        int max_steps = 1;
        for (const auto &d : p.opts)
            if (d.model == m) max_steps = std::max(max_steps, (int)d.steps);
        const int nv = m == OFG_MODEL_HESTON ? 2 : 1;
        const int seg = ofg::adj_segment(m);
        const int64_t nseg = (max_steps + seg - 1) / seg;
        ofg::AdjArgs &a = la[m];
        a = ofg::AdjArgs{};
        a.opts = ctx->opts.as<ofg::OptDev>();
        a.items = ctx->items.as<ofg::WorkItem>();
        a.order = ctx->order.as<int32_t>();
        a.order_begin = p.model_begin[m];
        a.order_end = p.model_end[m];
        a.queue_head = ctx->head.as<uint32_t>() + m;
        a.partials = ctx->partials.as<ofg::AdjPartial>();
        a.obs_w = ctx->obs_w.as<double>();
        a.normals1 = n1_dev;
        a.normals2 = n2_dev;
        a.wave_stride = nseg * nv * 64;
        a.rng_mode = c.rng_mode;
        a.model = m;
        a.antithetic = p.antithetic;
        a.block = t.block;
        const int wpb = t.block / 64;
        const int full = forward ? ofg::pathwise_grid(t.block, c.rng_mode, m, p.antithetic, ctx->device)
                                 : ofg::adjoint_grid(t.block, c.rng_mode, m, p.antithetic, ctx->device);
        a.grid = std::max(1, (int)std::min<int64_t>(full, ((int64_t)cnt + wpb - 1) / wpb));
        if (!forward)   // no checkpoints in the forward sweep
            scratch_doubles =
                std::max(scratch_doubles, (size_t)a.grid * wpb * (size_t)a.wave_stride);
    }
    ctx->adj.ensure(scratch_doubles * sizeof(double));
    if (staged) staged(s);
    hip_check(ofg::launch_reset_queue(ctx->head.as<uint32_t>(), 2, s), "reset queue");
    hip_check(hipEventRecord(ctx->ev0, s), "event");
    for (int m = 0; m < 2; ++m) {
        if (p.model_end[m] - p.model_begin[m] <= 0) continue;
        la[m].scratch = ctx->adj.as<double>();
        if (forward)
            hip_check(ofg::launch_pathwise(la[m], s), "pathwise kernel launch");
        else
            hip_check(ofg::launch_adjoint(la[m], s), "adjoint kernel launch");
    }
    hip_check(hipEventRecord(ctx->ev1, s), "event");
    hip_check(ofg::launch_adjoint_reduce(ctx->opts.as<ofg::OptDev>(), ctx->item_first.as<int64_t>(),
                                         ctx->item_count.as<int32_t>(),
                                         ctx->partials.as<ofg::AdjPartial>(),
                                         ctx->outs.as<ofg::AdjOut>(), (int32_t)nopt, s),
              "adjoint reduce launch");
    hip_check(hipEventRecord(ctx->ev2, s), "event");
}

// Launch shape of the Greeks kernels (built for 64, 128 and 256 threads).
Tuning adjoint_tuning(const ofg_engine_cfg &cfg) {
    Tuning t;
    t.block = cfg.tune_block > 0 ? std::min(cfg.tune_block, 256) : 256;
    t.item_steps = cfg.tune_item_steps > 0 ? cfg.tune_item_steps : 4096;
    t.auto_items = cfg.tune_item_steps <= 0;
    return t;
}

// Pathwise adjoint Greeks: forward + backward sweeps per sample in one
// persistent launch per model (ofg_adjoint.hip), then a segmented reduce.
int32_t adjoint_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                     ofg_result *out) {
    if (cfg->rng_mode == OFG_RNG_PCG64)
        fail(OFG_INVALID_INPUT,
             "pathwise Greeks draw OFG_RNG_REFERENCE, OFG_RNG_EXTERNAL or OFG_RNG_FAST normals");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ofg_engine_cfg c = *cfg;
    c.control_variate = 0;  // mc_aad.rs: plain estimator of price and Greeks
    const bool forward = cfg->greek_mode == OFG_GREEKS_FORWARD;
    c.greek_mode = OFG_GREEKS_NONE;
    const Tuning t = adjoint_tuning(*cfg);
    Plan p = build_plan(opts, n, c, t);
    const size_t nopt = p.opts.size();
    std::vector<ofg::AdjOut> outs(nopt);
    double ms = 0.0;
    if (nopt) {
        enqueue_adjoint(ctx, p, c, forward, t, nullptr);
        hip_check(hipMemcpyAsync(outs.data(), ctx->outs.p, nopt * sizeof(ofg::AdjOut),
                                 hipMemcpyDeviceToHost, ctx->stream), "D2H adjoint results");
        hip_check(hipStreamSynchronize(ctx->stream), "adjoint execution");
        float f = 0.f;
        hip_check(hipEventElapsedTime(&f, ctx->ev0, ctx->ev2), "event time");
        ms = f;
    }
    std::vector<const ofg::AdjOut *> map(n, nullptr);
    for (size_t k = 0; k < p.active.size(); ++k) map[(size_t)p.active[k]] = &outs[k];
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        ofg_result &r = out[i];
        fill_result(r, o, nullptr, ms);
        r.has_greeks = 1;
        const ofg::AdjOut *a = map[i];
        if (!a) {  // maturity zero: intrinsic value, zero Greeks (mc_aad.rs:96-112)
            r.kernel_ms = ms;
            continue;
        }
        r.price = a->price;
        r.stderr_ = a->stderr_;
        r.has_stderr = 1;
        r.num_paths = (double)o.num_paths;
        r.num_steps = (double)o.num_steps;
        r.vol = o.vol;
        r.delta = a->g[0];
        r.gamma = 0.0;
        r.rho = a->g[2];
        r.theta = -a->g[3];
        if (o.model == OFG_MODEL_GBM) {
            r.vega = a->g[1];
        } else {
            r.vega = a->g[1] * 2.0 * std::sqrt(o.v0);  // d price / d sqrt(v0)
            r.heston_grad[0] = a->g[0];
            r.heston_grad[1] = a->g[1];
            r.heston_grad[2] = a->g[4];
            r.heston_grad[3] = a->g[5];
            r.heston_grad[4] = a->g[6];
            r.heston_grad[5] = a->g[7];
        }
    }
    return OFG_OK;
}

// ofg_price_device: the plan is built on the host from placeholder values of
// the fields the caller's device arrays replace (they only have to pass the
// host checks: the work queue depends on paths, steps, model and payoff
// alone), then the staging kernel overwrites every market- and model-
// dependent descriptor field from the device arrays and validates them.
int32_t price_device_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                          const ofg_engine_cfg *cfg, const ofg_device_io *io) {
    if (!ctx || !cfg || !io || (!opts && n)) fail(OFG_INVALID_INPUT, "null argument");
    if (n && !io->price) fail(OFG_INVALID_INPUT, "null argument");
    if (cfg->rng_mode < 0 || cfg->rng_mode > 3) fail(OFG_INVALID_INPUT, "unknown rng mode");
    check_tuning(*cfg);
    if (cfg->control_variate)
        fail(OFG_INVALID_INPUT,
             "ofg_price_device: control variates take host closed forms; use ofg_price");
    if (cfg->rng_mode == OFG_RNG_EXTERNAL)
        fail(OFG_INVALID_INPUT, "ofg_price_device: external normals are host arrays; use ofg_price");
    const int32_t gm = cfg->greek_mode;
    if (gm != OFG_GREEKS_NONE && gm != OFG_GREEKS_ADJOINT && gm != OFG_GREEKS_FORWARD)
        fail(OFG_INVALID_INPUT, "ofg_price_device takes greek_mode NONE, ADJOINT or FORWARD");
    const bool greeks = gm != OFG_GREEKS_NONE;
    if (greeks && cfg->rng_mode == OFG_RNG_PCG64)
        fail(OFG_INVALID_INPUT,
             "pathwise Greeks draw OFG_RNG_REFERENCE, OFG_RNG_EXTERNAL or OFG_RNG_FAST normals");
    if (n > (size_t)INT32_MAX) fail(OFG_INVALID_INPUT, "too many options");
    std::vector<ofg_option> tpl(opts, opts + n);
    std::vector<std::vector<double>> every_step(n);
    for (size_t i = 0; i < n; ++i) {
        ofg_option &o = tpl[i];
        if (io->expiry) {
            if (o.payoff == OFG_PAYOFF_ASIAN) {
                if (o.obs_times)
                    fail(OFG_INVALID_INPUT,
                         "ofg_price_device: with device maturities an Asian option observes at "
                         "every step (obs_times NULL)");
                if (o.num_steps <= 0 || o.num_steps > 1000000)
                    fail(OFG_INVALID_INPUT, "num_steps must be > 0");
                for (int64_t j = 1; j <= o.num_steps; ++j)
                    every_step[i].push_back((double)j / (double)o.num_steps);
                o.obs_times = every_step[i].data();
                o.n_obs = (int32_t)o.num_steps;
            }
            o.expiry = 1.0;
        } else if (!(o.expiry > 0.0)) {
            fail(OFG_INVALID_INPUT, "ofg_price_device prices maturities > 0");
        }
        if (io->spot) o.spot = 100.0;
        if (io->strike) o.strike = 100.0;
        if (io->rate) o.rate = 0.0;
        if (io->vol) o.vol = 0.2;
        if (io->v0) o.v0 = 0.04;
        if (io->kappa) o.kappa = 1.0;
        if (io->theta) o.theta = 0.04;
        if (io->xi) o.xi = 0.5;
        if (io->rho) o.rho = 0.0;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ofg_engine_cfg c = *cfg;
    c.greek_mode = OFG_GREEKS_NONE;
    Plan p;
    Tuning t;
    if (greeks) {
        t = adjoint_tuning(*cfg);
        p = build_plan(tpl.data(), n, c, t);
    } else {
        Plan probe = build_plan(tpl.data(), n, c, Tuning{});
        t = tuning_for(ctx, probe.opts, c);
        p = (t.item_steps == Tuning{}.item_steps) ? std::move(probe) : build_plan(tpl.data(), n, c, t);
    }
    const int32_t nplan = (int32_t)p.opts.size();
    if (nplan == 0) return OFG_OK;
    std::vector<double> host(2 * (size_t)nplan);
    for (int32_t k = 0; k < nplan; ++k) {
        const ofg_option &o = opts[(size_t)p.active[(size_t)k]];
        host[2 * (size_t)k] = o.div_yield;
        host[2 * (size_t)k + 1] = o.expiry;
    }
    hipStream_t s = ctx->stream;
    upload(ctx->stage, host, s);
    upload(ctx->active, p.active, s);
    ctx->err.ensure(2 * sizeof(int32_t));
    ofg::StageArgs sa{};
    sa.active = ctx->active.as<int32_t>();
    sa.n_plan = nplan;
    sa.rng_mode = c.rng_mode;
    sa.host = ctx->stage.as<double>();
    sa.err = ctx->err.as<int32_t>();
    auto staged = [&](hipStream_t st) {
        sa.opts = ctx->opts.as<ofg::OptDev>();
        hip_check(hipMemsetAsync(sa.err, 0, 2 * sizeof(int32_t), st), "clear error flag");
        hip_check(ofg::launch_stage(sa, *io, st), "staging kernel launch");
    };
    if (greeks) {
        enqueue_adjoint(ctx, p, c, gm == OFG_GREEKS_FORWARD, t, staged);
        hip_check(ofg::launch_scatter_adjoint(ctx->outs.as<ofg::AdjOut>(), sa.active, nplan, *io,
                                              sa.err, s),
                  "scatter launch");
    } else {
        upload(ctx->opts, p.opts, s);
        upload(ctx->obs_w, p.obs_w, s);
        upload(ctx->items, p.items, s);
        upload(ctx->order, p.order, s);
        upload(ctx->item_first, p.item_first, s);
        upload(ctx->item_count, p.item_count, s);
        ctx->partials.ensure(p.items.size() * sizeof(ofg::Partial));
        ctx->outs.ensure((size_t)nplan * sizeof(ofg::OptOut));
        ctx->head.ensure(256);
        staged(s);
        ofg::LaunchArgs a{};
        a.opts = ctx->opts.as<ofg::OptDev>();
        a.items = ctx->items.as<ofg::WorkItem>();
        a.order = ctx->order.as<int32_t>();
        a.partials = ctx->partials.as<ofg::Partial>();
        a.obs_w = ctx->obs_w.as<double>();
        a.rng_mode = c.rng_mode;
        enqueue_pricing(a, p, t, ctx->device, ctx->head.as<uint32_t>(), s);
        hip_check(ofg::launch_reduce(a.opts, ctx->item_first.as<int64_t>(),
                                     ctx->item_count.as<int32_t>(), a.partials,
                                     ctx->outs.as<ofg::OptOut>(), nplan, s),
                  "reduce kernel launch");
        hip_check(ofg::launch_scatter_price(ctx->outs.as<ofg::OptOut>(), sa.active, nplan, *io,
                                            sa.err, s),
                  "scatter launch");
    }
    int32_t err[2] = {0, 0};
    hip_check(hipMemcpyAsync(err, sa.err, sizeof err, hipMemcpyDeviceToHost, s), "D2H status");
    hip_check(hipStreamSynchronize(s), "device pricing");
    if (err[0] != ofg::DIE_NONE) {
        const ofg_option &o = opts[(size_t)err[1]];
        const char *kind = o.payoff == OFG_PAYOFF_ASIAN     ? "asian"
                           : o.payoff == OFG_PAYOFF_BARRIER ? "barrier"
                                                            : "vanilla";
        std::string m;
        switch (err[0]) {
        case ofg::DIE_STRIKE: m = std::string(kind) + " strike must be > 0"; break;
        case ofg::DIE_EXPIRY: m = "ofg_price_device prices maturities > 0"; break;
        case ofg::DIE_VOL: m = "market volatility must be > 0"; break;
        case ofg::DIE_HESTON: m = "invalid Heston parameters"; break;
        case ofg::DIE_FAST_RANGE:
            m = "fast rng mode needs per-step log-returns below 2.9e6 in magnitude";
            break;
        default: m = "device inputs must be finite"; break;
        }
        fail(OFG_INVALID_INPUT, m + " (option " + std::to_string(err[1]) + ")");
    }
    return OFG_OK;
}

int32_t price_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                   ofg_result *out) {
    if (!ctx || !cfg || (!opts && n) || (!out && n)) fail(OFG_INVALID_INPUT, "null argument");
    if (cfg->rng_mode < 0 || cfg->rng_mode > 3) fail(OFG_INVALID_INPUT, "unknown rng mode");
    check_tuning(*cfg);
    if (cfg->greek_mode == OFG_GREEKS_FD_CRN) return price_fd(ctx, opts, n, cfg, out);
    if (cfg->greek_mode == OFG_GREEKS_ADJOINT || cfg->greek_mode == OFG_GREEKS_FORWARD)
        return adjoint_impl(ctx, opts, n, cfg, out);
    std::lock_guard<std::mutex> lock(ctx->mu);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    // First pass with default tuning only to classify the workload.
    Plan probe = build_plan(opts, n, *cfg, Tuning{});
    const Tuning t = tuning_for(ctx, probe.opts, *cfg);
    Plan p = (t.item_steps == Tuning{}.item_steps) ? std::move(probe) : build_plan(opts, n, *cfg, t);
    std::vector<ofg::OptOut> outs;
    double pricing_ms = 0;
    const double ms = run_plan(ctx, p, *cfg, t, outs, nullptr, nullptr, cfg->normals1,
                               cfg->normals2, &pricing_ms);
    std::vector<const ofg::OptOut *> map(n, nullptr);
    for (size_t k = 0; k < p.active.size(); ++k) map[(size_t)p.active[k]] = &outs[k];
    for (size_t i = 0; i < n; ++i) fill_result(out[i], opts[i], map[i], ms);
    return OFG_OK;
}

// ---------------------------------------------------------------------------
// Pricing without device allocation (ofg_price_noalloc): every device buffer
// of the pricing launch -- the staged plan, the work queue and its counters,
// the per-item partial sums and the per-option results -- lives in one fixed
// region per GPU, the file-scope array `scratch`. Host code never uses the
// array's name as an address: ofg_create resolves each GPU's copy with
// hipGetSymbolAddress and the context keeps that pointer.
// ---------------------------------------------------------------------------
constexpr size_t NA_MAX_OPTIONS = OFG_NOALLOC_MAX_OPTIONS;
constexpr size_t NA_MAX_ITEMS = OFG_NOALLOC_MAX_WORK_ITEMS;
constexpr size_t NA_MAX_OBS = OFG_NOALLOC_MAX_OBS_WEIGHTS;
constexpr size_t na_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_OPTS = 0;
constexpr size_t NA_OBS = NA_OPTS + na_align(NA_MAX_OPTIONS * sizeof(ofg::OptDev));
constexpr size_t NA_ITEMS = NA_OBS + na_align((NA_MAX_OBS + OBS_PAD) * sizeof(double));
constexpr size_t NA_ORDER = NA_ITEMS + na_align(NA_MAX_ITEMS * sizeof(ofg::WorkItem));
constexpr size_t NA_FIRST = NA_ORDER + na_align(NA_MAX_ITEMS * sizeof(int32_t));
constexpr size_t NA_COUNT = NA_FIRST + na_align(NA_MAX_OPTIONS * sizeof(int64_t));
constexpr size_t NA_PARTIALS = NA_COUNT + na_align(NA_MAX_OPTIONS * sizeof(int32_t));
constexpr size_t NA_OUTS = NA_PARTIALS + na_align(NA_MAX_ITEMS * sizeof(ofg::Partial));
constexpr size_t NA_HEAD = NA_OUTS + na_align(NA_MAX_OPTIONS * sizeof(ofg::OptOut));
constexpr size_t NA_BYTES = NA_HEAD + 256;
static_assert(NA_BYTES <= ((size_t)34 << 20), "no-allocation region larger than stated");

// One copy per GPU the library runs on, held until the process ends.
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_init_mu;
void *g_na_ptr[NA_MAX_DEVICES] = {};
// Calls on one GPU share its region: they are serialized per GPU.
std::mutex g_na_mu[NA_MAX_DEVICES];

// This GPU's copy of `scratch` (the device must be current); nullptr if the
// address cannot be resolved.
void *noalloc_region(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_na_init_mu);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_na_ptr[device] = p;
    }
    return g_na_ptr[device];
}

template <class T>
void stage(unsigned char *base, size_t offset, const std::vector<T> &v, hipStream_t s) {
    if (!v.empty())
        hip_check(hipMemcpyAsync(base + offset, v.data(), v.size() * sizeof(T),
                                 hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync H2D");
}

int32_t price_noalloc_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                           const ofg_engine_cfg *cfg, ofg_result *out) {
    if (!ctx || !cfg || (!opts && n) || (!out && n)) fail(OFG_INVALID_INPUT, "null argument");
    if (cfg->rng_mode != OFG_RNG_REFERENCE && cfg->rng_mode != OFG_RNG_FAST)
        fail(OFG_INVALID_INPUT,
             "ofg_price_noalloc takes generated draws (OFG_RNG_REFERENCE or OFG_RNG_FAST)");
    if (cfg->greek_mode != OFG_GREEKS_NONE)
        fail(OFG_INVALID_INPUT, "ofg_price_noalloc prices only (greek_mode OFG_GREEKS_NONE)");
    check_tuning(*cfg);
    if (n > NA_MAX_OPTIONS)
        fail(OFG_INVALID_INPUT, "ofg_price_noalloc takes at most 4096 options per call");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    unsigned char *base = static_cast<unsigned char *>(ctx->noalloc);
    if (!base) fail(OFG_DEVICE_ERROR, "the no-allocation region is not available on this device");
    Plan probe = build_plan(opts, n, *cfg, Tuning{});
    Tuning t = tuning_for(ctx, probe.opts, *cfg);
    Plan p = (t.item_steps == Tuning{}.item_steps) ? std::move(probe) : build_plan(opts, n, *cfg, t);
    // larger work items until the queue fits its part of the region
    while (p.items.size() > NA_MAX_ITEMS) {
        if (t.item_steps >= (1 << 26))
            fail(OFG_INVALID_INPUT, "ofg_price_noalloc: the work needs more than 262144 work items");
        t.item_steps *= 2;
        t.auto_items = false;
        p = build_plan(opts, n, *cfg, t);
    }
    if (p.obs_w.size() > NA_MAX_OBS + OBS_PAD)
        fail(OFG_INVALID_INPUT,
             "ofg_price_noalloc takes at most 1048576 observation weights per call");
    const size_t nopt = p.opts.size();
    std::vector<ofg::OptOut> outs(nopt);
    double ms = 0.0;
    if (nopt) {
        std::lock_guard<std::mutex> region_lock(g_na_mu[ctx->device]);
        hipStream_t s = ctx->stream;
        stage(base, NA_OPTS, p.opts, s);
        stage(base, NA_OBS, p.obs_w, s);
        stage(base, NA_ITEMS, p.items, s);
        stage(base, NA_ORDER, p.order, s);
        stage(base, NA_FIRST, p.item_first, s);
        stage(base, NA_COUNT, p.item_count, s);
        ofg::LaunchArgs a{};
        a.opts = reinterpret_cast<const ofg::OptDev *>(base + NA_OPTS);
        a.items = reinterpret_cast<const ofg::WorkItem *>(base + NA_ITEMS);
        a.order = reinterpret_cast<const int32_t *>(base + NA_ORDER);
        a.partials = reinterpret_cast<ofg::Partial *>(base + NA_PARTIALS);
        a.obs_w = reinterpret_cast<const double *>(base + NA_OBS);
        a.rng_mode = cfg->rng_mode;
        auto *d_outs = reinterpret_cast<ofg::OptOut *>(base + NA_OUTS);
        hip_check(hipEventRecord(ctx->ev0, s), "event");
        // the queue counters are zeroed by a kernel; every item's partial and
        // every option's result is written before it is read
        enqueue_pricing(a, p, t, ctx->device, reinterpret_cast<uint32_t *>(base + NA_HEAD), s);
        hip_check(ofg::launch_reduce(a.opts, reinterpret_cast<const int64_t *>(base + NA_FIRST),
                                     reinterpret_cast<const int32_t *>(base + NA_COUNT),
                                     a.partials, d_outs, (int32_t)nopt, s),
                  "reduce kernel launch");
        hip_check(hipEventRecord(ctx->ev2, s), "event");
        hip_check(hipMemcpyAsync(outs.data(), d_outs, nopt * sizeof(ofg::OptOut),
                                 hipMemcpyDeviceToHost, s),
                  "D2H results");
        hip_check(hipStreamSynchronize(s), "kernel execution");
        float f = 0.f;
        hip_check(hipEventElapsedTime(&f, ctx->ev0, ctx->ev2), "event time");
        ms = f;
    }
    std::vector<const ofg::OptOut *> map(n, nullptr);
    for (size_t k = 0; k < p.active.size(); ++k) map[(size_t)p.active[k]] = &outs[k];
    for (size_t i = 0; i < n; ++i) fill_result(out[i], opts[i], map[i], ms);
    return OFG_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *ofg_last_error(void) { return g_last_error.c_str(); }

const char *ofg_version(void) { return "openferric_amd 0.1.0 (gfx950)"; }

void ofg_inject_alloc_failure(int32_t after) { g_alloc_fault.store(after > 0 ? after : 0); }

ofg_ctx *ofg_create(int32_t device) {
    ofg_ctx *ctx = nullptr;
    const int32_t rc = guarded([&]() -> int32_t {
        int count = 0;
        trace("create: hipGetDeviceCount");
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        trace("create: device count ok");
        if (device < 0 || device >= count) fail(OFG_DEVICE_ERROR, "no such HIP device");
        hip_check(hipSetDevice(device), "hipSetDevice");
        auto c = std::make_unique<ofg_ctx>();
        c->device = device;
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&c->ev0), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev1), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev2), "hipEventCreate");
        hip_check(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device),
                  "hipDeviceGetAttribute");
        // ofg_price_noalloc's region on this GPU (nullptr: that entry point
        // reports a device error; the others are unaffected)
        c->noalloc = noalloc_region(device);
        ctx = c.release();
        trace("create: done");
        return OFG_OK;
    });
    return rc == OFG_OK ? ctx : nullptr;
}

void ofg_destroy(ofg_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    lsm_release(ctx);
    qmc_release(ctx);
    multi_release(ctx);
    basket_release(ctx);
    dsl_release(ctx);
    exact_release(ctx);
    spread_release(ctx);
    for (DevBuf *b : {&ctx->opts, &ctx->obs_w, &ctx->items, &ctx->order, &ctx->item_first,
                      &ctx->item_count, &ctx->partials, &ctx->outs, &ctx->head, &ctx->xs, &ctx->ys,
                      &ctx->n1, &ctx->n2, &ctx->adj, &ctx->fill, &ctx->stage, &ctx->active,
                      &ctx->err})
        b->release();
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->ev2);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int32_t ofg_price(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                  ofg_result *out) {
    return guarded([&]() { return price_impl(ctx, opts, n, cfg, out); });
}

int32_t ofg_finite_difference_greeks(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                                     const ofg_engine_cfg *cfg, const ofg_fd_bumps *bumps,
                                     ofg_fd_greeks *out) {
    return guarded([&]() { return fd_greeks_impl(ctx, opts, n, cfg, bumps, out); });
}

int32_t ofg_bump_and_reprice(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                             const ofg_engine_cfg *cfg, const ofg_fd_bumps *bumps, double *out) {
    return guarded([&]() { return bump_reprice_impl(ctx, opts, n, cfg, bumps, out); });
}

int32_t ofg_fill_normals(ofg_ctx *ctx, int32_t rng_mode, uint64_t seed, uint64_t first_stream,
                         int64_t n_streams, int64_t n_per_stream, int32_t step_major,
                         double *out, double *kernel_ms) {
    return guarded([&]() -> int32_t {
        if (!ctx || (!out && n_streams > 0 && n_per_stream > 0))
            fail(OFG_INVALID_INPUT, "null argument");
        if (rng_mode != OFG_RNG_REFERENCE && rng_mode != OFG_RNG_PCG64)
            fail(OFG_INVALID_INPUT, "ofg_fill_normals draws OFG_RNG_REFERENCE or OFG_RNG_PCG64 streams");
        if (n_streams < 0 || n_per_stream < 0)
            fail(OFG_INVALID_INPUT, "stream counts must be >= 0");
        if (kernel_ms) *kernel_ms = 0.0;
        if (n_streams == 0 || n_per_stream == 0) return OFG_OK;
        if (n_streams > ((int64_t)1 << 40) || n_per_stream > ((int64_t)1 << 40) / n_streams)
            fail(OFG_INVALID_INPUT, "at most 2^40 normals per call");
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        const size_t bytes = (size_t)n_streams * (size_t)n_per_stream * sizeof(double);
        ctx->fill.ensure(bytes);
        ofg::NormalsArgs a{};
        a.out = ctx->fill.as<double>();
        a.seed = seed;
        a.first = first_stream;
        a.n_streams = n_streams;
        a.n_per_stream = n_per_stream;
        a.step_major = step_major ? 1 : 0;
        hipStream_t s = ctx->stream;
        hip_check(hipEventRecord(ctx->ev0, s), "event");
        hip_check(ofg::launch_normals(a, rng_mode, s), "normals kernel launch");
        hip_check(hipEventRecord(ctx->ev1, s), "event");
        hip_check(hipMemcpyAsync(out, a.out, bytes, hipMemcpyDeviceToHost, s), "D2H normals");
        hip_check(hipStreamSynchronize(s), "normals kernel execution");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
        if (kernel_ms) *kernel_ms = ms;
        return OFG_OK;
    });
}

int32_t ofg_price_device(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                         const ofg_engine_cfg *cfg, const ofg_device_io *io) {
    return guarded([&]() { return price_device_impl(ctx, opts, n, cfg, io); });
}

int32_t ofg_price_noalloc(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                          const ofg_engine_cfg *cfg, ofg_result *out) {
    return guarded([&]() { return price_noalloc_impl(ctx, opts, n, cfg, out); });
}

int32_t ofg_price_greeks(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                         const ofg_engine_cfg *cfg, ofg_result *out) {
    return guarded([&]() -> int32_t {
        if (!cfg) fail(OFG_INVALID_INPUT, "null argument");
        ofg_engine_cfg c = *cfg;
        c.greek_mode = OFG_GREEKS_ADJOINT;
        return price_impl(ctx, opts, n, &c, out);
    });
}

int32_t ofg_mc_greeks_estimate(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                               const ofg_engine_cfg *cfg, double spot_bump_rel,
                               ofg_mc_greeks *out) {
    return guarded([&]() { return mc_greeks_impl(ctx, opts, n, cfg, spot_bump_rel, out); });
}

int32_t ofg_sample_payoffs(ofg_ctx *ctx, const ofg_option *opt, const ofg_engine_cfg *cfg,
                           double *xs, double *ys) {
    return guarded([&]() -> int32_t {
        if (!ctx || !opt || !cfg || !xs) fail(OFG_INVALID_INPUT, "null argument");
        if (cfg->rng_mode < 0 || cfg->rng_mode > 3) fail(OFG_INVALID_INPUT, "unknown rng mode");
        check_tuning(*cfg);
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        if (opt->expiry == 0.0) fail(OFG_INVALID_INPUT, "maturity-zero option has no paths");
        Plan p = build_plan(opt, 1, *cfg, tuning_for(ctx, {}, *cfg));
        std::vector<ofg::OptOut> outs;
        run_plan(ctx, p, *cfg, tuning_for(ctx, p.opts, *cfg), outs, xs, ys, cfg->normals1,
                 cfg->normals2, nullptr);
        return OFG_OK;
    });
}

ofg_batch *ofg_batch_create(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                            const ofg_engine_cfg *cfg, int32_t use_graph) {
    ofg_batch *out = nullptr;
    guarded([&]() -> int32_t {
        if (!ctx || !opts || !cfg || n == 0) fail(OFG_INVALID_INPUT, "null argument");
        if (cfg->rng_mode < 0 || cfg->rng_mode > 3) fail(OFG_INVALID_INPUT, "unknown rng mode");
        check_tuning(*cfg);
        if (cfg->rng_mode == OFG_RNG_EXTERNAL)
            fail(OFG_INVALID_INPUT, "batches use generated normals");
        if (cfg->greek_mode != OFG_GREEKS_NONE)
            fail(OFG_INVALID_INPUT, "batches price only; use ofg_price_greeks for Greeks");
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto b = std::make_unique<ofg_batch>();
        b->ctx = ctx;
        b->cfg = *cfg;
        Plan probe = build_plan(opts, n, *cfg, Tuning{});
        const Tuning t = tuning_for(ctx, probe.opts, *cfg);
        b->plan = build_plan(opts, n, *cfg, t);
        b->results.resize(n);
        for (size_t i = 0; i < n; ++i) fill_result(b->results[i], opts[i], nullptr, 0.0);
        const Plan &p = b->plan;
        hipStream_t s = ctx->stream;
        upload(b->opts, p.opts, s);
        upload(b->obs_w, p.obs_w, s);
        upload(b->items, p.items, s);
        upload(b->order, p.order, s);
        upload(b->item_first, p.item_first, s);
        upload(b->item_count, p.item_count, s);
        b->partials.ensure(std::max<size_t>(p.items.size(), 1) * sizeof(ofg::Partial));
        b->outs.ensure(std::max<size_t>(p.opts.size(), 1) * sizeof(ofg::OptOut));
        b->head.ensure(256);
        hip_check(hipStreamSynchronize(s), "upload");
        ofg::LaunchArgs &a = b->args;
        a.opts = b->opts.as<ofg::OptDev>();
        a.items = b->items.as<ofg::WorkItem>();
        a.order = b->order.as<int32_t>();
        a.partials = b->partials.as<ofg::Partial>();
        a.obs_w = b->obs_w.as<double>();
        a.rng_mode = cfg->rng_mode;
        b->tune = t;
        hip_check(hipEventCreate(&b->ev0), "hipEventCreate");
        hip_check(hipEventCreate(&b->ev1), "hipEventCreate");
        hip_check(hipEventCreate(&b->ev2), "hipEventCreate");
        b->use_graph = use_graph != 0 && !p.opts.empty();
        if (b->use_graph) {
            // Queue reset is a kernel, not a memset node, so replays stay exact.
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
            hipError_t e1 = hipSuccess, e2 = hipSuccess;
            try {
                enqueue_pricing(a, p, t, ctx->device, b->head.as<uint32_t>(), s);
            } catch (const OfgError &) {
                e2 = hipErrorLaunchFailure;
            }
            hipError_t e3 = ofg::launch_reduce(a.opts, b->item_first.as<int64_t>(),
                                               b->item_count.as<int32_t>(), a.partials,
                                               b->outs.as<ofg::OptOut>(), (int32_t)p.opts.size(), s);
            hipError_t e4 = hipStreamEndCapture(s, &b->graph);
            hip_check(e1, "capture reset");
            hip_check(e2, "capture pricing");
            hip_check(e3, "capture reduce");
            hip_check(e4, "end capture");
            hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
        }
        out = b.release();
        return OFG_OK;
    });
    return out;
}

int32_t ofg_batch_run(ofg_batch *b, double *kernel_ms) {
    return guarded([&]() -> int32_t {
        if (!b) fail(OFG_INVALID_INPUT, "null batch");
        ofg_ctx *ctx = b->ctx;
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t s = ctx->stream;
        const Plan &p = b->plan;
        if (p.opts.empty()) {
            if (kernel_ms) *kernel_ms = 0.0;
            return OFG_OK;
        }
        hip_check(hipEventRecord(b->ev0, s), "event");
        if (b->use_graph) {
            hip_check(hipGraphLaunch(b->exec, s), "graph launch");
            hip_check(hipEventRecord(b->ev2, s), "event");
        } else {
            enqueue_pricing(b->args, p, b->tune, ctx->device, b->head.as<uint32_t>(), s);
            hip_check(hipEventRecord(b->ev1, s), "event");
            hip_check(ofg::launch_reduce(b->args.opts, b->item_first.as<int64_t>(),
                                         b->item_count.as<int32_t>(), b->args.partials,
                                         b->outs.as<ofg::OptOut>(), (int32_t)p.opts.size(), s),
                      "reduce kernel");
            hip_check(hipEventRecord(b->ev2, s), "event");
        }
        hip_check(hipStreamSynchronize(s), "batch execution");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, b->ev0, b->ev2), "event time");
        b->last_ms = ms;
        if (!b->use_graph) {
            float mp = 0.f;
            hip_check(hipEventElapsedTime(&mp, b->ev0, b->ev1), "event time");
            b->last_pricing_ms = mp;
        } else {
            b->last_pricing_ms = ms;
        }
        if (kernel_ms) *kernel_ms = ms;
        b->ran = true;
        return OFG_OK;
    });
}

int32_t ofg_batch_results(ofg_batch *b, ofg_result *out) {
    return guarded([&]() -> int32_t {
        if (!b || !out) fail(OFG_INVALID_INPUT, "null argument");
        ofg_ctx *ctx = b->ctx;
        std::lock_guard<std::mutex> lock(ctx->mu);
        const Plan &p = b->plan;
        b->host_out.resize(p.opts.size());
        if (!p.opts.empty()) {
            if (!b->ran) fail(OFG_INVALID_INPUT, "batch has not been run");
            hip_check(hipSetDevice(ctx->device), "hipSetDevice");
            hip_check(hipMemcpy(b->host_out.data(), b->outs.p, p.opts.size() * sizeof(ofg::OptOut),
                                hipMemcpyDeviceToHost),
                      "D2H batch results");
        }
        for (size_t i = 0; i < b->results.size(); ++i) out[i] = b->results[i];
        for (size_t k = 0; k < p.active.size(); ++k) {
            ofg_result &r = out[(size_t)p.active[k]];
            r.price = b->host_out[k].price;
            r.stderr_ = b->host_out[k].stderr_;
            r.has_stderr = 1;
            r.kernel_ms = b->last_ms;
        }
        return OFG_OK;
    });
}

double ofg_batch_pricing_kernel_ms(ofg_batch *b) { return b ? b->last_pricing_ms : 0.0; }

void ofg_batch_destroy(ofg_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (DevBuf *d : {&b->opts, &b->obs_w, &b->items, &b->order, &b->item_first, &b->item_count,
                      &b->partials, &b->outs, &b->head})
        d->release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev2) (void)hipEventDestroy(b->ev2);
    delete b;
}

int32_t ofg_autotune(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                     const char *cache_path, int32_t *best_block, int32_t *best_item_steps) {
    return guarded([&]() -> int32_t {
        if (!ctx || !opts || !cfg || n == 0) fail(OFG_INVALID_INPUT, "null argument");
        if (cfg->rng_mode < 0 || cfg->rng_mode > 3) fail(OFG_INVALID_INPUT, "unknown rng mode");
        check_tuning(*cfg);
        if (cfg->rng_mode == OFG_RNG_EXTERNAL) fail(OFG_INVALID_INPUT, "autotune needs generated normals");
        // Scaled-down copy: same options, path counts scaled so the sample
        // workload is ~2e10 path-steps (or the full workload if smaller):
        // large enough that the launch's tail behaves like the full batch's.
        double total = 0;
        for (size_t i = 0; i < n; ++i) total += (double)opts[i].num_paths * (double)opts[i].num_steps;
        const double scale = std::min(1.0, 2e10 / std::max(total, 1.0));
        std::vector<ofg_option> small(opts, opts + n);
        for (auto &o : small) o.num_paths = std::max<int64_t>(64, (int64_t)(o.num_paths * scale));
        ofg_engine_cfg c = *cfg;
        c.greek_mode = OFG_GREEKS_NONE;
        const int32_t blocks[] = {64, 128, 256, 512};
        const int32_t items[] = {2048, 8192, 32768};
        Tuning best;
        double best_ms = 1e300;
        std::vector<ofg_result> res(n);
        std::string cls;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            cls = workload_class(build_plan(small.data(), n, c, Tuning{}).opts, c.rng_mode);
        }
        for (int32_t bl : blocks) {
            for (int32_t is : items) {
                c.tune_block = bl;
                c.tune_item_steps = is;
                double ms = 1e300;
                for (int rep = 0; rep < 3; ++rep) {
                    const int32_t rc = price_impl(ctx, small.data(), n, &c, res.data());
                    if (rc != OFG_OK) return rc;
                    ms = std::min(ms, res[0].kernel_ms);
                }
                if (ms < best_ms) {
                    best_ms = ms;
                    best.block = bl;
                    best.item_steps = is;
                }
            }
        }
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->tuned[cls] = best;
        }
        if (cache_path) {
            std::ofstream f(cache_path, std::ios::app);
            f << cls << " " << best.block << " " << best.item_steps << " " << best_ms << "\n";
        }
        if (best_block) *best_block = best.block;
        if (best_item_steps) *best_item_steps = best.item_steps;
        return OFG_OK;
    });
}

}  // extern "C"
