// Host-side launch plans of the pricing kernels (declarations and scope in
// ofg_plan.hpp). Host only: no HIP calls.
#include "ofg_plan.hpp"

#include "ofg_math.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <sstream>

using namespace ofg_host;

namespace ofg_plan {

// ---------------------------------------------------------------------------
// Closed forms used by the built-in control variates. These restate the
// reference formulas so that the CV expectation is the reference's own:
// hart_norm_cdf (src/math/fast_norm.rs:27-50), bs_price generic branch
// (src/engines/analytic/black_scholes.rs:62-107) and
// geometric_asian_discrete_fixed_closed_form (src/pricing/asian.rs:86-136).
// ---------------------------------------------------------------------------
double hart_norm_cdf(double x) {
    const double P = 0.2316419;
    const double A1 = 0.319381530, A2 = -0.356563782, A3 = 1.781477937, A4 = -1.821255978,
                 A5 = 1.330274429;
    const double z = std::fabs(x);
    const double t = 1.0 / std::fma(P, z, 1.0);
    const double poly = std::fma(std::fma(std::fma(std::fma(A5, t, A4), t, A3), t, A2), t, A1) * t;
    const double pdf = 0.3989422804014327 * std::exp((-0.5) * z * z);
    const double cdf_pos = std::fma(pdf, -poly, 1.0);
    const double sign = std::signbit(x) ? 1.0 : 0.0;
    return std::fma(sign, 1.0 - 2.0 * cdf_pos, cdf_pos);
}

double vanilla_payoff(int32_t type, double spot, double strike) {
    return type == OFG_CALL ? std::fmax(spot - strike, 0.0) : std::fmax(strike - spot, 0.0);
}

double bs_price(int32_t type, double spot, double strike, double rate, double q, double vol,
                double expiry) {
    if (expiry <= 0.0) return vanilla_payoff(type, spot, strike);
    const double df_r = std::exp(-rate * expiry);
    const double df_q = std::exp(-q * expiry);
    if (vol <= 0.0) {
        return type == OFG_CALL ? std::fmax(spot * df_q - strike * df_r, 0.0)
                                : std::fmax(strike * df_r - spot * df_q, 0.0);
    }
    const double sig_sqrt_t = vol * std::sqrt(expiry);
    const double d1 =
        (std::log(spot / strike) + std::fma(0.5 * vol, vol, rate - q) * expiry) / sig_sqrt_t;
    const double d2 = d1 - sig_sqrt_t;
    const double nd1 = hart_norm_cdf(d1), nd2 = hart_norm_cdf(d2);
    const double s_df_q = spot * df_q;
    const double k_df_r = strike * df_r;
    const double call = std::fma(s_df_q, nd1, -(k_df_r * nd2));
    return type == OFG_CALL ? call : call - s_df_q + k_df_r;
}

double geometric_asian_discrete(int32_t type, double s0, double k, double r, double q,
                                double sigma, const double *obs, int64_t n_obs) {
    if (n_obs == 0) return vanilla_payoff(type, s0, k);
    double maturity = -INFINITY;
    for (int64_t i = 0; i < n_obs; ++i) maturity = std::fmax(maturity, obs[i]);
    maturity = std::fmax(maturity, 0.0);
    if (maturity <= 0.0 || sigma <= 0.0) return std::exp(-r * maturity) * vanilla_payoff(type, s0, k);
    const double n = (double)n_obs;
    double sum_t = 0.0;
    for (int64_t i = 0; i < n_obs; ++i) sum_t += obs[i];
    const double mean_t = sum_t / n;
    double cov_sum = 0.0;
    for (int64_t i = 0; i < n_obs; ++i)
        for (int64_t j = 0; j < n_obs; ++j) cov_sum += std::fmin(obs[i], obs[j]);
    const double var = sigma * sigma * cov_sum / (n * n);
    const double m = std::log(s0) + std::fma(-0.5 * sigma, sigma, r - q) * mean_t;
    const double df = std::exp(-r * maturity);
    if (var <= 0.0) return df * vanilla_payoff(type, std::exp(m), k);
    const double sqrt_v = std::sqrt(var);
    const double d1 = (m - std::log(k) + var) / sqrt_v;
    const double d2 = d1 - sqrt_v;
    const double eg = std::exp(m + 0.5 * var);
    return type == OFG_CALL ? df * (eg * hart_norm_cdf(d1) - k * hart_norm_cdf(d2))
                            : df * (k * hart_norm_cdf(-d2) - eg * hart_norm_cdf(-d1));
}

// Heston::validate (src/models/stochastic.rs:47-54)
bool heston_valid(const ofg_option &o) {
    return o.kappa > 0.0 && o.theta >= 0.0 && o.xi >= 0.0 && o.v0 >= 0.0 && o.rho > -1.0 &&
           o.rho < 1.0;
}

// payoff_from_path(&[spot]) of the maturity-zero shortcut (mc_engine.rs:896-904)
double payoff_at_spot(const ofg_option &o) {
    if (o.payoff == OFG_PAYOFF_BARRIER) {
        const bool hit = o.barrier_dir == OFG_UP ? o.spot >= o.barrier_level : o.spot <= o.barrier_level;
        const bool active = o.barrier_style == OFG_IN ? hit : !hit;
        return active ? vanilla_payoff(o.option_type, o.spot, o.strike) : o.rebate;
    }
    // Vanilla, and Asian whose average over a one-point path is the spot.
    if (o.payoff == OFG_PAYOFF_ASIAN && o.strike_type == OFG_FLOATING)
        return vanilla_payoff(o.option_type, o.spot, o.spot);
    return vanilla_payoff(o.option_type, o.spot, o.strike);
}

std::string workload_class(const std::vector<ofg::OptDev> &opts, int32_t rng_mode) {
    // Keyed by model mix, payoff mix, RNG mode and a step bucket.
    int gbm = 0, hes = 0, kinds[4] = {0, 0, 0, 0};
    double steps = 0;
    for (const auto &o : opts) {
        (o.model == OFG_MODEL_GBM ? gbm : hes)++;
        kinds[o.kind & 3]++;
        steps += o.steps;
    }
    const double mean_steps = opts.empty() ? 0 : steps / (double)opts.size();
    const int bucket = mean_steps < 64 ? 0 : mean_steps < 256 ? 1 : mean_steps < 1024 ? 2 : 3;
    std::ostringstream k;
    k << "rng" << rng_mode << "_m" << (gbm ? "g" : "") << (hes ? "h" : "") << "_k"
      << (kinds[0] ? "v" : "") << (kinds[1] ? "a" : "") << (kinds[2] ? "l" : "")
      << (kinds[3] ? "b" : "") << "_s" << bucket;
    return k.str();
}

// Builds the device descriptors, observation weights and work items.
// `vols` is Market::vol_for (flat) per option; for Heston options ignored.
Plan build_plan(const ofg_option *in, size_t n, const ofg_engine_cfg &cfg, const Tuning &tune) {
    Plan p;
    // the call's whole work in lane-steps, weighted as the items below
    double work = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = in[i];
        if (o.num_paths <= 0 || o.num_steps <= 0 || !(o.expiry > 0.0)) continue;
        const double samples = cfg.antithetic ? (double)((o.num_paths + 1) / 2) : (double)o.num_paths;
        work += samples * (double)o.num_steps * (o.model == OFG_MODEL_HESTON ? 1.9 : 1.0) *
                (cfg.antithetic ? 1.6 : 1.0);
    }
    const double item_steps = tune.auto_items
                                  ? std::min<double>(tune.item_steps, work / (64.0 * MIN_ITEMS))
                                  : (double)tune.item_steps;
    p.vols.resize(n);
    int64_t out_offset = 0, normals_offset = 0;
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = in[i];
        validate_instrument(o);
        if (o.model != OFG_MODEL_GBM && o.model != OFG_MODEL_HESTON)
            fail(OFG_INVALID_INPUT, "unknown model");
        if (o.num_paths <= 0) fail(OFG_INVALID_INPUT, "num_paths must be > 0");
        if (o.num_steps <= 0) fail(OFG_INVALID_INPUT, "num_steps must be > 0");
        if (o.num_steps > 1000000) fail(OFG_INVALID_INPUT, "num_steps must be <= 1000000");
        if (o.expiry < 0.0) fail(OFG_INVALID_INPUT, "instrument maturity must be >= 0");
        if (o.model == OFG_MODEL_HESTON && !heston_valid(o))
            fail(OFG_INVALID_INPUT, "invalid Heston parameters");
        p.vols[i] = o.vol;
        if (o.expiry == 0.0) continue;  // maturity-zero shortcut, no simulation
        if (o.model == OFG_MODEL_GBM && o.vol <= 0.0)
            fail(OFG_INVALID_INPUT, "market volatility must be > 0");

        ofg::OptDev d{};
        d.model = o.model;
        d.steps = (int32_t)o.num_steps;
        const int64_t samples = cfg.antithetic ? (o.num_paths + 1) / 2 : o.num_paths;
        d.samples = samples;
        d.out_offset = out_offset;
        d.normals_offset = normals_offset;
        d.obs_offset = (int64_t)p.obs_w.size();
        d.seed = o.seed;
        d.s0 = o.spot;
        d.strike = o.strike;
        d.mu = o.rate - o.div_yield;  // mc_engine.rs:916
        d.sigma = o.vol;
        // GbmPathGenerator::generate_into constants (simulation.rs:59-62)
        d.dt = o.expiry / (double)o.num_steps;
        d.sqrt_dt = std::sqrt(d.dt);
        d.drift = (d.mu - 0.5 * d.sigma * d.sigma) * d.dt;
        d.diffusion = d.sigma * d.sqrt_dt;
        d.kappa = o.kappa;
        d.theta = o.theta;
        d.xi = o.xi;
        d.rho = o.rho;
        d.rho_c = std::sqrt(1.0 - o.rho * o.rho);
        d.v0 = o.v0;
        d.barrier = o.barrier_level;
        d.rebate = o.rebate;
        d.nobs = (double)std::max<int32_t>(o.n_obs, 0);
        d.log_s0 = std::log(std::fmax(o.spot, 1e-12));
        d.df = std::exp(-o.rate * o.expiry);  // mc_engine.rs:945
        d.rate = o.rate;
        d.kdt = o.kappa * d.dt;
        d.kthdt = o.kappa * o.theta * d.dt;
        d.xsdt = o.xi * d.sqrt_dt;
        d.mudt = d.mu * d.dt;
        d.hdt = 0.5 * d.dt;
        d.drift2 = d.drift * OFG_INV_LN2_256;
        d.diffusion2 = d.diffusion * OFG_INV_LN2_256;
        d.sqrt_dt2 = d.sqrt_dt * OFG_INV_LN2_256;
        d.mudt2 = d.mudt * OFG_INV_LN2_256;
        d.hdt2 = d.hdt * OFG_INV_LN2_256;
        // fast GBM steps evaluate 2^(y/256), y = diffusion2 z + drift2 with
        // |z| <= sqrt(106 ln 2) < 8.6; keep |y| inside its domain
        if (cfg.rng_mode == OFG_RNG_FAST && o.model == OFG_MODEL_GBM &&
            !(std::fabs(d.drift2) + 8.6 * d.diffusion2 < OFG_EXP2_256_YMAX))
            fail(OFG_INVALID_INPUT,
                 "fast rng mode needs per-step log-returns below 2.9e6 in magnitude");
        d.antithetic = cfg.antithetic ? 1 : 0;
        int32_t flags = 0;
        if (o.option_type == OFG_PUT) flags |= ofg::F_PUT;

        // observation weights over path indices 0..steps
        const int64_t len = o.num_steps + 1;
        if (o.payoff == OFG_PAYOFF_ASIAN) {
            std::vector<double> w((size_t)len, 0.0);
            const double last_idx = (double)(len - 1);
            for (int32_t k = 0; k < o.n_obs; ++k) {
                // average_for_observations index map (mc_engine.rs:103-104)
                const double x = std::round((o.obs_times[k] / o.expiry) * last_idx);
                int64_t idx = !(x > 0.0) ? 0 : (x >= 9.2e18 ? INT64_MAX : (int64_t)x);
                idx = std::min<int64_t>(idx, len - 1);
                w[(size_t)idx] += 1.0;
            }
            double jsum = 0.0;
            for (int64_t j = 0; j < len; ++j) jsum += (double)j * w[(size_t)j];
            d.obs_jsum = jsum;
            p.obs_w.insert(p.obs_w.end(), w.begin(), w.end());
        } else {
            p.obs_w.push_back(0.0);  // keep offsets valid
        }

        // payoff class and control variate (mc_engine.rs:441-584, 934-943)
        bool need_log = false;
        if (o.payoff == OFG_PAYOFF_VANILLA) {
            d.kind = ofg::PK_VANILLA;
        } else if (o.payoff == OFG_PAYOFF_BARRIER) {
            d.kind = o.barrier_dir == OFG_DOWN ? ofg::PK_BARRIER_DOWN : ofg::PK_BARRIER;
            if (o.barrier_dir == OFG_DOWN) flags |= ofg::F_DOWN;
            if (o.barrier_style == OFG_IN) flags |= ofg::F_KNOCK_IN;
        } else {
            d.kind = ofg::PK_ASIAN;
            if (o.strike_type == OFG_FLOATING) flags |= ofg::F_FLOATING;
            if (o.averaging == OFG_GEOMETRIC) {
                flags |= ofg::F_GEOMETRIC;
                need_log = true;
            }
        }
        if (cfg.control_variate && o.model == OFG_MODEL_GBM) {
            const double df = d.df;
            if (o.payoff == OFG_PAYOFF_VANILLA) {
                // black_scholes_price_with_dividend (mc_engine.rs:66-77, 466-474)
                const double adj = o.spot * std::exp(-o.div_yield * o.expiry);
                const double e = bs_price(o.option_type, adj, o.strike, o.rate, 0.0, o.vol, o.expiry);
                d.cv_expected = e / df;
                d.has_cv = 1;
                flags |= ofg::F_CV;
            } else if (o.payoff == OFG_PAYOFF_ASIAN && o.averaging == OFG_ARITHMETIC &&
                       o.strike_type == OFG_FIXED) {
                const double e = geometric_asian_discrete(o.option_type, o.spot, o.strike, o.rate,
                                                          o.div_yield, o.vol, o.obs_times, o.n_obs);
                d.cv_expected = e / df;
                d.has_cv = 1;
                flags |= ofg::F_CV | ofg::F_CV_GEO;
                need_log = true;
            }
        }
        if (need_log) d.kind = ofg::PK_ASIAN_LOG;
        d.flags = flags;
        if (o.model == OFG_MODEL_HESTON) p.any_heston = true;

        const int32_t plan_idx = (int32_t)p.opts.size();
        p.opts.push_back(d);
        p.active.push_back((int32_t)i);

        // work items: samples_per_lane chosen so an item is ~item_steps
        // lane-steps of work (Heston and antithetic legs weigh more).
        const double leg_w = (o.model == OFG_MODEL_HESTON ? 1.9 : 1.0) * (cfg.antithetic ? 1.6 : 1.0);
        const double per_sample = (double)o.num_steps * leg_w;
        int64_t spl = (int64_t)std::llround(item_steps / per_sample);
        spl = std::max<int64_t>(1, std::min<int64_t>(spl, 1024));
        const int64_t item_samples = 64 * spl;
        // refuse before building the queue: 2^27 items is already 4 GiB of host plan
        if ((samples + item_samples - 1) / item_samples > ((int64_t)1 << 27) - (int64_t)p.items.size())
            fail(OFG_INVALID_INPUT, "too many work items");
        p.item_first.push_back((int64_t)p.items.size());
        int32_t count = 0;
        for (int64_t b = 0; b < samples; b += item_samples) {
            ofg::WorkItem wi{};
            wi.option = plan_idx;
            wi.begin = b;
            wi.end = std::min(samples, b + item_samples);
            p.items.push_back(wi);
            ++count;
        }
        p.item_count.push_back(count);
        out_offset += samples;
        normals_offset += samples * o.num_steps;
    }
    if (p.items.size() > (size_t)INT32_MAX) fail(OFG_INVALID_INPUT, "too many work items");
    p.total_samples = out_offset;
    p.normals_elems = normals_offset;
    // cost-descending queue order (longest-processing-time first)
    std::vector<double> cost(p.items.size());
    for (size_t k = 0; k < p.items.size(); ++k) {
        const auto &wi = p.items[k];
        const auto &d = p.opts[(size_t)wi.option];
        const double leg_w = (d.model == OFG_MODEL_HESTON ? 1.9 : 1.0) * (d.antithetic ? 1.6 : 1.0);
        const double lanes_steps = std::ceil((double)(wi.end - wi.begin) / 64.0) * d.steps;
        cost[k] = lanes_steps * leg_w;
    }
    p.order.resize(p.items.size());
    std::iota(p.order.begin(), p.order.end(), 0);
    auto model_of = [&](int32_t k) { return p.opts[(size_t)p.items[(size_t)k].option].model; };
    std::stable_sort(p.order.begin(), p.order.end(), [&](int32_t a, int32_t b) {
        const int ma = model_of(a), mb = model_of(b);
        if (ma != mb) return ma < mb;
        return cost[(size_t)a] > cost[(size_t)b];
    });
    p.antithetic = cfg.antithetic ? 1 : 0;
    for (int m = 0; m < 2; ++m) {
        auto lo = std::find_if(p.order.begin(), p.order.end(), [&](int32_t k) { return model_of(k) == m; });
        auto hi = std::find_if(lo, p.order.end(), [&](int32_t k) { return model_of(k) != m; });
        p.model_begin[m] = (int32_t)(lo - p.order.begin());
        p.model_end[m] = (int32_t)(hi - p.order.begin());
    }
    p.obs_w.insert(p.obs_w.end(), OBS_PAD, 0.0);
    return p;
}

}  // namespace ofg_plan

// ---------------------------------------------------------------------------
// Validation (reference error strings)
// ---------------------------------------------------------------------------
void ofg_host::validate_instrument(const ofg_option &o) {
    switch (o.payoff) {
    case OFG_PAYOFF_VANILLA:  // instruments/vanilla.rs:69-96
        if (!(o.strike > 0.0)) fail(OFG_INVALID_INPUT, "vanilla strike must be > 0");
        if (o.expiry < 0.0) fail(OFG_INVALID_INPUT, "vanilla expiry must be >= 0");
        break;
    case OFG_PAYOFF_ASIAN:  // instruments/asian.rs:39-65
        if (o.expiry < 0.0) fail(OFG_INVALID_INPUT, "asian expiry must be >= 0");
        if (!(o.strike > 0.0)) fail(OFG_INVALID_INPUT, "asian strike must be > 0");
        if (o.n_obs <= 0 || o.obs_times == nullptr)
            fail(OFG_INVALID_INPUT, "asian observation_times cannot be empty");
        for (int32_t i = 0; i < o.n_obs; ++i) {
            if (o.obs_times[i] < 0.0 || o.obs_times[i] > o.expiry)
                fail(OFG_INVALID_INPUT, "asian observation_times must lie in [0, expiry]");
        }
        if (o.averaging != OFG_ARITHMETIC && o.averaging != OFG_GEOMETRIC)
            fail(OFG_INVALID_INPUT, "unknown averaging");
        if (o.strike_type != OFG_FIXED && o.strike_type != OFG_FLOATING)
            fail(OFG_INVALID_INPUT, "unknown strike type");
        break;
    case OFG_PAYOFF_BARRIER:  // instruments/barrier.rs:36-58
        if (!(o.strike > 0.0)) fail(OFG_INVALID_INPUT, "barrier strike must be > 0");
        if (o.expiry < 0.0) fail(OFG_INVALID_INPUT, "barrier expiry must be >= 0");
        if (!(o.barrier_level > 0.0)) fail(OFG_INVALID_INPUT, "barrier level must be > 0");
        if (o.rebate < 0.0) fail(OFG_INVALID_INPUT, "barrier rebate must be >= 0");
        if (o.barrier_dir != OFG_UP && o.barrier_dir != OFG_DOWN)
            fail(OFG_INVALID_INPUT, "unknown barrier direction");
        if (o.barrier_style != OFG_IN && o.barrier_style != OFG_OUT)
            fail(OFG_INVALID_INPUT, "unknown barrier style");
        break;
    default:
        fail(OFG_INVALID_INPUT, "unknown payoff kind");
    }
    if (o.option_type != OFG_CALL && o.option_type != OFG_PUT)
        fail(OFG_INVALID_INPUT, "unknown option type");
}
