// Host side of ofg_autocall_price: price_autocallable /
// price_phoenix_autocallable and their bump-and-reprice sensitivities
// (src/pricing/autocallable.rs:52-231, 366-418) for n contracts in one launch.
//
// Validation follows the reference's order and messages (Autocallable /
// PhoenixAutocallable::validate, instruments/autocallable.rs:37-98, 136-156;
// prepare_common and validate_market_inputs, pricing/autocallable.rs:
// 301-355, 602-657; simulate_autocallable_paths' counts, :449-459). As the
// reference's price_autocallable, an invalid contract yields a NaN price and
// no stderr (the message is left in ofg_last_error); the sensitivities entry
// point reports errors.
#include "../../include/openferric_amd.h"
#include "ofg_host.hpp"
#include "ofg_internal.hpp"
#include "ofg_math.hpp"
#include "ofg_multi.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <vector>

using namespace ofg_host;

namespace ofg_host {

struct MultiBuffers {
    DevBuf scen, tiles, obs, partials, out, normals, gct, gtiles, gpartials, gout;
};

void multi_release(ofg_ctx *ctx) {
    if (!ctx || !ctx->multi) return;
    MultiBuffers *b = ctx->multi;
    for (DevBuf *d : {&b->scen, &b->tiles, &b->obs, &b->partials, &b->out, &b->normals, &b->gct,
                      &b->gtiles, &b->gpartials, &b->gout})
        d->release();
    delete b;
    ctx->multi = nullptr;
}

}  // namespace ofg_host

namespace {

constexpr uint64_t AUTOCALL_MC_SEED = 94210;   // pricing/autocallable.rs:18
constexpr double SPOT_BUMP_REL = 0.01, VOL_BUMP_ABS = 0.01, CORR_BUMP = 0.01;

// cholesky_lower (autocallable.rs:669-700); false when not PSD.
bool cholesky_lower(const std::vector<double> &a, int n, std::vector<double> &l) {
    const double tol = 1.0e-12;
    l.assign((size_t)(n * n), 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double sum = a[(size_t)(i * n + j)];
            for (int k = 0; k < j; ++k) sum -= l[(size_t)(i * n + k)] * l[(size_t)(j * n + k)];
            if (i == j) {
                if (sum < -tol) return false;
                l[(size_t)(i * n + j)] = std::sqrt(std::fmax(sum, tol));
            } else if (l[(size_t)(j * n + j)] > tol) {
                l[(size_t)(i * n + j)] = sum / l[(size_t)(j * n + j)];
            }
        }
    return true;
}

void validate_contract(const ofg_autocall &c) {
    if (c.n_underlyings <= 0 || !c.underlyings)
        fail(OFG_INVALID_INPUT, "autocallable underlyings cannot be empty");
    if (c.notional <= 0.0) fail(OFG_INVALID_INPUT, "autocallable notional must be > 0");
    if (c.maturity <= 0.0) fail(OFG_INVALID_INPUT, "autocallable maturity must be > 0");
    if (c.n_dates <= 0 || !c.autocall_dates)
        fail(OFG_INVALID_INPUT, "autocallable autocall_dates cannot be empty");
    if (c.coupon_rate < 0.0) fail(OFG_INVALID_INPUT, "autocallable coupon_rate must be >= 0");
    if (c.autocall_barrier <= 0.0)
        fail(OFG_INVALID_INPUT, "autocallable autocall_barrier must be > 0");
    if (c.ki_barrier <= 0.0) fail(OFG_INVALID_INPUT, "autocallable ki_barrier must be > 0");
    if (c.ki_strike <= 0.0) fail(OFG_INVALID_INPUT, "autocallable ki_strike must be > 0");
    for (int32_t k = 0; k < c.n_dates; ++k)
        if (c.autocall_dates[k] <= 0.0 || c.autocall_dates[k] > c.maturity)
            fail(OFG_INVALID_INPUT, "autocallable dates must lie in (0, maturity]");
    for (int32_t k = 1; k < c.n_dates; ++k)
        if (c.autocall_dates[k] <= c.autocall_dates[k - 1])
            fail(OFG_INVALID_INPUT, "autocallable dates must be strictly increasing");
    std::set<int32_t> uniq(c.underlyings, c.underlyings + c.n_underlyings);
    if ((int32_t)uniq.size() != c.n_underlyings)
        fail(OFG_INVALID_INPUT, "autocallable underlyings must be unique");
    if (c.phoenix && c.coupon_barrier <= 0.0)
        fail(OFG_INVALID_INPUT, "phoenix coupon_barrier must be > 0");
}

void validate_market(const ofg_autocall &c, const std::vector<double> &spots,
                     const std::vector<double> &vols, const std::vector<double> &corr) {
    const int n = c.n_assets;
    if (n <= 0) fail(OFG_INVALID_INPUT, "autocallable spots cannot be empty");
    for (double s : spots)
        if (s <= 0.0) fail(OFG_INVALID_INPUT, "autocallable spots must be > 0");
    for (double v : vols)
        if (v < 0.0) fail(OFG_INVALID_INPUT, "autocallable vols must be >= 0");
    for (int i = 0; i < n; ++i) {
        if (std::fabs(corr[(size_t)(i * n + i)] - 1.0) > 1.0e-10)
            fail(OFG_INVALID_INPUT, "autocallable correlation matrix diagonal must be 1");
        for (int j = 0; j < n; ++j) {
            const double rho = corr[(size_t)(i * n + j)];
            if (!(rho >= -1.0 && rho <= 1.0))
                fail(OFG_INVALID_INPUT, "autocallable correlation entries must be in [-1, 1]");
            if (std::fabs(rho - corr[(size_t)(j * n + i)]) > 1.0e-10)
                fail(OFG_INVALID_INPUT, "autocallable correlation matrix must be symmetric");
        }
    }
}

// One scenario's device descriptor from (possibly bumped) market inputs;
// throws on the reference's errors for those inputs.
ofg::AcDev make_scenario(const ofg_autocall &c, const std::vector<double> &spots,
                         const std::vector<double> &vols, const std::vector<double> &corr) {
    if (c.num_steps <= 0) fail(OFG_INVALID_INPUT, "autocallable n_steps must be > 0");
    validate_market(c, spots, vols, corr);
    const int n = c.n_underlyings, na = c.n_assets;
    for (int k = 0; k < n; ++k)
        if (c.underlyings[k] < 0 || c.underlyings[k] >= na)
            fail(OFG_INVALID_INPUT, "autocallable underlying index out of range");
    if (c.num_paths <= 0) fail(OFG_INVALID_INPUT, "autocallable n_paths must be > 0");
    if (n > ofg::MA_MAX) fail(OFG_INVALID_INPUT, "at most 8 underlyings per contract");
    std::vector<double> sel((size_t)(n * n)), l;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            sel[(size_t)(i * n + j)] =
                corr[(size_t)(c.underlyings[i] * na + c.underlyings[j])];
    if (!cholesky_lower(sel, n, l))
        fail(OFG_INVALID_INPUT, "autocallable correlation matrix is not PSD");
    ofg::AcDev d{};
    d.n = n;
    d.phoenix = c.phoenix != 0;
    d.memory = c.memory != 0;
    d.steps = (int32_t)c.num_steps;
    d.paths = c.num_paths;
    const double dt = c.maturity / (double)c.num_steps;
    const double sqrt_dt = std::sqrt(dt);
    for (int i = 0; i < n; ++i) {
        const double s0 = spots[(size_t)c.underlyings[i]], v = vols[(size_t)c.underlyings[i]];
        d.s0[i] = s0;
        d.yfloor[i] = std::log(1.0e-12 / s0) * OFG_INV_LN2_256;
        d.drift[i] = (c.rate - c.div_yield - 0.5 * v * v) * dt;
        d.vol_dt[i] = v * sqrt_dt;
        d.drift2[i] = d.drift[i] * OFG_INV_LN2_256;
        d.vol_dt2[i] = d.vol_dt[i] * OFG_INV_LN2_256;
        for (int j = 0; j <= i; ++j) d.chol[i * (i + 1) / 2 + j] = l[(size_t)(i * n + j)];
    }
    d.notional = c.notional;
    d.maturity = c.maturity;
    d.autocall_barrier = c.autocall_barrier;
    d.coupon_rate = c.coupon_rate;
    d.coupon_barrier = c.coupon_barrier;
    d.ki_barrier = c.ki_barrier;
    d.ki_strike = c.ki_strike;
    d.autocall_y = std::log(c.autocall_barrier) * OFG_INV_LN2_256;
    d.coupon_y = c.coupon_barrier > 0.0 ? std::log(c.coupon_barrier) * OFG_INV_LN2_256
                                        : -INFINITY;
    d.ki_y = std::log(c.ki_barrier) * OFG_INV_LN2_256;
    d.df_maturity = std::exp(-c.rate * c.maturity);
    d.redeem_plain = c.notional * (1.0 + c.coupon_rate * c.maturity);
    d.shift = c.notional;
    d.seed = AUTOCALL_MC_SEED;
    return d;
}

// observation_schedule (autocallable.rs:582-592): step -> latest date mapped
// to it, in step order.
std::vector<ofg::AcObs> schedule(const ofg_autocall &c) {
    std::map<int64_t, double> m;
    for (int32_t k = 0; k < c.n_dates; ++k) {
        const double t = c.autocall_dates[k];
        const double raw = std::round((t / c.maturity) * (double)c.num_steps);
        int64_t step = raw > 0.0 ? (int64_t)raw : 0;
        step = std::min<int64_t>(std::max<int64_t>(step, 1), c.num_steps);
        auto it = m.find(step);
        if (it == m.end()) m[step] = t;
        else it->second = std::max(it->second, t);
    }
    std::vector<ofg::AcObs> out;
    for (const auto &e : m) {
        ofg::AcObs o{};
        o.step = (int32_t)e.first;
        o.time = e.second;
        o.df = std::exp(-c.rate * e.second);
        o.call_pv = o.df * (c.notional * (1.0 + c.coupon_rate * e.second));
        out.push_back(o);
    }
    return out;
}

// bump_corr_matrix (autocallable.rs:702-714)
std::vector<double> bump_corr(const std::vector<double> &c, int n, double b) {
    std::vector<double> out = c;
    for (int i = 0; i < n; ++i) {
        out[(size_t)(i * n + i)] = 1.0;
        for (int j = i + 1; j < n; ++j) {
            const double v = std::min(0.999, std::max(-0.999, c[(size_t)(i * n + j)] + b));
            out[(size_t)(i * n + j)] = v;
            out[(size_t)(j * n + i)] = v;
        }
    }
    return out;
}

struct ContractPlan {
    bool valid = false;
    int32_t base = -1;                       // scenario index of the price
    std::vector<int32_t> delta_up, delta_dn; // per underlying
    std::vector<double> delta_bump;
    int32_t vega_up = -1, vega_dn = -1;
    int32_t cega_up = -1, cega_dn = -1;
    double cega_bump = 0.0;
    bool sens_ok = false;
};

int32_t autocall_impl(ofg_ctx *ctx, const ofg_autocall *in, size_t n, const ofg_engine_cfg *cfg,
                      int32_t with_greeks, ofg_result *out, ofg_autocall_sens *sens) {
    if (!ctx || !cfg || (n && (!in || !out))) fail(OFG_INVALID_INPUT, "null argument");
    const int32_t rng = cfg->rng_mode;
    if (rng != OFG_RNG_EXTERNAL && rng != OFG_RNG_FAST)
        fail(OFG_INVALID_INPUT,
             "autocallables take external normals or fast draws: a called path ends early, so "
             "the reference's single stream cannot be split across paths");
    std::vector<ofg::AcDev> scen;
    std::vector<ofg::AcObs> obs;
    std::vector<ContractPlan> plans(n);
    std::string last_msg;
    int64_t normals_elems = 0;
    for (size_t i = 0; i < n; ++i) {
        const ofg_autocall &c = in[i];
        std::memset(&out[i], 0, sizeof out[i]);
        out[i].price = NAN;
        out[i].num_paths = out[i].num_steps = out[i].vol = NAN;
        if (sens) std::memset(&sens[i], 0, sizeof sens[i]);
        ContractPlan &pl = plans[i];
        const int64_t contract_normals = c.num_steps > 0 && c.num_paths > 0 && c.n_underlyings > 0
                                             ? c.num_steps * c.n_underlyings * c.num_paths
                                             : 0;
        const int64_t my_normals = normals_elems;
        normals_elems += contract_normals;
        try {
            validate_contract(c);
            if (c.n_assets <= 0 || !c.spots || !c.vols || !c.corr)
                fail(OFG_INVALID_INPUT, "autocallable spots cannot be empty");
            const int na = c.n_assets;
            std::vector<double> spots(c.spots, c.spots + na), vols(c.vols, c.vols + na),
                corr(c.corr, c.corr + (size_t)na * na);
            const int32_t obs_off = (int32_t)obs.size();
            auto push = [&](ofg::AcDev d) {
                d.obs_offset = obs_off;
                d.normals_offset = my_normals;
                scen.push_back(d);
                return (int32_t)scen.size() - 1;
            };
            ofg::AcDev base = make_scenario(c, spots, vols, corr);
            std::vector<ofg::AcObs> sch = schedule(c);
            for (auto &s : sch) obs.push_back(s);
            base.n_obs = (int32_t)sch.size();
            const int32_t n_obs = base.n_obs;
            pl.base = push(base);
            pl.valid = true;
            if (with_greeks) {
                // bump_and_reprice_sensitivities (autocallable.rs:366-418)
                try {
                    for (int k = 0; k < c.n_underlyings; ++k) {
                        const int idx = c.underlyings[k];
                        const double bump = std::max(std::fabs(spots[(size_t)idx]) * SPOT_BUMP_REL, 1.0e-4);
                        std::vector<double> up = spots, dn = spots;
                        up[(size_t)idx] += bump;
                        dn[(size_t)idx] = std::max(dn[(size_t)idx] - bump, 1.0e-8);
                        ofg::AcDev u = make_scenario(c, up, vols, corr), d = make_scenario(c, dn, vols, corr);
                        u.n_obs = d.n_obs = n_obs;
                        pl.delta_up.push_back(push(u));
                        pl.delta_dn.push_back(push(d));
                        pl.delta_bump.push_back(bump);
                    }
                    std::vector<double> vu = vols, vd = vols;
                    for (int k = 0; k < c.n_underlyings; ++k) {
                        const int idx = c.underlyings[k];
                        vu[(size_t)idx] += VOL_BUMP_ABS;
                        vd[(size_t)idx] = std::max(vd[(size_t)idx] - VOL_BUMP_ABS, 1.0e-6);
                    }
                    ofg::AcDev u = make_scenario(c, spots, vu, corr), d = make_scenario(c, spots, vd, corr);
                    u.n_obs = d.n_obs = n_obs;
                    pl.vega_up = push(u);
                    pl.vega_dn = push(d);
                    double b = CORR_BUMP;
                    for (int tries = 0; tries < 6; ++tries, b *= 0.5) {
                        try {
                            ofg::AcDev cu = make_scenario(c, spots, vols, bump_corr(corr, na, b));
                            ofg::AcDev cd = make_scenario(c, spots, vols, bump_corr(corr, na, -b));
                            cu.n_obs = cd.n_obs = n_obs;
                            pl.cega_up = push(cu);
                            pl.cega_dn = push(cd);
                            pl.cega_bump = b;
                            break;
                        } catch (const OfgError &) {
                        }
                    }
                    pl.sens_ok = true;
                } catch (const OfgError &) {
                    pl.sens_ok = false;
                }
            }
        } catch (const OfgError &e) {
            last_msg = e.what();
            pl.valid = false;
        }
    }
    if (rng == OFG_RNG_EXTERNAL && !scen.empty() && !cfg->normals1)
        fail(OFG_INVALID_INPUT, "external rng mode needs normals1");

    // sensitivities fused per path: every scenario of a contract on the same
    // draws in one pass (OFG_AUTOCALL_FUSED=0 prices them one by one)
    std::vector<int32_t> fused_slot(scen.size(), -1);
    std::vector<ofg::AcGreeksDev> gdev;
    std::vector<int32_t> gtiles;
    int32_t g_out = 0, max_n = 0;
    const char *fenv = std::getenv("OFG_AUTOCALL_FUSED");
    const bool fuse = with_greeks && !(fenv && std::atoi(fenv) == 0);
    for (size_t i = 0; fuse && i < n; ++i) {
        const ContractPlan &pl = plans[i];
        if (!pl.valid || !pl.sens_ok) continue;
        const ofg::AcDev &b0 = scen[(size_t)pl.base];
        const ofg::AcDev &vu = scen[(size_t)pl.vega_up], &vd = scen[(size_t)pl.vega_dn];
        ofg::AcGreeksDev g{};
        g.n = b0.n;
        g.phoenix = b0.phoenix;
        g.memory = b0.memory;
        g.n_obs = b0.n_obs;
        g.steps = b0.steps;
        g.obs_offset = b0.obs_offset;
        g.has_cega = pl.cega_up >= 0;
        g.n_scen = 5 + 2 * g.n;
        g.paths = b0.paths;
        g.normals_offset = b0.normals_offset;
        g.seed = b0.seed;
        const ofg::AcDev *sets[3] = {&b0, &vu, &vd};
        const ofg::AcDev *csets[3] = {&b0, g.has_cega ? &scen[(size_t)pl.cega_up] : &b0,
                                      g.has_cega ? &scen[(size_t)pl.cega_dn] : &b0};
        for (int a = 0; a < g.n; ++a) {
            g.s0[a] = b0.s0[a];
            g.s0_up[a] = scen[(size_t)pl.delta_up[(size_t)a]].s0[a];
            g.s0_dn[a] = scen[(size_t)pl.delta_dn[(size_t)a]].s0[a];
            g.yfloor[a] = b0.yfloor[a];
            g.yfloor_up[a] = scen[(size_t)pl.delta_up[(size_t)a]].yfloor[a];
            g.yfloor_dn[a] = scen[(size_t)pl.delta_dn[(size_t)a]].yfloor[a];
            for (int v = 0; v < 3; ++v) {
                g.drift[v][a] = sets[v]->drift[a];
                g.vol_dt[v][a] = sets[v]->vol_dt[a];
                g.drift2[v][a] = sets[v]->drift2[a];
                g.vol_dt2[v][a] = sets[v]->vol_dt2[a];
            }
        }
        for (int v = 0; v < 3; ++v) std::memcpy(g.chol[v], csets[v]->chol, sizeof g.chol[v]);
        g.notional = b0.notional;
        g.autocall_barrier = b0.autocall_barrier;
        g.coupon_rate = b0.coupon_rate;
        g.coupon_barrier = b0.coupon_barrier;
        g.ki_barrier = b0.ki_barrier;
        g.ki_strike = b0.ki_strike;
        g.autocall_y = b0.autocall_y;
        g.coupon_y = b0.coupon_y;
        g.ki_y = b0.ki_y;
        g.df_maturity = b0.df_maturity;
        g.redeem_plain = b0.redeem_plain;
        g.shift = b0.shift;
        const int64_t blk = ofg::MA_BLOCK;
        g.ppt = (int32_t)std::max<int64_t>(
            1, (g.paths + blk * ofg::MA_MAX_TILES - 1) / (blk * ofg::MA_MAX_TILES));
        g.n_tiles = (int32_t)((g.paths + blk * g.ppt - 1) / (blk * g.ppt));
        g.tile_first = (int32_t)gtiles.size();
        for (int32_t t = 0; t < g.n_tiles; ++t) gtiles.push_back((int32_t)gdev.size());
        g.out_first = g_out;
        g_out += ofg::ACG_MAXS;
        fused_slot[(size_t)pl.base] = g.out_first;
        fused_slot[(size_t)pl.vega_up] = g.out_first + 1;
        fused_slot[(size_t)pl.vega_dn] = g.out_first + 2;
        if (g.has_cega) {
            fused_slot[(size_t)pl.cega_up] = g.out_first + 3;
            fused_slot[(size_t)pl.cega_dn] = g.out_first + 4;
        }
        for (int k = 0; k < g.n; ++k) {
            fused_slot[(size_t)pl.delta_up[(size_t)k]] = g.out_first + 5 + 2 * k;
            fused_slot[(size_t)pl.delta_dn[(size_t)k]] = g.out_first + 6 + 2 * k;
        }
        max_n = std::max(max_n, g.n);
        gdev.push_back(g);
    }
    // the remaining scenarios one by one; tiles: the size depends on the
    // path count alone
    std::vector<int32_t> reg_index(scen.size(), -1);
    std::vector<ofg::AcDev> rscen;
    std::vector<int32_t> tiles;
    for (size_t s = 0; s < scen.size(); ++s) {
        if (fused_slot[s] >= 0) continue;
        ofg::AcDev d = scen[s];
        const int64_t blk = ofg::MA_BLOCK;
        d.ppt = (int32_t)std::max<int64_t>(1, (d.paths + blk * ofg::MA_MAX_TILES - 1) /
                                                  (blk * ofg::MA_MAX_TILES));
        d.n_tiles = (int32_t)((d.paths + blk * d.ppt - 1) / (blk * d.ppt));
        d.tile_first = (int32_t)tiles.size();
        for (int32_t t = 0; t < d.n_tiles; ++t) tiles.push_back((int32_t)rscen.size());
        reg_index[s] = (int32_t)rscen.size();
        rscen.push_back(d);
    }
    std::vector<ofg::AcOut> rres(rscen.size()), gres((size_t)g_out);
    float ms = 0.f;
    if (!rscen.empty() || !gdev.empty()) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!ctx->multi) ctx->multi = new MultiBuffers();
        MultiBuffers &b = *ctx->multi;
        hipStream_t st = ctx->stream;
        auto up = [&](DevBuf &buf, const void *src, size_t bytes, const char *what) {
            buf.ensure(std::max<size_t>(bytes, 1));
            if (bytes) hip_check(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, st), what);
        };
        up(b.scen, rscen.data(), rscen.size() * sizeof(rscen[0]), "H2D autocall scenarios");
        up(b.tiles, tiles.data(), tiles.size() * sizeof(int32_t), "H2D autocall tiles");
        up(b.obs, obs.data(), obs.size() * sizeof(obs[0]), "H2D autocall schedule");
        b.partials.ensure(tiles.size() * sizeof(ofg::AcPartial));
        b.out.ensure(rscen.size() * sizeof(ofg::AcOut));
        ofg::AcArgs a{};
        if (rng == OFG_RNG_EXTERNAL) {
            up(b.normals, cfg->normals1, (size_t)normals_elems * sizeof(double), "H2D normals1");
            a.normals = b.normals.as<double>();
        }
        a.scen = b.scen.as<ofg::AcDev>();
        a.tile_scen = b.tiles.as<int32_t>();
        a.obs = b.obs.as<ofg::AcObs>();
        a.partials = b.partials.as<ofg::AcPartial>();
        a.out = b.out.as<ofg::AcOut>();
        a.n_tiles = (int32_t)tiles.size();
        a.n_scen = (int32_t)rscen.size();
        a.rng_mode = rng;
        ofg::AcgArgs ga{};
        if (!gdev.empty()) {
            up(b.gct, gdev.data(), gdev.size() * sizeof(gdev[0]), "H2D autocall sensitivities");
            up(b.gtiles, gtiles.data(), gtiles.size() * sizeof(int32_t), "H2D autocall tiles");
            b.gpartials.ensure(gtiles.size() * ofg::ACG_MAXS * sizeof(ofg::AcPartial));
            b.gout.ensure((size_t)g_out * sizeof(ofg::AcOut));
            ga.ct = b.gct.as<ofg::AcGreeksDev>();
            ga.tile_ct = b.gtiles.as<int32_t>();
            ga.obs = a.obs;
            ga.partials = b.gpartials.as<ofg::AcPartial>();
            ga.out = b.gout.as<ofg::AcOut>();
            ga.normals = a.normals;
            ga.n_tiles = (int32_t)gtiles.size();
            ga.n_ct = (int32_t)gdev.size();
            ga.rng_mode = rng;
            ga.max_n = max_n;
        }
        hip_check(hipEventRecord(ctx->ev0, st), "event");
        if (!rscen.empty()) hip_check(ofg::launch_autocall(a, st), "autocall kernel launch");
        if (!gdev.empty())
            hip_check(ofg::launch_autocall_greeks(ga, st), "autocall sensitivity kernel launch");
        hip_check(hipEventRecord(ctx->ev1, st), "event");
        if (!rscen.empty())
            hip_check(hipMemcpyAsync(rres.data(), a.out, rres.size() * sizeof(rres[0]),
                                     hipMemcpyDeviceToHost, st),
                      "D2H autocall results");
        if (!gdev.empty())
            hip_check(hipMemcpyAsync(gres.data(), ga.out, gres.size() * sizeof(gres[0]),
                                     hipMemcpyDeviceToHost, st),
                      "D2H autocall sensitivity results");
        hip_check(hipStreamSynchronize(st), "autocall kernel execution");
        hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
    }
    std::vector<ofg::AcOut> res(scen.size());
    for (size_t sidx = 0; sidx < scen.size(); ++sidx)
        res[sidx] = fused_slot[sidx] >= 0 ? gres[(size_t)fused_slot[sidx]] : rres[(size_t)reg_index[sidx]];
    for (size_t i = 0; i < n; ++i) {
        const ContractPlan &pl = plans[i];
        out[i].kernel_ms = ms;
        if (!pl.valid) continue;
        const ofg::AcOut &r = res[(size_t)pl.base];
        out[i].price = r.price;
        out[i].stderr_ = r.stderr_;
        out[i].has_stderr = 1;
        // diagnostics num_paths, num_steps (observation_count is the caller's
        // n_dates); no vol key (autocallable.rs:86-89)
        out[i].num_paths = (double)in[i].num_paths;
        out[i].num_steps = (double)in[i].num_steps;
        if (with_greeks && pl.sens_ok) {
            double dsum = 0.0;
            ofg_autocall_sens s{};
            s.n_delta = (int32_t)pl.delta_up.size();
            for (size_t k = 0; k < pl.delta_up.size(); ++k) {
                const double d = (res[(size_t)pl.delta_up[k]].price - res[(size_t)pl.delta_dn[k]].price) /
                                 (2.0 * pl.delta_bump[k]);
                s.delta[k] = d;
                dsum += d;
            }
            s.vega = (res[(size_t)pl.vega_up].price - res[(size_t)pl.vega_dn].price) / (2.0 * VOL_BUMP_ABS);
            s.cega = pl.cega_up >= 0 ? (res[(size_t)pl.cega_up].price - res[(size_t)pl.cega_dn].price) /
                                           (2.0 * pl.cega_bump)
                                     : NAN;
            s.ok = 1;
            out[i].has_greeks = 1;
            out[i].delta = dsum;
            out[i].vega = s.vega;
            out[i].rho = s.cega;   // Greeks { rho: cega } (autocallable.rs:78-84)
            if (sens) sens[i] = s;
        }
    }
    if (!last_msg.empty()) g_last_error = last_msg;
    return OFG_OK;
}

}  // namespace

extern "C" int32_t ofg_autocall_price(ofg_ctx *ctx, const ofg_autocall *contracts, size_t n,
                                      const ofg_engine_cfg *cfg, int32_t with_greeks,
                                      ofg_result *out, ofg_autocall_sens *sens) {
    return guarded([&]() { return autocall_impl(ctx, contracts, n, cfg, with_greeks, out, sens); });
}
