// Host side of ofg_basket_price: price_basket_mc and basket_sensitivities
// (src/pricing/basket.rs:99-160, 456-597) for n baskets in one launch, and
// ofg_repair_correlation (validate_or_repair_correlation_matrix and
// nearest_correlation_matrix_higham, src/math/correlation.rs:271-358).
//
// Validation follows the reference's order and messages (BasketOption::
// validate, instruments/basket.rs:36-60; validate_inputs_matrix and
// validate_market_inputs, pricing/basket.rs:802-860; validate_correlation_
// matrix, math/correlation.rs:220-248). As price_basket_mc, an invalid basket
// yields a NaN price and no stderr (the message is left in ofg_last_error).
//
// Every scenario (the base price, each bump) is prepared as the reference's
// simulate_basket_once_matrix prepares it: a zero maturity prices the payoff
// of the spots on the host; otherwise the correlation matrix is repaired when
// it is not PSD and factored by cholesky_lower_psd (tolerance 1e-12).
#include "../../include/openferric_amd.h"
#include "ofg_host.hpp"
#include "ofg_internal.hpp"
#include "ofg_corr.hpp"
#include "ofg_jump.hpp"
#include "ofg_math.hpp"
#include "ofg_multi.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace ofg_host;

// the seed kernel's thread-start mode assumes 256-thread tiles
static_assert(ofg::MA_BLOCK == 256, "tile width of ofg::thread_stream_start");
using namespace ofg_corr;

namespace ofg_host {

struct BasketBuffers {
    DevBuf scen, tiles, partials, out, normals, seeds, polys, jobs, gbk, gtiles, gpartials, gout;
};

void basket_release(ofg_ctx *ctx) {
    if (!ctx || !ctx->basket) return;
    BasketBuffers *b = ctx->basket;
    for (DevBuf *d : {&b->scen, &b->tiles, &b->partials, &b->out, &b->normals, &b->seeds,
                      &b->polys, &b->jobs, &b->gbk, &b->gtiles, &b->gpartials, &b->gout})
        d->release();
    delete b;
    ctx->basket = nullptr;
}

}  // namespace ofg_host

namespace {

constexpr double SPOT_BUMP_REL = 0.01, VOL_BUMP_ABS = 0.01, CORR_BUMP = 0.01;  // basket.rs:28-30

// validate_inputs_matrix (basket.rs:842-860) for one basket's inputs; throws.
void validate_inputs(const ofg_basket &b, const std::vector<double> &spots,
                     const std::vector<double> &vols, const Mat &corr) {
    // BasketOption::validate (instruments/basket.rs:38-60)
    if (b.strike < 0.0) fail(OFG_INVALID_INPUT, "basket strike must be >= 0");
    if (b.maturity < 0.0) fail(OFG_INVALID_INPUT, "basket maturity must be >= 0");
    for (int32_t k = 0; k < b.n_weights; ++k)
        if (!std::isfinite(b.weights[k])) fail(OFG_INVALID_INPUT, "basket weights must be finite");
    if (b.basket_type == OFG_BASKET_AVERAGE && b.n_weights <= 0)
        fail(OFG_INVALID_INPUT, "average basket requires non-empty weights");
    // validate_market_inputs (basket.rs:802-840)
    if (b.num_paths <= 0) fail(OFG_INVALID_INPUT, "basket n_paths must be > 0");
    if (b.n_assets <= 0) fail(OFG_INVALID_INPUT, "basket spots cannot be empty");
    for (double s : spots)
        if (s <= 0.0 || !std::isfinite(s)) fail(OFG_INVALID_INPUT, "basket spots must be finite and > 0");
    for (double v : vols)
        if (v < 0.0 || !std::isfinite(v)) fail(OFG_INVALID_INPUT, "basket vols must be finite and >= 0");
    for (int32_t i = 0; i < b.n_assets; ++i)
        if (!std::isfinite(b.dividends[i])) fail(OFG_INVALID_INPUT, "basket dividends must be finite");
    if (b.basket_type == OFG_BASKET_AVERAGE && b.n_weights != b.n_assets)
        fail(OFG_INVALID_INPUT, "average basket weights length must match number of assets");
    validate_corr(corr, b.n_assets);
}

// validate_outperformance_inputs (basket.rs:891-915) with
// OutperformanceBasketOption::validate (instruments/basket.rs:82-110).
void validate_outperformance(const ofg_basket &b, const std::vector<double> &spots,
                             const std::vector<double> &vols, const Mat &corr) {
    if (b.n_weights <= 0) fail(OFG_INVALID_INPUT, "outperformance lagger weights cannot be empty");
    if (b.strike < 0.0) fail(OFG_INVALID_INPUT, "outperformance strike must be >= 0");
    if (b.maturity < 0.0) fail(OFG_INVALID_INPUT, "outperformance maturity must be >= 0");
    for (int32_t k = 0; k < b.n_weights; ++k)
        if (!std::isfinite(b.weights[k]))
            fail(OFG_INVALID_INPUT, "outperformance lagger weights must be finite");
    if (b.num_paths <= 0) fail(OFG_INVALID_INPUT, "basket n_paths must be > 0");
    if (b.n_assets <= 0) fail(OFG_INVALID_INPUT, "basket spots cannot be empty");
    for (double s : spots)
        if (s <= 0.0 || !std::isfinite(s)) fail(OFG_INVALID_INPUT, "basket spots must be finite and > 0");
    for (double v : vols)
        if (v < 0.0 || !std::isfinite(v)) fail(OFG_INVALID_INPUT, "basket vols must be finite and >= 0");
    for (int32_t i = 0; i < b.n_assets; ++i)
        if (!std::isfinite(b.dividends[i])) fail(OFG_INVALID_INPUT, "basket dividends must be finite");
    validate_corr(corr, b.n_assets);
    if (b.n_weights != b.n_assets)
        fail(OFG_INVALID_INPUT, "outperformance lagger weights length must match assets");
    if (b.leader_index < 0 || b.leader_index >= b.n_assets)
        fail(OFG_INVALID_INPUT, "outperformance leader_index out of bounds");
}

// validate_quanto_inputs (basket.rs:917-943) with QuantoBasketOption::validate
// (instruments/basket.rs:131-166).
void validate_quanto(const ofg_basket &b, const std::vector<double> &spots,
                     const std::vector<double> &vols, const Mat &corr) {
    if (b.strike < 0.0) fail(OFG_INVALID_INPUT, "basket strike must be >= 0");
    if (b.maturity < 0.0) fail(OFG_INVALID_INPUT, "basket maturity must be >= 0");
    for (int32_t k = 0; k < b.n_weights; ++k)
        if (!std::isfinite(b.weights[k])) fail(OFG_INVALID_INPUT, "basket weights must be finite");
    if (b.basket_type == OFG_BASKET_AVERAGE && b.n_weights <= 0)
        fail(OFG_INVALID_INPUT, "average basket requires non-empty weights");
    if (b.fx_rate <= 0.0 || !std::isfinite(b.fx_rate))
        fail(OFG_INVALID_INPUT, "quanto fx_rate must be finite and > 0");
    if (b.fx_vol < 0.0 || !std::isfinite(b.fx_vol))
        fail(OFG_INVALID_INPUT, "quanto fx_vol must be finite and >= 0");
    for (int32_t k = 0; k < b.n_asset_fx_corr; ++k) {
        const double rho = b.asset_fx_corr[k];
        if (!std::isfinite(rho) || !(rho >= -1.0 && rho <= 1.0))
            fail(OFG_INVALID_INPUT, "quanto asset_fx_corr entries must be finite and in [-1, 1]");
    }
    if (b.num_paths <= 0) fail(OFG_INVALID_INPUT, "basket n_paths must be > 0");
    if (b.n_assets <= 0) fail(OFG_INVALID_INPUT, "basket spots cannot be empty");
    for (double s : spots)
        if (s <= 0.0 || !std::isfinite(s)) fail(OFG_INVALID_INPUT, "basket spots must be finite and > 0");
    for (double v : vols)
        if (v < 0.0 || !std::isfinite(v)) fail(OFG_INVALID_INPUT, "basket vols must be finite and >= 0");
    for (int32_t i = 0; i < b.n_assets; ++i)
        if (!std::isfinite(b.dividends[i])) fail(OFG_INVALID_INPUT, "basket dividends must be finite");
    validate_corr(corr, b.n_assets);
    if (b.n_asset_fx_corr != b.n_assets)
        fail(OFG_INVALID_INPUT, "quanto asset_fx_corr length must match assets");
    if (b.basket_type == OFG_BASKET_AVERAGE && b.n_weights != b.n_assets)
        fail(OFG_INVALID_INPUT, "quanto average basket weights length must match assets");
}

// validate_inputs_factor (basket.rs:862-889) with FactorCorrelationModel::
// validate (math/correlation.rs:84-121).
void validate_factor(const ofg_basket &b, const std::vector<double> &spots,
                     const std::vector<double> &vols) {
    if (b.strike < 0.0) fail(OFG_INVALID_INPUT, "basket strike must be >= 0");
    if (b.maturity < 0.0) fail(OFG_INVALID_INPUT, "basket maturity must be >= 0");
    for (int32_t k = 0; k < b.n_weights; ++k)
        if (!std::isfinite(b.weights[k])) fail(OFG_INVALID_INPUT, "basket weights must be finite");
    if (b.basket_type == OFG_BASKET_AVERAGE && b.n_weights <= 0)
        fail(OFG_INVALID_INPUT, "average basket requires non-empty weights");
    if (b.num_paths <= 0) fail(OFG_INVALID_INPUT, "basket n_paths must be > 0");
    if (b.n_assets <= 0) fail(OFG_INVALID_INPUT, "basket spots cannot be empty");
    for (double s : spots)
        if (s <= 0.0 || !std::isfinite(s)) fail(OFG_INVALID_INPUT, "basket spots must be finite and > 0");
    for (double v : vols)
        if (v < 0.0 || !std::isfinite(v)) fail(OFG_INVALID_INPUT, "basket vols must be finite and >= 0");
    for (int32_t i = 0; i < b.n_assets; ++i)
        if (!std::isfinite(b.dividends[i])) fail(OFG_INVALID_INPUT, "basket dividends must be finite");
    const int rows = b.n_loading_rows, k = b.n_factors;
    if (b.factor_kind == 1) {
        if (rows <= 0) fail(OFG_INVALID_INPUT, "one-factor model requires at least one loading");
        for (int i = 0; i < rows; ++i) {
            const double beta = b.loadings[(size_t)i * k];
            if (!std::isfinite(beta) || std::fabs(beta) > 1.0 + 1.0e-12)
                fail(OFG_INVALID_INPUT, "one-factor loadings must be finite and in [-1, 1]");
        }
    } else {
        if (rows <= 0) fail(OFG_INVALID_INPUT, "multi-factor model requires at least one asset row");
        if (k <= 0) fail(OFG_INVALID_INPUT, "multi-factor model requires at least one factor");
        for (int i = 0; i < rows; ++i) {
            double norm2 = -0.0;
            for (int j = 0; j < k; ++j) {
                const double x = b.loadings[(size_t)i * k + j];
                if (!std::isfinite(x)) fail(OFG_INVALID_INPUT, "multi-factor loadings must be finite");
            }
            for (int j = 0; j < k; ++j) norm2 += b.loadings[(size_t)i * k + j] * b.loadings[(size_t)i * k + j];
            if (norm2 > 1.0 + 1.0e-10)
                fail(OFG_INVALID_INPUT, "multi-factor row norm must be <= 1 for unit-variance assets");
        }
    }
    if (rows != b.n_assets)
        fail(OFG_INVALID_INPUT, "factor-model asset count must match spots length");
    if (b.basket_type == OFG_BASKET_AVERAGE && b.n_weights != b.n_assets)
        fail(OFG_INVALID_INPUT, "average basket weights length must match number of assets");
}

// ln Gamma(x) for x >= 0.5 as statrs 0.18's ln_gamma (the crate of
// StudentsT::cdf): the Lanczos sum with Pugh's r and d_k, the table of
// src/math/gamma.rs:17-31.
double ln_gamma_lanczos(double x) {
    static const double R = 10.900511;
    static const double DK[11] = {
        2.4857408913875355e-5, 1.0514237858172197, -3.4568709722201624, 4.512277094668948,
        -2.9828522532357666, 1.056397115771267, -1.9542877319164587e-1, 1.709705434044412e-2,
        -5.719261174043057e-4, 4.633994733599057e-6, -2.7199490848860772e-9};
    const double LN_2_SQRT_E_OVER_PI = 0.6207822376352452223455184457816472122518527279025978;
    const double E = 2.7182818284590452353602874713526624977572470937;
    double s = DK[0];
    for (int i = 1; i < 11; ++i) s += DK[i] / (x + (double)i - 1.0);
    return std::log(s) + LN_2_SQRT_E_OVER_PI + (x - 0.5) * std::log((x - 0.5 + R) / E);
}

// terminal_payoff (basket.rs:759-784) of the spots themselves, times the
// quanto FX rate; outperformance_payoff (:786-800).
double payoff_at_spots(const ofg_basket &b, const std::vector<double> &spots) {
    if (b.product == 1) {
        double lag = -0.0;
        for (int32_t i = 0; i < b.n_assets; ++i) lag += b.weights[i] * spots[(size_t)i];
        const double ratio = spots[(size_t)b.leader_index] / std::fmax(lag, 1.0e-12);
        return b.is_call ? std::fmax(ratio - b.strike, 0.0) : std::fmax(b.strike - ratio, 0.0);
    }
    const double scale = b.product == 2 ? b.fx_rate : 1.0;
    double u;
    if (b.basket_type == OFG_BASKET_AVERAGE) {
        u = -0.0;
        for (int32_t i = 0; i < b.n_assets; ++i) u += b.weights[i] * spots[(size_t)i];
    } else if (b.basket_type == OFG_BASKET_BEST_OF) {
        u = -INFINITY;
        for (int32_t i = 0; i < b.n_assets; ++i) u = std::fmax(u, spots[(size_t)i] / spots[(size_t)i]);
    } else {
        u = INFINITY;
        for (int32_t i = 0; i < b.n_assets; ++i) u = std::fmin(u, spots[(size_t)i] / spots[(size_t)i]);
    }
    return scale * (b.is_call ? std::fmax(u - b.strike, 0.0) : std::fmax(b.strike - u, 0.0));
}

// One scenario: a device descriptor or a host value (zero maturity).
struct Scenario {
    bool on_device = false;
    double host_price = 0.0;
    bool repaired = false;
    ofg::BkDev d{};
};

// simulate_basket_once_matrix's preparation (basket.rs:610-641) and those
// of price_outperformance_basket_mc (:299-326) and price_quanto_basket_mc
// (:376-403), which repair and factor before the zero-maturity shortcut;
// throws.
Scenario prepare(const ofg_basket &b, const std::vector<double> &spots,
                 const std::vector<double> &vols, const Mat &corr) {
    Scenario s;
    const int n = b.n_assets;
    Mat l;
    if (b.product == 0 && b.factor_kind != 0) {
        // simulate_basket_once_factor (basket.rs:686-757): no matrix to repair
        validate_factor(b, spots, vols);
        if (b.maturity <= 0.0) {
            s.host_price = payoff_at_spots(b, spots);
            return s;
        }
        if (n > ofg::BK_MAX) fail(OFG_INVALID_INPUT, "at most 16 assets per basket");
        if (b.n_factors > ofg::BK_FMAX) fail(OFG_INVALID_INPUT, "at most 4 factors per basket");
    } else if (b.product == 0) {
        validate_inputs(b, spots, vols, corr);
        if (b.maturity <= 0.0) {
            s.host_price = payoff_at_spots(b, spots);
            return s;
        }
        if (n > ofg::BK_MAX) fail(OFG_INVALID_INPUT, "at most 16 assets per basket");
        const Mat c = validate_or_repair(corr, n, s.repaired);
        if (!cholesky_psd(c, n, l)) fail(OFG_INVALID_INPUT, "basket correlation matrix is not PSD");
    } else {
        if (b.product == 1) validate_outperformance(b, spots, vols, corr);
        else validate_quanto(b, spots, vols, corr);
        const Mat c = validate_or_repair(corr, n, s.repaired);
        if (!cholesky_psd(c, n, l)) fail(OFG_INVALID_INPUT, "correlation matrix is not PSD");
        if (b.maturity <= 0.0) {
            s.host_price = payoff_at_spots(b, spots);
            return s;
        }
        if (n > ofg::BK_MAX) fail(OFG_INVALID_INPUT, "at most 16 assets per basket");
    }
    // the Student-t copula's dof check runs when the first path is drawn
    // (map_copula_normals_to_uniforms, math/correlation.rs:559-562), after
    // the zero-maturity shortcut, the repair and the factorisation
    if (b.copula == OFG_COPULA_STUDENT_T && b.copula_dof < 2)
        fail(OFG_INVALID_INPUT, "student-t copula requires degrees_of_freedom >= 2");
    ofg::BkDev &d = s.d;
    if (b.copula == OFG_COPULA_STUDENT_T) {
        const double nu = (double)b.copula_dof, a = nu / 2.0;
        d.t_dof = b.copula_dof;
        d.t_nu = nu;
        d.t_a = a;
        d.t_lbeta = ln_gamma_lanczos(a + 0.5) - ln_gamma_lanczos(a) - ln_gamma_lanczos(0.5);
    }
    d.n = n;
    d.type = b.basket_type;
    d.is_call = b.is_call != 0;
    d.paths = b.num_paths;
    d.seed = b.seed;
    d.strike = b.strike;
    d.payoff = b.product == 1 ? ofg::BK_OUTPERFORMANCE : ofg::BK_TERMINAL;
    d.leader = b.product == 1 ? b.leader_index : -1;
    d.scale = b.product == 2 ? b.fx_rate : 1.0;
    const double t = b.maturity, sqrt_t = std::sqrt(t);
    // discount exp(-r T); quanto: exp(-r_d T) (basket.rs:436)
    d.discount = std::exp(-(b.product == 2 ? b.domestic_rate : b.rate) * t);
    for (int i = 0; i < n; ++i) {
        const double v = vols[(size_t)i];
        d.s0[i] = spots[(size_t)i];
        d.inv_s0[i] = 1.0 / spots[(size_t)i];
        if (b.product == 2) {   // quanto drift (basket.rs:408-417)
            const double quanto_adj = b.asset_fx_corr[i] * v * b.fx_vol;
            d.drift[i] = (b.foreign_rate - b.dividends[i] - quanto_adj - 0.5 * v * v) * t;
        } else {
            d.drift[i] = (b.rate - b.dividends[i] - 0.5 * v * v) * t;
        }
        d.vol_t[i] = v * sqrt_t;
        d.drift2[i] = d.drift[i] * OFG_INV_LN2_256;
        d.vol_t2[i] = d.vol_t[i] * OFG_INV_LN2_256;
        d.w[i] = (b.basket_type == OFG_BASKET_AVERAGE || b.product == 1) ? b.weights[i] : 0.0;
        if (b.factor_kind == 0)
            for (int j = 0; j <= i; ++j) d.chol[i * (i + 1) / 2 + j] = l[(size_t)i * n + j];
    }
    if (b.factor_kind != 0) {
        const int k = b.n_factors;
        d.n_factors = b.factor_kind == 1 ? 1 : k;
        d.one_factor = b.factor_kind == 1;
        for (int i = 0; i < n; ++i) {
            if (b.factor_kind == 1) {   // beta clamped to [-1, 1], idio sqrt(max(1 - beta^2, 0))
                const double beta = std::min(1.0, std::max(-1.0, b.loadings[(size_t)i * k]));
                d.load[i * ofg::BK_FMAX] = beta;
                d.idio[i] = std::sqrt(std::max(1.0 - beta * beta, 0.0));
            } else {
                double norm2 = -0.0;
                for (int j = 0; j < k; ++j) {
                    const double x = b.loadings[(size_t)i * k + j];
                    d.load[i * ofg::BK_FMAX + j] = x;
                    norm2 += x * x;
                }
                d.idio[i] = std::sqrt(std::max(1.0 - norm2, 0.0));
            }
        }
    }
    s.on_device = true;
    return s;
}

// bump_corr_matrix (basket.rs:1025-1037)
Mat bump_corr(const Mat &c, int n, double bump) {
    Mat out = c;
    for (int i = 0; i < n; ++i) {
        out[(size_t)i * n + i] = 1.0;
        for (int j = i + 1; j < n; ++j) {
            const double v = std::min(0.999, std::max(-0.999, c[(size_t)i * n + j] + bump));
            out[(size_t)i * n + j] = out[(size_t)j * n + i] = v;
        }
    }
    return out;
}

// average_off_diagonal (basket.rs:1039-1056)
double average_off_diagonal(const Mat &c, int n) {
    if (n < 2) return 0.0;
    double sum = 0.0;
    int64_t count = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            sum += c[(size_t)i * n + j];
            ++count;
        }
    return sum / (double)count;
}

struct BasketPlan {
    bool valid = false;
    int32_t base = -1;
    std::vector<int32_t> delta_up, delta_dn;
    std::vector<double> delta_bump;
    int32_t vega_up = -1, vega_dn = -1, cega_up = -1, cega_dn = -1;
    double cega_bump = 0.0;
    bool sens_ok = false;
    double rho = 0.0;
    bool repaired = false;
};

int32_t basket_impl(ofg_ctx *ctx, const ofg_basket *in, size_t n, const ofg_engine_cfg *cfg,
                    int32_t with_sens, ofg_result *out, ofg_basket_sens *sens) {
    if (!ctx || !cfg || (n && (!in || !out))) fail(OFG_INVALID_INPUT, "null argument");
    const int32_t rng = cfg->rng_mode;
    if (rng != OFG_RNG_EXTERNAL && rng != OFG_RNG_FAST && rng != OFG_RNG_REFERENCE)
        fail(OFG_INVALID_INPUT, "unknown rng mode");
    std::vector<Scenario> scen;
    std::vector<BasketPlan> plans(n);
    std::string last_msg;
    int64_t normals_elems = 0, seed_words = 0, seed_threads = 0;
    std::vector<ofg::LsmSeedJob> jobs;
    std::vector<uint64_t> polys;
    std::map<uint64_t, std::vector<ofg_jump::Poly>> poly_cache;
    for (size_t i = 0; i < n; ++i) {
        const ofg_basket &b = in[i];
        std::memset(&out[i], 0, sizeof out[i]);
        out[i].price = NAN;
        out[i].num_paths = out[i].num_steps = out[i].vol = NAN;
        if (sens) std::memset(&sens[i], 0, sizeof sens[i]);
        BasketPlan &pl = plans[i];
        // a basket's external normals: [asset][path], n_assets * num_paths
        // draws per path: one per asset, plus the factors of a factor model
        // and the chi-square draws of a Student-t copula
        const int64_t draws =
            (int64_t)b.n_assets +
            (b.factor_kind == 1 ? 1 : b.factor_kind == 2 ? std::max(b.n_factors, 0) : 0) +
            (b.copula == OFG_COPULA_STUDENT_T ? std::max(b.copula_dof, 0) : 0);
        const int64_t my_normals = normals_elems;
        if (b.n_assets > 0 && b.num_paths > 0) normals_elems += draws * b.num_paths;
        try {
            if (b.n_assets <= 0 || !b.spots || !b.vols || !b.dividends || (!b.corr && !b.factor_kind))
                fail(OFG_INVALID_INPUT, "basket spots cannot be empty");
            if (b.n_weights < 0 || (b.n_weights > 0 && !b.weights))
                fail(OFG_INVALID_INPUT, "basket weights must be finite");
            if (b.product < 0 || b.product > 2) fail(OFG_INVALID_INPUT, "unknown basket product");
            if (b.factor_kind < 0 || b.factor_kind > 2 || (b.factor_kind && b.product != 0))
                fail(OFG_INVALID_INPUT, "factor models apply to BasketOption only");
            if (b.factor_kind && (b.n_loading_rows < 0 || b.n_factors < 0 ||
                                  (b.factor_kind == 1 && b.n_factors != 1) ||
                                  (b.n_loading_rows * b.n_factors > 0 && !b.loadings)))
                fail(OFG_INVALID_INPUT, "factor-model loadings must be n_loading_rows x n_factors "
                                        "(one factor for OneFactor)");
            if (b.copula != OFG_COPULA_GAUSSIAN && b.copula != OFG_COPULA_STUDENT_T)
                fail(OFG_INVALID_INPUT, "unknown basket copula");
            if (b.copula != OFG_COPULA_GAUSSIAN && b.product != 0)
                fail(OFG_INVALID_INPUT, "the Student-t copula applies to BasketOption only");
            if (b.product == 2 && b.n_asset_fx_corr > 0 && !b.asset_fx_corr)
                fail(OFG_INVALID_INPUT, "quanto asset_fx_corr length must match assets");
            const int na = b.n_assets;
            std::vector<double> spots(b.spots, b.spots + na), vols(b.vols, b.vols + na);
            const Mat corr = b.corr ? Mat(b.corr, b.corr + (size_t)na * na) : Mat((size_t)na * na, 0.0);
            int64_t seed_off = -1;
            auto push = [&](Scenario s) {
                if (s.on_device) {
                    s.d.normals_offset = my_normals;
                    if (rng == OFG_RNG_REFERENCE) {
                        if (seed_off < 0) {   // one stream per basket, shared by its bumps
                            // one state per kernel thread (every scenario of
                            // the basket has these tiles), its paths
                            // consecutive in the stream
                            const int64_t blk = ofg::MA_BLOCK;
                            const int32_t ppt = (int32_t)std::max<int64_t>(
                                1, (b.num_paths + blk * ofg::MA_MAX_TILES - 1) /
                                       (blk * ofg::MA_MAX_TILES));
                            const ofg::LsmSeedJob j = ofg_jump::make_thread_seed_job(
                                b.seed, b.num_paths, draws, ppt, seed_words, seed_threads, polys,
                                poly_cache);
                            jobs.push_back(j);
                            seed_threads += ofg_jump::seed_job_threads(j);
                            seed_off = seed_words;
                            seed_words += 4 * ((b.num_paths + blk * ppt - 1) / (blk * ppt)) * blk;
                        }
                        s.d.seed_offset = seed_off;
                    }
                }
                scen.push_back(s);
                return (int32_t)scen.size() - 1;
            };
            Scenario base = prepare(b, spots, vols, corr);
            pl.repaired = base.repaired;
            pl.base = push(base);
            pl.valid = true;
            // diagnostics rho: the matrix's average off-diagonal, 0 under a
            // factor model (basket.rs:151, 195)
            pl.rho = b.factor_kind ? 0.0 : average_off_diagonal(corr, na);
            if (with_sens && b.product == 0 && b.factor_kind == 0 &&
                b.copula == OFG_COPULA_GAUSSIAN) {
                // basket_sensitivities (basket.rs:456-539)
                try {
                    for (int k = 0; k < na; ++k) {
                        const double bump = std::max(std::fabs(spots[(size_t)k]) * SPOT_BUMP_REL, 1.0e-4);
                        std::vector<double> up = spots, dn = spots;
                        up[(size_t)k] += bump;
                        dn[(size_t)k] = std::max(dn[(size_t)k] - bump, 1.0e-8);
                        Scenario su = prepare(b, up, vols, corr), sd = prepare(b, dn, vols, corr);
                        pl.delta_up.push_back(push(su));
                        pl.delta_dn.push_back(push(sd));
                        pl.delta_bump.push_back(bump);
                    }
                    std::vector<double> vu = vols, vd = vols;
                    for (int k = 0; k < na; ++k) {
                        vu[(size_t)k] += VOL_BUMP_ABS;
                        vd[(size_t)k] = std::max(vd[(size_t)k] - VOL_BUMP_ABS, 1.0e-6);
                    }
                    Scenario su = prepare(b, spots, vu, corr), sd = prepare(b, spots, vd, corr);
                    pl.vega_up = push(su);
                    pl.vega_dn = push(sd);
                    // correlation_sensitivity (basket.rs:551-597): halve the bump
                    // while either bumped matrix fails to prepare
                    double bump = CORR_BUMP;
                    for (int tries = 0; tries < 6; ++tries, bump *= 0.5) {
                        try {
                            Scenario cu = prepare(b, spots, vols, bump_corr(corr, na, bump));
                            Scenario cd = prepare(b, spots, vols, bump_corr(corr, na, -bump));
                            pl.cega_up = push(cu);
                            pl.cega_dn = push(cd);
                            pl.cega_bump = bump;
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
    // basket_sensitivities fused per path: every scenario of a basket on the
    // same draws in one pass (OFG_BASKET_FUSED=0 prices them one by one)
    std::vector<int32_t> fused_slot(scen.size(), -1);
    std::vector<ofg::BkGreeksDev> gdev;
    std::vector<int32_t> gtiles;
    int32_t g_out = 0, max_n = 0;
    const char *fenv = std::getenv("OFG_BASKET_FUSED");
    const bool fuse = with_sens && !(fenv && std::atoi(fenv) == 0);
    for (size_t i = 0; fuse && i < n; ++i) {
        const BasketPlan &pl = plans[i];
        if (!pl.valid || !pl.sens_ok || in[i].product != 0 || in[i].factor_kind != 0 ||
            in[i].copula != OFG_COPULA_GAUSSIAN)
            continue;
        std::vector<int32_t> ids{pl.base, pl.vega_up, pl.vega_dn};
        for (size_t k = 0; k < pl.delta_up.size(); ++k) {
            ids.push_back(pl.delta_up[k]);
            ids.push_back(pl.delta_dn[k]);
        }
        if (pl.cega_up >= 0) {
            ids.push_back(pl.cega_up);
            ids.push_back(pl.cega_dn);
        }
        bool all_dev = true;
        for (int32_t id : ids) all_dev = all_dev && scen[(size_t)id].on_device;
        if (!all_dev) continue;
        const ofg::BkDev &b0 = scen[(size_t)pl.base].d;
        ofg::BkGreeksDev g{};
        g.n = b0.n;
        g.type = b0.type;
        g.is_call = b0.is_call;
        g.has_cega = pl.cega_up >= 0;
        g.n_scen = 1 + 2 * g.n + 2 + 2 * g.has_cega;
        g.paths = b0.paths;
        g.normals_offset = b0.normals_offset;
        g.seed_offset = b0.seed_offset;
        g.seed = b0.seed;
        g.strike = b0.strike;
        g.discount = b0.discount;
        const ofg::BkDev *sets[3] = {&b0, &scen[(size_t)pl.vega_up].d, &scen[(size_t)pl.vega_dn].d};
        const ofg::BkDev *csets[3] = {&b0, g.has_cega ? &scen[(size_t)pl.cega_up].d : &b0,
                                      g.has_cega ? &scen[(size_t)pl.cega_dn].d : &b0};
        for (int a = 0; a < g.n; ++a) {
            g.s0[a] = b0.s0[a];
            g.w[a] = b0.w[a];
            g.s0_up[a] = scen[(size_t)pl.delta_up[(size_t)a]].d.s0[a];
            g.s0_dn[a] = scen[(size_t)pl.delta_dn[(size_t)a]].d.s0[a];
            g.inv_s0[a] = b0.inv_s0[a];
            g.inv_up[a] = scen[(size_t)pl.delta_up[(size_t)a]].d.inv_s0[a];
            g.inv_dn[a] = scen[(size_t)pl.delta_dn[(size_t)a]].d.inv_s0[a];
            for (int v = 0; v < 3; ++v) {
                g.drift[v][a] = sets[v]->drift[a];
                g.vol_t[v][a] = sets[v]->vol_t[a];
                g.drift2[v][a] = sets[v]->drift2[a];
                g.vol_t2[v][a] = sets[v]->vol_t2[a];
            }
        }
        for (int v = 0; v < 3; ++v)
            std::memcpy(g.chol[v], csets[v]->chol, sizeof g.chol[v]);
        const int64_t blk = ofg::MA_BLOCK;
        g.ppt = (int32_t)std::max<int64_t>(
            1, (g.paths + blk * ofg::MA_MAX_TILES - 1) / (blk * ofg::MA_MAX_TILES));
        g.n_tiles = (int32_t)((g.paths + blk * g.ppt - 1) / (blk * g.ppt));
        g.tile_first = (int32_t)gtiles.size();
        for (int32_t t = 0; t < g.n_tiles; ++t) gtiles.push_back((int32_t)gdev.size());
        g.out_first = g_out;
        g_out += ofg::BKG_MAXS;
        fused_slot[(size_t)pl.base] = g.out_first;
        for (int k = 0; k < g.n; ++k) {
            fused_slot[(size_t)pl.delta_up[(size_t)k]] = g.out_first + 1 + 2 * k;
            fused_slot[(size_t)pl.delta_dn[(size_t)k]] = g.out_first + 2 + 2 * k;
        }
        fused_slot[(size_t)pl.vega_up] = g.out_first + 1 + 2 * g.n;
        fused_slot[(size_t)pl.vega_dn] = g.out_first + 2 + 2 * g.n;
        if (g.has_cega) {
            fused_slot[(size_t)pl.cega_up] = g.out_first + 3 + 2 * g.n;
            fused_slot[(size_t)pl.cega_dn] = g.out_first + 4 + 2 * g.n;
        }
        max_n = std::max(max_n, g.n);
        gdev.push_back(g);
    }
    std::vector<int32_t> dev_index(scen.size(), -1);
    std::vector<ofg::BkDev> dev;
    std::vector<int32_t> tiles;
    for (size_t s = 0; s < scen.size(); ++s) {
        if (!scen[s].on_device || fused_slot[s] >= 0) continue;
        ofg::BkDev d = scen[s].d;
        const int64_t blk = ofg::MA_BLOCK;
        d.ppt = (int32_t)std::max<int64_t>(
            1, (d.paths + blk * ofg::MA_MAX_TILES - 1) / (blk * ofg::MA_MAX_TILES));
        d.n_tiles = (int32_t)((d.paths + blk * d.ppt - 1) / (blk * d.ppt));
        d.tile_first = (int32_t)tiles.size();
        for (int32_t t = 0; t < d.n_tiles; ++t) tiles.push_back((int32_t)dev.size());
        dev_index[s] = (int32_t)dev.size();
        dev.push_back(d);
    }
    if (rng == OFG_RNG_EXTERNAL && (!dev.empty() || !gdev.empty()) && !cfg->normals1)
        fail(OFG_INVALID_INPUT, "external rng mode needs normals1");
    std::vector<ofg::AcOut> res(dev.size()), gres((size_t)g_out);
    float ms = 0.f;
    if (!dev.empty() || !gdev.empty()) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!ctx->basket) ctx->basket = new BasketBuffers();
        BasketBuffers &bb = *ctx->basket;
        hipStream_t st = ctx->stream;
        auto up = [&](DevBuf &buf, const void *src, size_t bytes, const char *what) {
            buf.ensure(std::max<size_t>(bytes, 1));
            if (bytes) hip_check(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, st), what);
        };
        up(bb.scen, dev.data(), dev.size() * sizeof(dev[0]), "H2D basket scenarios");
        up(bb.tiles, tiles.data(), tiles.size() * sizeof(int32_t), "H2D basket tiles");
        bb.partials.ensure(tiles.size() * sizeof(ofg::AcPartial));
        bb.out.ensure(dev.size() * sizeof(ofg::AcOut));
        ofg::BkArgs a{};
        if (rng == OFG_RNG_EXTERNAL) {
            up(bb.normals, cfg->normals1, (size_t)normals_elems * sizeof(double), "H2D normals1");
            a.normals = bb.normals.as<double>();
        }
        if (rng == OFG_RNG_REFERENCE) {
            up(bb.jobs, jobs.data(), jobs.size() * sizeof(jobs[0]), "H2D seed jobs");
            up(bb.polys, polys.data(), polys.size() * sizeof(uint64_t), "H2D jump polynomials");
            bb.seeds.ensure((size_t)seed_words * sizeof(uint64_t));
            a.seeds = bb.seeds.as<uint64_t>();
        }
        a.scen = bb.scen.as<ofg::BkDev>();
        a.tile_scen = bb.tiles.as<int32_t>();
        a.partials = bb.partials.as<ofg::AcPartial>();
        a.out = bb.out.as<ofg::AcOut>();
        a.n_tiles = (int32_t)tiles.size();
        a.n_scen = (int32_t)dev.size();
        a.rng_mode = rng;
        for (const ofg::BkDev &d : dev) a.any_t |= d.t_dof > 0;
        hip_check(hipEventRecord(ctx->ev0, st), "event");
        if (rng == OFG_RNG_REFERENCE)
            hip_check(ofg::launch_lsm_seeds(bb.jobs.as<ofg::LsmSeedJob>(), (int32_t)jobs.size(),
                                            seed_threads, bb.polys.as<uint64_t>(),
                                            bb.seeds.as<uint64_t>(), st),
                      "basket seed kernel launch");
        if (!dev.empty()) hip_check(ofg::launch_basket(a, st), "basket kernel launch");
        ofg::BkgArgs ga{};
        if (!gdev.empty()) {
            up(bb.gbk, gdev.data(), gdev.size() * sizeof(gdev[0]), "H2D basket sensitivities");
            up(bb.gtiles, gtiles.data(), gtiles.size() * sizeof(int32_t), "H2D basket tiles");
            bb.gpartials.ensure(gtiles.size() * ofg::BKG_MAXS * sizeof(ofg::AcPartial));
            bb.gout.ensure((size_t)g_out * sizeof(ofg::AcOut));
            ga.bk = bb.gbk.as<ofg::BkGreeksDev>();
            ga.tile_bk = bb.gtiles.as<int32_t>();
            ga.partials = bb.gpartials.as<ofg::AcPartial>();
            ga.out = bb.gout.as<ofg::AcOut>();
            ga.normals = a.normals;
            ga.seeds = a.seeds;
            ga.n_tiles = (int32_t)gtiles.size();
            ga.n_bk = (int32_t)gdev.size();
            ga.rng_mode = rng;
            ga.max_n = max_n;
            hip_check(ofg::launch_basket_greeks(ga, st), "basket sensitivity kernel launch");
        }
        hip_check(hipEventRecord(ctx->ev1, st), "event");
        if (!dev.empty())
            hip_check(hipMemcpyAsync(res.data(), a.out, res.size() * sizeof(res[0]),
                                     hipMemcpyDeviceToHost, st),
                      "D2H basket results");
        if (!gdev.empty())
            hip_check(hipMemcpyAsync(gres.data(), ga.out, gres.size() * sizeof(gres[0]),
                                     hipMemcpyDeviceToHost, st),
                      "D2H basket sensitivity results");
        hip_check(hipStreamSynchronize(st), "basket kernel execution");
        hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
    }
    auto price_of = [&](int32_t s, double *se) {
        if (fused_slot[(size_t)s] >= 0) {
            const ofg::AcOut &r = gres[(size_t)fused_slot[(size_t)s]];
            if (se) *se = r.stderr_;
            return r.price;
        }
        if (dev_index[(size_t)s] < 0) {
            if (se) *se = 0.0;
            return scen[(size_t)s].host_price;
        }
        const ofg::AcOut &r = res[(size_t)dev_index[(size_t)s]];
        if (se) *se = r.stderr_;
        return r.price;
    };
    for (size_t i = 0; i < n; ++i) {
        const BasketPlan &pl = plans[i];
        out[i].kernel_ms = ms;
        if (!pl.valid) continue;
        double se = 0.0;
        out[i].price = price_of(pl.base, &se);
        out[i].stderr_ = se;
        out[i].has_stderr = 1;
        out[i].num_paths = (double)in[i].num_paths;   // diagnostics (basket.rs:149-151)
        ofg_basket_sens s{};
        s.rho = pl.rho;
        s.repaired = pl.repaired ? 1 : 0;
        if (with_sens && pl.sens_ok) {
            s.n_delta = (int32_t)pl.delta_up.size();
            for (size_t k = 0; k < pl.delta_up.size(); ++k)
                s.delta[k] = (price_of(pl.delta_up[k], nullptr) - price_of(pl.delta_dn[k], nullptr)) /
                             (2.0 * pl.delta_bump[k]);
            s.vega = (price_of(pl.vega_up, nullptr) - price_of(pl.vega_dn, nullptr)) /
                     (2.0 * VOL_BUMP_ABS);
            s.cega = pl.cega_up >= 0 ? (price_of(pl.cega_up, nullptr) - price_of(pl.cega_dn, nullptr)) /
                                           (2.0 * pl.cega_bump)
                                     : NAN;
            s.ok = 1;
        }
        if (sens) sens[i] = s;
    }
    if (!last_msg.empty()) g_last_error = last_msg;
    return OFG_OK;
}

}  // namespace

extern "C" int32_t ofg_basket_price(ofg_ctx *ctx, const ofg_basket *baskets, size_t n,
                                    const ofg_engine_cfg *cfg, int32_t with_sens, ofg_result *out,
                                    ofg_basket_sens *sens) {
    return guarded([&]() { return basket_impl(ctx, baskets, n, cfg, with_sens, out, sens); });
}

extern "C" int32_t ofg_repair_correlation(const double *corr, int32_t n, double *out,
                                          int32_t *was_repaired) {
    return guarded([&]() {
        if (!corr || !out || n <= 0) fail(OFG_INVALID_INPUT, "null argument");
        bool rep = false;
        const Mat c = validate_or_repair(Mat(corr, corr + (size_t)n * n), n, rep);
        std::copy(c.begin(), c.end(), out);
        if (was_repaired) *was_repaired = rep ? 1 : 0;
        return OFG_OK;
    });
}

// apply_correlation_stress with repair_to_psd (math/correlation.rs:442-543),
// as stressed_correlation_matrix calls it (pricing/basket.rs:447-454).
extern "C" int32_t ofg_stress_correlation(const double *corr, int32_t n,
                                          const ofg_corr_stress *scenarios, int32_t n_scenarios,
                                          double *out) {
    return guarded([&]() {
        if (!corr || !out || n <= 0 || n_scenarios < 0 || (n_scenarios > 0 && !scenarios))
            fail(OFG_INVALID_INPUT, "null argument");
        Mat c(corr, corr + (size_t)n * n);
        validate_corr(c, n);
        const double lim = 0.999999;
        auto clamp = [lim](double x) { return std::min(lim, std::max(-lim, x)); };
        auto each_pair = [&](auto &&f) {
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j) {
                    const double v = clamp(f(c[(size_t)i * n + j]));
                    c[(size_t)i * n + j] = c[(size_t)j * n + i] = v;
                }
        };
        for (int32_t k = 0; k < n_scenarios; ++k) {
            const ofg_corr_stress &sc = scenarios[k];
            const double x = sc.value;
            switch (sc.kind) {
            case OFG_STRESS_SCALE_OFF_DIAGONAL:
                if (!std::isfinite(x)) fail(OFG_INVALID_INPUT, "scale factor must be finite");
                each_pair([x](double r) { return r * x; });
                break;
            case OFG_STRESS_ADDITIVE_SHIFT:
                if (!std::isfinite(x)) fail(OFG_INVALID_INPUT, "additive shift must be finite");
                each_pair([x](double r) { return r + x; });
                break;
            case OFG_STRESS_FLOOR_OFF_DIAGONAL:
                if (!std::isfinite(x) || x < -1.0 || x > 1.0)
                    fail(OFG_INVALID_INPUT, "floor must be finite and in [-1, 1]");
                each_pair([x](double r) { return std::fmax(r, x); });
                break;
            case OFG_STRESS_CAP_OFF_DIAGONAL:
                if (!std::isfinite(x) || x < -1.0 || x > 1.0)
                    fail(OFG_INVALID_INPUT, "cap must be finite and in [-1, 1]");
                each_pair([x](double r) { return std::fmin(r, x); });
                break;
            case OFG_STRESS_OVERRIDE_PAIR: {
                if (sc.i < 0 || sc.j < 0 || sc.i >= n || sc.j >= n || sc.i == sc.j)
                    fail(OFG_INVALID_INPUT, "override pair indices must be distinct and in-range");
                if (!std::isfinite(x) || !(x >= -1.0 && x <= 1.0))
                    fail(OFG_INVALID_INPUT, "override pair value must be finite and in [-1, 1]");
                const double v = clamp(x);
                c[(size_t)sc.i * n + sc.j] = c[(size_t)sc.j * n + sc.i] = v;
                break;
            }
            default: fail(OFG_INVALID_INPUT, "unknown correlation stress scenario");
            }
            for (int i = 0; i < n; ++i) c[(size_t)i * n + i] = 1.0;
        }
        if (!(min_eigenvalue(c, n) >= -PSD_TOL)) c = nearest_correlation(c, n);
        validate_corr(c, n);
        std::copy(c.begin(), c.end(), out);
        return OFG_OK;
    });
}
