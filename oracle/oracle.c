/*
 * CPU oracle for the OpenFerric Monte Carlo pricing path -- TEST
 * INFRASTRUCTURE ONLY (the tests' reference results and the timed CPU baseline; never part of
 * the product path). See oracle.h for the list of functions and how each one
 * is pinned. Reference file:line citations point into rosssaunders/openferric.
 *
 * Compile with -ffp-contract=off: Rust never contracts a*b+c into an FMA, so
 * the only fused operations here are the explicit fma() calls that restate
 * the reference's `mul_add`.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* RNG: src/math/fast_rng.rs                                              */
/* ===================================================================== */

/* SplitMix64::next_u64, fast_rng.rs:192-199 */
uint64_t or_splitmix64_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Xoshiro256PlusPlus::seed_from_u64, fast_rng.rs:35-48 */
void or_xoshiro_seed(uint64_t seed, uint64_t s[4]) {
    uint64_t sm = seed;
    for (int i = 0; i < 4; ++i) s[i] = or_splitmix64_next(&sm);
    if ((s[0] | s[1] | s[2] | s[3]) == 0) s[0] = 1;
}

static inline uint64_t rotl64(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
}

/* Xoshiro256PlusPlus::next_u64, fast_rng.rs:50-67 */
uint64_t or_xoshiro_next_u64(uint64_t s[4]) {
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* k successive next_u64 calls (brute-force stepping for jump checks) */
void or_xoshiro_advance(uint64_t s[4], uint64_t k) {
    for (uint64_t i = 0; i < k; ++i) (void)or_xoshiro_next_u64(s);
}

/* Xoshiro256PlusPlus::next_f64, fast_rng.rs:69-73 */
double or_xoshiro_next_f64(uint64_t s[4]) {
    const uint64_t x = or_xoshiro_next_u64(s) >> 11;
    return (double)x * (1.0 / 9007199254740992.0);
}

/* stream_seed, fast_rng.rs:202-205 (wrapping arithmetic) */
uint64_t or_stream_seed(uint64_t base, uint64_t index) {
    return base + index * 7919ull;
}

/* uniform_open01, fast_rng.rs:218-221: f64::clamp(EPSILON, 1 - EPSILON) */
double or_uniform_open01(double u) {
    const double lo = 2.220446049250313e-16;
    const double hi = 1.0 - 2.220446049250313e-16;
    if (u < lo) return lo;
    if (u > hi) return hi;
    return u;
}

/* ===================================================================== */
/* Normal distribution: src/math/fast_norm.rs                            */
/* ===================================================================== */

/* fast_norm_pdf, fast_norm.rs:14-19 */
double or_fast_norm_pdf(double x) {
    const double INV_SQRT_2PI = 0.3989422804014327;
    return INV_SQRT_2PI * exp((-0.5) * x * x);
}

/* hart_norm_cdf, fast_norm.rs:27-50 */
double or_hart_norm_cdf(double x) {
    const double P = 0.2316419;
    const double A1 = 0.319381530, A2 = -0.356563782, A3 = 1.781477937,
                 A4 = -1.821255978, A5 = 1.330274429;
    const double z = fabs(x);
    const double t = 1.0 / fma(P, z, 1.0);
    const double poly = fma(fma(fma(fma(A5, t, A4), t, A3), t, A2), t, A1) * t;
    const double cdf_pos = fma(or_fast_norm_pdf(z), -poly, 1.0);
    uint64_t bits;
    memcpy(&bits, &x, sizeof bits);
    const double sign = (double)(bits >> 63);
    return fma(sign, 1.0 - 2.0 * cdf_pos, cdf_pos);
}

/* Acklam rational approximation (named beasley_springer_moro_inv_cdf in the
 * reference), fast_norm.rs:53-139. Coefficients are Acklam's published ones
 * (fast_norm.rs:66-94). */
static const double AK_A[6] = {-3.969683028665376e1, 2.209460984245205e2,
                               -2.759285104469687e2, 1.383577518672690e2,
                               -3.066479806614716e1, 2.506628277459239};
static const double AK_B[5] = {-5.447609879822406e1, 1.615858368580409e2,
                               -1.556989798598866e2, 6.680131188771972e1,
                               -1.328068155288572e1};
static const double AK_C[6] = {-7.784894002430293e-3, -3.223964580411365e-1,
                               -2.400758277161838, -2.549732539343734,
                               4.374664141464968, 2.938163982698783};
static const double AK_D[4] = {7.784695709041462e-3, 3.224671290700398e-1,
                               2.445134137142996, 3.754408661907416};

double or_acklam_inv_cdf(double p) {
    if (isnan(p) || !(p >= 0.0 && p <= 1.0)) return NAN;
    if (p <= 0.0) return -INFINITY;
    if (p >= 1.0) return INFINITY;
    const double P_LOW = 0.02425;
    const double P_HIGH = 1.0 - P_LOW;
    if (p < P_LOW) {
        const double q = sqrt(-2.0 * log(p));
        const double num = fma(fma(fma(fma(fma(AK_C[0], q, AK_C[1]), q, AK_C[2]), q, AK_C[3]), q, AK_C[4]), q, AK_C[5]);
        const double den = fma(fma(fma(fma(AK_D[0], q, AK_D[1]), q, AK_D[2]), q, AK_D[3]), q, 1.0);
        return num / den;
    } else if (p <= P_HIGH) {
        const double q = p - 0.5;
        const double r = q * q;
        const double num = fma(fma(fma(fma(fma(AK_A[0], r, AK_A[1]), r, AK_A[2]), r, AK_A[3]), r, AK_A[4]), r, AK_A[5]);
        const double den = fma(fma(fma(fma(fma(AK_B[0], r, AK_B[1]), r, AK_B[2]), r, AK_B[3]), r, AK_B[4]), r, 1.0);
        return num * q / den;
    } else {
        const double q = sqrt(-2.0 * log(1.0 - p));
        const double num = fma(fma(fma(fma(fma(AK_C[0], q, AK_C[1]), q, AK_C[2]), q, AK_C[3]), q, AK_C[4]), q, AK_C[5]);
        const double den = fma(fma(fma(fma(AK_D[0], q, AK_D[1]), q, AK_D[2]), q, AK_D[3]), q, 1.0);
        /* In Rust `-x.f()` is `-(x.f())`: the numerator chain is negated as a
           whole before the division (fast_norm.rs:127-137). */
        return -num / den;
    }
}

/* ===================================================================== */
/* Models and generators                                                  */
/* ===================================================================== */

/* GbmPathGenerator::generate_into, simulation.rs:58-71 */
void or_gbm_generate(double mu, double sigma, double s0, double maturity,
                     int64_t steps, const double *z, double *out) {
    const double dt = maturity / (double)steps;
    const double sqrt_dt = sqrt(dt);
    const double drift = (mu - 0.5 * sigma * sigma) * dt;
    const double diffusion = sigma * sqrt_dt;
    double s = s0;
    out[0] = s;
    for (int64_t j = 0; j < steps; ++j) {
        s *= exp(fma(diffusion, z[j], drift));
        out[j + 1] = s;
    }
}

/* Heston::step_euler, stochastic.rs:56-71 */
void or_heston_step_euler(double mu, double kappa, double theta, double xi,
                          double rho, double s, double v, double dt,
                          double z1, double z2, double *s_next, double *v_next) {
    const double v_pos = fmax(v, 0.0);
    const double sqrt_dt = sqrt(dt);
    const double zv = z1;
    const double zs = rho * z1 + sqrt(1.0 - rho * rho) * z2;
    const double vn = v + kappa * (theta - v_pos) * dt + xi * sqrt(v_pos) * sqrt_dt * zv;
    *v_next = fmax(vn, 0.0);
    *s_next = s * exp((mu - 0.5 * v_pos) * dt + sqrt(v_pos) * sqrt_dt * zs);
}

/* HestonPathGenerator::generate_into, simulation.rs:98-116 */
void or_heston_generate(double mu, double kappa, double theta, double xi,
                        double rho, double v0, double s0, double maturity,
                        int64_t steps, const double *z1, const double *z2,
                        double *out) {
    const double dt = maturity / (double)steps;
    double s = s0, v = v0;
    out[0] = s;
    for (int64_t j = 0; j < steps; ++j) {
        double sn, vn;
        or_heston_step_euler(mu, kappa, theta, xi, rho, s, v, dt, z1[j], z2[j], &sn, &vn);
        s = fmax(sn, 1e-12);
        v = fmax(vn, 0.0);
        out[j + 1] = s;
    }
}

/* ===================================================================== */
/* Payoffs: src/engines/monte_carlo/mc_engine.rs                          */
/* ===================================================================== */

/* vanilla_payoff, mc_engine.rs:59-64 */
static inline double vanilla_payoff(int32_t type, double spot, double strike) {
    return type == OR_CALL ? fmax(spot - strike, 0.0) : fmax(strike - spot, 0.0);
}

/* `((t / maturity) * last_idx).round() as usize` clamped to path.len()-1,
 * mc_engine.rs:103-104. Rust's `as usize` saturates (negative/NaN -> 0). */
static inline int64_t obs_index(double t, double maturity, int64_t len) {
    const double last_idx = (double)(len - 1);
    const double x = round((t / maturity) * last_idx);
    int64_t idx;
    if (!(x > 0.0)) idx = 0;
    else if (x >= 9.2e18) idx = INT64_MAX;
    else idx = (int64_t)x;
    return idx < len - 1 ? idx : len - 1;
}

/* average_for_observations, mc_engine.rs:86-121 */
static double average_for_observations(const double *path, int64_t len,
                                       double maturity, const double *obs,
                                       int64_t n_obs, int32_t averaging) {
    if (n_obs == 0 || maturity <= 0.0) return path[len - 1];
    if (averaging == OR_ARITH) {
        double sum = 0.0;
        for (int64_t i = 0; i < n_obs; ++i) sum += path[obs_index(obs[i], maturity, len)];
        return sum / (double)n_obs;
    } else {
        double sum = 0.0;
        for (int64_t i = 0; i < n_obs; ++i)
            sum += log(fmax(path[obs_index(obs[i], maturity, len)], 1e-12));
        const double mean_log = sum / (double)n_obs;
        return exp(mean_log);
    }
}

/* path_hits_barrier, mc_engine.rs:79-84 (inclusive, includes path[0]) */
static int path_hits_barrier(const double *path, int64_t len, double level, int32_t dir) {
    for (int64_t i = 0; i < len; ++i) {
        if (dir == OR_UP ? path[i] >= level : path[i] <= level) return 1;
    }
    return 0;
}

/* MonteCarloInstrument::payoff_from_path for Vanilla (:454-456),
 * Barrier (:501-513) and Asian (:532-544). */
double or_payoff(const or_option *o, const double *path, int64_t len) {
    const double st = path[len - 1];
    switch (o->payoff) {
    case OR_VANILLA:
        return vanilla_payoff(o->option_type, st, o->strike);
    case OR_BARRIER: {
        const int hit = path_hits_barrier(path, len, o->barrier_level, o->barrier_dir);
        const int active = o->barrier_style == OR_IN ? hit : !hit;
        return active ? vanilla_payoff(o->option_type, st, o->strike) : o->rebate;
    }
    case OR_ASIAN: {
        const double avg = average_for_observations(path, len, o->expiry, o->obs_times,
                                                    o->n_obs, o->averaging);
        return o->strike_type == OR_FIXED ? vanilla_payoff(o->option_type, avg, o->strike)
                                          : vanilla_payoff(o->option_type, st, avg);
    }
    default:
        return NAN;
    }
}

/* Control-variate evaluators: vanilla terminal payoff (mc_engine.rs:481-483)
 * and geometric Asian payoff (mc_engine.rs:573-581). */
double or_cv_eval(const or_option *o, int32_t cv_kind, const double *path, int64_t len) {
    if (cv_kind == 0) return vanilla_payoff(o->option_type, path[len - 1], o->strike);
    const double g = average_for_observations(path, len, o->expiry, o->obs_times, o->n_obs, OR_GEO);
    return vanilla_payoff(o->option_type, g, o->strike);
}

/* ===================================================================== */
/* Closed forms                                                           */
/* ===================================================================== */

/* d1_d2 and bs_price (generic, non-asm branch), black_scholes.rs:55-107 */
static double bs_price_with_cdf(int32_t type, double spot, double strike, double rate,
                                double q, double vol, double expiry,
                                double (*cdf)(double)) {
    if (expiry <= 0.0) return vanilla_payoff(type, spot, strike);
    const double df_r = exp(-rate * expiry);
    const double df_q = exp(-q * expiry);
    if (vol <= 0.0) {
        return type == OR_CALL ? fmax(spot * df_q - strike * df_r, 0.0)
                               : fmax(strike * df_r - spot * df_q, 0.0);
    }
    const double sig_sqrt_t = vol * sqrt(expiry);
    const double d1 = (log(spot / strike) + fma(0.5 * vol, vol, rate - q) * expiry) / sig_sqrt_t;
    const double d2 = d1 - sig_sqrt_t;
    const double nd1 = cdf(d1), nd2 = cdf(d2);
    const double s_df_q = spot * df_q;
    const double k_df_r = strike * df_r;
    const double call = fma(s_df_q, nd1, -(k_df_r * nd2));
    return type == OR_CALL ? call : call - s_df_q + k_df_r;
}

static double exact_cdf(double x) { return 0.5 * erfc(-x / sqrt(2.0)); }

double or_bs_price(int32_t type, double spot, double strike, double rate, double q,
                   double vol, double expiry) {
    return bs_price_with_cdf(type, spot, strike, rate, q, vol, expiry, or_hart_norm_cdf);
}

double or_bs_price_exact(int32_t type, double spot, double strike, double rate, double q,
                         double vol, double expiry) {
    return bs_price_with_cdf(type, spot, strike, rate, q, vol, expiry, exact_cdf);
}

/* bs_delta/bs_vega/bs_theta/bs_rho/bs_gamma, black_scholes.rs:110-215, with the
 * exact normal CDF. out = delta, gamma, vega, theta, rho. */
void or_bs_greeks(int32_t type, double spot, double strike, double rate, double q,
                  double vol, double expiry, double out[5]) {
    if (expiry <= 0.0 || vol <= 0.0 || spot <= 0.0) {
        for (int i = 0; i < 5; ++i) out[i] = 0.0;
        return;
    }
    const double sqrt_t = sqrt(expiry);
    const double sig_sqrt_t = vol * sqrt_t;
    const double d1 = (log(spot / strike) + fma(0.5 * vol, vol, rate - q) * expiry) / sig_sqrt_t;
    const double d2 = d1 - sig_sqrt_t;
    const double df_q = exp(-q * expiry), df_r = exp(-rate * expiry);
    const double nd1 = exact_cdf(d1), nd2 = exact_cdf(d2);
    const double pdf1 = 0.3989422804014327 * exp(-0.5 * d1 * d1);
    out[0] = type == OR_CALL ? df_q * nd1 : df_q * (nd1 - 1.0);
    out[1] = df_q * pdf1 / (spot * vol * sqrt_t);
    out[2] = spot * df_q * pdf1 * sqrt_t;
    const double theta_common = -spot * df_q * pdf1 * vol / (2.0 * sqrt_t);
    out[3] = type == OR_CALL
                 ? theta_common + q * spot * df_q * nd1 - rate * strike * df_r * nd2
                 : theta_common - q * spot * df_q * (1.0 - nd1) + rate * strike * df_r * (1.0 - nd2);
    out[4] = type == OR_CALL ? strike * expiry * df_r * nd2
                             : -strike * expiry * df_r * (1.0 - nd2);
}

/* geometric_asian_discrete_fixed_closed_form, pricing/asian.rs:86-136
 * (normal_cdf = Hart approximation, math/functions.rs:56-58). */
double or_geometric_asian_discrete(int32_t type, double s0, double k, double r,
                                   double q, double sigma, const double *obs,
                                   int64_t n_obs) {
    if (n_obs == 0) return vanilla_payoff(type, s0, k);
    double maturity = -INFINITY;
    for (int64_t i = 0; i < n_obs; ++i) maturity = fmax(maturity, obs[i]);
    maturity = fmax(maturity, 0.0);
    if (maturity <= 0.0 || sigma <= 0.0) return exp(-r * maturity) * vanilla_payoff(type, s0, k);
    const double n = (double)n_obs;
    double sum_t = 0.0;
    for (int64_t i = 0; i < n_obs; ++i) sum_t += obs[i];
    const double mean_t = sum_t / n;
    double cov_sum = 0.0;
    for (int64_t i = 0; i < n_obs; ++i)
        for (int64_t j = 0; j < n_obs; ++j) cov_sum += fmin(obs[i], obs[j]);
    const double sigma_sq = sigma * sigma;
    const double var = sigma_sq * cov_sum / (n * n);
    const double m = log(s0) + fma(-0.5 * sigma, sigma, r - q) * mean_t;
    const double df = exp(-r * maturity);
    if (var <= 0.0) return df * vanilla_payoff(type, exp(m), k);
    const double sqrt_v = sqrt(var);
    const double d1 = (m - log(k) + var) / sqrt_v;
    const double d2 = d1 - sqrt_v;
    const double eg = exp(m + 0.5 * var);
    return type == OR_CALL
               ? df * (eg * or_hart_norm_cdf(d1) - k * or_hart_norm_cdf(d2))
               : df * (k * or_hart_norm_cdf(-d2) - eg * or_hart_norm_cdf(-d1));
}

/* Built-in control variates (mc_engine.rs:458-485, 546-583): returns the
 * undiscounted expectation and the evaluator kind, or NAN when the instrument
 * has none. */
double or_builtin_cv_expected(const or_option *o, int32_t *cv_kind) {
    const double df = exp(-o->rate * o->expiry);
    if (o->payoff == OR_VANILLA) {
        /* black_scholes_price_with_dividend, mc_engine.rs:66-77 */
        const double adj = o->spot * exp(-o->div_yield * o->expiry);
        const double e = or_bs_price(o->option_type, adj, o->strike, o->rate, 0.0, o->vol, o->expiry);
        *cv_kind = 0;
        return e / df;
    }
    if (o->payoff == OR_ASIAN && o->averaging == OR_ARITH && o->strike_type == OR_FIXED) {
        const double e = or_geometric_asian_discrete(o->option_type, o->spot, o->strike, o->rate,
                                                     o->div_yield, o->vol, o->obs_times, o->n_obs);
        *cv_kind = 1;
        return e / df;
    }
    *cv_kind = -1;
    return NAN;
}

/* ===================================================================== */
/* MonteCarloEngine::run, simulation.rs:182-348                           */
/* ===================================================================== */

void or_sample_normals(uint64_t base_seed, int64_t sample, int64_t steps,
                       int32_t streams, double *z1, double *z2) {
    or_fast_rng g;
    or_fast_rng_seed(&g, or_rng_kind(), or_stream_seed(base_seed, (uint64_t)sample));
    for (int64_t j = 0; j < steps; ++j) {
        z1[j] = or_fast_rng_normal(&g);
        if (streams >= 2) z2[j] = or_fast_rng_normal(&g);
    }
}

typedef struct { double sx, sx2, sy, sxy, sy2; uint64_t n; } acc6;

static void generate(const or_option *o, int64_t steps, const double *z1,
                     const double *z2, double *path) {
    const double mu = o->rate - o->div_yield; /* mc_engine.rs:916 */
    if (o->model == OR_GBM)
        or_gbm_generate(mu, o->vol, o->spot, o->expiry, steps, z1, path);
    else
        or_heston_generate(mu, o->kappa, o->theta, o->xi, o->rho, o->v0, o->spot,
                           o->expiry, steps, z1, z2, path);
}

/* fold_fn body, simulation.rs:212-262 */
static void fold_sample(const or_option *o, const or_cfg *cfg, int64_t i,
                        const double *normals1, const double *normals2,
                        double *z1, double *z2, double *path, acc6 *acc,
                        double *xs, double *ys) {
    const int64_t steps = cfg->num_steps;
    const int32_t streams = o->model == OR_GBM ? 1 : 2;
    if (normals1) {
        memcpy(z1, normals1 + i * steps, sizeof(double) * (size_t)steps);
        if (streams >= 2) memcpy(z2, normals2 + i * steps, sizeof(double) * (size_t)steps);
    } else {
        or_sample_normals(cfg->seed, i, steps, streams, z1, z2);
    }
    const int64_t len = steps + 1;
    generate(o, steps, z1, z2, path);
    double x = or_payoff(o, path, len);
    double y = cfg->control_variate ? or_cv_eval(o, cfg->cv_kind, path, len) : 0.0;
    if (cfg->antithetic) {
        for (int64_t j = 0; j < steps; ++j) z1[j] = -z1[j];
        if (streams >= 2)
            for (int64_t j = 0; j < steps; ++j) z2[j] = -z2[j];
        generate(o, steps, z1, z2, path);
        const double xa = or_payoff(o, path, len);
        const double ya = cfg->control_variate ? or_cv_eval(o, cfg->cv_kind, path, len) : 0.0;
        x = 0.5 * (x + xa);
        y = 0.5 * (y + ya);
    }
    acc->sx += x;
    acc->sx2 += x * x;
    acc->sy += y;
    acc->sxy += x * y;
    acc->sy2 += y * y;
    acc->n += 1;
    if (xs) xs[i] = x;
    if (ys) ys[i] = y;
}

int or_mc_run(const or_option *o, const or_cfg *cfg, int32_t nthreads,
              const double *normals1, const double *normals2,
              double *price, double *stderr_out, double *xs, double *ys) {
    if (cfg->num_paths <= 0 || cfg->num_steps <= 0) return 1;
    const int64_t samples = cfg->antithetic ? (cfg->num_paths + 1) / 2 : cfg->num_paths;
    const int64_t steps = cfg->num_steps;
    if (nthreads < 1) nthreads = 1;

    acc6 total = {0, 0, 0, 0, 0, 0};
    acc6 *parts = (acc6 *)calloc((size_t)nthreads, sizeof(acc6));
    if (!parts) return 2;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) if (nthreads > 1)
#endif
    {
        int tid = 0, nt = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        double *z1 = (double *)malloc(sizeof(double) * (size_t)steps);
        double *z2 = (double *)malloc(sizeof(double) * (size_t)steps);
        double *path = (double *)malloc(sizeof(double) * (size_t)(steps + 1));
        acc6 a = {0, 0, 0, 0, 0, 0};
        /* contiguous chunks per worker, as rayon's fold splits the range */
        const int64_t lo = samples * tid / nt, hi = samples * (tid + 1) / nt;
        for (int64_t i = lo; i < hi; ++i)
            fold_sample(o, cfg, i, normals1, normals2, z1, z2, path, &a, xs, ys);
        parts[tid] = a;
        free(z1);
        free(z2);
        free(path);
    }
    /* reduce (simulation.rs:266-290): pairwise tree over worker partials */
    for (int width = 1; width < nthreads; width *= 2) {
        for (int i = 0; i + width < nthreads; i += 2 * width) {
            parts[i].sx += parts[i + width].sx;
            parts[i].sx2 += parts[i + width].sx2;
            parts[i].sy += parts[i + width].sy;
            parts[i].sxy += parts[i + width].sxy;
            parts[i].sy2 += parts[i + width].sy2;
            parts[i].n += parts[i + width].n;
        }
    }
    total = parts[0];
    free(parts);

    const double df = exp(-o->rate * o->expiry);
    const double n = (double)total.n;
    if (cfg->control_variate) {
        /* simulation.rs:307-336 */
        const double denom = fmax(n - 1.0, 1.0);
        const double cov_xy = (total.sxy - total.sx * total.sy / n) / denom;
        const double var_y = (total.sy2 - total.sy * total.sy / n) / denom;
        const double beta = var_y > 1e-16 ? cov_xy / var_y : 0.0;
        const double e = cfg->cv_expected;
        const double sum_adj = total.sx + beta * (n * e - total.sy);
        const double sum_adj_sq = total.sx2 + 2.0 * beta * e * total.sx - 2.0 * beta * total.sxy +
                                  beta * beta * (n * e * e - 2.0 * e * total.sy + total.sy2);
        const double mean = sum_adj / n;
        const double var = n > 1.0 ? (sum_adj_sq - sum_adj * sum_adj / n) / (n - 1.0) : 0.0;
        *price = df * mean;
        *stderr_out = df * sqrt(var / n);
    } else {
        /* simulation.rs:337-347 */
        const double mean = total.sx / n;
        const double var = n > 1.0 ? (total.sx2 - total.sx * total.sx / n) / (n - 1.0) : 0.0;
        *price = df * mean;
        *stderr_out = df * sqrt(var / n);
    }
    return 0;
}

/* ===================================================================== */
/* Pathwise AAD: src/engines/monte_carlo/mc_aad.rs, src/math/aad.rs       */
/* ===================================================================== */

/* Minimal restatement of AadTape (aad.rs:424-680): nodes hold a value and up
 * to two (parent, weight) edges; reverse() sweeps from the output down and
 * skips zero adjoints (aad.rs:630-658). */
typedef struct { int32_t l, r; double wl, wr; } tape_op;
typedef struct { double *val, *adj; tape_op *op; int64_t n, cap; } tape;

static int64_t tp_push(tape *t, double v, int32_t l, double wl, int32_t r, double wr) {
    t->val[t->n] = v;
    t->op[t->n].l = l;
    t->op[t->n].r = r;
    t->op[t->n].wl = wl;
    t->op[t->n].wr = wr;
    return t->n++;
}
static int64_t tp_leaf(tape *t, double v) { return tp_push(t, v, -1, 0, -1, 0); }
static int64_t tp_add(tape *t, int64_t a, int64_t b) { return tp_push(t, t->val[a] + t->val[b], (int32_t)a, 1.0, (int32_t)b, 1.0); }
static int64_t tp_sub(tape *t, int64_t a, int64_t b) { return tp_push(t, t->val[a] - t->val[b], (int32_t)a, 1.0, (int32_t)b, -1.0); }
static int64_t tp_mul(tape *t, int64_t a, int64_t b) {
    const double av = t->val[a], bv = t->val[b];
    return tp_push(t, av * bv, (int32_t)a, bv, (int32_t)b, av);
}
static int64_t tp_div(tape *t, int64_t a, int64_t b) {
    const double av = t->val[a], bv = t->val[b];
    const double inv = 1.0 / bv;
    return tp_push(t, av * inv, (int32_t)a, inv, (int32_t)b, -av * inv * inv);
}
static int64_t tp_neg(tape *t, int64_t a) { return tp_push(t, -t->val[a], (int32_t)a, -1.0, -1, 0); }
static int64_t tp_exp(tape *t, int64_t a) {
    const double v = exp(t->val[a]);
    return tp_push(t, v, (int32_t)a, v, -1, 0);
}
static int64_t tp_sqrt(tape *t, int64_t a) {
    const double v = sqrt(t->val[a]);
    return tp_push(t, v, (int32_t)a, 0.5 / v, -1, 0);
}
static int64_t tp_pos(tape *t, int64_t a) {
    const double av = t->val[a];
    return tp_push(t, fmax(av, 0.0), (int32_t)a, av > 0.0 ? 1.0 : 0.0, -1, 0);
}
static void tp_reverse(tape *t, int64_t out) {
    memset(t->adj, 0, sizeof(double) * (size_t)t->n);
    t->adj[out] = 1.0;
    for (int64_t i = out; i >= 0; --i) {
        const double bar = t->adj[i];
        if (bar == 0.0) continue;
        if (t->op[i].l >= 0) t->adj[t->op[i].l] += bar * t->op[i].wl;
        if (t->op[i].r >= 0) t->adj[t->op[i].r] += bar * t->op[i].wr;
    }
}

/* gbm_single_path_reverse, mc_aad.rs:16-72 */
static double gbm_single_path_reverse(tape *t, int32_t type, double strike, double spot,
                                      double rate, double q, double vol, double expiry,
                                      const double *normals, int64_t n, double g[4]) {
    t->n = 0;
    const int64_t s0 = tp_leaf(t, spot), r = tp_leaf(t, rate), sigma = tp_leaf(t, vol),
                  tt = tp_leaf(t, expiry);
    const int64_t qv = tp_leaf(t, q), k = tp_leaf(t, strike), steps = tp_leaf(t, (double)n),
                  half = tp_leaf(t, 0.5);
    const int64_t dt = tp_div(t, tt, steps);
    const int64_t sqrt_dt = tp_sqrt(t, dt);
    const int64_t sigma_sq = tp_mul(t, sigma, sigma);
    const int64_t half_sigma_sq = tp_mul(t, half, sigma_sq);
    const int64_t r_minus_q = tp_sub(t, r, qv);
    const int64_t drift_inner = tp_sub(t, r_minus_q, half_sigma_sq);
    const int64_t drift = tp_mul(t, drift_inner, dt);
    const int64_t diffusion = tp_mul(t, sigma, sqrt_dt);
    int64_t s = s0;
    for (int64_t j = 0; j < n; ++j) {
        const int64_t zv = tp_leaf(t, normals[j]);
        const int64_t diff_term = tp_mul(t, diffusion, zv);
        const int64_t expo = tp_add(t, drift, diff_term);
        const int64_t growth = tp_exp(t, expo);
        s = tp_mul(t, s, growth);
    }
    const int64_t intrinsic = type == OR_CALL ? tp_sub(t, s, k) : tp_sub(t, k, s);
    const int64_t payoff = tp_pos(t, intrinsic);
    const int64_t rt = tp_mul(t, r, tt);
    const int64_t minus_rt = tp_neg(t, rt);
    const int64_t discount = tp_exp(t, minus_rt);
    const int64_t pv = tp_mul(t, discount, payoff);
    tp_reverse(t, pv);
    g[0] = t->adj[s0];
    g[1] = t->adj[r];
    g[2] = t->adj[sigma];
    g[3] = t->adj[tt];
    return t->val[pv];
}

/* mc_european_pathwise_aad, mc_aad.rs:75-219 (Vanilla European; the dividend
 * yield passed in is the market's effective yield, equal to the continuous
 * one without discrete dividends). */
int or_mc_european_pathwise_aad(const or_option *o, int64_t num_paths, int64_t num_steps,
                                uint64_t seed, int32_t antithetic, double out[6]) {
    if (num_paths <= 0 || num_steps <= 0) return 1;
    if (o->expiry <= 0.0) {
        out[0] = vanilla_payoff(o->option_type, o->spot, o->strike);
        for (int i = 1; i < 6; ++i) out[i] = 0.0;
        return 0;
    }
    const int64_t samples = antithetic ? (num_paths + 1) / 2 : num_paths;
    tape t;
    t.cap = num_steps * 8 + 64;
    t.val = (double *)malloc(sizeof(double) * (size_t)t.cap);
    t.adj = (double *)malloc(sizeof(double) * (size_t)t.cap);
    t.op = (tape_op *)malloc(sizeof(tape_op) * (size_t)t.cap);
    double *normals = (double *)malloc(sizeof(double) * (size_t)num_steps);
    double sum = 0.0, sum_sq = 0.0, gs[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < samples; ++i) {
        or_sample_normals(seed, i, num_steps, 1, normals, NULL);
        double g0[4], g1[4], g[4];
        const double pv0 = gbm_single_path_reverse(&t, o->option_type, o->strike, o->spot, o->rate,
                                                   o->div_yield, o->vol, o->expiry, normals,
                                                   num_steps, g0);
        double pv;
        if (antithetic) {
            for (int64_t j = 0; j < num_steps; ++j) normals[j] = -normals[j];
            const double pv1 = gbm_single_path_reverse(&t, o->option_type, o->strike, o->spot,
                                                       o->rate, o->div_yield, o->vol, o->expiry,
                                                       normals, num_steps, g1);
            for (int j = 0; j < 4; ++j) g[j] = 0.5 * (g0[j] + g1[j]);
            pv = 0.5 * (pv0 + pv1);
        } else {
            for (int j = 0; j < 4; ++j) g[j] = g0[j];
            pv = pv0;
        }
        sum += pv;
        sum_sq += pv * pv;
        for (int j = 0; j < 4; ++j) gs[j] += g[j];
    }
    const double n = (double)samples;
    const double mean = sum / n;
    const double variance = samples > 1 ? (sum_sq - sum * sum / n) / (n - 1.0) : 0.0;
    out[0] = mean;
    out[1] = sqrt(variance / n);
    out[2] = gs[0] / n;     /* delta */
    out[3] = gs[2] / n;     /* vega */
    out[4] = -(gs[3] / n);  /* theta */
    out[5] = gs[1] / n;     /* rho */
    free(t.val);
    free(t.adj);
    free(t.op);
    free(normals);
    return 0;
}

/* heston_single_path_price_delta + heston_price_delta_aad,
 * mc_aad.rs:239-342 (forward-mode dual on the spot; no S clamp). */
int or_heston_price_delta_aad(const or_option *o, int64_t num_paths, int64_t num_steps,
                              uint64_t seed, double out[2]) {
    if (num_paths <= 0 || num_steps <= 0) return 1;
    if (o->expiry <= 0.0) {
        out[0] = vanilla_payoff(o->option_type, o->spot, o->strike);
        out[1] = 0.0;
        return 0;
    }
    const double mu = o->rate - o->div_yield;
    double *z1 = (double *)malloc(sizeof(double) * (size_t)num_steps);
    double *z2 = (double *)malloc(sizeof(double) * (size_t)num_steps);
    double sum_price = 0.0, sum_delta = 0.0;
    const double dt = o->expiry / (double)num_steps;
    const double sqrt_dt = sqrt(dt);
    const double rho2 = sqrt(1.0 - o->rho * o->rho);
    const double discount = exp(-o->rate * o->expiry);
    for (int64_t i = 0; i < num_paths; ++i) {
        or_sample_normals(seed, i, num_steps, 2, z1, z2);
        double s = o->spot, ds = 1.0, v = o->v0;
        for (int64_t j = 0; j < num_steps; ++j) {
            const double v_pos = fmax(v, 0.0);
            const double sqrt_v = sqrt(v_pos);
            const double zs = o->rho * z1[j] + rho2 * z2[j];
            const double v_next = fmax(v + o->kappa * (o->theta - v_pos) * dt + o->xi * sqrt_v * sqrt_dt * z1[j], 0.0);
            const double growth = exp((mu - 0.5 * v_pos) * dt + sqrt_v * sqrt_dt * zs);
            s = s * growth;
            ds = ds * growth;
            v = v_next;
        }
        double pv, dpv;
        if (o->option_type == OR_CALL) {
            const double val = s - o->strike;
            pv = val > 0.0 ? val : 0.0;
            dpv = val > 0.0 ? ds : 0.0;
        } else {
            const double val = o->strike - s;
            pv = val > 0.0 ? val : 0.0;
            dpv = val > 0.0 ? (0.0 - ds) : 0.0;
        }
        sum_price += pv * discount;
        sum_delta += dpv * discount;
    }
    out[0] = sum_price / (double)num_paths;
    out[1] = sum_delta / (double)num_paths;
    free(z1);
    free(z2);
    return 0;
}

/* ===================================================================== */
/* MonteCarloGreeksEngine, mc_greeks.rs:131-314                            */
/* ===================================================================== */

/* pathwise_delta, mc_greeks.rs:293-310 */
static double mcg_pathwise_delta(int32_t type, double st, double strike, double growth) {
    if (type == OR_CALL) return st > strike ? growth : 0.0;
    return st < strike ? -growth : 0.0;
}

/* single_path_contribution, mc_greeks.rs:245-284 */
static void mcg_single_path(int32_t type, double z, double spot, double strike,
                            double spot_up, double spot_dn, double spot_span, double vol,
                            double maturity, double drift, double sig_sqrt_t, double e[5]) {
    const double growth = exp(fma(sig_sqrt_t, z, drift));
    const double st = spot * growth;
    const double payoff = vanilla_payoff(type, st, strike);
    const double delta_pw = mcg_pathwise_delta(type, st, strike, growth);
    const double delta_up = mcg_pathwise_delta(type, spot_up * growth, strike, growth);
    const double delta_dn = mcg_pathwise_delta(type, spot_dn * growth, strike, growth);
    const double gamma_pw = (delta_up - delta_dn) / spot_span;
    const double sqrt_t = sqrt(maturity);
    const double inv_spot_vol_sqrt_t = 1.0 / (spot * vol * sqrt_t);
    const double z2 = z * z;
    const double z2m1 = z2 - 1.0;
    const double w_delta = z * inv_spot_vol_sqrt_t;
    const double inv_spot2_vol2_t = inv_spot_vol_sqrt_t * inv_spot_vol_sqrt_t;
    const double w_gamma = z2m1 * inv_spot2_vol2_t - z * inv_spot_vol_sqrt_t / spot;
    const double w_vega = (z2m1 - vol * sqrt_t * z) / vol;
    e[0] = delta_pw;
    e[1] = gamma_pw;
    e[2] = payoff * w_delta;
    e[3] = payoff * w_gamma;
    e[4] = payoff * w_vega;
}

/* run_estimators, mc_greeks.rs:131-243 (the non-parallel fold) */
int or_mc_greeks_estimate(const or_option *o, int64_t num_paths, uint64_t seed,
                          int32_t antithetic, double spot_bump_rel, double out[5]) {
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
    if (!(o->strike > 0.0) || o->expiry < 0.0) return 1;
    if (num_paths <= 0) return 1;
    if (!(spot_bump_rel > 0.0) || !isfinite(spot_bump_rel)) return 1;
    if (o->expiry <= 0.0) return 0;
    const double spot = o->spot;
    if (spot <= 0.0) return 1;
    const double strike = o->strike, maturity = o->expiry, vol = o->vol;
    if (vol <= 0.0 || !isfinite(vol)) return 1;
    const double rate = o->rate, dividend = o->div_yield;
    const double drift = (rate - dividend - 0.5 * vol * vol) * maturity;
    const double sig_sqrt_t = vol * sqrt(maturity);
    const double discount = exp(-rate * maturity);
    double h = spot * spot_bump_rel;
    if (h < 1.0e-6) h = 1.0e-6;
    const double spot_up = spot + h;
    double spot_dn = spot - h;
    if (spot_dn < 1.0e-8) spot_dn = 1.0e-8;
    const double spot_span = spot_up - spot_dn;
    const int64_t samples = antithetic ? (num_paths + 1) / 2 : num_paths;
    double sums[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t i = 0; i < samples; ++i) {
        or_fast_rng g;
        or_fast_rng_seed(&g, or_rng_kind(), or_stream_seed(seed, (uint64_t)i));
        const double z = or_fast_rng_normal(&g);
        double b[5], a[5];
        mcg_single_path(o->option_type, z, spot, strike, spot_up, spot_dn, spot_span, vol,
                        maturity, drift, sig_sqrt_t, b);
        if (antithetic) {
            mcg_single_path(o->option_type, -z, spot, strike, spot_up, spot_dn, spot_span, vol,
                            maturity, drift, sig_sqrt_t, a);
            for (int k = 0; k < 5; ++k) b[k] = 0.5 * (b[k] + a[k]);
        }
        for (int k = 0; k < 5; ++k) sums[k] = sums[k] + b[k];
    }
    const double n = (double)samples;
    for (int k = 0; k < 5; ++k) out[k] = discount * sums[k] / n;
    return 0;
}
