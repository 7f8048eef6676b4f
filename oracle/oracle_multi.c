/*
 * CPU oracle, multi-asset part -- TEST INFRASTRUCTURE ONLY (the tests' reference for
 * the HIP multi-asset path and its timed CPU baseline).
 *
 * Restates src/pricing/autocallable.rs of rosssaunders/openferric:
 * prepare_common (:301-355, selection of the contract's underlyings),
 * observation_schedule (:582-592), cholesky_lower (:669-700),
 * correlate_normals (:659-667), worst_of_ratio (:594-600) and
 * simulate_autocallable_paths (:441-579, standard and Phoenix coupons with
 * memory, knock-in monitored every step, early call), and the bump helpers of
 * bump_and_reprice_sensitivities (:366-418, bump_corr_matrix :702-714).
 * Compiled with -ffp-contract=off (Rust never contracts a*b+c).
 *
 * Draws: normals == NULL walks the engine's single FastRng stream
 * (Xoshiro256PlusPlus seeded with `seed`, sample_standard_normal), exactly
 * as the reference; otherwise normals[path][step][asset] (a called path
 * leaves the rest of its slice unused).
 *
 * Also restates src/pricing/basket.rs: terminal_payoff (:759-784) and
 * simulate_basket_once_matrix under the Gaussian copula (:599-683: one exact
 * terminal step per asset, correlate_normals, RunningStats' Welford mean and
 * variance, discount exp(-rT), zero maturity = payoff of the spots) on a PSD
 * correlation matrix (the reference repairs others first, which the tests do
 * with an independent Higham projection).
 *
 * Pinned by (tests/test_oracle_multi.py): zero-volatility paths with closed
 * form present values (call dates, standard and Phoenix coupons, memory);
 * never-called / always-called limits; certain knock-in, where the price is
 * e^{-rT} N E[min_i S_i(T)/S_i(0)] / K_ki with E[S(T)/S(0)] = e^{(r-q)T} for
 * one asset and Margrabe's exchange-option formula for two correlated
 * assets; the reference's own monotonicity and single-underlying tests.
 * Factor-model baskets (or_basket_factor_price): pinned by the model's
 * implied correlation -- a two-asset best-of under a one-factor model
 * equals Margrabe with rho = beta1 beta2, a multi-factor basket matches
 * the Cholesky basket on the implied matrix -- and the reference's test.
 * The outperformance and quanto variants (basket.rs:288-445) share the
 * loop: pinned by the lognormal ratio's closed forms (E[S1/S2] and the
 * Black formula on it) and the quanto-adjusted Black-Scholes call.
 * The Student-t copula (t_dof >= 2; map_copula_normals_to_uniforms and
 * copula_uniforms_to_normals, math/correlation.rs:545-582, 624-633) draws
 * t_dof more normals per path after the asset (and factor) draws, scales
 * the correlated normals by sqrt(max(chi2 / dof, 1e-16)) and maps them
 * through StudentsT::cdf and the Acklam inverse. The t CDF restates
 * statrs 0.18 (a dependency absent from the reference tree): beta_reg's
 * Lentz continued fraction and ln_gamma's Lanczos sum, whose coefficients
 * are the reference's own (src/math/gamma.rs:17-31). Pinned by scipy's
 * t CDF and regularized incomplete beta, the closed forms at 2 and 4
 * degrees of freedom, symmetry, scipy's gammaln; the copula by a one-asset
 * price equal to Black-Scholes (the t-copula marginal is exactly normal),
 * the Gaussian limit at large dof, and an independent numpy simulation.
 * Baskets (tests/test_oracle_basket.py): zero volatility (forward basket
 * intrinsic), one asset = Black-Scholes, a zero weight = the one-asset price
 * on the same first column of normals, best-of / worst-of of two = Margrabe,
 * the single stream = the same draws passed path-major, zero maturity, and
 * the reference's tests (Gentle moment matching within 2% on five assets).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* cholesky_lower (autocallable.rs:669-700), row-major n x n; returns 0 when
 * the matrix is not PSD (a diagonal sum below -1e-12). */
int or_cholesky_lower(const double *a, int32_t n, double *l) {
    const double tol = 1.0e-12;
    memset(l, 0, sizeof(double) * (size_t)(n * n));
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t j = 0; j <= i; ++j) {
            double sum = a[i * n + j];
            for (int32_t k = 0; k < j; ++k) sum -= l[i * n + k] * l[j * n + k];
            if (i == j) {
                if (sum < -tol) return 0;
                l[i * n + j] = sqrt(fmax(sum, tol));
            } else if (l[j * n + j] > tol) {
                l[i * n + j] = sum / l[j * n + j];
            }
        }
    }
    return 1;
}

/* observation_schedule: BTreeMap step -> max date (autocallable.rs:582-592).
 * Writes up to n_dates (step, time) pairs sorted by step; returns the count. */
int32_t or_observation_schedule(const double *dates, int32_t n_dates, double maturity,
                                int64_t n_steps, int64_t *steps_out, double *times_out) {
    int32_t m = 0;
    for (int32_t k = 0; k < n_dates; ++k) {
        const double raw = round((dates[k] / maturity) * (double)n_steps);
        int64_t step = raw > 0.0 ? (int64_t)raw : 0;   /* `as usize` saturates at 0 */
        if (step < 1) step = 1;
        if (step > n_steps) step = n_steps;
        int32_t pos = 0;
        while (pos < m && steps_out[pos] < step) ++pos;
        if (pos < m && steps_out[pos] == step) {
            times_out[pos] = fmax(times_out[pos], dates[k]);
            continue;
        }
        for (int32_t q = m; q > pos; --q) {
            steps_out[q] = steps_out[q - 1];
            times_out[q] = times_out[q - 1];
        }
        steps_out[pos] = step;
        times_out[pos] = dates[k];
        ++m;
    }
    return m;
}

int or_autocall_price(const or_autocall *a, const double *normals, double *price,
                      double *stderr_out) {
    const int32_t n = a->n_assets;
    if (a->num_paths <= 0 || a->num_steps <= 0 || n <= 0) return 1;
    const int64_t steps = a->num_steps;
    const double dt = a->maturity / (double)steps;
    const double sqrt_dt = sqrt(dt);
    double *l = (double *)malloc(sizeof(double) * (size_t)(n * n));
    if (!or_cholesky_lower(a->corr, n, l)) {
        free(l);
        return 2;
    }
    int64_t *obs_step = (int64_t *)malloc(sizeof(int64_t) * (size_t)(a->n_dates + 1));
    double *obs_time = (double *)malloc(sizeof(double) * (size_t)(a->n_dates + 1));
    const int32_t n_obs = or_observation_schedule(a->dates, a->n_dates, a->maturity, steps,
                                                  obs_step, obs_time);
    double drift[64], vol_dt[64], state[64], indep[64], corr[64];
    for (int32_t i = 0; i < n; ++i) {
        drift[i] = (a->rate - a->div_yield - 0.5 * a->vols[i] * a->vols[i]) * dt;
        vol_dt[i] = a->vols[i] * sqrt_dt;
    }
    uint64_t st[4];
    or_xoshiro_seed(a->seed, st);
    double *pvs = (double *)malloc(sizeof(double) * (size_t)a->num_paths);
    for (int64_t p = 0; p < a->num_paths; ++p) {
        for (int32_t i = 0; i < n; ++i) state[i] = a->spots[i];
        int32_t obs_idx = 0;
        int called = 0, ki_breached = 0;
        double worst_final = 1.0, pv = 0.0, pending = 0.0, prev_obs_time = 0.0;
        const double *zp = normals ? normals + p * steps * n : NULL;
        for (int64_t step = 1; step <= steps; ++step) {
            for (int32_t i = 0; i < n; ++i)
                indep[i] = zp ? zp[(step - 1) * n + i]
                              : or_acklam_inv_cdf(or_uniform_open01(or_xoshiro_next_f64(st)));
            for (int32_t i = 0; i < n; ++i) {
                double sum = 0.0;
                for (int32_t j = 0; j <= i; ++j) sum += l[i * n + j] * indep[j];
                corr[i] = sum;
            }
            for (int32_t i = 0; i < n; ++i) {
                state[i] *= exp(drift[i] + vol_dt[i] * corr[i]);
                state[i] = fmax(state[i], 1.0e-12);
            }
            double worst = INFINITY;
            for (int32_t i = 0; i < n; ++i) worst = fmin(worst, state[i] / a->spots[i]);
            worst_final = worst;
            if (worst <= a->ki_barrier) ki_breached = 1;
            while (obs_idx < n_obs && obs_step[obs_idx] == step) {
                const double t = obs_time[obs_idx];
                if (a->phoenix) {
                    const double accrual = a->notional * a->coupon_rate * fmax(t - prev_obs_time, 0.0);
                    if (worst >= a->coupon_barrier) {
                        const double due = a->memory ? pending + accrual : accrual;
                        pv += exp(-a->rate * t) * due;
                        pending = 0.0;
                    } else if (a->memory) {
                        pending += accrual;
                    }
                    prev_obs_time = t;
                }
                if (worst >= a->autocall_barrier) {
                    if (!a->phoenix)
                        pv = exp(-a->rate * t) * (a->notional * (1.0 + a->coupon_rate * t));
                    else
                        pv += exp(-a->rate * t) * a->notional;
                    called = 1;
                    break;
                }
                ++obs_idx;
            }
            if (called) break;
        }
        if (!called) {
            double redemption;
            if (ki_breached)
                redemption = a->notional * (worst_final / a->ki_strike);
            else if (!a->phoenix)
                redemption = a->notional * (1.0 + a->coupon_rate * a->maturity);
            else
                redemption = a->notional;
            pv += exp(-a->rate * a->maturity) * redemption;
        }
        pvs[p] = pv;
    }
    /* two-pass mean and variance (autocallable.rs:566-578) */
    const double nn = (double)a->num_paths;
    double s = 0.0;
    for (int64_t p = 0; p < a->num_paths; ++p) s += pvs[p];
    const double mean = s / nn;
    double v = 0.0;
    if (a->num_paths > 1) {
        for (int64_t p = 0; p < a->num_paths; ++p) v += (pvs[p] - mean) * (pvs[p] - mean);
        v /= (nn - 1.0);
    }
    *price = mean;
    *stderr_out = sqrt(v / nn);
    free(pvs);
    free(l);
    free(obs_step);
    free(obs_time);
    return 0;
}

/* bump_corr_matrix (autocallable.rs:702-714) */
void or_bump_corr_matrix(const double *c, int32_t n, double bump, double *out) {
    memcpy(out, c, sizeof(double) * (size_t)(n * n));
    for (int32_t i = 0; i < n; ++i) {
        out[i * n + i] = 1.0;
        for (int32_t j = i + 1; j < n; ++j) {
            double b = c[i * n + j] + bump;
            b = b < -0.999 ? -0.999 : (b > 0.999 ? 0.999 : b);
            out[i * n + j] = b;
            out[j * n + i] = b;
        }
    }
}

/* ===================================================================== */
/* Basket Monte Carlo: src/pricing/basket.rs                              */
/* ===================================================================== */

/* statrs 0.18 function::gamma::ln_gamma: Lanczos with Pugh's r and d_k
 * (the table of src/math/gamma.rs:17-31, which the reference ports from the
 * same crate). */
static const double OR_GAMMA_R = 10.900511;
static const double OR_GAMMA_DK[11] = {
    2.4857408913875355e-5, 1.0514237858172197, -3.4568709722201624, 4.512277094668948,
    -2.9828522532357666, 1.056397115771267, -1.9542877319164587e-1, 1.709705434044412e-2,
    -5.719261174043057e-4, 4.633994733599057e-6, -2.7199490848860772e-9};
static const double OR_LN_2_SQRT_E_OVER_PI = 0.6207822376352452223455184457816472122518527279025978;
static const double OR_LN_PI = 1.1447298858494001741434273513530587116472948129153;
static const double OR_E = 2.7182818284590452353602874713526624977572470937;
static const double OR_PI = 3.1415926535897932384626433832795028841971693993751;

double or_ln_gamma(double x) {
    double s = OR_GAMMA_DK[0];
    if (x < 0.5) {
        for (int i = 1; i < 11; ++i) s += OR_GAMMA_DK[i] / ((double)i - x);
        return OR_LN_PI - log(sin(OR_PI * x)) - log(s) - OR_LN_2_SQRT_E_OVER_PI -
               (0.5 - x) * log((0.5 - x + OR_GAMMA_R) / OR_E);
    }
    for (int i = 1; i < 11; ++i) s += OR_GAMMA_DK[i] / (x + (double)i - 1.0);
    return log(s) + OR_LN_2_SQRT_E_OVER_PI + (x - 0.5) * log((x - 0.5 + OR_GAMMA_R) / OR_E);
}

/* statrs 0.18 function::beta::beta_reg: I_x(a, b) by the modified Lentz
 * continued fraction (Numerical Recipes' betacf), the symmetry transform
 * when x >= (a + 1) / (a + b + 2), at most 140 terms, tolerance 2^-53,
 * prefactor 0 when x == 0 or x is within 4 ulps of 1. NaN outside the
 * domain (the crate panics there). */
double or_beta_reg(double a, double b, double x) {
    if (!(a > 0.0) || !(b > 0.0) || !(x >= 0.0 && x <= 1.0)) return NAN;
    const double bt = (x == 0.0 || x >= 1.0 - 4.0 * 0x1p-53)
                          ? 0.0
                          : exp(or_ln_gamma(a + b) - or_ln_gamma(a) - or_ln_gamma(b) + a * log(x) +
                                b * log(1.0 - x));
    const int symm = x >= (a + 1.0) / (a + b + 2.0);
    const double eps = 0x1p-53, fpmin = 0x1p-1022 / eps;
    if (symm) {
        const double t = a;
        x = 1.0 - x;
        a = b;
        b = t;
    }
    const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0, d = 1.0 - qab * x / qap;
    if (fabs(d) < fpmin) d = fpmin;
    d = 1.0 / d;
    double h = d;
    for (int mi = 1; mi < 141; ++mi) {
        const double m = (double)mi, m2 = m * 2.0;
        double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < fpmin) d = fpmin;
        c = 1.0 + aa / c;
        if (fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        h = h * d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < fpmin) d = fpmin;
        c = 1.0 + aa / c;
        if (fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) <= eps) break;
    }
    return symm ? 1.0 - bt * h / a : bt * h / a;
}

/* statrs 0.18 StudentsT::cdf at location 0, scale 1 (finite freedom) */
double or_student_t_cdf(double x, double nu) {
    const double h = nu / (nu + x * x);
    const double ib = 0.5 * or_beta_reg(nu / 2.0, 0.5, h);
    return x <= 0.0 ? ib : 1.0 - ib;
}

static double clamp_unit(double u) {   /* f64::clamp(1e-12, 1 - 1e-12) */
    const double lo = 1.0e-12, hi = 1.0 - 1.0e-12;
    return u < lo ? lo : (u > hi ? hi : u);
}

/* map_copula_normals_to_uniforms, Student-t (correlation.rs:559-570), then
 * copula_uniforms_to_normals (:624-633) on the n correlated normals. */
static void student_t_remap(int32_t dof, double chi2, double *z, int32_t n) {
    const double nu = (double)dof;
    const double scale = sqrt(fmax(chi2 / nu, 1.0e-16));
    for (int32_t i = 0; i < n; ++i) {
        const double u = clamp_unit(or_student_t_cdf(z[i] / scale, nu));
        z[i] = or_acklam_inv_cdf(clamp_unit(u));
    }
}

/* terminal_payoff (basket.rs:759-784) */
static double basket_payoff(const or_basket *b, const double *s0, const double *st) {
    double u;
    if (b->basket_type == OR_BASKET_AVERAGE) {
        u = -0.0;
        for (int32_t i = 0; i < b->n_assets; ++i) u += b->weights[i] * st[i];
    } else if (b->basket_type == OR_BASKET_BEST_OF) {
        u = -INFINITY;
        for (int32_t i = 0; i < b->n_assets; ++i) u = fmax(u, st[i] / s0[i]);
    } else {
        u = INFINITY;
        for (int32_t i = 0; i < b->n_assets; ++i) u = fmin(u, st[i] / s0[i]);
    }
    return b->is_call ? fmax(u - b->strike, 0.0) : fmax(b->strike - u, 0.0);
}

/* simulate_basket_once_matrix (basket.rs:599-683), Gaussian or Student-t
 * copula, on a PSD correlation matrix (the reference repairs others first).
 * normals NULL: the engine's single stream from `seed`; else
 * [path][asset (+ t_dof)]. Returns 0, 2 when the matrix is not PSD. */
int or_basket_price(const or_basket *b, const double *normals, double *price, double *stderr_out) {
    const int32_t n = b->n_assets;
    if (b->maturity <= 0.0) {
        *price = basket_payoff(b, b->spots, b->spots);
        *stderr_out = 0.0;
        return 0;
    }
    const double t = b->maturity, sqrt_t = sqrt(t);
    const double discount = exp(-b->rate * t);
    double *l = (double *)malloc(sizeof(double) * (size_t)(n * n));
    if (!or_cholesky_lower(b->corr, n, l)) {
        free(l);
        return 2;
    }
    double drift[64], vol_t[64], indep[64], corr[64], term[64];
    for (int32_t i = 0; i < n; ++i) {
        drift[i] = (b->rate - b->dividends[i] - 0.5 * b->vols[i] * b->vols[i]) * t;
        vol_t[i] = b->vols[i] * sqrt_t;
    }
    const int64_t stride = n + (b->t_dof > 0 ? b->t_dof : 0);
    uint64_t st[4];
    or_xoshiro_seed(b->seed, st);
    /* RunningStats (basket.rs:66-97): Welford */
    double mean = 0.0, m2 = 0.0;
    int64_t cnt = 0;
    for (int64_t p = 0; p < b->num_paths; ++p) {
        int64_t d = 0;
#define OR_DRAW() (normals ? normals[p * stride + (d++)] \
                           : (d++, or_acklam_inv_cdf(or_uniform_open01(or_xoshiro_next_f64(st)))))
        for (int32_t i = 0; i < n; ++i) indep[i] = OR_DRAW();
        for (int32_t i = 0; i < n; ++i) {   /* correlate_normals (correlation.rs:374-388) */
            double sum = 0.0;
            for (int32_t j = 0; j <= i; ++j) sum += l[i * n + j] * indep[j];
            corr[i] = sum;
        }
        if (b->t_dof > 0) {   /* sample_chi_square_integer_dof (correlation.rs:575-582) */
            double chi2 = 0.0;
            for (int32_t k = 0; k < b->t_dof; ++k) {
                const double z = OR_DRAW();
                chi2 += z * z;
            }
            student_t_remap(b->t_dof, chi2, corr, n);
        }
#undef OR_DRAW
        for (int32_t i = 0; i < n; ++i) term[i] = b->spots[i] * exp(drift[i] + vol_t[i] * corr[i]);
        const double x = basket_payoff(b, b->spots, term);
        ++cnt;
        const double delta = x - mean;
        mean += delta / (double)cnt;
        m2 += delta * (x - mean);
    }
    const double var = cnt > 1 ? m2 / ((double)cnt - 1.0) : 0.0;
    *price = discount * mean;
    *stderr_out = discount * sqrt(var / (double)b->num_paths);
    free(l);
    return 0;
}

/* The common loop of the two variants: terminal spots from `drift`, the
 * payoff of each path by `kind` (1 outperformance, 2 quanto), Welford. */
static int basket_variant(const or_basket *b, int kind, int32_t leader, double fx_rate,
                          const double *drift, double discount, const double *normals,
                          double *price, double *stderr_out) {
    const int32_t n = b->n_assets;
    const double t = b->maturity, sqrt_t = sqrt(t);
    double *l = (double *)malloc(sizeof(double) * (size_t)(n * n));
    if (!or_cholesky_lower(b->corr, n, l)) {
        free(l);
        return 2;
    }
    if (t <= 0.0) {   /* after the repair and factorisation, as the reference */
        free(l);
        if (kind == 1) {
            double lag = -0.0;
            for (int32_t i = 0; i < n; ++i) lag += b->weights[i] * b->spots[i];
            const double ratio = b->spots[leader] / fmax(lag, 1.0e-12);
            *price = b->is_call ? fmax(ratio - b->strike, 0.0) : fmax(b->strike - ratio, 0.0);
        } else {
            *price = fx_rate * basket_payoff(b, b->spots, b->spots);
        }
        *stderr_out = 0.0;
        return 0;
    }
    double vol_t[64], indep[64], term[64];
    for (int32_t i = 0; i < n; ++i) vol_t[i] = b->vols[i] * sqrt_t;
    uint64_t st[4];
    or_xoshiro_seed(b->seed, st);
    double mean = 0.0, m2 = 0.0;
    int64_t cnt = 0;
    for (int64_t p = 0; p < b->num_paths; ++p) {
        for (int32_t i = 0; i < n; ++i)
            indep[i] = normals ? normals[p * n + i]
                               : or_acklam_inv_cdf(or_uniform_open01(or_xoshiro_next_f64(st)));
        for (int32_t i = 0; i < n; ++i) {
            double sum = 0.0;
            for (int32_t j = 0; j <= i; ++j) sum += l[i * n + j] * indep[j];
            term[i] = b->spots[i] * exp(drift[i] + vol_t[i] * sum);
        }
        double x;
        if (kind == 1) {   /* outperformance_payoff (basket.rs:786-800) */
            double lag = -0.0;
            for (int32_t i = 0; i < n; ++i) lag += b->weights[i] * term[i];
            const double ratio = term[leader] / fmax(lag, 1.0e-12);
            x = b->is_call ? fmax(ratio - b->strike, 0.0) : fmax(b->strike - ratio, 0.0);
        } else {
            x = fx_rate * basket_payoff(b, b->spots, term);
        }
        ++cnt;
        const double delta = x - mean;
        mean += delta / (double)cnt;
        m2 += delta * (x - mean);
    }
    const double var = cnt > 1 ? m2 / ((double)cnt - 1.0) : 0.0;
    *price = discount * mean;
    *stderr_out = discount * sqrt(var / (double)b->num_paths);
    free(l);
    return 0;
}

int or_outperformance_price(const or_basket *b, int32_t leader, const double *normals,
                            double *price, double *stderr_out) {
    double drift[64];
    const double t = b->maturity;
    for (int32_t i = 0; i < b->n_assets; ++i)
        drift[i] = (b->rate - b->dividends[i] - 0.5 * b->vols[i] * b->vols[i]) * t;
    return basket_variant(b, 1, leader, 1.0, drift, exp(-b->rate * t), normals, price, stderr_out);
}

int or_quanto_price(const or_basket *b, double fx_rate, double fx_vol, const double *asset_fx_corr,
                    double domestic_rate, double foreign_rate, const double *normals,
                    double *price, double *stderr_out) {
    double drift[64];
    const double t = b->maturity;
    for (int32_t i = 0; i < b->n_assets; ++i) {
        const double v = b->vols[i];
        const double quanto_adj = asset_fx_corr[i] * v * fx_vol;
        drift[i] = (foreign_rate - b->dividends[i] - quanto_adj - 0.5 * v * v) * t;
    }
    return basket_variant(b, 2, 0, fx_rate, drift, exp(-domestic_rate * t), normals, price,
                          stderr_out);
}

/* simulate_basket_once_factor (basket.rs:686-757) with
 * FactorCorrelationModel::sample_correlated_normals (correlation.rs:173-212):
 * one_factor: loadings[i] = beta_i (clamped), out = fma(beta, m, idio eps);
 * else loadings [n][k], out = sum_j l_ij f_j + idio eps. Draws per path: the
 * k factors, then one per asset; normals [path][k + n] or NULL. */
int or_basket_factor_price(const or_basket *b, int one_factor, int32_t k, const double *loadings,
                           const double *normals, double *price, double *stderr_out) {
    const int32_t n = b->n_assets;
    if (b->maturity <= 0.0) {
        *price = basket_payoff(b, b->spots, b->spots);
        *stderr_out = 0.0;
        return 0;
    }
    if (one_factor) k = 1;
    const double t = b->maturity, sqrt_t = sqrt(t);
    const double discount = exp(-b->rate * t);
    double drift[64], vol_t[64], term[64], f[64], idio[64];
    for (int32_t i = 0; i < n; ++i) {
        drift[i] = (b->rate - b->dividends[i] - 0.5 * b->vols[i] * b->vols[i]) * t;
        vol_t[i] = b->vols[i] * sqrt_t;
        if (one_factor) {
            const double beta = fmin(1.0, fmax(-1.0, loadings[i]));
            idio[i] = sqrt(fmax(1.0 - beta * beta, 0.0));
        } else {
            double norm2 = -0.0;
            for (int32_t j = 0; j < k; ++j) norm2 += loadings[i * k + j] * loadings[i * k + j];
            idio[i] = sqrt(fmax(1.0 - norm2, 0.0));
        }
    }
    const int64_t stride = k + n + (b->t_dof > 0 ? b->t_dof : 0);
    double zc[64];
    uint64_t st[4];
    or_xoshiro_seed(b->seed, st);
    double mean = 0.0, m2 = 0.0;
    int64_t cnt = 0, d;
    for (int64_t p = 0; p < b->num_paths; ++p) {
        d = 0;
#define OR_DRAW() (normals ? normals[p * stride + (d++)] \
                           : (d++, or_acklam_inv_cdf(or_uniform_open01(or_xoshiro_next_f64(st)))))
        for (int32_t j = 0; j < k; ++j) f[j] = OR_DRAW();
        for (int32_t i = 0; i < n; ++i) {
            double z;
            if (one_factor) {
                const double beta = fmin(1.0, fmax(-1.0, loadings[i]));
                const double eps = OR_DRAW();
                z = fma(beta, f[0], idio[i] * eps);
            } else {
                double sys = -0.0;
                for (int32_t j = 0; j < k; ++j) sys += loadings[i * k + j] * f[j];
                const double eps = OR_DRAW();
                z = sys + idio[i] * eps;
            }
            zc[i] = z;
        }
        if (b->t_dof > 0) {
            double chi2 = 0.0;
            for (int32_t j = 0; j < b->t_dof; ++j) {
                const double z = OR_DRAW();
                chi2 += z * z;
            }
            student_t_remap(b->t_dof, chi2, zc, n);
        }
#undef OR_DRAW
        for (int32_t i = 0; i < n; ++i) term[i] = b->spots[i] * exp(drift[i] + vol_t[i] * zc[i]);
        const double x = basket_payoff(b, b->spots, term);
        ++cnt;
        const double delta = x - mean;
        mean += delta / (double)cnt;
        m2 += delta * (x - mean);
    }
    const double var = cnt > 1 ? m2 / ((double)cnt - 1.0) : 0.0;
    *price = discount * mean;
    *stderr_out = discount * sqrt(var / (double)b->num_paths);
    return 0;
}
