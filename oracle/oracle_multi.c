/*
 * CPU oracle, multi-asset part -- TEST INFRASTRUCTURE ONLY (the checker of
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

/* simulate_basket_once_matrix, Gaussian copula (basket.rs:599-683), on a PSD
 * correlation matrix (the reference repairs others first). normals NULL: the
 * engine's single stream from `seed`; else [path][asset]. Returns 0, 2 when
 * the matrix is not PSD. */
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
    double drift[64], vol_t[64], indep[64], term[64];
    for (int32_t i = 0; i < n; ++i) {
        drift[i] = (b->rate - b->dividends[i] - 0.5 * b->vols[i] * b->vols[i]) * t;
        vol_t[i] = b->vols[i] * sqrt_t;
    }
    uint64_t st[4];
    or_xoshiro_seed(b->seed, st);
    /* RunningStats (basket.rs:66-97): Welford */
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
    uint64_t st[4];
    or_xoshiro_seed(b->seed, st);
    double mean = 0.0, m2 = 0.0;
    int64_t cnt = 0, d;
    for (int64_t p = 0; p < b->num_paths; ++p) {
        d = 0;
#define OR_DRAW() (normals ? normals[p * (k + n) + (d++)] \
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
            term[i] = b->spots[i] * exp(drift[i] + vol_t[i] * z);
        }
#undef OR_DRAW
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
