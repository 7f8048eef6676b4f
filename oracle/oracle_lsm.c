/*
 * CPU oracle, Longstaff-Schwartz part -- TEST INFRASTRUCTURE ONLY (the tests' reference
 * for the HIP least-squares Monte Carlo path and its timed CPU baseline).
 *
 * Restates src/engines/lsm/longstaff_schwartz.rs of rosssaunders/openferric:
 *   - path simulation from ONE xoshiro256++ stream drawn path-major
 *     (:195-283 for BermudanOption dynamics, :506-556 for VanillaOption),
 *   - the backward induction with the quadratic regression on in-the-money
 *     paths (:291-417, :558-660),
 *   - the 3x3 normal equations solved by nalgebra's LU with partial pivoting
 *     (nalgebra 0.34, src/linalg/lu.rs and src/linalg/solve.rs, restated in
 *     lu3_solve below), zeros when the matrix is singular,
 *   - the plain Monte Carlo barrier pricer of the same engine (:674-735).
 * Compiled with -ffp-contract=off (Rust never contracts a*b+c).
 *
 * Pinned by (tests/test_oracle_lsm.py): the Longstaff-Schwartz (2001) Table 1
 * values the reference's tests/lsm_reference.rs asserts, the QuantLib FD
 * Bermudan prices of tests/bermudan_quantlib.rs, an independent numpy
 * least-squares restatement (lstsq, not normal equations) on small cases,
 * European exercise against Black-Scholes, and a CRR binomial tree.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* intrinsic, longstaff_schwartz.rs:114-119 */
static inline double lsm_intrinsic(int32_t type, double spot, double strike) {
    return type == OR_CALL ? fmax(spot - strike, 0.0) : fmax(strike - spot, 0.0);
}

/* nalgebra LU::new (partial pivoting by icamax, multipliers scaled by the
 * reciprocal of the pivot, rank-1 update by axpy) followed by LU::solve
 * (row permutation, unit-lower forward substitution, upper back substitution
 * by columns). Returns 0 when a pivot of U is exactly zero (solve -> None). */
int or_lu3_solve(const double a_in[9], const double b_in[3], double x[3]) {
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = a_in[3 * r + c];
    int perm_from[3], perm_to[3], n_perm = 0;
    for (int i = 0; i < 3; ++i) {
        int piv = i;
        double best = fabs(a[i][i]);
        for (int r = i + 1; r < 3; ++r) {
            const double v = fabs(a[r][i]);
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        const double diag = a[piv][i];
        if (diag == 0.0) continue;
        if (piv != i) {
            perm_from[n_perm] = i;
            perm_to[n_perm] = piv;
            ++n_perm;
            for (int c = 0; c < 3; ++c) {
                const double t = a[i][c];
                a[i][c] = a[piv][c];
                a[piv][c] = t;
            }
        }
        const double inv_diag = 1.0 / diag;
        for (int r = i + 1; r < 3; ++r) a[r][i] = a[r][i] * inv_diag;
        for (int c = i + 1; c < 3; ++c)
            for (int r = i + 1; r < 3; ++r) a[r][c] = (-a[i][c]) * a[r][i] + a[r][c];
    }
    double b[3] = {b_in[0], b_in[1], b_in[2]};
    for (int k = 0; k < n_perm; ++k) {
        const double t = b[perm_from[k]];
        b[perm_from[k]] = b[perm_to[k]];
        b[perm_to[k]] = t;
    }
    for (int i = 0; i < 2; ++i) {
        const double coeff = b[i];
        for (int r = i + 1; r < 3; ++r) b[r] = (-coeff) * a[r][i] + b[r];
    }
    for (int i = 2; i >= 0; --i) {
        const double diag = a[i][i];
        if (diag == 0.0) return 0;
        const double coeff = b[i] / diag;
        b[i] = coeff;
        for (int r = 0; r < i; ++r) b[r] = (-coeff) * a[r][i] + b[r];
    }
    x[0] = b[0];
    x[1] = b[1];
    x[2] = b[2];
    return 1;
}

/* mean_and_stderr, longstaff_schwartz.rs:128-143 */
static void lsm_mean_stderr(const double *v, int64_t n, double *mean, double *se) {
    double sum = 0.0, sum_sq = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        sum += v[i];
        sum_sq += v[i] * v[i];
    }
    const double nn = (double)n;
    *mean = sum / nn;
    const double var = n > 1 ? (sum_sq - sum * sum / nn) / (nn - 1.0) : 0.0;
    *se = sqrt(var / nn);
}

/* One draw of the engine's stream: beasley_springer_moro_inv_cdf (Acklam)
 * of uniform_open01(next_f64), longstaff_schwartz.rs:218. */
static inline double lsm_draw(uint64_t s[4], const double *ext, int64_t *cursor) {
    if (ext) return ext[(*cursor)++];
    return or_acklam_inv_cdf(or_uniform_open01(or_xoshiro_next_f64(s)));
}

/* Paths [path][step] (steps+1 per path). External normals, when given, are
 * consumed in the stream's own order: per path, per step, z1 (then z2 for
 * Heston). Returns 0, or 1 for invalid dynamics parameters. */
static int lsm_simulate(const or_lsm_option *o, int32_t dynamics, double vol,
                        const double *ext, double *paths) {
    const int64_t n = o->num_steps, np = o->num_paths, stride = n + 1;
    const double dt = o->expiry / (double)n;
    const double sqrt_dt = sqrt(dt);
    const double drift_rn = o->rate - o->div_yield;
    uint64_t st[4];
    or_xoshiro_seed(o->seed, st);
    int64_t cursor = 0;
    if (dynamics == OR_LSM_GBM) {
        /* :214-222 (and :544-553 for VanillaOption) */
        const double drift = (drift_rn - 0.5 * vol * vol) * dt;
        const double step_vol = vol * sqrt_dt;
        for (int64_t p = 0; p < np; ++p) {
            double *row = paths + p * stride;
            row[0] = o->spot;
            for (int64_t t = 1; t <= n; ++t) {
                const double z = lsm_draw(st, ext, &cursor);
                row[t] = row[t - 1] * exp(fma(step_vol, z, drift));
            }
        }
    } else if (dynamics == OR_LSM_LOCAL_VOL) {
        /* :224-243 with a flat surface: vol_for(S, t) = vol */
        for (int64_t p = 0; p < np; ++p) {
            double *row = paths + p * stride;
            double s = o->spot;
            row[0] = s;
            for (int64_t t = 1; t <= n; ++t) {
                const double sigma = vol;
                const double z = lsm_draw(st, ext, &cursor);
                const double drift = (drift_rn - 0.5 * sigma * sigma) * dt;
                s *= exp(drift + sigma * sqrt_dt * z);
                row[t] = fmax(s, 1.0e-12);
                s = row[t];
            }
        }
    } else {
        /* :245-276, Heston::validate (models/stochastic.rs:47-54) */
        if (!(o->kappa > 0.0 && o->theta >= 0.0 && o->xi >= 0.0 && o->v0 >= 0.0 &&
              o->rho > -1.0 && o->rho < 1.0))
            return 1;
        for (int64_t p = 0; p < np; ++p) {
            double *row = paths + p * stride;
            double s = o->spot, v = o->v0;
            row[0] = s;
            for (int64_t t = 1; t <= n; ++t) {
                const double z1 = lsm_draw(st, ext, &cursor);
                const double z2 = lsm_draw(st, ext, &cursor);
                double sn, vn;
                or_heston_step_euler(drift_rn, o->kappa, o->theta, o->xi, o->rho, s, v, dt, z1,
                                     z2, &sn, &vn);
                s = fmax(sn, 1.0e-12);
                v = fmax(vn, 0.0);
                row[t] = s;
            }
        }
    }
    return 0;
}

/* Regression of the discounted values on {1, S, S^2} over the in-the-money
 * paths at step ti, regression_beta (:145-179) and the inlined copy
 * (:589-616). Returns the ITM count; beta is valid when it is >= 3. */
static int64_t lsm_regress(const double *paths, int64_t stride, int64_t np, int64_t ti,
                           int32_t type, double strike, const double *values, double beta[3]) {
    double s1 = 0.0, s_s = 0.0, s_s2 = 0.0, s_s3 = 0.0, s_s4 = 0.0;
    double s_y = 0.0, s_sy = 0.0, s_s2y = 0.0;
    int64_t itm = 0;
    for (int64_t p = 0; p < np; ++p) {
        const double s = paths[p * stride + ti];
        if (!(lsm_intrinsic(type, s, strike) > 0.0)) continue;
        ++itm;
        const double s2 = s * s;
        const double y = values[p];
        s1 += 1.0;
        s_s += s;
        s_s2 += s2;
        s_s3 += s2 * s;
        s_s4 += s2 * s2;
        s_y += y;
        s_sy += s * y;
        s_s2y += s2 * y;
    }
    if (itm < 3) return itm;
    const double xtx[9] = {s1, s_s, s_s2, s_s, s_s2, s_s3, s_s2, s_s3, s_s4};
    const double xty[3] = {s_y, s_sy, s_s2y};
    if (!or_lu3_solve(xtx, xty, beta)) beta[0] = beta[1] = beta[2] = 0.0;
    return itm;
}

/* Exercise decision on the ITM paths (:377-387, :618-626). Returns the
 * number exercised; *extreme receives the largest (put) / smallest (call)
 * exercised spot (boundary_from_exercised, :182-191). */
static int64_t lsm_exercise(const double *paths, int64_t stride, int64_t np, int64_t ti,
                            int32_t type, double strike, const double beta[3], double *values,
                            double *extreme) {
    int64_t ex_n = 0;
    double ext = type == OR_PUT ? -INFINITY : INFINITY;
    for (int64_t p = 0; p < np; ++p) {
        const double s = paths[p * stride + ti];
        const double exercise = lsm_intrinsic(type, s, strike);
        if (!(exercise > 0.0)) continue;
        const double continuation = beta[0] + beta[1] * s + beta[2] * s * s;
        if (exercise > continuation) {
            values[p] = exercise;
            ++ex_n;
            ext = type == OR_PUT ? fmax(ext, s) : fmin(ext, s);
        }
    }
    *extreme = ext;
    return ex_n;
}

int or_lsm_price(const or_lsm_option *o, const double *normals, double *price,
                 double *stderr_out, or_lsm_boundary *boundary, int64_t *n_boundary) {
    if (n_boundary) *n_boundary = 0;
    const int64_t n = o->num_steps, np = o->num_paths;
    if (o->instrument == OR_LSM_BARRIER) {
        /* :674-735: plain Monte Carlo, one stream */
        if (n <= 0 || np <= 0 || !(o->vol > 0.0)) return 1;
        const double dt = o->expiry / (double)n;
        const double drift = (o->rate - o->div_yield - 0.5 * o->vol * o->vol) * dt;
        const double step_vol = o->vol * sqrt(dt);
        const double discount = exp(-o->rate * o->expiry);
        uint64_t st[4];
        or_xoshiro_seed(o->seed, st);
        int64_t cursor = 0;
        double *pv = (double *)malloc(sizeof(double) * (size_t)np);
        if (!pv) return 2;
        for (int64_t p = 0; p < np; ++p) {
            double s = o->spot;
            int hit = o->barrier_dir == OR_UP ? s >= o->barrier_level : s <= o->barrier_level;
            for (int64_t t = 0; t < n; ++t) {
                const double z = lsm_draw(st, normals, &cursor);
                s = s * exp(fma(step_vol, z, drift));
                if (o->barrier_dir == OR_UP ? s >= o->barrier_level : s <= o->barrier_level)
                    hit = 1;
            }
            const int active = o->barrier_style == OR_IN ? hit : !hit;
            const double payoff = active ? lsm_intrinsic(o->option_type, s, o->strike) : o->rebate;
            pv[p] = discount * payoff;
        }
        lsm_mean_stderr(pv, np, price, stderr_out);
        free(pv);
        return 0;
    }

    if (n < 2 || np < 3) return 1;
    const int bermudan = o->instrument == OR_LSM_BERMUDAN;
    /* schedule: (step index, strike) per exercisable step in [1, n-1] */
    double *strike_at = (double *)calloc((size_t)(n + 1), sizeof(double));
    double *time_at = (double *)calloc((size_t)(n + 1), sizeof(double));
    char *can = (char *)calloc((size_t)(n + 1), 1);
    if (!strike_at || !time_at || !can) return 2;
    double terminal_strike = o->strike;
    int32_t dynamics = OR_LSM_GBM;
    if (bermudan) {
        /* effective_schedule (instruments/bermudan.rs:85-103) and the step
           mapping (:314-320): later entries overwrite earlier ones */
        int64_t m = o->n_dates;
        terminal_strike = o->strikes[m - 1];
        for (int64_t k = 0; k < m; ++k) {
            int64_t idx = (int64_t)round(o->dates[k] / o->expiry * (double)n);
            if (idx < 1) idx = 1;
            if (idx > n) idx = n;
            can[idx] = 1;
            strike_at[idx] = o->strikes[k];
            time_at[idx] = o->dates[k];
        }
        dynamics = o->dynamics;
    } else {
        /* can_exercise (:567-584) */
        if (o->exercise == OR_AMERICAN) {
            for (int64_t t = 1; t < n; ++t) can[t] = 1;
        } else if (o->exercise == OR_BERMUDAN) {
            for (int64_t k = 0; k < o->n_dates; ++k) {
                int64_t idx = (int64_t)round(o->dates[k] / o->expiry * (double)n);
                if (idx > n) idx = n;
                can[idx] = 1;
            }
        }
        for (int64_t t = 0; t <= n; ++t) strike_at[t] = o->strike;
    }
    const double vol = o->vol;
    if (dynamics != OR_LSM_HESTON && !(vol > 0.0)) {
        free(strike_at);
        free(time_at);
        free(can);
        return 1;
    }
    const int64_t stride = n + 1;
    double *paths = (double *)malloc(sizeof(double) * (size_t)(np * stride));
    double *values = (double *)malloc(sizeof(double) * (size_t)np);
    if (!paths || !values) return 2;
    if (lsm_simulate(o, dynamics, vol, normals, paths)) {
        free(paths);
        free(values);
        free(strike_at);
        free(time_at);
        free(can);
        return 1;
    }
    const double dt = o->expiry / (double)n;
    const double disc = exp(-o->rate * dt);
    int64_t term_itm = 0;
    for (int64_t p = 0; p < np; ++p) {
        values[p] = lsm_intrinsic(o->option_type, paths[p * stride + n], terminal_strike);
        if (values[p] > 0.0) ++term_itm;
    }
    /* boundary points collected in reverse time order (:327-334) */
    int64_t nb = 0;
    if (boundary) {
        boundary[nb].time = o->expiry;
        boundary[nb].strike = terminal_strike;
        boundary[nb].boundary_spot = terminal_strike;
        boundary[nb].has_boundary = 1;
        boundary[nb].itm_paths = term_itm;
        boundary[nb].exercised_paths = term_itm;
    }
    ++nb;
    for (int64_t ti = n - 1; ti >= 1; --ti) {
        for (int64_t p = 0; p < np; ++p) values[p] *= disc;
        if (!can[ti]) continue;
        double beta[3];
        const int64_t itm = lsm_regress(paths, stride, np, ti, o->option_type, strike_at[ti],
                                        values, beta);
        int64_t ex_n = 0;
        double extreme = 0.0;
        if (itm >= 3)
            ex_n = lsm_exercise(paths, stride, np, ti, o->option_type, strike_at[ti], beta,
                                values, &extreme);
        if (boundary) {
            or_lsm_boundary *b = &boundary[nb];
            b->time = time_at[ti];
            b->strike = strike_at[ti];
            b->has_boundary = ex_n > 0;
            b->boundary_spot = ex_n > 0 ? extreme : 0.0;
            b->itm_paths = itm;
            b->exercised_paths = ex_n;
        }
        ++nb;
    }
    for (int64_t p = 0; p < np; ++p) values[p] *= disc;
    lsm_mean_stderr(values, np, price, stderr_out);
    if (boundary) {
        /* chronological order (:412) */
        for (int64_t i = 0, j = nb - 1; i < j; ++i, --j) {
            or_lsm_boundary t = boundary[i];
            boundary[i] = boundary[j];
            boundary[j] = t;
        }
    }
    if (n_boundary) *n_boundary = bermudan ? nb : 0;
    free(paths);
    free(values);
    free(strike_at);
    free(time_at);
    free(can);
    return 0;
}
