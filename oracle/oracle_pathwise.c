/*
 * Oracle: pathwise sensitivities of MonteCarloEngine::run's estimator.
 *
 * TEST INFRASTRUCTURE ONLY (the tests' reference for the GPU adjoint and forward-mode
 * Greeks kernels); the product never links or loads it.
 *
 * The reference differentiates two cases itself: the European vanilla under
 * GBM by a reverse-mode tape (mc_european_pathwise_aad,
 * src/engines/monte_carlo/mc_aad.rs:16-219) and the Heston delta by a forward
 * dual number on S0 (heston_price_delta_aad, mc_aad.rs:239-342). This file
 * carries the same pathwise estimator to every payoff of MonteCarloEngine::run
 * (vanilla, barrier, arithmetic / geometric Asian with fixed or floating
 * strike: mc_engine.rs:59-121, 441-584) and every model input, by forward
 * tangents: each path propagates, next to (S, v), the derivative of (S, v)
 * with respect to eight inputs
 *
 *   slot 0  S0
 *   slot 1  sigma (GBM) | v0 (Heston)
 *   slot 2  r   (through mu = r - q, and the discount exp(-r T))
 *   slot 3  T   (through dt = T / n, sqrt(dt) and the discount; observation
 *                indices held fixed)
 *   slot 4  kappa   slot 5 theta   slot 6 xi   slot 7 rho   (Heston)
 *
 * through the reference's own steps: GbmPathGenerator::generate_into
 * (src/mc/simulation.rs:58-71) and Heston::step_euler with the generator's
 * clamps (src/models/stochastic.rs:56-71, simulation.rs:112-113). Path values
 * use the reference's operation order, so every pv equals the payoff
 * or_mc_run computes for the same sample. Conventions at the non-smooth
 * points, as the reference's own tape (positive_part has derivative
 * 1{x > 0}, src/math/aad.rs:616-626): d max(x, 0) = 1{x > 0} dx, the spot
 * floor d max(s, 1e-12) = 1{s >= 1e-12} ds, d sqrt(v) = dv / (2 sqrt(v)) for
 * v > 0 and 0 at v = 0, and zero derivative for the barrier indicator.
 *
 * Pinned by (tests/test_oracle_pathwise.py): the reference's own estimators
 * on the same draws (or_mc_european_pathwise_aad for GBM Europeans,
 * or_heston_price_delta_aad for the Heston delta), and per-path central
 * differences of or_mc_run's payoffs with common random numbers for every
 * slot, payoff and model, which a dropped term fails (shown by planting
 * each mistake through `mutate`).
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PW_NG 8

/* mc_engine.rs:103-104 index map (as obs_index in oracle.c) */
static int64_t pw_obs_index(double t, double maturity, int64_t len) {
    const double last_idx = (double)(len - 1);
    const double x = round((t / maturity) * last_idx);
    int64_t idx;
    if (!(x > 0.0)) idx = 0;
    else if (x >= 9.2e18) idx = INT64_MAX;
    else idx = (int64_t)x;
    return idx < len - 1 ? idx : len - 1;
}

/* One leg: the path S[0..n] and its tangents dS[j][8]; returns the
 * undiscounted payoff and writes dx[8]. `mutate` (test use only) drops one
 * term of the tangent recursion: 1 GBM d drift/d sigma, 2 the Heston
 * variance tangent's kappa term, 3 the d sqrt(v) chain, 4 the rho slot's
 * z2 term, 5 the spot tangent's growth factor, 6 the T slot's sqrt(dt) term. */
static double pw_leg(const or_option *o, int64_t n, const double *z1, const double *z2,
                     int32_t mutate, double *S, double *dS, double dx[PW_NG]) {
    const double mu = o->rate - o->div_yield; /* mc_engine.rs:916 */
    const double T = o->expiry;
    const double dt = T / (double)n;
    const double sqrt_dt = sqrt(dt);
    const double ddt_dT = 1.0 / (double)n;
    const double dsq_dT = mutate == 6 ? 0.0 : 0.5 / ((double)n * sqrt_dt);
    double ds[PW_NG], dv[PW_NG];
    memset(ds, 0, sizeof ds);
    memset(dv, 0, sizeof dv);
    ds[0] = 1.0;
    double s = o->spot;
    S[0] = s;
    memcpy(dS, ds, sizeof ds);
    if (o->model == OR_GBM) {
        const double sigma = o->vol;
        const double drift = (mu - 0.5 * sigma * sigma) * dt;
        const double diffusion = sigma * sqrt_dt;
        /* d drift, d diffusion per slot (sigma, r, T) */
        double ddrift[PW_NG] = {0}, ddiff[PW_NG] = {0};
        ddrift[1] = mutate == 1 ? 0.0 : -sigma * dt;
        ddiff[1] = sqrt_dt;
        ddrift[2] = dt;
        ddrift[3] = (mu - 0.5 * sigma * sigma) * ddt_dT;
        ddiff[3] = sigma * dsq_dT;
        for (int64_t j = 0; j < n; ++j) {
            const double g = exp(fma(diffusion, z1[j], drift));
            s *= g;
            for (int k = 0; k < PW_NG; ++k) {
                const double dxk = ddrift[k] + ddiff[k] * z1[j];
                ds[k] = (mutate == 5 ? 1.0 : g) * ds[k] + s * dxk;
            }
            S[j + 1] = s;
            memcpy(dS + (j + 1) * PW_NG, ds, sizeof ds);
        }
    } else {
        const double kappa = o->kappa, theta = o->theta, xi = o->xi, rho = o->rho;
        double v = o->v0;
        dv[1] = 1.0;
        for (int64_t j = 0; j < n; ++j) {
            /* Heston::step_euler (stochastic.rs:56-71), operation order as
               or_heston_step_euler */
            const double v_pos = fmax(v, 0.0);
            const double rho_c = sqrt(1.0 - rho * rho);
            const double zs = rho * z1[j] + rho_c * z2[j];
            const double sv = sqrt(v_pos);
            const double vn = v + kappa * (theta - v_pos) * dt + xi * sv * sqrt_dt * z1[j];
            const double e = (mu - 0.5 * v_pos) * dt + sv * sqrt_dt * zs;
            const double growth = exp(e);
            const double sn = s * growth;
            for (int k = 0; k < PW_NG; ++k) {
                const double dvpos = v > 0.0 ? dv[k] : 0.0;
                const double dsv = (sv > 0.0 && mutate != 3) ? dvpos / (2.0 * sv) : 0.0;
                const double d_dt = k == 3 ? ddt_dT : 0.0;
                const double d_sq = k == 3 ? dsq_dT : 0.0;
                const double d_mu = k == 2 ? 1.0 : 0.0;
                const double d_kappa = k == 4 ? 1.0 : 0.0;
                const double d_theta = k == 5 ? 1.0 : 0.0;
                const double d_xi = k == 6 ? 1.0 : 0.0;
                const double d_zs =
                    k == 7 ? z1[j] - (mutate == 4 ? 0.0 : rho / rho_c * z2[j]) : 0.0;
                const double dvn = dv[k] +
                                   (mutate == 2 ? 0.0 : d_kappa * (theta - v_pos) * dt) +
                                   kappa * (d_theta - dvpos) * dt + kappa * (theta - v_pos) * d_dt +
                                   d_xi * sv * sqrt_dt * z1[j] + xi * dsv * sqrt_dt * z1[j] +
                                   xi * sv * d_sq * z1[j];
                const double de = (d_mu - 0.5 * dvpos) * dt + (mu - 0.5 * v_pos) * d_dt +
                                  dsv * sqrt_dt * zs + sv * d_sq * zs + sv * sqrt_dt * d_zs;
                const double dsn = (mutate == 5 ? 1.0 : growth) * ds[k] + sn * de;
                ds[k] = sn >= 1e-12 ? dsn : 0.0;
                dv[k] = vn > 0.0 ? dvn : 0.0;
            }
            s = fmax(sn, 1e-12);
            v = fmax(fmax(vn, 0.0), 0.0);
            S[j + 1] = s;
            memcpy(dS + (j + 1) * PW_NG, ds, sizeof ds);
        }
    }

    const int64_t len = n + 1;
    const double x = or_payoff(o, S, len);
    const double *dT = dS + n * PW_NG;
    memset(dx, 0, sizeof(double) * PW_NG);
    const int put = o->option_type == OR_PUT;
    if (o->payoff == OR_VANILLA || o->payoff == OR_BARRIER) {
        int active = 1;
        if (o->payoff == OR_BARRIER) {
            int hit = 0;
            for (int64_t i = 0; i < len && !hit; ++i)
                hit = o->barrier_dir == OR_UP ? S[i] >= o->barrier_level : S[i] <= o->barrier_level;
            active = o->barrier_style == OR_IN ? hit : !hit;
        }
        const double st = S[n];
        const double ind = put ? (o->strike - st > 0.0 ? -1.0 : 0.0) : (st - o->strike > 0.0 ? 1.0 : 0.0);
        if (active)
            for (int k = 0; k < PW_NG; ++k) dx[k] = ind * dT[k];
    } else {
        /* average_for_observations (mc_engine.rs:86-121) and its tangent */
        double avg, davg[PW_NG] = {0};
        const double m = (double)o->n_obs;
        if (o->n_obs == 0 || o->expiry <= 0.0) {
            avg = S[n];
            memcpy(davg, dT, sizeof davg);
        } else if (o->averaging == OR_ARITH) {
            double sum = 0.0;
            for (int32_t i = 0; i < o->n_obs; ++i) {
                const int64_t idx = pw_obs_index(o->obs_times[i], o->expiry, len);
                sum += S[idx];
                for (int k = 0; k < PW_NG; ++k) davg[k] += dS[idx * PW_NG + k];
            }
            avg = sum / m;
            for (int k = 0; k < PW_NG; ++k) davg[k] /= m;
        } else {
            double sum = 0.0;
            for (int32_t i = 0; i < o->n_obs; ++i) {
                const int64_t idx = pw_obs_index(o->obs_times[i], o->expiry, len);
                sum += log(fmax(S[idx], 1e-12));
                if (S[idx] > 1e-12)
                    for (int k = 0; k < PW_NG; ++k) davg[k] += dS[idx * PW_NG + k] / S[idx];
            }
            avg = exp(sum / m);
            for (int k = 0; k < PW_NG; ++k) davg[k] *= avg / m;
        }
        if (o->strike_type == OR_FIXED) {
            const double ind = put ? (o->strike - avg > 0.0 ? -1.0 : 0.0) : (avg - o->strike > 0.0 ? 1.0 : 0.0);
            for (int k = 0; k < PW_NG; ++k) dx[k] = ind * davg[k];
        } else {
            const double st = S[n];
            const double on = put ? (avg - st > 0.0) : (st - avg > 0.0);
            for (int k = 0; k < PW_NG; ++k)
                dx[k] = on ? (put ? davg[k] - dT[k] : dT[k] - davg[k]) : 0.0;
        }
    }
    return x;
}

int or_pathwise_greeks(const or_option *o, int64_t num_paths, int64_t num_steps, uint64_t seed,
                       int32_t antithetic, const double *normals1, const double *normals2,
                       int32_t nthreads, int32_t mutate, double out[2 + PW_NG],
                       double *per_sample) {
    if (num_paths <= 0 || num_steps <= 0 || !(o->expiry > 0.0)) return 1;
    const int64_t samples = antithetic ? (num_paths + 1) / 2 : num_paths;
    const int64_t n = num_steps;
    const int32_t streams = o->model == OR_GBM ? 1 : 2;
    const double T = o->expiry;
    const double df = exp(-o->rate * T);
    if (nthreads < 1) nthreads = 1;
    double *parts = (double *)calloc((size_t)nthreads * (2 + PW_NG), sizeof(double));
    if (!parts) return 2;
    int failed = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) if (nthreads > 1)
#endif
    {
        int tid = 0, nt = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        double *z1 = (double *)malloc(sizeof(double) * (size_t)n);
        double *z2 = (double *)malloc(sizeof(double) * (size_t)n);
        double *S = (double *)malloc(sizeof(double) * (size_t)(n + 1));
        double *dS = (double *)malloc(sizeof(double) * (size_t)(n + 1) * PW_NG);
        double *acc = parts + (size_t)tid * (2 + PW_NG);
        if (!z1 || !z2 || !S || !dS) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            failed = 1;
        } else {
            const int64_t lo = samples * tid / nt, hi = samples * (tid + 1) / nt;
            for (int64_t i = lo; i < hi; ++i) {
                if (normals1) {
                    memcpy(z1, normals1 + i * n, sizeof(double) * (size_t)n);
                    if (streams >= 2) memcpy(z2, normals2 + i * n, sizeof(double) * (size_t)n);
                } else {
                    or_sample_normals(seed, i, n, streams, z1, z2);
                }
                double dx[PW_NG], g[PW_NG];
                double x = pw_leg(o, n, z1, z2, mutate, S, dS, dx);
                for (int k = 0; k < PW_NG; ++k) g[k] = dx[k];
                if (antithetic) {
                    for (int64_t j = 0; j < n; ++j) z1[j] = -z1[j];
                    if (streams >= 2)
                        for (int64_t j = 0; j < n; ++j) z2[j] = -z2[j];
                    double dxa[PW_NG];
                    const double xa = pw_leg(o, n, z1, z2, mutate, S, dS, dxa);
                    /* fold_fn's pair average (simulation.rs:235-252) */
                    x = 0.5 * (x + xa);
                    for (int k = 0; k < PW_NG; ++k) g[k] = 0.5 * (g[k] + dxa[k]);
                }
                const double pv = df * x;
                double gp[PW_NG];
                for (int k = 0; k < PW_NG; ++k) gp[k] = df * g[k];
                gp[2] += -T * pv;       /* d exp(-r T) / dr */
                gp[3] += -o->rate * pv; /* d exp(-r T) / dT */
                acc[0] += pv;
                acc[1] += pv * pv;
                for (int k = 0; k < PW_NG; ++k) acc[2 + k] += gp[k];
                if (per_sample) {
                    double *row = per_sample + (size_t)i * (1 + PW_NG);
                    row[0] = pv;
                    for (int k = 0; k < PW_NG; ++k) row[1 + k] = gp[k];
                }
            }
        }
        free(z1);
        free(z2);
        free(S);
        free(dS);
    }
    if (failed) {
        free(parts);
        return 2;
    }
    double tot[2 + PW_NG] = {0};
    for (int t = 0; t < nthreads; ++t)
        for (int k = 0; k < 2 + PW_NG; ++k) tot[k] += parts[(size_t)t * (2 + PW_NG) + k];
    free(parts);
    const double m = (double)samples;
    /* mc_aad.rs:193-200: mean, sample variance (n - 1), stderr */
    out[0] = tot[0] / m;
    const double var = m > 1.0 ? (tot[1] - tot[0] * tot[0] / m) / (m - 1.0) : 0.0;
    out[1] = sqrt(var / m);
    for (int k = 0; k < PW_NG; ++k) out[2 + k] = tot[2 + k] / m;
    return 0;
}
