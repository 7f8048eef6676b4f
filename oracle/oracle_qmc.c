/*
 * CPU oracle, quasi-Monte Carlo part -- TEST INFRASTRUCTURE ONLY (the tests' reference
 * for the HIP Sobol path and its timed CPU baseline).
 *
 * Restates src/math/sobol.rs (SobolSequence: direction numbers, scramblers,
 * Gray-code stepping, the u64 -> (0,1) map) and
 * src/engines/monte_carlo/mc_qmc.rs:40-144 (mc_european_qmc_with_seed)
 * of rosssaunders/openferric, sequentially as the reference runs them.
 * Compiled with -ffp-contract=off; the fma() calls restate Rust mul_add.
 *
 * Pinned by (tests/test_oracle_qmc.py): the reference's assertions (points
 * in [0,1), reproducibility, dimension-0 mean, QMC within 0.1 % of
 * Black-Scholes at 10k points, QMC stderr below PRNG MC's); the (0,m,1)-net
 * property of every dimension (the first 2^m Gray-code points of one
 * coordinate fall one per dyadic interval), which fails for a wrong
 * direction-number mask or shift; and the Gray-code identity (the n-th point
 * is the XOR of the direction numbers at the set bits of n ^ (n >> 1)),
 * checked against this sequential recurrence.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* splitmix64 (sobol.rs:147-152): the stateless mix of x */
static uint64_t sob_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* build_direction_numbers (sobol.rs:121-144) */
void or_sobol_directions(uint64_t dim, uint64_t seed, uint64_t v[64]) {
    if (dim == 0) {
        for (int j = 0; j < 64; ++j) v[j] = 1ull << (63 - j);
        return;
    }
    for (int j = 0; j < 64; ++j) {
        const uint64_t hash = sob_mix(seed ^ ((dim + 1) << 40) ^ (uint64_t)j);
        const uint64_t mask = j == 63 ? ~0ull : ((1ull << (j + 1)) - 1);
        const uint64_t m = (hash | 1ull) & mask;
        v[j] = m << (63 - j);
    }
}

/* scramblers (sobol.rs:37) */
uint64_t or_sobol_scrambler(uint64_t dim, uint64_t seed) {
    return sob_mix(seed ^ ((dim + 1) << 32));
}

/* SobolSequence::new + next_into repeated: writes points 1..n (index n in the
 * reference's counting) as raw scrambled words and as uniforms, [point][dim].
 * Either output may be NULL. */
void or_sobol_points(int64_t dims, uint64_t seed, int64_t n, uint64_t *words, double *u) {
    uint64_t *v = (uint64_t *)malloc(sizeof(uint64_t) * 64 * (size_t)dims);
    uint64_t *scr = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)dims);
    uint64_t *x = (uint64_t *)calloc((size_t)dims, sizeof(uint64_t));
    for (int64_t d = 0; d < dims; ++d) {
        or_sobol_directions((uint64_t)d, seed, v + 64 * d);
        scr[d] = or_sobol_scrambler((uint64_t)d, seed);
    }
    const double inv = 1.0 / 18446744073709551616.0;
    const double half = 0.5 * inv;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t next = (uint64_t)i + 1;
        const int c = __builtin_ctzll(next);
        for (int64_t d = 0; d < dims; ++d) {
            x[d] ^= v[64 * d + c];
            const uint64_t s = x[d] ^ scr[d];
            if (words) words[i * dims + d] = s;
            if (u) u[i * dims + d] = fma((double)s, inv, half);
        }
    }
    free(v);
    free(scr);
    free(x);
}

/* mc_european_qmc_with_seed (mc_qmc.rs:40-144). o: GBM vanilla, European
 * exercise, div_yield = the market's effective yield. *has_stderr = 0 where
 * the reference returns stderr None (price NaN). */
void or_mc_european_qmc(const or_option *o, int64_t n_paths, int64_t n_steps, uint64_t seed,
                        double *price, double *stderr_out, int32_t *has_stderr) {
    *has_stderr = 0;
    *stderr_out = 0.0;
    if (n_paths <= 0 || n_steps <= 0) {
        *price = NAN;
        return;
    }
    if (o->expiry <= 0.0) {
        *price = o->option_type == OR_CALL ? fmax(o->spot - o->strike, 0.0)
                                           : fmax(o->strike - o->spot, 0.0);
        *has_stderr = 1;
        return;
    }
    const double vol = o->vol;
    if (vol <= 0.0 || !isfinite(vol)) {
        *price = NAN;
        return;
    }
    const double dt = o->expiry / (double)n_steps;
    const double dt_drift = (o->rate - o->div_yield - 0.5 * vol * vol) * dt;
    const double dt_vol = vol * sqrt(dt);
    const double discount = exp(-o->rate * o->expiry);
    double *u = (double *)malloc(sizeof(double) * (size_t)n_steps);
    uint64_t *v = (uint64_t *)malloc(sizeof(uint64_t) * 64 * (size_t)n_steps);
    uint64_t *scr = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n_steps);
    uint64_t *x = (uint64_t *)calloc((size_t)n_steps, sizeof(uint64_t));
    for (int64_t d = 0; d < n_steps; ++d) {
        or_sobol_directions((uint64_t)d, seed, v + 64 * d);
        scr[d] = or_sobol_scrambler((uint64_t)d, seed);
    }
    const double inv = 1.0 / 18446744073709551616.0;
    const double half = 0.5 * inv;
    double sum = 0.0, sum_sq = 0.0;
    for (int64_t i = 0; i < n_paths; ++i) {
        const uint64_t next = (uint64_t)i + 1;
        const int c = __builtin_ctzll(next);
        for (int64_t d = 0; d < n_steps; ++d) {
            x[d] ^= v[64 * d + c];
            u[d] = fma((double)(x[d] ^ scr[d]), inv, half);
        }
        double spot = o->spot;
        for (int64_t d = 0; d < n_steps; ++d) {
            const double z = or_acklam_inv_cdf(or_uniform_open01(u[d]));
            spot *= exp(fma(dt_vol, z, dt_drift));
        }
        const double px = o->option_type == OR_CALL ? fmax(spot - o->strike, 0.0)
                                                    : fmax(o->strike - spot, 0.0);
        sum += px;
        sum_sq += px * px;
    }
    const double n = (double)n_paths;
    const double mean = sum / n;
    const double variance = n_paths > 1 ? fmax((sum_sq - sum * sum / n) / (n - 1.0), 0.0) : 0.0;
    *price = discount * mean;
    *stderr_out = discount * sqrt(variance / n);
    *has_stderr = 1;
    free(u);
    free(v);
    free(scr);
    free(x);
}
