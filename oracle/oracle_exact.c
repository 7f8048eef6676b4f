/*
 * Oracle: the exact-terminal European pricers of the reference's default
 * build (Cargo feature "simd" off), restated as the tests' reference.
 *
 * TEST INFRASTRUCTURE ONLY; the product never links or loads it.
 *
 *   simulate_chunk_exact_scalar  src/engines/monte_carlo/mc_parallel.rs:98-148
 *   split_paths                  mc_parallel.rs:37-46
 *   mc_european_parallel         mc_parallel.rs:264-362
 *   mc_european_sequential       mc_parallel.rs:365-434
 *   mc_european_with_arena       src/engines/monte_carlo/mc_engine.rs:186-343
 *
 * Each path takes ONE draw, sample_standard_normal, from a single
 * xoshiro256++ stream (per chunk for the parallel pricer), and the terminal
 * spot spot * exp(fma(total_diffusion, z, total_drift)). The sums follow
 * the reference's grouping (blocks of 8 payoffs, Kahan for the arena); the
 * parallel pricer's chunk sums are added in chunk order (rayon's reduce
 * tree is not reproduced: that changes only the last bits of the price).
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double ex_payoff(int32_t type, double spot, double strike) {
    return type == OR_CALL ? fmax(spot - strike, 0.0) : fmax(strike - spot, 0.0);
}

static double ex_normal(uint64_t s[4]) {
    return or_acklam_inv_cdf(or_uniform_open01(or_xoshiro_next_f64(s)));
}

/* simulate_chunk_exact_scalar: sums over n paths of one stream */
static void chunk_exact(int32_t type, double strike, double spot0, double drift, double diff,
                        int64_t n, uint64_t seed, double *sum_out, double *sq_out,
                        double *pay) {
    uint64_t s[4];
    or_xoshiro_seed(seed, s);
    double sum = 0.0, sum_sq = 0.0;
    int64_t remaining = n, i = 0;
    while (remaining >= 8) {
        double z[8], p[8];
        for (int k = 0; k < 8; ++k) z[k] = ex_normal(s);
        for (int k = 0; k < 8; ++k) p[k] = ex_payoff(type, spot0 * exp(fma(diff, z[k], drift)), strike);
        sum += p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7];
        sum_sq += p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3] + p[4] * p[4] +
                  p[5] * p[5] + p[6] * p[6] + p[7] * p[7];
        if (pay) memcpy(pay + i, p, sizeof p);
        i += 8;
        remaining -= 8;
    }
    while (remaining > 0) {
        const double z = ex_normal(s);
        const double px = ex_payoff(type, spot0 * exp(fma(diff, z, drift)), strike);
        sum += px;
        sum_sq += px * px;
        if (pay) pay[i] = px;
        ++i;
        --remaining;
    }
    *sum_out = sum;
    *sq_out = sum_sq;
}

/* kind 0 sequential, 1 parallel (n_chunks = rayon::current_num_threads()),
 * 2 with_arena. q: the effective dividend yield (0, 1) or dividend_yield
 * (2). out: price, stderr, has_stderr (0: stderr None). payoffs (may be
 * NULL): n_paths, in path order. Returns 0, or 1 for an unknown kind. */
int or_mc_european_exact(int32_t kind, int32_t type, double spot, double strike, double rate,
                         double q, double vol, double expiry, int64_t n_paths, int32_t n_chunks,
                         double out[3], double *payoffs) {
    out[0] = NAN;
    out[1] = 0.0;
    out[2] = 0.0;
    if (kind < 0 || kind > 2) return 1;
    if (n_paths == 0) return 0;
    if (kind == 2 && (!(strike > 0.0) || expiry < 0.0)) return 0; /* validate() */
    if (expiry <= 0.0) {
        out[0] = ex_payoff(type, spot, strike);
        out[2] = 1.0;
        return 0;
    }
    if (vol <= 0.0 || !isfinite(vol)) return 0;
    const double t = expiry;
    const double total_drift = (rate - q - 0.5 * vol * vol) * t;
    const double total_diffusion = vol * sqrt(t);
    const double discount = exp(-rate * t);
    const double n = (double)n_paths;
    double sum = 0.0, sum_sq = 0.0, variance;
    if (kind == 0) {
        chunk_exact(type, strike, spot, total_drift, total_diffusion, n_paths,
                    0xDEADBEEFCAFEBABEull, &sum, &sum_sq, payoffs);
    } else if (kind == 1) {
        const int64_t chunks = n_chunks < 1 ? 1 : n_chunks;
        const int64_t base = n_paths / chunks, rem = n_paths % chunks;
        int64_t first = 0, idx = 0;
        for (int64_t c = 0; c < chunks; ++c) {
            const int64_t len = c < rem ? base + 1 : base;
            if (len == 0) continue; /* split_paths filters empty chunks */
            const uint64_t seed = 0xDEADBEEFCAFEBABEull + (uint64_t)idx * 6364136223846793005ull;
            double cs, cq;
            chunk_exact(type, strike, spot, total_drift, total_diffusion, len, seed, &cs, &cq,
                        payoffs ? payoffs + first : NULL);
            sum += cs;
            sum_sq += cq;
            first += len;
            ++idx;
        }
    } else {
        uint64_t s[4];
        or_xoshiro_seed(42, s); /* ARENA_MC_SEED */
        double *buf = payoffs ? payoffs : (double *)malloc(sizeof(double) * (size_t)n_paths);
        if (!buf) return 2;
        for (int64_t i = 0; i < n_paths; ++i) {
            const double z = ex_normal(s);
            buf[i] = ex_payoff(type, spot * exp(fma(total_diffusion, z, total_drift)), strike);
        }
        /* Kahan summation (mc_engine.rs:296-312) */
        double comp = 0.0, comp_sq = 0.0;
        for (int64_t i = 0; i < n_paths; ++i) {
            const double v = buf[i];
            const double y = v - comp;
            const double tt = sum + y;
            comp = (tt - sum) - y;
            sum = tt;
            const double v2 = v * v;
            const double y2 = v2 - comp_sq;
            const double t2 = sum_sq + y2;
            comp_sq = (t2 - sum_sq) - y2;
            sum_sq = t2;
        }
        if (!payoffs) free(buf);
    }
    const double mean = sum / n;
    if (kind == 2) {
        variance = n_paths > 1 ? (sum_sq - sum * sum / n) / (n - 1.0) : 0.0;
    } else {
        variance = n_paths > 1 ? fmax((sum_sq - sum * sum / n) / (n - 1.0), 0.0) : 0.0;
    }
    out[0] = discount * mean;
    out[1] = discount * sqrt(variance / n);
    out[2] = 1.0;
    return 0;
}
