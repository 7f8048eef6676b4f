/*
 * Oracle restatement of SpreadMonteCarloEngine::price
 * (src/engines/monte_carlo/spread_mc.rs:83-200) and of the reference's Pcg64
 * stream (src/math/fast_rng.rs:93-137).
 *
 * TEST INFRASTRUCTURE ONLY: the tests' reference for ofg_spread_price; the product
 * library never links, calls or falls back to it.
 *
 * Pins (tests/test_oracle_spread.py):
 *   or_pcg64_*      numpy's PCG64 (PCG XSL-RR 128/64 with the same
 *                   multiplier), whose first output from a state is the
 *                   reference's output after one step: exact u64 equality
 *                   over thousands of draws and many seeds.
 *   or_spread_mc    Margrabe's exchange-option closed form at k = 0 (4 SE),
 *                   the zero-vol legs as Black-Scholes calls and puts on the
 *                   other asset (4 SE), Kirk's approximation within 1 % on
 *                   the reference's own test case, each sample's payoff from
 *                   the pinned streams and Acklam inverse, the early returns
 *                   and the validation messages.
 */
#include "oracle.h"

#include <math.h>

/* ---------------------------------------------------------------- Pcg64 */

typedef unsigned __int128 u128;

static const u128 PCG_MULT =
    ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;   /* fast_rng.rs:100 */

static u128 pcg_state(const uint64_t st[4]) { return ((u128)st[1] << 64) | st[0]; }
static u128 pcg_inc(const uint64_t st[4]) { return ((u128)st[3] << 64) | st[2]; }

/* fast_rng.rs:114-126: output of the OLD state, XSL-RR 128/64. */
uint64_t or_pcg64_next_u64(uint64_t st[4]) {
    const u128 old = pcg_state(st);
    const u128 next = old * PCG_MULT + pcg_inc(st);
    st[0] = (uint64_t)next;
    st[1] = (uint64_t)(next >> 64);
    const uint64_t xorshifted = (uint64_t)((old >> 64) ^ old);
    const unsigned rot = (unsigned)(old >> 122);
    return rot ? (xorshifted >> rot) | (xorshifted << (64u - rot)) : xorshifted;
}

/* fast_rng.rs:102-112: state (hi, lo) and stream from SplitMix64, inc =
 * (stream << 1) | 1 over 128 bits, one draw discarded. st = {state lo, state
 * hi, inc lo, inc hi}. */
void or_pcg64_seed(uint64_t seed, uint64_t st[4]) {
    uint64_t sm = seed;
    const uint64_t hi = or_splitmix64_next(&sm);
    const uint64_t lo = or_splitmix64_next(&sm);
    const uint64_t stream = or_splitmix64_next(&sm);
    const u128 inc = ((u128)stream << 1) | 1u;
    st[0] = lo;
    st[1] = hi;
    st[2] = (uint64_t)inc;
    st[3] = (uint64_t)(inc >> 64);
    (void)or_pcg64_next_u64(st);
}

/* fast_rng.rs:128-132 */
double or_pcg64_next_f64(uint64_t st[4]) {
    const uint64_t x = or_pcg64_next_u64(st) >> 11;
    return (double)x * (1.0 / 9007199254740992.0);
}

/* --------------------------------------------------------- spread engine */

/* FastRng::from_seed + sample_standard_normal (fast_rng.rs:143-152, 223-226)
 * for the two reproducible kinds. */
void or_fast_rng_seed(or_fast_rng *g, int32_t kind, uint64_t seed) {
    g->kind = kind;
    if (kind == OR_RNG_PCG64) or_pcg64_seed(seed, g->s);
    else or_xoshiro_seed(seed, g->s);
}

double or_fast_rng_normal(or_fast_rng *g) {
    const double u = g->kind == OR_RNG_PCG64 ? or_pcg64_next_f64(g->s) : or_xoshiro_next_f64(g->s);
    return or_acklam_inv_cdf(or_uniform_open01(u));
}

/* The stream kind of MonteCarloEngine's per-sample streams
 * (simulation.rs:215-225 with FastRngKind, mc_engine.rs:596-633). */
static int32_t g_rng_kind = OR_RNG_XOSHIRO;
void or_set_rng_kind(int32_t kind) { g_rng_kind = kind; }
int32_t or_rng_kind(void) { return g_rng_kind; }

/* terminal_spread_payoff (spread_mc.rs:178-197) */
static double terminal_payoff(const double *p, double z1, double z2, double corr_tail,
                              double drift1, double drift2, double vol_term1, double vol_term2) {
    const double s1 = p[0], s2 = p[1], k = p[2], rho = p[5];
    const double w1 = z1;
    const double w2 = fma(rho, z1, corr_tail * z2);
    const double s1_t = s1 * exp(fma(vol_term1, w1, drift1));
    const double s2_t = s2 * exp(fma(vol_term2, w2, drift2));
    return fmax(s1_t - s2_t - k, 0.0);
}

/* SpreadOption::validate (src/instruments/spread.rs:31-58): 0 or the
 * negative code of the first failing check, in the reference's order. */
int32_t or_spread_validate(const double p[10]) {
    const double s1 = p[0], s2 = p[1], k = p[2], vol1 = p[3], vol2 = p[4], rho = p[5], t = p[9];
    if (s1 <= 0.0 || s2 <= 0.0) return OR_SPREAD_BAD_SPOTS;
    if (k < 0.0) return OR_SPREAD_BAD_STRIKE;
    if (vol1 < 0.0 || vol2 < 0.0) return OR_SPREAD_BAD_VOLS;
    if (!(rho >= -1.0 && rho <= 1.0)) return OR_SPREAD_BAD_RHO;
    if (t < 0.0) return OR_SPREAD_BAD_MATURITY;
    return 0;
}

/* SpreadMonteCarloEngine::price (spread_mc.rs:83-175) with reproducible
 * streams. p = {s1, s2, k, vol1, vol2, rho, q1, q2, r, t}. out = {price,
 * stderr}. payoffs (optional) receives each sample's payoff (antithetic:
 * the pair's average), in sample order. Returns 0, or a negative code:
 * validation (or_spread_validate), OR_SPREAD_NO_PATHS, OR_SPREAD_BAD_RNG. */
int32_t or_spread_mc(const double p[10], int64_t num_paths, uint64_t seed, int32_t antithetic,
                     int32_t rng_kind, double out[2], double *payoffs) {
    const int32_t v = or_spread_validate(p);
    if (v) return v;
    if (num_paths == 0) return OR_SPREAD_NO_PATHS;
    if (rng_kind != OR_RNG_XOSHIRO && rng_kind != OR_RNG_PCG64) return OR_SPREAD_BAD_RNG;
    const double s1 = p[0], s2 = p[1], k = p[2], vol1 = p[3], vol2 = p[4], rho = p[5];
    const double q1 = p[6], q2 = p[7], r = p[8], t = p[9];
    if (t <= 0.0) {
        out[0] = fmax(s1 - s2 - k, 0.0);
        out[1] = 0.0;
        return 0;
    }
    const double sqrt_t = sqrt(t);
    const double corr_tail = sqrt(fmax(1.0 - rho * rho, 0.0));
    const double drift1 = (r - q1 - 0.5 * vol1 * vol1) * t;
    const double drift2 = (r - q2 - 0.5 * vol2 * vol2) * t;
    const double vol_term1 = vol1 * sqrt_t;
    const double vol_term2 = vol2 * sqrt_t;
    const double discount = exp(-r * t);
    const int64_t samples = antithetic ? (num_paths + 1) / 2 : num_paths;
    double sum = 0.0, sum_sq = 0.0;
    for (int64_t i = 0; i < samples; ++i) {
        or_fast_rng g;
        or_fast_rng_seed(&g, rng_kind, or_stream_seed(seed, (uint64_t)i));
        const double z1 = or_fast_rng_normal(&g);
        const double z2 = or_fast_rng_normal(&g);
        double x = terminal_payoff(p, z1, z2, corr_tail, drift1, drift2, vol_term1, vol_term2);
        if (antithetic) {
            const double a = terminal_payoff(p, -z1, -z2, corr_tail, drift1, drift2, vol_term1,
                                             vol_term2);
            x = 0.5 * (x + a);
        }
        if (payoffs) payoffs[i] = x;
        sum += x;
        sum_sq += x * x;
    }
    const double n = (double)samples;
    const double mean = sum / n;
    const double variance = samples > 1 ? (sum_sq - sum * sum / n) / (n - 1.0) : 0.0;
    out[0] = discount * mean;
    out[1] = discount * sqrt(variance / n);
    return 0;
}
