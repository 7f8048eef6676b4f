/*
 * openferric_amd -- C ABI of the MI355X (gfx950) Monte Carlo pricing engine.
 *
 * A drop-in for OpenFerric's Monte Carlo pricing hot path: plain C types,
 * caller-owned host arrays, no exceptions across the boundary. Status codes
 * map 1:1 onto the reference's PricingError variants
 * (src/core/engine.rs:292-302) plus OFG_DEVICE_ERROR for HIP failures; the
 * message of the last failure on the calling thread is ofg_last_error().
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef OPENFERRIC_AMD_H
#define OPENFERRIC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: PricingError (src/core/engine.rs:292-302) ---------- */
enum ofg_status {
    OFG_OK = 0,
    OFG_INVALID_INPUT = 1,          /* PricingError::InvalidInput */
    OFG_CONVERGENCE_FAILURE = 2,    /* PricingError::ConvergenceFailure */
    OFG_MARKET_DATA_MISSING = 3,    /* PricingError::MarketDataMissing */
    OFG_NUMERICAL_ERROR = 4,        /* PricingError::NumericalError */
    OFG_DEVICE_ERROR = 5            /* HIP runtime / device failure */
};

/* ---- enumerations ------------------------------------------------------ */
enum ofg_model { OFG_MODEL_GBM = 0, OFG_MODEL_HESTON = 1 };
/* MonteCarloInstrument impls (src/engines/monte_carlo/mc_engine.rs:441-584) */
enum ofg_payoff { OFG_PAYOFF_VANILLA = 0, OFG_PAYOFF_ASIAN = 1, OFG_PAYOFF_BARRIER = 2 };
enum ofg_option_type { OFG_CALL = 0, OFG_PUT = 1 };              /* core/types.rs:13-19 */
enum ofg_averaging { OFG_ARITHMETIC = 0, OFG_GEOMETRIC = 1 };     /* core/types.rs:98-104 */
enum ofg_strike_type { OFG_FIXED = 0, OFG_FLOATING = 1 };        /* core/types.rs:106-112 */
enum ofg_barrier_dir { OFG_UP = 0, OFG_DOWN = 1 };               /* core/types.rs:50-56 */
enum ofg_barrier_style { OFG_IN = 0, OFG_OUT = 1 };              /* core/types.rs:58-64 */

/* Where a path's standard normals come from. */
enum ofg_rng_mode {
    /* The reference's own draws: per-sample xoshiro256++ stream seeded with
       SplitMix64(seed + 7919*i), Acklam inverse CDF of the clamped uniform
       (src/mc/simulation.rs:215-225, src/math/fast_rng.rs). Per-path parity
       with the reference. */
    OFG_RNG_REFERENCE = 0,
    /* Caller-supplied normals, step-major [step][sample] per option. */
    OFG_RNG_EXTERNAL = 1,
    /* Throughput mode: per-sample xoroshiro128+ streams (same SplitMix64
       seeding) feeding a branch-free, table-driven Box-Muller transform; path
       exponentials as 2^(y/256). Statistically equivalent draws. */
    OFG_RNG_FAST = 2,
    /* The reference's Pcg64 kind (FastRngKind::Pcg64, fast_rng.rs:93-137):
       per-sample PCG XSL-RR 128/64 streams seeded from SplitMix64 of
       seed + 7919*i (one draw discarded), Acklam inverse CDF as REFERENCE.
       Per-path parity with the reference's Pcg64 engines. Taken by
       ofg_price (greek modes NONE and FD_CRN), ofg_mc_greeks_estimate and
       ofg_spread_price. */
    OFG_RNG_PCG64 = 3
};

/* How ofg_price fills the Greeks.
   ADJOINT  pathwise reverse sweep (checkpointed adjoint kernel), GBM and Heston;
   FD_CRN   central bump-and-revalue with common random numbers, all bumps in
            one batched launch (greeks/bsm.rs:130-212);
   FORWARD  the same pathwise estimator as ADJOINT evaluated in one forward
            sweep, no reverse pass or checkpoints: GBM options by running
            sums of the path, Heston options by a forward tangent of the
            Euler step (all six Heston sensitivities). */
enum ofg_greek_mode {
    OFG_GREEKS_NONE = 0,
    OFG_GREEKS_ADJOINT = 1,
    OFG_GREEKS_FD_CRN = 2,
    OFG_GREEKS_FORWARD = 3
};

/* One priced contract: instrument + market + model + path/step counts.
 * Mirrors (VanillaOption | AsianOption | BarrierOption, Market, Gbm | Heston)
 * as consumed by MonteCarloPricingEngine::price (mc_engine.rs:875-975). */
typedef struct ofg_option {
    int32_t model;          /* ofg_model */
    int32_t payoff;         /* ofg_payoff */
    int32_t option_type;    /* ofg_option_type */
    int32_t averaging;      /* Asian: ofg_averaging */
    int32_t strike_type;    /* Asian: ofg_strike_type */
    int32_t barrier_dir;    /* Barrier: ofg_barrier_dir */
    int32_t barrier_style;  /* Barrier: ofg_barrier_style */
    int32_t n_obs;          /* Asian: number of observation times */
    double spot, strike, expiry;
    double rate, div_yield; /* GBM/Heston drift mu = rate - div_yield */
    double vol;             /* GBM sigma (Market::vol_for, flat) */
    double kappa, theta, xi, rho, v0;   /* Heston (models/stochastic.rs:37-44) */
    double barrier_level, rebate;
    const double *obs_times;            /* Asian observation times (host) */
    int64_t num_paths;      /* MonteCarloPricingEngine::num_paths */
    int64_t num_steps;      /* MonteCarloPricingEngine::num_steps */
    uint64_t seed;          /* MonteCarloPricingEngine::seed */
} ofg_option;

/* Engine configuration shared by all options of one call. */
typedef struct ofg_engine_cfg {
    int32_t antithetic;       /* VarianceReduction::Antithetic */
    int32_t control_variate;  /* VarianceReduction::ControlVariate (built-in CVs) */
    int32_t rng_mode;         /* ofg_rng_mode */
    int32_t greek_mode;       /* ofg_greek_mode */
    double bump_rel;          /* FD_CRN relative bump (e.g. 1e-4) */
    /* OFG_RNG_EXTERNAL: host arrays, options concatenated in call order,
       each option laid out [step][sample] (samples = ceil(paths/2) with
       antithetic). normals2 only for Heston. */
    const double *normals1;
    const double *normals2;
    int32_t tune_block;       /* 0 = autotuned/default; else threads/block */
    int32_t tune_item_steps;  /* 0 = default; else lane-steps per work item */
} ofg_engine_cfg;

/* PricingResult (src/core/engine.rs:277-287) flattened, plus the Heston
 * parameter gradient and timing side channel. */
typedef struct ofg_result {
    double price;
    double stderr_;           /* PricingResult::stderr */
    int32_t has_stderr;
    int32_t has_greeks;
    double delta, gamma, vega, theta, rho;   /* Greeks (engine.rs:15-28) */
    /* dPrice/d{S0, v0, kappa, theta, xi, rho} (Heston adjoint) */
    double heston_grad[6];
    /* Diagnostics keys NumPaths, NumSteps, Vol (mc_engine.rs:964-967) */
    double num_paths, num_steps, vol;
    double kernel_ms;         /* device time of this call's kernels */
} ofg_result;

/* The five estimates of MonteCarloGreeksEngine::run_estimators
 * (src/engines/monte_carlo/mc_greeks.rs:131-243, MonteCarloGreekEstimate
 * :235-242): discounted means over the samples. */
typedef struct ofg_mc_greeks {
    double pathwise_delta, pathwise_gamma;
    double lr_delta, lr_gamma, lr_vega;
    double num_samples;       /* ceil(paths/2) with antithetic, else paths */
    double kernel_ms;
} ofg_mc_greeks;

/* ---- Longstaff-Schwartz least-squares Monte Carlo --------------------- */
/* Instruments priced by LongstaffSchwartzEngine
 * (src/engines/lsm/longstaff_schwartz.rs:448-735). */
enum ofg_lsm_instrument {
    OFG_LSM_VANILLA = 0,    /* VanillaOption (instruments/vanilla.rs) */
    OFG_LSM_BERMUDAN = 1,   /* BermudanOption (instruments/bermudan.rs) */
    OFG_LSM_BARRIER = 2     /* BarrierOption: the engine's plain MC pricer */
};
/* ExerciseStyle (core/types.rs) of a VanillaOption */
enum ofg_exercise { OFG_EUROPEAN = 0, OFG_AMERICAN = 1, OFG_BERMUDAN = 2 };
/* LsmDynamics (longstaff_schwartz.rs:25-39); BermudanOption only -- the
 * VanillaOption and BarrierOption paths always simulate GBM. LOCAL_VOL uses
 * the market's flat volatility as the surface. */
enum ofg_lsm_dynamics { OFG_LSM_GBM = 0, OFG_LSM_LOCAL_VOL = 1, OFG_LSM_HESTON = 2 };

/* One LongstaffSchwartzEngine::price call: instrument + market + engine. */
typedef struct ofg_lsm_option {
    int32_t instrument;     /* ofg_lsm_instrument */
    int32_t option_type;    /* ofg_option_type */
    int32_t exercise;       /* VanillaOption: ofg_exercise */
    int32_t dynamics;       /* BermudanOption: ofg_lsm_dynamics */
    int32_t n_dates;        /* exercise dates (Bermudan exercise / BermudanOption) */
    int32_t barrier_dir;    /* BarrierOption: ofg_barrier_dir */
    int32_t barrier_style;  /* BarrierOption: ofg_barrier_style */
    int32_t pad;
    double spot, strike, expiry;
    double rate, div_yield; /* div_yield: Market::effective_dividend_yield(expiry) */
    double vol;             /* Market::vol_for (flat) */
    double kappa, theta, xi, rho, v0;   /* Heston dynamics (models/stochastic.rs:37-44) */
    double barrier_level, rebate;
    const double *dates;    /* exercise dates in years (host) */
    const double *strikes;  /* BermudanOption::strike_schedule (host, n_dates) */
    int64_t num_paths;      /* LongstaffSchwartzEngine::num_paths */
    int64_t num_steps;      /* LongstaffSchwartzEngine::num_steps */
    uint64_t seed;          /* LongstaffSchwartzEngine::seed */
} ofg_lsm_option;

/* ExerciseBoundaryPoint (longstaff_schwartz.rs:42-54). */
typedef struct ofg_lsm_boundary {
    double time, strike;
    double boundary_spot;   /* valid when has_boundary */
    int32_t has_boundary;
    int32_t pad;
    int64_t itm_paths, exercised_paths;
} ofg_lsm_boundary;

/* ---- multi-asset: worst-of autocallables ------------------------------ */
/* Autocallable / PhoenixAutocallable (src/instruments/autocallable.rs) with
 * the market inputs of price_autocallable (src/pricing/autocallable.rs:52). */
typedef struct ofg_autocall {
    int32_t phoenix;          /* 0 Autocallable, 1 PhoenixAutocallable */
    int32_t memory;           /* Phoenix: memory coupons */
    int32_t n_underlyings;    /* contract underlyings (at most 8) */
    int32_t n_dates;
    const int32_t *underlyings;      /* indices into the market arrays */
    const double *autocall_dates;
    double notional, maturity, autocall_barrier, coupon_rate;
    double coupon_barrier;    /* Phoenix only */
    double ki_barrier, ki_strike;
    int32_t n_assets;         /* market arrays */
    int32_t pad;
    const double *spots, *vols;
    const double *corr;       /* n_assets x n_assets, row-major */
    double rate, div_yield;   /* r, q */
    int64_t num_paths, num_steps;
} ofg_autocall;

/* AutocallableSensitivities (pricing/autocallable.rs:25-33). */
typedef struct ofg_autocall_sens {
    double delta[8];          /* per contract underlying */
    double vega, cega;        /* cega NaN when no correlation bump was valid */
    int32_t n_delta;
    int32_t ok;               /* 0: the reference's sensitivities would fail */
} ofg_autocall_sens;

/* BasketOption (src/instruments/basket.rs:14-34) with its market inputs
 * (price_basket_mc, src/pricing/basket.rs:99-120). */
typedef enum ofg_basket_type {
    OFG_BASKET_AVERAGE = 0,   /* sum_i w_i S_i(T) */
    OFG_BASKET_BEST_OF = 1,   /* max_i S_i(T)/S_i(0) */
    OFG_BASKET_WORST_OF = 2   /* min_i S_i(T)/S_i(0) */
} ofg_basket_type;

typedef struct ofg_basket {
    int32_t basket_type;      /* ofg_basket_type */
    int32_t is_call;
    int32_t n_weights;        /* Average: must equal n_assets */
    int32_t n_assets;         /* at most 16 */
    const double *weights;
    double strike, maturity;
    const double *spots, *vols, *dividends;
    const double *corr;       /* n_assets x n_assets, row-major */
    double rate;
    int64_t num_paths;
    uint64_t seed;            /* the reference's MC_SEED is 3148159 */
    /* The contract: 0 BasketOption (above), 1 OutperformanceBasketOption
     * (instruments/basket.rs:70-118: weights = lagger_weights, is_call =
     * option_type, basket_type ignored), 2 QuantoBasketOption (:120-165:
     * the basket above plus the FX fields; `rate` is ignored). */
    int32_t product;
    int32_t leader_index;     /* outperformance */
    double fx_rate, fx_vol, domestic_rate, foreign_rate;   /* quanto */
    const double *asset_fx_corr;                             /* quanto */
    int32_t n_asset_fx_corr;
    /* BasketOption under a FactorCorrelationModel instead of `corr`
     * (price_basket_mc_with_factor_model, pricing/basket.rs:162-205;
     * math/correlation.rs:62-212): 0 none, 1 OneFactor, 2 MultiFactor. */
    int32_t factor_kind;
    int32_t n_loading_rows, n_factors;   /* rows = the model's asset count; factors <= 4 */
    const double *loadings;              /* n_loading_rows x n_factors, row-major */
    /* BasketCopula (pricing/basket.rs:32-46) of a BasketOption, matrix or
     * factor model (price_basket_mc_with_copula, :120-160): 0 Gaussian,
     * 1 StudentT with copula_dof degrees of freedom (at least 2, checked when
     * the first path is drawn, as math/correlation.rs:559-562). */
    int32_t copula;
    int32_t copula_dof;
} ofg_basket;

enum { OFG_COPULA_GAUSSIAN = 0, OFG_COPULA_STUDENT_T = 1 };

/* BasketSensitivities (pricing/basket.rs:57-65) and the diagnostics of
 * price_basket_mc (:149-151). */
typedef struct ofg_basket_sens {
    double delta[16];         /* per asset */
    double vega, cega;        /* cega NaN when no correlation bump was valid */
    double rho;               /* average off-diagonal input correlation */
    int32_t n_delta;
    int32_t ok;               /* 0: basket_sensitivities would fail */
    int32_t repaired;         /* 1: the input matrix was not PSD and was projected */
    int32_t pad;
} ofg_basket_sens;

/* Structured-product language (src/dsl): a compiled product, owned by the
 * caller (ofg_dsl_compile / ofg_dsl_free). */
typedef struct ofg_dsl_product ofg_dsl_product;

/* AssetMarketData (src/dsl/market.rs:4-30): the kind selects the stepper
 * (Equity, Fx: GBM; Commodity: Schwartz one-factor; Rate: Vasicek). */
typedef enum ofg_dsl_asset_kind {
    OFG_DSL_EQUITY = 0,
    OFG_DSL_FX = 1,
    OFG_DSL_COMMODITY = 2,
    OFG_DSL_RATE = 3
} ofg_dsl_asset_kind;

typedef struct ofg_dsl_asset {
    int32_t kind;             /* ofg_dsl_asset_kind */
    int32_t pad;
    double spot;              /* Equity / Fx / Commodity spot; Rate initial_rate */
    double vol;
    double dividend_yield;    /* Equity */
    double domestic_rate, foreign_rate;       /* Fx */
    double convenience_yield, kappa, mu;      /* Commodity */
    double mean_reversion, long_run_mean;     /* Rate */
} ofg_dsl_asset;

/* MultiAssetMarket (src/dsl/market.rs:158-168). */
typedef struct ofg_dsl_market {
    int32_t n_assets;         /* at most 16 */
    int32_t pad;
    const ofg_dsl_asset *assets;
    const double *corr;       /* n_assets x n_assets, row-major */
    double rate;
} ofg_dsl_market;

/* DslMonteCarloEngine (src/dsl/engine.rs:237-253). */
typedef struct ofg_dsl_engine {
    int64_t num_paths, num_steps;
    uint64_t seed;
} ofg_dsl_engine;

/* CompiledProduct's header fields (src/dsl/ir.rs:141-151). */
typedef struct ofg_dsl_info {
    double notional, maturity;
    int32_t num_underlyings, n_state, n_schedules, n_locals;
    char name[128];           /* truncated, NUL-terminated */
} ofg_dsl_info;

typedef struct ofg_ctx ofg_ctx;
typedef struct ofg_batch ofg_batch;

/* Context: owns the HIP stream, device work buffers and tuning cache for one
 * GPU. Replaces the process-wide wgpu context of the existing GPU engine
 * (src/engines/gpu/gpu_mc.rs:276-283). NULL on failure (see ofg_last_error). */
ofg_ctx *ofg_create(int32_t device);
void ofg_destroy(ofg_ctx *ctx);

/* Message of the last failed call on this thread ("" if none). */
const char *ofg_last_error(void);

/* Library version string. */
const char *ofg_version(void);

/* Prices n options. Replaces PricingEngine::price for
 * MonteCarloPricingEngine (src/engines/monte_carlo/mc_engine.rs:875-975)
 * and the simulate+reduce loop MonteCarloEngine::run
 * (src/mc/simulation.rs:182-348); a Heston option replaces
 * MonteCarloEngine::run with HestonPathGenerator (simulation.rs:98-116).
 * With greek_mode != NONE also fills the Greeks
 * (PricingEngine::price_with_greeks_aad, mc_engine.rs:977-987). out[n]. */
int32_t ofg_price(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                  const ofg_engine_cfg *cfg, ofg_result *out);

/* Absolute bumps of finite_difference_greeks / bump_and_reprice
 * (src/greeks/bsm.rs:130-212): spot, rate, vol, maturity. */
typedef struct ofg_fd_bumps {
    double ds, dr, dv, dt;
} ofg_fd_bumps;

/* FiniteDifferenceGreeks (src/greeks/bsm.rs:16-24), plus the device time. */
typedef struct ofg_fd_greeks {
    double delta, gamma, vega, theta, rho, vanna, volga;
    double kernel_ms;
} ofg_fd_greeks;

/* finite_difference_greeks (src/greeks/bsm.rs:165-212) with the Monte Carlo
 * engine as the GenericPricer (bsm.rs:44-55): price(s, k, r, sigma, t) is
 * the price of opts[i] with spot s, strike k, rate r, vol sigma and expiry t
 * (num_steps unchanged; an Asian option's fixings scaled by t / expiry, so
 * each stays at the same path index). GBM options only (sigma is the
 * market vol). The reference's 13 revaluations -- s +- ds (s - ds floored at
 * 1e-8), r +- dr, sigma +- dv (sigma - dv floored at 1e-8), t +- dt (t - dt
 * floored at 1e-8), the four (s, sigma) corners -- run on the option's own
 * seed, so every bump shares the base price's draws, and all n options'
 * revaluations are priced in one batched launch. The Greeks combine them
 * with the reference's formulas (theta = -dV/dT; vanna from the four
 * corners; volga). cfg: rng_mode, antithetic, control_variate as ofg_price
 * (greek_mode ignored). out[n]. */
int32_t ofg_finite_difference_greeks(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                                     const ofg_engine_cfg *cfg, const ofg_fd_bumps *bumps,
                                     ofg_fd_greeks *out);

/* bump_and_reprice (src/greeks/bsm.rs:130-163) with the same pricer:
 * out[6 i .. 6 i + 5] = (base, s + ds, r + dr, sigma + dv, expiry
 * max(t + dt, 1e-8), s - ds) prices of opts[i] (the spot-down bump is not
 * floored, as in the reference), all in one batched launch. out[6 n]. */
int32_t ofg_bump_and_reprice(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                             const ofg_engine_cfg *cfg, const ofg_fd_bumps *bumps, double *out);

/* The exact-terminal European pricers of the reference's default build
 * (Cargo feature "simd" off). */
enum ofg_exact_kind {
    OFG_EXACT_SEQUENTIAL = 0,   /* mc_european_sequential (mc_parallel.rs:365-434) */
    OFG_EXACT_PARALLEL = 1,     /* mc_european_parallel (mc_parallel.rs:264-362) */
    OFG_EXACT_ARENA = 2         /* mc_european_with_arena (mc_engine.rs:186-343) */
};

/* Replaces mc_european_sequential, mc_european_parallel and
 * mc_european_with_arena (src/engines/monte_carlo/mc_parallel.rs:98-148,
 * 264-434; mc_engine.rs:186-343) on the reference's own draws: one standard
 * normal per path, drawn in path order from a single xoshiro256++ stream --
 * seed 0xDEAD_BEEF_CAFE_BABE (sequential), 42 (arena), or per chunk of
 * split_paths(num_paths, n_chunks) seed 0xDEAD_BEEF_CAFE_BABE + i *
 * 6364136223846793005 (parallel; n_chunks is the reference run's
 * rayon::current_num_threads()) -- split across GPU threads by jump-ahead.
 * opts[i]: European GBM vanillas; spot, strike, expiry, rate, vol,
 * num_paths; div_yield = the market's effective dividend yield (sequential,
 * parallel) or dividend_yield (arena); num_steps and seed are ignored (the
 * reference ignores n_steps and fixes the seeds). As the reference, degenerate
 * inputs give results, not errors: zero paths, a non-positive or non-finite
 * vol, or (arena) a strike <= 0 or expiry < 0 give price NaN and
 * has_stderr 0; expiry <= 0 gives the payoff at the spot with stderr 0.
 * Diagnostics num_paths, num_steps = 1, vol. payoffs (may be NULL): every
 * path's undiscounted payoff, options concatenated (num_paths each, in path
 * order). out[n]. */
int32_t ofg_european_exact(ofg_ctx *ctx, const ofg_option *opts, size_t n, int32_t kind,
                           int32_t n_chunks, ofg_result *out, double *payoffs);

/* SpreadOption (src/instruments/spread.rs:15-26) with the path count and
 * seed of the SpreadMonteCarloEngine pricing it (spread_mc.rs:21-33). */
typedef struct ofg_spread_option {
    double s1, s2, k, vol1, vol2, rho, q1, q2, r, t;
    int64_t num_paths;
    uint64_t seed;
} ofg_spread_option;

/* Replaces SpreadMonteCarloEngine::price (src/engines/monte_carlo/
 * spread_mc.rs:83-175) with reproducible streams: sample i draws z1 then z2
 * from its own stream seeded with seed + 7919 i -- xoshiro256++
 * (OFG_RNG_REFERENCE, the reference's default kind) or Pcg64
 * (OFG_RNG_PCG64) per path as the reference, or the fast generator
 * (OFG_RNG_FAST); antithetic averages each sample with (-z1, -z2) over
 * ceil(num_paths / 2) samples. Errors, in the reference's order per option:
 * SpreadOption::validate's messages, then "spread MC num_paths must be > 0";
 * any failing option fails the call. t <= 0 gives max(s1 - s2 - k, 0) with
 * stderr 0 and num_paths 0 (no diagnostics). Otherwise price and stderr
 * (discounted sample mean, n - 1 variance over the samples, unclamped) and
 * the num_paths diagnostic. payoffs (may be NULL): each simulated option's
 * sample payoffs, options concatenated, samples slots per option in call
 * order. out[n]. */
int32_t ofg_spread_price(ofg_ctx *ctx, const ofg_spread_option *opts, size_t n,
                         int32_t antithetic, int32_t rng_mode, ofg_result *out, double *payoffs);

/* Per-sample payoffs of ONE option: x[i] (and control variate y[i]) exactly
 * as the fold body of MonteCarloEngine::run computes them
 * (simulation.rs:212-252), antithetic legs averaged. xs/ys hold
 * samples = antithetic ? ceil(paths/2) : paths doubles; ys may be NULL. */
int32_t ofg_sample_payoffs(ofg_ctx *ctx, const ofg_option *opt,
                           const ofg_engine_cfg *cfg, double *xs, double *ys);

/* The reference's per-sample normal streams, drawn on the GPU. Replaces the
 * draw loop of MonteCarloEngine::run's fold body (src/mc/simulation.rs:
 * 214-225: FastRng::from_seed(kind, resolve_stream_seed(seed, i)), then
 * sample_standard_normal per value; fill_standard_normals,
 * src/math/fast_rng.rs:228-240, over a whole stream): stream i = 0 ..
 * n_streams-1 is seeded with stream_seed(seed, first_stream + i) and yields
 * n_per_stream standard normals, Acklam's inverse CDF of the clamped uniform
 * (a Heston path's z1, z2 interleave in one stream). rng_mode
 * OFG_RNG_REFERENCE (xoshiro256++) or OFG_RNG_PCG64. out: host array of
 * n_streams * n_per_stream doubles, out[i * n_per_stream + j], or with
 * step_major out[j * n_streams + i] (the layout OFG_RNG_EXTERNAL reads).
 * kernel_ms (may be NULL) receives the device time. */
int32_t ofg_fill_normals(ofg_ctx *ctx, int32_t rng_mode, uint64_t seed, uint64_t first_stream,
                         int64_t n_streams, int64_t n_per_stream, int32_t step_major,
                         double *out, double *kernel_ms);

/* Device-resident parameters and results of ofg_price_device. Each input is
 * a device array of n doubles (option i at [i]) that replaces the
 * corresponding ofg_option field, or NULL to keep the field; the outputs are
 * device arrays the call writes. */
typedef struct ofg_device_io {
    const double *spot, *strike, *expiry, *rate, *vol;   /* vol: GBM sigma */
    const double *v0, *kappa, *theta, *xi, *rho;         /* Heston */
    double *price;      /* n values; required */
    double *stderr_;    /* n values; may be NULL */
    /* n x 8, row-major: d price / d (spot, vol | v0, rate, expiry, kappa,
       theta, xi, rho) with greek_mode OFG_GREEKS_ADJOINT or FORWARD (the
       pathwise estimator of ofg_price_greeks; GBM rows fill the first four);
       may be NULL, ignored with OFG_GREEKS_NONE */
    double *grads;
} ofg_device_io;

/* ofg_price with the market and model parameters read from, and the prices
 * and gradients written to, device memory: the zero-copy form of
 * PricingEngine::price / price_with_greeks_aad (mc_engine.rs:875-987) for
 * callers whose data already live on the GPU (the torch autograd path).
 * opts supplies each contract (payoff, type, barrier, observation times,
 * dividend yield, paths, steps, seed) and the fields io leaves NULL; the
 * descriptors are formed and validated on the device (the reference's
 * rules: strike > 0, volatility > 0 under GBM, Heston::validate; every
 * maturity must be > 0 here) and a violation returns OFG_INVALID_INPUT
 * naming the offending option, with no output written. With device
 * maturities an Asian option observes at every step (obs_times NULL, as
 * its default schedule).
 * greek_mode OFG_GREEKS_NONE, ADJOINT or FORWARD; control variates,
 * OFG_RNG_EXTERNAL (host normals) and FD_CRN are refused: use ofg_price.
 * The call reads its inputs on the library's own stream: work on other
 * streams that produces them must be complete; the outputs are complete
 * when it returns. Results equal ofg_price's on the same values to the
 * rounding of the derived constants (exp, log and sqrt of the device). */
int32_t ofg_price_device(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                         const ofg_engine_cfg *cfg, const ofg_device_io *io);

/* Limits of the fixed device region ofg_price_noalloc prices in. */
#define OFG_NOALLOC_MAX_OPTIONS 4096
#define OFG_NOALLOC_MAX_WORK_ITEMS 262144
#define OFG_NOALLOC_MAX_OBS_WEIGHTS 1048576

/* ofg_price without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The staged
 * contracts, the work queue and its counters, the partial sums and the
 * results live in one fixed region of the library per GPU (33 MiB, resolved
 * by ofg_create, held until the process ends); the call makes no device
 * allocation or free. Limits: at most OFG_NOALLOC_MAX_OPTIONS options and
 * OFG_NOALLOC_MAX_OBS_WEIGHTS observation weights (num_steps + 1 per Asian
 * option, 1 per other option); the work-item size is doubled until the queue
 * holds at most OFG_NOALLOC_MAX_WORK_ITEMS items. Beyond them, or with
 * cfg->rng_mode OFG_RNG_EXTERNAL or greek_mode other than OFG_GREEKS_NONE,
 * the call returns OFG_INVALID_INPUT before anything runs. The region keeps
 * the last call's contents; calls on one GPU share it, so the library
 * serializes them (a call from another context on the same GPU waits).
 * Results equal ofg_price's for the same launch shape. out[n]. */
int32_t ofg_price_noalloc(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                          const ofg_engine_cfg *cfg, ofg_result *out);

/* Pathwise adjoint Greeks. Replaces mc_european_pathwise_aad
 * (src/engines/monte_carlo/mc_aad.rs:75-219) and heston_price_delta_aad
 * (mc_aad.rs:279-342), generalised to Asian/barrier payoffs and the full
 * Heston parameter gradient. Equivalent to ofg_price with
 * greek_mode = OFG_GREEKS_ADJOINT. */
int32_t ofg_price_greeks(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                         const ofg_engine_cfg *cfg, ofg_result *out);

/* Pathwise and likelihood-ratio Greeks of European GBM vanillas from one
 * exact-terminal normal per sample, n options in one persistent launch.
 * Replaces MonteCarloGreeksEngine::estimate_pathwise and
 * ::estimate_likelihood_ratio (src/engines/monte_carlo/mc_greeks.rs:93-129;
 * estimators :131-284). opts[i] must be GBM vanillas (num_steps is ignored:
 * the estimator steps once over T; div_yield is the market's effective
 * yield); cfg supplies antithetic and rng_mode (control_variate and
 * greek_mode are ignored); spot_bump_rel is the relative spot bump of the
 * pathwise gamma (h = max(spot * rel, 1e-6)). Validation and messages follow
 * run_estimators; expiry 0 gives all-zero estimates. out[n]. */
int32_t ofg_mc_greeks_estimate(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                               const ofg_engine_cfg *cfg, double spot_bump_rel,
                               ofg_mc_greeks *out);

/* Batched pricing for repeated replay: builds the work queue, uploads the
 * contracts once and captures [queue reset, persistent pricing kernel,
 * segmented reduce] into a hipGraph. Replaces a loop of
 * PricingEngine::price calls (mc_engine.rs:875). */
ofg_batch *ofg_batch_create(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                            const ofg_engine_cfg *cfg, int32_t use_graph);
/* Replays the batch; kernel_ms (may be NULL) receives its device time. */
int32_t ofg_batch_run(ofg_batch *batch, double *kernel_ms);
/* Copies the last run's results (n entries). */
int32_t ofg_batch_results(ofg_batch *batch, ofg_result *out);
/* Device time of the persistent pricing kernel alone in the last run. */
double ofg_batch_pricing_kernel_ms(ofg_batch *batch);
void ofg_batch_destroy(ofg_batch *batch);

/* Autotuner: times candidate launch shapes (threads per block, work-item
 * size) on a scaled-down copy of the given workload, stores the winner in
 * the context (and in cache_path when non-NULL) keyed by the workload class.
 * best_block / best_item_steps may be NULL. */
int32_t ofg_autotune(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                     const ofg_engine_cfg *cfg, const char *cache_path,
                     int32_t *best_block, int32_t *best_item_steps);

/* Least-squares Monte Carlo: prices n options in one batched sequence of
 * launches (path simulation, then one launch per exercise-date round).
 * Replaces LongstaffSchwartzEngine's PricingEngine::price for VanillaOption
 * (src/engines/lsm/longstaff_schwartz.rs:448-670), BermudanOption and
 * price_bermudan_with_boundary (:291-417, :672-681), and BarrierOption
 * (:683-735). cfg supplies rng_mode (REFERENCE: the engine's own single
 * xoshiro256++ stream, drawn path after path exactly as the reference does;
 * EXTERNAL: normals1 (and normals2, Heston) per option [step][path], options
 * concatenated; FAST: per-path xoroshiro128+ Box-Muller streams);
 * antithetic, control_variate and greek_mode are ignored (the engine has no
 * variance reduction). out[n]: price, stderr and the diagnostics num_paths,
 * num_steps, vol (NaN where the reference records no vol). boundary (may be
 * NULL) holds num_steps slots per option, options in call order; for a
 * BermudanOption it receives the exercise boundary in chronological order and
 * boundary_len[i] its length (0 for other instruments); boundary_len may be
 * NULL only with boundary. */
int32_t ofg_lsm_price(ofg_ctx *ctx, const ofg_lsm_option *opts, size_t n,
                      const ofg_engine_cfg *cfg, ofg_result *out,
                      ofg_lsm_boundary *boundary, int64_t *boundary_len);

/* Quasi-Monte Carlo European vanillas under GBM on Sobol points: replaces
 * mc_european_qmc / mc_european_qmc_with_seed
 * (src/engines/monte_carlo/mc_qmc.rs:29-144), n options in one launch.
 * opts[i] must be GBM vanillas (the reference takes a European
 * VanillaOption); num_paths = Sobol points, num_steps = dimensions (at most
 * 21201), seed = the Sobol seed, div_yield = the market's effective yield.
 * Each path's point is the reference sequence's point of the same index, so
 * the draws are the reference's. As the reference, degenerate inputs return
 * results, not errors: price NaN and has_stderr 0 for zero paths or steps or
 * a vol that is not finite and positive; the payoff at the spot with
 * stderr 0 for expiry <= 0; diagnostics (num_paths, num_steps, vol) NaN
 * where the reference records none. out[n]. */
int32_t ofg_qmc_price(ofg_ctx *ctx, const ofg_option *opts, size_t n, ofg_result *out);

/* Worst-of autocallables and Phoenix notes on correlated GBM paths: replaces
 * price_autocallable / price_phoenix_autocallable and, with with_greeks,
 * autocallable_sensitivities / phoenix_autocallable_sensitivities
 * (src/pricing/autocallable.rs:52-231, bump-and-reprice :366-418), n contracts
 * in one launch, every bump a scenario on the base price's draws.
 * cfg->rng_mode: EXTERNAL (normals1 per contract [step][underlying][path],
 * contracts concatenated) or FAST (per-path streams from the reference's
 * seed 94210); the reference's single stream cannot be split across paths
 * because a called path stops drawing. As price_autocallable, an invalid
 * contract gives price NaN and has_stderr 0 (ofg_last_error holds the
 * reference's message); greeks: delta = sum of deltas, vega, rho = cega.
 * sens (may be NULL): the per-underlying sensitivities. out[n], sens[n]. */
int32_t ofg_autocall_price(ofg_ctx *ctx, const ofg_autocall *contracts, size_t n,
                           const ofg_engine_cfg *cfg, int32_t with_greeks, ofg_result *out,
                           ofg_autocall_sens *sens);

/* Basket options under the Gaussian copula on one exact terminal step:
 * replaces price_basket_mc and, with with_sens, basket_sensitivities
 * (src/pricing/basket.rs:99-160, 456-597; simulation :599-683),
 * price_outperformance_basket_mc (:288-366) and price_quanto_basket_mc
 * (:368-445) (sensitivities only for BasketOption), n contracts in
 * one launch, every bump a scenario on the base price's draws. A correlation
 * matrix that is valid but not PSD is projected to the nearest correlation
 * matrix (Higham 2002 alternating projections, as
 * validate_or_repair_correlation_matrix, src/math/correlation.rs:340-358).
 * cfg->rng_mode: REFERENCE (the engine's single xoshiro256++ stream from
 * `seed`, split across paths by jump-ahead: the reference's draws),
 * EXTERNAL (normals1 per basket [asset][path], baskets concatenated) or
 * FAST. As price_basket_mc, an invalid basket gives price NaN and
 * has_stderr 0 (ofg_last_error holds the reference's message); diagnostics
 * num_paths in out, rho in sens. out[n], sens[n] (sens may be NULL). */
int32_t ofg_basket_price(ofg_ctx *ctx, const ofg_basket *baskets, size_t n,
                         const ofg_engine_cfg *cfg, int32_t with_sens, ofg_result *out,
                         ofg_basket_sens *sens);

/* Nearest correlation matrix (nearest_correlation_matrix_higham,
 * src/math/correlation.rs:271-335): Higham's alternating projections with
 * the reference's defaults (tol 1e-10, 100 iterations) when min_eig < -1e-10,
 * else a copy. Writes out (n x n), *was_repaired. Host-only. */
int32_t ofg_repair_correlation(const double *corr, int32_t n, double *out,
                               int32_t *was_repaired);

/* CorrelationStressScenario (src/math/correlation.rs:49-60). value: the
 * factor, shift, floor, cap or pair value; i, j: OverridePair only. */
typedef enum ofg_corr_stress_kind {
    OFG_STRESS_SCALE_OFF_DIAGONAL = 0,
    OFG_STRESS_ADDITIVE_SHIFT = 1,
    OFG_STRESS_FLOOR_OFF_DIAGONAL = 2,
    OFG_STRESS_CAP_OFF_DIAGONAL = 3,
    OFG_STRESS_OVERRIDE_PAIR = 4
} ofg_corr_stress_kind;

typedef struct ofg_corr_stress {
    int32_t kind;             /* ofg_corr_stress_kind */
    int32_t i, j;
    int32_t pad;
    double value;
} ofg_corr_stress;

/* stressed_correlation_matrix (src/pricing/basket.rs:447-454, i.e.
 * apply_correlation_stress with repair, math/correlation.rs:442-543): the
 * scenarios in order, off-diagonal entries clamped to +-0.999999 and the
 * diagonal reset after each, then the Higham repair when the minimum
 * eigenvalue is below -1e-10, then validation. Writes out (n x n). Host
 * only; errors carry the reference's messages. */
int32_t ofg_stress_correlation(const double *corr, int32_t n, const ofg_corr_stress *scenarios,
                               int32_t n_scenarios, double *out);

/* Structured-product language: parse_and_compile (src/dsl/mod.rs:24-28,
 * lexer.rs, parser.rs, compiler.rs). On failure *out is NULL, the status is
 * OFG_INVALID_INPUT and ofg_last_error holds the DslError text. */
int32_t ofg_dsl_compile(const char *source, ofg_dsl_product **out);
void ofg_dsl_free(ofg_dsl_product *product);
int32_t ofg_dsl_get_info(const ofg_dsl_product *product, ofg_dsl_info *info);
/* Observation dates of schedule `schedule` (ScheduleFreq::generate_dates,
 * src/dsl/ast.rs:64-80): writes min(count, cap) dates, *count = all. */
int32_t ofg_dsl_schedule_dates(const ofg_dsl_product *product, int32_t schedule,
                               double *dates, int32_t cap, int32_t *count);

/* DslMonteCarloEngine::price_multi_asset (src/dsl/engine.rs:255-353) of one
 * product under n_markets markets in one launch, all on the same draws (the
 * reference reprices every bump of greeks_multi_asset /
 * extended_greeks_multi_asset / cross_greeks_multi_asset, :785-1000, with the
 * same seed): out[i] is the price under markets[i]. Correlated paths with
 * the reference's steppers; the compiled product runs per path on the GPU.
 * cfg->rng_mode: REFERENCE (the engine's single xoshiro256++ stream split
 * across paths by jump-ahead; the default build's draws), EXTERNAL
 * (normals1 [step][asset][path], shared by all markets) or FAST. Errors are
 * the reference's (InvalidInput for the market and engine, NumericalError
 * "eval error: ..." for the product); the markets must share their asset
 * count. Diagnostics num_paths, num_steps. */
int32_t ofg_dsl_price(ofg_ctx *ctx, const ofg_dsl_product *product, const ofg_dsl_market *markets,
                      size_t n_markets, const ofg_dsl_engine *engine, const ofg_engine_cfg *cfg,
                      ofg_result *out);

/* Compiles the product's programs, for markets of n_assets assets, into a
 * native gfx950 evaluation kernel at run time (hipRTC) and attaches it to the
 * context: later ofg_dsl_price calls of this product with that asset count
 * run it in place of the predicated-program interpreter (split form, which
 * the path and market counts of a call select as before), with prices equal
 * bit for bit. The reference's counterpart is the cranelift JIT of the
 * product bytecode (src/dsl/jit.rs:168-200, JitCompiledProgram::compile,
 * feature "jit"). Compiling takes a few hundred milliseconds; a product is
 * compiled once per context and asset count (later calls return at once).
 * compile_ms (may be NULL) receives the compile and load time.
 * OFG_DEVICE_ERROR with hipRTC's log when compiling fails. */
int32_t ofg_dsl_jit_compile(ofg_ctx *ctx, const ofg_dsl_product *product, int32_t n_assets,
                            double *compile_ms);

/* The HIP source that ofg_dsl_jit_compile compiles for this product and
 * asset count (no device needed): *len receives its length in bytes; when
 * buf is not NULL, up to cap - 1 bytes and a terminating NUL are copied. No
 * counterpart in the reference (whose JIT emits machine code directly); for
 * inspection and tests. */
int32_t ofg_dsl_jit_source(const ofg_dsl_product *product, int32_t n_assets, char *buf, size_t cap,
                           size_t *len);

/* Native evaluation kernels attached to the context and how many times
 * ofg_dsl_price has launched one. No counterpart in the reference; for tests
 * and measurement. Either output may be NULL. */
int32_t ofg_dsl_jit_stats(ofg_ctx *ctx, int64_t *kernels, int64_t *launches);

/* Device time of the context's last ofg_lsm_price call, split into path
 * simulation (seed and path kernels) and the backward rounds, and the number
 * of round launches. No counterpart in the reference; for measurement.
 * Any output pointer may be NULL. */
int32_t ofg_lsm_last_timing(ofg_ctx *ctx, double *paths_ms, double *rounds_ms,
                            int32_t *rounds);

/* Positions a xoshiro256++ state k draws ahead (state <- T^k state, T the
 * generator's GF(2)-linear transition) in O(log k) polynomial steps. No
 * counterpart in the reference (whose engine steps one draw at a time); it
 * is how ofg_lsm_price's reference draws split the engine's single stream
 * across paths, exported for callers that partition a stream the same way.
 * Host-only, needs no GPU. */
int32_t ofg_stream_jump(const uint64_t state[4], uint64_t k, uint64_t out[4]);

/* Fault injection for callers' error-path tests. No counterpart in the
 * reference. The after-th device allocation made by any context from now on
 * (1 = the next one) fails as an out-of-memory hipMalloc would: the entry
 * point that made it returns OFG_DEVICE_ERROR with a message naming the
 * allocation, and the context stays usable (the buffer is allocated again
 * by the next call). 0 disarms. Process-wide; needs no GPU. */
void ofg_inject_alloc_failure(int32_t after);

#ifdef __cplusplus
}
#endif
#endif /* OPENFERRIC_AMD_H */
