// Device layouts and launch interface of the multi-asset path kernels
// (ofg_multi.hip): worst-of autocallables and Phoenix notes on correlated GBM
// paths (src/pricing/autocallable.rs). Shared with ofg_multi_api.cpp.
//
// A scenario is one contract under one set of market inputs (the base price,
// or one bump of the reference's bump-and-reprice sensitivities); scenarios
// of a contract share its draws, as the reference reprices every bump with
// the same seed. Up to MA_MAX underlyings are held in registers per path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ofg {

constexpr int MA_MAX = 8;
constexpr int MA_BLOCK = 256;
constexpr int MA_MAX_TILES = 1024;

struct alignas(16) AcDev {
    int32_t n;             // underlyings of the contract
    int32_t phoenix;
    int32_t memory;
    int32_t n_obs;         // observation schedule entries
    int32_t steps;
    int32_t obs_offset;    // into the AcObs array
    int32_t tile_first;
    int32_t n_tiles;
    int32_t ppt;           // paths per thread
    int32_t pad;
    int64_t paths;
    int64_t normals_offset;   // external normals of the contract, [step][asset][path]
    uint64_t seed;            // fast draws: stream_seed(seed, path)
    double s0[MA_MAX], drift[MA_MAX], vol_dt[MA_MAX];
    // fast mode: each underlying's log performance is carried as
    // Y = 256 log2(S / s0), advanced by drift2 + vol_dt2 c (the increments
    // scaled by 256/ln2) and held at yfloor = 256 log2(1e-12 / s0), the
    // spot floor of the step; the barriers are compared as 256 log2(b)
    double yfloor[MA_MAX];
    double drift2[MA_MAX], vol_dt2[MA_MAX];
    double chol[MA_MAX * (MA_MAX + 1) / 2];   // lower factor, row i at i(i+1)/2
    double notional, maturity, autocall_barrier, coupon_rate, coupon_barrier;
    double ki_barrier, ki_strike;
    double autocall_y, coupon_y, ki_y;        // fast mode: 256 log2 of the barriers
    double df_maturity;       // exp(-r maturity)
    double redeem_plain;      // notional (1 + c maturity): standard note, no knock-in
    double shift;             // payoffs are summed as pv - shift (variance accuracy)
};

// One observation date: its path step, time, exp(-r t) and, for a standard
// note, the called payoff exp(-r t) N (1 + c t).
struct AcObs {
    int32_t step;
    int32_t pad;
    double time, df, call_pv;
};

struct AcPartial {
    double sum, sum_sq;
};

struct AcOut {
    double price, stderr_;
};

struct AcArgs {
    const AcDev *scen;
    const int32_t *tile_scen;
    const AcObs *obs;
    AcPartial *partials;
    AcOut *out;
    const double *normals;
    int32_t n_tiles;
    int32_t n_scen;
    int32_t rng_mode;
    int32_t pad;
};

hipError_t launch_autocall(const AcArgs &a, hipStream_t stream);

// ---- autocallable sensitivities fused per path ----
// The 1 + 2n + 2 (+ 2) scenarios of a contract's bump-and-reprice
// sensitivities on each path's shared draws: the spot bumps reuse the base
// spots of the other underlyings and the base factors (a spot bump scales
// one underlying's spot and leaves every exponent unchanged), the vol bumps
// reuse the base correlated normals, the correlation bumps the independent
// draws. Scenario order: 0 base; 1 / 2 vol up / down; 3 / 4 correlation up
// / down (when present); 5 + 2k / 6 + 2k spot up / down of underlying k.
constexpr int ACG_MAXS = 5 + 2 * MA_MAX;

struct alignas(16) AcGreeksDev {
    int32_t n, phoenix, memory, n_obs;
    int32_t steps, obs_offset, tile_first, n_tiles;
    int32_t ppt, n_scen, has_cega, out_first;
    int64_t paths, normals_offset;
    uint64_t seed;
    double s0[MA_MAX], s0_up[MA_MAX], s0_dn[MA_MAX];
    // fast mode: the log-performance floors of the three spots (AcDev)
    double yfloor[MA_MAX], yfloor_up[MA_MAX], yfloor_dn[MA_MAX];
    double drift[3][MA_MAX], vol_dt[3][MA_MAX];     // base, vol up, vol down
    double drift2[3][MA_MAX], vol_dt2[3][MA_MAX];   // fast mode: scaled by 256/ln2
    double chol[3][MA_MAX * (MA_MAX + 1) / 2];     // base, correlation up, down
    double notional, autocall_barrier, coupon_rate, coupon_barrier;
    double ki_barrier, ki_strike, df_maturity, redeem_plain, shift;
    double autocall_y, coupon_y, ki_y;
};

struct AcgArgs {
    const AcGreeksDev *ct;
    const int32_t *tile_ct;
    const AcObs *obs;
    AcPartial *partials;     // [tile][ACG_MAXS]
    AcOut *out;
    const double *normals;
    int32_t n_tiles;
    int32_t n_ct;
    int32_t rng_mode;
    int32_t max_n;
};

hipError_t launch_autocall_greeks(const AcgArgs &a, hipStream_t stream);

// ---- basket options (src/pricing/basket.rs:599-757, Gaussian and Student-t copulas) ----
constexpr int BK_MAX = 16;
constexpr int BK_FMAX = 4;   // factors of a factor-model basket
enum BasketType : int32_t { BK_AVERAGE = 0, BK_BEST_OF = 1, BK_WORST_OF = 2 };

enum BasketPayoff : int32_t { BK_TERMINAL = 0, BK_OUTPERFORMANCE = 1 };

struct alignas(16) BkDev {
    int32_t n;             // assets
    int32_t type;          // BasketType
    int32_t is_call;
    int32_t tile_first;
    int32_t n_tiles;
    int32_t ppt;
    int32_t payoff;        // BasketPayoff
    int32_t leader;        // outperformance: leader index
    double scale;          // payoff multiplier (quanto: the fixed FX rate)
    int64_t paths;
    int64_t normals_offset;   // external normals, [asset][path]
    int64_t seed_offset;      // reference-stream start states, [path][word]
    uint64_t seed;            // fast draws: stream_seed(seed, path)
    double strike, discount;
    double s0[BK_MAX], drift[BK_MAX], vol_t[BK_MAX], w[BK_MAX];   // w: weights or lagger weights
    double inv_s0[BK_MAX];                     // fast mode: performances s * (1 / s0)
    double drift2[BK_MAX], vol_t2[BK_MAX];     // fast mode: scaled by 256/ln2
    double chol[BK_MAX * (BK_MAX + 1) / 2];   // lower factor, row i at i(i+1)/2
    // factor model (n_factors > 0): normals from loadings instead of chol
    int32_t n_factors;
    int32_t one_factor;        // OneFactor: fma(beta, m, idio eps)
    double load[BK_MAX * BK_FMAX];   // [asset][factor]; OneFactor: the clamped beta
    double idio[BK_MAX];
    // Student-t copula (t_dof > 0): t_dof more draws per path after the
    // asset and factor draws; t_nu = dof, t_a = dof / 2 and t_lbeta =
    // lnG(t_a + 1/2) - lnG(t_a) - lnG(1/2), the log-gamma terms of the
    // incomplete beta's prefactor, evaluated on the host
    int32_t t_dof;
    int32_t pad_t;
    double t_nu, t_a, t_lbeta;
};

struct BkArgs {
    const BkDev *scen;
    const int32_t *tile_scen;
    AcPartial *partials;
    AcOut *out;
    const double *normals;
    const uint64_t *seeds;
    int32_t n_tiles;
    int32_t n_scen;
    int32_t rng_mode;
    int32_t any_t;    // some scenario uses the Student-t copula
};

hipError_t launch_basket(const BkArgs &a, hipStream_t stream);

// ---- basket_sensitivities fused per path ----
// All scenarios of one basket's sensitivities (base, spot up / down per
// asset, vol up / down, correlation up / down) evaluated on each path's
// shared draws: the spot bumps reuse the base terminal factors (a spot bump
// scales S_i(T) = S_i(0) exp(...) and leaves the exponent unchanged), the
// vol bumps reuse the base correlated normals, the correlation bumps reuse
// the independent draws. Scenario order: 0 base; 1 + 2k / 2 + 2k spot up /
// down of asset k; 1 + 2n / 2 + 2n vol up / down; 3 + 2n / 4 + 2n
// correlation up / down (when present).
constexpr int BKG_MAXS = 1 + 2 * BK_MAX + 4;

struct alignas(16) BkGreeksDev {
    int32_t n, type, is_call, has_cega;
    int32_t tile_first, n_tiles, ppt, n_scen;
    int32_t out_first;     // first result slot
    int32_t pad;
    int64_t paths, normals_offset, seed_offset;
    uint64_t seed;
    double strike, discount;
    double s0[BK_MAX], s0_up[BK_MAX], s0_dn[BK_MAX], w[BK_MAX];
    // fast mode: performances s * (1 / s0), the reciprocals of the three
    double inv_s0[BK_MAX], inv_up[BK_MAX], inv_dn[BK_MAX];
    double drift[3][BK_MAX], vol_t[3][BK_MAX];     // base, vol up, vol down
    double drift2[3][BK_MAX], vol_t2[3][BK_MAX];   // fast mode: scaled by 256/ln2
    double chol[3][BK_MAX * (BK_MAX + 1) / 2];    // base, correlation up, down
};

struct BkgArgs {
    const BkGreeksDev *bk;
    const int32_t *tile_bk;
    AcPartial *partials;     // [tile][BKG_MAXS]
    AcOut *out;
    const double *normals;
    const uint64_t *seeds;
    int32_t n_tiles;
    int32_t n_bk;
    int32_t rng_mode;
    int32_t max_n;           // largest asset count of the launch
};

hipError_t launch_basket_greeks(const BkgArgs &a, hipStream_t stream);

}  // namespace ofg
