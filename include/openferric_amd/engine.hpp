// C++ mirror of OpenFerric's pricing contract over the openferric_amd C ABI.
//
// Same names, argument meaning and error behaviour as the reference's Rust
// types, so the GPU engine is a drop-in beside MonteCarloPricingEngine:
//   Greeks, PricingResult, PricingError          src/core/engine.rs:15-315
//   OptionType, BarrierSpec, AsianSpec, ...      src/core/types.rs:13-122
//   VanillaOption / AsianOption / BarrierOption / BermudanOption
//                                                src/instruments/*.rs
//   LongstaffSchwartzEngine                      src/engines/lsm/longstaff_schwartz.rs
//   mc_european_qmc(_with_seed)                  src/engines/monte_carlo/mc_qmc.rs
//   Market (flat vol)                            src/market/market.rs:202-414
//   Heston                                       src/models/stochastic.rs:37-72
//   Autocallable / PhoenixAutocallable, price_autocallable, ...
//                                                src/pricing/autocallable.rs
//   BasketOption, price_basket_mc(_with_copula), basket_sensitivities, ...
//                                                src/pricing/basket.rs
//   dsl::parse_and_compile, DslMonteCarloEngine, MultiAssetMarket
//                                                src/dsl/{mod,engine,market}.rs
//   PricingEngine<I>::price / price_with_greeks_aad (engine.rs:37-52)
// Errors are C++ exceptions carrying the PricingError variant (Rust returns
// Result<_, PricingError>); every call goes through libopenferric_amd.so.
#pragma once

#include "../openferric_amd.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <thread>
#include <vector>

namespace openferric_amd {

// ---------------------------------------------------------------- errors --
class PricingError : public std::runtime_error {
public:
    enum class Kind { InvalidInput, ConvergenceFailure, MarketDataMissing, NumericalError, Device };
    PricingError(Kind k, const std::string &msg) : std::runtime_error(prefix(k) + msg), kind_(k) {}
    Kind kind() const { return kind_; }

private:
    static std::string prefix(Kind k) {
        switch (k) {
        case Kind::InvalidInput: return "invalid input: ";
        case Kind::ConvergenceFailure: return "convergence failure: ";
        case Kind::MarketDataMissing: return "market data missing: ";
        case Kind::NumericalError: return "numerical error: ";
        default: return "device error: ";
        }
    }
    Kind kind_;
};

inline void check(int32_t rc) {
    if (rc == OFG_OK) return;
    const std::string msg = ofg_last_error();
    switch (rc) {
    case OFG_INVALID_INPUT: throw PricingError(PricingError::Kind::InvalidInput, msg);
    case OFG_CONVERGENCE_FAILURE: throw PricingError(PricingError::Kind::ConvergenceFailure, msg);
    case OFG_MARKET_DATA_MISSING: throw PricingError(PricingError::Kind::MarketDataMissing, msg);
    case OFG_NUMERICAL_ERROR: throw PricingError(PricingError::Kind::NumericalError, msg);
    default: throw PricingError(PricingError::Kind::Device, msg);
    }
}

// ------------------------------------------------------------ contract --
enum class OptionType { Call, Put };
enum class ExerciseStyle { European, American, Bermudan };
enum class BarrierDirection { Up, Down };
enum class BarrierStyle { In, Out };
enum class Averaging { Arithmetic, Geometric };
enum class StrikeType { Fixed, Floating };

struct Greeks {
    double delta = 0, gamma = 0, vega = 0, theta = 0, rho = 0;
};

// Diagnostics: the reference's 8-slot key/value store (engine.rs:191-274).
class Diagnostics {
public:
    static constexpr size_t CAPACITY = 8;
    std::optional<double> insert(const std::string &key, double v) {
        auto it = map_.find(key);
        if (it != map_.end()) {
            const double prev = it->second;
            it->second = v;
            return prev;
        }
        if (map_.size() >= CAPACITY)
            throw std::length_error("diagnostics capacity exceeded (8)");
        map_[key] = v;
        return std::nullopt;
    }
    std::optional<double> get(const std::string &key) const {
        auto it = map_.find(key);
        if (it == map_.end()) return std::nullopt;
        return it->second;
    }
    size_t len() const { return map_.size(); }
    bool is_empty() const { return map_.empty(); }

private:
    std::map<std::string, double> map_;
};

struct PricingResult {
    double price = 0;
    std::optional<double> stderr_;
    std::optional<Greeks> greeks;
    Diagnostics diagnostics;
    // dPrice/d{S0, v0, kappa, theta, xi, rho} for Heston adjoint runs.
    std::optional<std::vector<double>> heston_gradient;
    double kernel_ms = 0;
};

struct VanillaOption {
    OptionType option_type = OptionType::Call;
    double strike = 0, expiry = 0;
    ExerciseStyle exercise = ExerciseStyle::European;
    std::vector<double> bermudan_dates;   // ExerciseStyle::Bermudan { dates }
    static VanillaOption european_call(double k, double t) { return {OptionType::Call, k, t}; }
    static VanillaOption european_put(double k, double t) { return {OptionType::Put, k, t}; }
    static VanillaOption american_call(double k, double t) {
        return {OptionType::Call, k, t, ExerciseStyle::American, {}};
    }
    static VanillaOption american_put(double k, double t) {
        return {OptionType::Put, k, t, ExerciseStyle::American, {}};
    }
};

// BermudanOption (instruments/bermudan.rs:18-144): exercise dates with a
// strike per date; validated by the engine with the reference's messages.
struct BermudanOption {
    OptionType option_type = OptionType::Put;
    double expiry = 0;
    std::vector<double> exercise_dates, strike_schedule;
    static BermudanOption with_constant_strike(OptionType t, double k, double expiry,
                                               std::vector<double> dates) {
        std::vector<double> ks(dates.size(), k);
        return BermudanOption{t, expiry, std::move(dates), std::move(ks)};
    }
};

// ExerciseBoundaryPoint / BermudanLsmOutput (longstaff_schwartz.rs:42-63).
struct ExerciseBoundaryPoint {
    double time = 0, strike = 0;
    std::optional<double> boundary_spot;
    int64_t itm_paths = 0, exercised_paths = 0;
};

struct AsianSpec {
    Averaging averaging = Averaging::Arithmetic;
    StrikeType strike_type = StrikeType::Fixed;
    std::vector<double> observation_times;
};

struct AsianOption {
    OptionType option_type = OptionType::Call;
    double strike = 0, expiry = 0;
    AsianSpec asian;
};

struct BarrierSpec {
    BarrierDirection direction = BarrierDirection::Up;
    BarrierStyle style = BarrierStyle::Out;
    double level = 0, rebate = 0;
};

struct BarrierOption {
    OptionType option_type = OptionType::Call;
    double strike = 0, expiry = 0;
    BarrierSpec barrier;
};

// Market with a flat vol source; builder validation as MarketBuilder::build.
struct Market {
    double spot = 0, rate = 0, dividend_yield = 0, flat_vol = 0;
    double vol_for(double, double) const { return flat_vol; }
    // market.rs:275-282 with no discrete dividends (dividends.rs:151-217):
    // r - ln(F/S)/T, F = S exp((r - q) T); equal to q up to rounding.
    double effective_dividend_yield(double maturity) const {
        if (maturity <= 0.0 || spot <= 0.0) return dividend_yield;
        const double forward = spot * std::exp((rate - dividend_yield) * maturity);
        const double ratio = std::fmax(forward / spot, 1.0e-16);
        return rate - std::log(ratio) / maturity;
    }

    class Builder {
    public:
        Builder &spot(double s) { spot_ = s; return *this; }
        Builder &rate(double r) { rate_ = r; return *this; }
        Builder &dividend_yield(double q) { q_ = q; return *this; }
        Builder &flat_vol(double v) { vol_ = v; return *this; }
        Market build() const {
            if (!spot_) throw PricingError(PricingError::Kind::InvalidInput, "market spot is required");
            if (*spot_ <= 0.0) throw PricingError(PricingError::Kind::InvalidInput, "market spot must be > 0");
            if (!vol_)
                throw PricingError(PricingError::Kind::InvalidInput,
                                   "either market flat_vol or vol_surface is required");
            if (*vol_ <= 0.0) throw PricingError(PricingError::Kind::InvalidInput, "market flat_vol must be > 0");
            return Market{*spot_, rate_, q_, *vol_};
        }

    private:
        std::optional<double> spot_, vol_;
        double rate_ = 0, q_ = 0;
    };
    static Builder builder() { return Builder{}; }
};

struct Heston {
    double mu, kappa, theta, xi, rho, v0;
    bool validate() const {
        return kappa > 0.0 && theta >= 0.0 && xi >= 0.0 && v0 >= 0.0 && rho > -1.0 && rho < 1.0;
    }
};

enum class VarianceReduction { None, Antithetic, ControlVariate };
// Differentiation of price_with_greeks_aad: checkpointed reverse sweep, or
// the same pathwise estimator in one forward sweep (GBM; Heston uses the
// adjoint).
enum class GreekMethod { Adjoint = OFG_GREEKS_ADJOINT, Forward = OFG_GREEKS_FORWARD };
enum class RngMode { Reference = OFG_RNG_REFERENCE, Fast = OFG_RNG_FAST, Pcg64 = OFG_RNG_PCG64 };
// FastRngKind (fast_rng.rs:20-27). Xoshiro256PlusPlus and Pcg64 draw the
// reference's per-path streams; ThreadRng is served by xoshiro256++ streams on
// a fresh random base seed (non-reproducible, as there); StdRng (rand's
// ChaCha12) has no GPU stream and is refused.
enum class FastRngKind { Xoshiro256PlusPlus, Pcg64, ThreadRng, StdRng };

namespace detail {
inline RngMode kind_mode(FastRngKind k) {
    if (k == FastRngKind::StdRng)
        throw PricingError(PricingError::Kind::InvalidInput,
                           "FastRngKind::StdRng streams are not available on the GPU");
    return k == FastRngKind::Pcg64 ? RngMode::Pcg64 : RngMode::Reference;
}
inline uint64_t random_base_seed() {
    std::random_device rd;
    return ((uint64_t)rd() << 32) ^ (uint64_t)rd();
}
}  // namespace detail

// ------------------------------------------------------------- context --
// One process-wide context per device (the engine's OnceLock analogue).
inline ofg_ctx *device_context(int device = 0) {
    static std::mutex mu;
    static std::map<int, ofg_ctx *> ctxs;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second;
    ofg_ctx *c = ofg_create(device);
    if (!c) throw PricingError(PricingError::Kind::Device, ofg_last_error());
    ctxs[device] = c;
    return c;
}

namespace detail {
inline ofg_option base(OptionType t, double k, double T, const Market &m, int64_t paths,
                       int64_t steps, uint64_t seed) {
    ofg_option o{};
    o.model = OFG_MODEL_GBM;
    o.option_type = t == OptionType::Call ? OFG_CALL : OFG_PUT;
    o.spot = m.spot;
    o.strike = k;
    o.expiry = T;
    o.rate = m.rate;
    o.div_yield = m.dividend_yield;
    o.num_paths = paths;
    o.num_steps = steps;
    o.seed = seed;
    return o;
}
inline ofg_option to_ofg(const VanillaOption &i, const Market &m, int64_t p, int64_t s, uint64_t seed) {
    ofg_option o = base(i.option_type, i.strike, i.expiry, m, p, s, seed);
    o.payoff = OFG_PAYOFF_VANILLA;
    o.vol = m.vol_for(i.strike, i.expiry);
    return o;
}
inline ofg_option to_ofg(const AsianOption &i, const Market &m, int64_t p, int64_t s, uint64_t seed) {
    ofg_option o = base(i.option_type, i.strike, i.expiry, m, p, s, seed);
    o.payoff = OFG_PAYOFF_ASIAN;
    o.averaging = i.asian.averaging == Averaging::Arithmetic ? OFG_ARITHMETIC : OFG_GEOMETRIC;
    o.strike_type = i.asian.strike_type == StrikeType::Fixed ? OFG_FIXED : OFG_FLOATING;
    o.n_obs = (int32_t)i.asian.observation_times.size();
    o.obs_times = i.asian.observation_times.data();
    const double ref = i.asian.strike_type == StrikeType::Fixed ? i.strike : m.spot;
    o.vol = m.vol_for(std::fmax(ref, 1e-12), i.expiry);
    return o;
}
inline ofg_option to_ofg(const BarrierOption &i, const Market &m, int64_t p, int64_t s, uint64_t seed) {
    ofg_option o = base(i.option_type, i.strike, i.expiry, m, p, s, seed);
    o.payoff = OFG_PAYOFF_BARRIER;
    o.barrier_dir = i.barrier.direction == BarrierDirection::Up ? OFG_UP : OFG_DOWN;
    o.barrier_style = i.barrier.style == BarrierStyle::In ? OFG_IN : OFG_OUT;
    o.barrier_level = i.barrier.level;
    o.rebate = i.barrier.rebate;
    o.vol = m.vol_for(i.strike, i.expiry);
    return o;
}

inline PricingResult to_result(const ofg_result &r, bool greeks, bool heston) {
    PricingResult out;
    out.price = r.price;
    if (r.has_stderr) out.stderr_ = r.stderr_;
    if (r.num_paths != 0.0 || r.num_steps != 0.0) {
        out.diagnostics.insert("num_paths", r.num_paths);
        out.diagnostics.insert("num_steps", r.num_steps);
        out.diagnostics.insert("vol", r.vol);
    }
    if (greeks && r.has_greeks) {
        out.greeks = Greeks{r.delta, r.gamma, r.vega, r.theta, r.rho};
        if (heston) out.heston_gradient = std::vector<double>(r.heston_grad, r.heston_grad + 6);
    }
    out.kernel_ms = r.kernel_ms;
    return out;
}
}  // namespace detail

// PricingEngine<I> (engine.rs:37-52).
template <class I>
class PricingEngine {
public:
    virtual ~PricingEngine() = default;
    virtual PricingResult price(const I &instrument, const Market &market) const = 0;
    virtual PricingResult price_with_greeks_aad(const I &instrument, const Market &market) const {
        return price(instrument, market);
    }
};

// Drop-in for MonteCarloPricingEngine (mc_engine.rs:587-670, 871-988):
// every supported instrument, GBM paths on the GPU.
class GpuMonteCarloPricingEngine : public PricingEngine<VanillaOption>,
                                   public PricingEngine<AsianOption>,
                                   public PricingEngine<BarrierOption> {
public:
    GpuMonteCarloPricingEngine(int64_t num_paths, int64_t num_steps, uint64_t seed)
        : num_paths(num_paths), num_steps(num_steps), seed(seed) {}
    GpuMonteCarloPricingEngine &with_variance_reduction(VarianceReduction vr) {
        variance_reduction = vr;
        return *this;
    }
    GpuMonteCarloPricingEngine &with_seed(uint64_t s) {
        seed = s;
        reproducible = true;
        return *this;
    }
    GpuMonteCarloPricingEngine &with_rng_mode(RngMode m) {
        rng_mode = m;
        return *this;
    }
    // mc_engine.rs:626-656
    GpuMonteCarloPricingEngine &with_rng_kind(FastRngKind k) {
        rng_mode = detail::kind_mode(k);
        if (k == FastRngKind::ThreadRng) reproducible = false;
        return *this;
    }
    GpuMonteCarloPricingEngine &with_randomized_streams() {
        reproducible = false;
        return *this;
    }
    GpuMonteCarloPricingEngine &with_thread_rng() { return with_rng_kind(FastRngKind::ThreadRng); }
    GpuMonteCarloPricingEngine &with_greek_method(GreekMethod g) {
        greek_method = g;
        return *this;
    }

    PricingResult price(const VanillaOption &i, const Market &m) const override { return run(i, m, false); }
    PricingResult price(const AsianOption &i, const Market &m) const override { return run(i, m, false); }
    PricingResult price(const BarrierOption &i, const Market &m) const override { return run(i, m, false); }
    PricingResult price_with_greeks_aad(const VanillaOption &i, const Market &m) const override {
        return run(i, m, true);
    }
    PricingResult price_with_greeks_aad(const AsianOption &i, const Market &m) const override {
        return run(i, m, true);
    }
    PricingResult price_with_greeks_aad(const BarrierOption &i, const Market &m) const override {
        return run(i, m, true);
    }

    int64_t num_paths, num_steps;
    uint64_t seed;
    VarianceReduction variance_reduction = VarianceReduction::None;
    RngMode rng_mode = RngMode::Reference;
    GreekMethod greek_method = GreekMethod::Adjoint;
    bool reproducible = true;
    int device = 0;

private:
    template <class I>
    PricingResult run(const I &i, const Market &m, bool greeks) const {
        ofg_option o = detail::to_ofg(i, m, num_paths, num_steps,
                                      reproducible ? seed : detail::random_base_seed());
        ofg_engine_cfg c{};
        c.antithetic = variance_reduction == VarianceReduction::Antithetic;
        c.control_variate = variance_reduction == VarianceReduction::ControlVariate && !greeks;
        c.rng_mode = (int32_t)rng_mode;
        c.greek_mode = greeks ? (int32_t)greek_method : OFG_GREEKS_NONE;
        ofg_result r{};
        ofg_ctx *ctx = device_context(device);
        check(ofg_price(ctx, &o, 1, &c, &r));
        return detail::to_result(r, greeks, false);
    }
};

// PricingEngine for Heston paths (MonteCarloEngine::run with
// HestonPathGenerator, simulation.rs:79-117, behind the engine contract).
class HestonMonteCarloPricingEngine : public PricingEngine<VanillaOption>,
                                      public PricingEngine<BarrierOption> {
public:
    HestonMonteCarloPricingEngine(Heston model, int64_t num_paths, int64_t num_steps, uint64_t seed)
        : model(model), num_paths(num_paths), num_steps(num_steps), seed(seed) {}
    HestonMonteCarloPricingEngine &with_antithetic(bool a) {
        antithetic = a;
        return *this;
    }
    PricingResult price(const VanillaOption &i, const Market &m) const override { return run(i, m, false); }
    PricingResult price(const BarrierOption &i, const Market &m) const override { return run(i, m, false); }
    PricingResult price_with_greeks_aad(const VanillaOption &i, const Market &m) const override {
        return run(i, m, true);
    }
    PricingResult price_with_greeks_aad(const BarrierOption &i, const Market &m) const override {
        return run(i, m, true);
    }

    Heston model;
    int64_t num_paths, num_steps;
    uint64_t seed;
    bool antithetic = false;
    RngMode rng_mode = RngMode::Reference;
    int device = 0;

private:
    template <class I>
    PricingResult run(const I &i, const Market &m, bool greeks) const {
        ofg_option o = detail::to_ofg(i, m, num_paths, num_steps, seed);
        o.model = OFG_MODEL_HESTON;
        o.div_yield = m.rate - model.mu;   // drift mu = rate - div_yield
        o.kappa = model.kappa;
        o.theta = model.theta;
        o.xi = model.xi;
        o.rho = model.rho;
        o.v0 = model.v0;
        ofg_engine_cfg c{};
        c.antithetic = antithetic;
        c.rng_mode = (int32_t)rng_mode;
        ofg_result r{};
        ofg_ctx *ctx = device_context(device);
        check(greeks ? ofg_price_greeks(ctx, &o, 1, &c, &r) : ofg_price(ctx, &o, 1, &c, &r));
        return detail::to_result(r, greeks, true);
    }
};

// MonteCarloGreeksEngine (mc_greeks.rs:20-129): pathwise delta, CRN
// pathwise gamma and likelihood-ratio delta/gamma/vega of European GBM
// vanillas, one exact-terminal normal per sample (ofg_mc_greeks_estimate).
class MonteCarloGreeksEngine {
public:
    struct Estimates {
        double pathwise_delta, pathwise_gamma, lr_delta, lr_gamma, lr_vega, num_samples;
    };
    MonteCarloGreeksEngine(int64_t num_paths, uint64_t seed) : paths_(num_paths), seed_(seed) {}
    MonteCarloGreeksEngine &with_antithetic(bool a) { anti_ = a; return *this; }
    MonteCarloGreeksEngine &with_spot_bump_rel(double rel) { bump_ = std::fmax(rel, 1.0e-6); return *this; }
    MonteCarloGreeksEngine &with_seed(uint64_t seed) { seed_ = seed; return *this; }
    MonteCarloGreeksEngine &with_rng_mode(RngMode m) { rng_ = m; return *this; }

    Estimates estimates(const VanillaOption &i, const Market &m) const {
        if (i.exercise != ExerciseStyle::European)
            throw PricingError(PricingError::Kind::InvalidInput,
                               "MonteCarloGreeksEngine supports European exercise only");
        Market eff = m;
        eff.dividend_yield = m.effective_dividend_yield(i.expiry);
        ofg_option o = detail::to_ofg(i, eff, paths_, 1, seed_);
        ofg_engine_cfg c{};
        c.antithetic = anti_ ? 1 : 0;
        c.rng_mode = (int32_t)rng_;
        ofg_mc_greeks out{};
        check(ofg_mc_greeks_estimate(device_context(), &o, 1, &c, bump_, &out));
        return {out.pathwise_delta, out.pathwise_gamma, out.lr_delta, out.lr_gamma, out.lr_vega,
                out.num_samples};
    }
    Greeks estimate_pathwise(const VanillaOption &i, const Market &m) const {
        const Estimates e = estimates(i, m);
        return Greeks{e.pathwise_delta, e.pathwise_gamma, 0.0, 0.0, 0.0};
    }
    Greeks estimate_likelihood_ratio(const VanillaOption &i, const Market &m) const {
        const Estimates e = estimates(i, m);
        return Greeks{e.lr_delta, e.lr_gamma, e.lr_vega, 0.0, 0.0};
    }

private:
    int64_t paths_;
    uint64_t seed_;
    bool anti_ = true;
    double bump_ = 1.0e-2;
    RngMode rng_ = RngMode::Reference;
};

// mc_european_parallel / mc_european_sequential (mc_parallel.rs:264-434) and
// mc_european_with_arena (mc_engine.rs:186-343): exact terminal sampling on
// the reference's own single streams (ofg_european_exact), one normal per
// path; invalid inputs give a NaN price (the reference returns, it does not
// fail). The early returns follow each function's own order.
namespace detail {
inline PricingResult exact_intrinsic(const VanillaOption &i, const Market &m) {
    PricingResult r;
    r.price = i.option_type == OptionType::Call ? std::fmax(m.spot - i.strike, 0.0)
                                                : std::fmax(i.strike - m.spot, 0.0);
    r.stderr_ = 0.0;
    return r;
}
inline PricingResult exact_simulate(const VanillaOption &i, const Market &m, int64_t n_paths,
                                    int32_t kind, int32_t n_chunks) {
    Market eff = m;
    if (kind != OFG_EXACT_ARENA) eff.dividend_yield = m.effective_dividend_yield(i.expiry);
    ofg_option o = detail::to_ofg(i, eff, n_paths, 1, 0);
    ofg_result res{};
    check(ofg_european_exact(device_context(), &o, 1, kind, n_chunks, &res, nullptr));
    PricingResult r;
    r.price = res.price;
    if (!res.has_stderr) return r;
    r.stderr_ = res.stderr_;
    r.diagnostics.insert("num_paths", (double)n_paths);
    r.diagnostics.insert("num_steps", 1.0);
    if (kind == OFG_EXACT_PARALLEL) r.diagnostics.insert("num_threads", (double)n_chunks);
    r.diagnostics.insert("vol", m.vol_for(i.strike, i.expiry));
    r.kernel_ms = res.kernel_ms;
    return r;
}
}  // namespace detail

// n_chunks: the reference run's rayon::current_num_threads() (default: the
// host's hardware threads, rayon's default pool)
inline PricingResult mc_european_parallel(const VanillaOption &i, const Market &m, int64_t n_paths,
                                          int64_t /*n_steps*/ = 1, int32_t n_chunks = 0) {
    PricingResult r;
    if (n_paths == 0 || i.exercise != ExerciseStyle::European) {
        r.price = std::nan("");
        return r;
    }
    if (i.expiry <= 0.0) return detail::exact_intrinsic(i, m);
    if (n_chunks <= 0) n_chunks = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    return detail::exact_simulate(i, m, n_paths, OFG_EXACT_PARALLEL, n_chunks);
}
inline PricingResult mc_european_sequential(const VanillaOption &i, const Market &m,
                                            int64_t n_paths, int64_t /*n_steps*/ = 1) {
    PricingResult r;
    if (n_paths == 0) {
        r.price = std::nan("");
        return r;
    }
    if (i.expiry <= 0.0) return detail::exact_intrinsic(i, m);
    if (i.exercise != ExerciseStyle::European) {
        r.price = std::nan("");
        return r;
    }
    return detail::exact_simulate(i, m, n_paths, OFG_EXACT_SEQUENTIAL, 1);
}
inline PricingResult mc_european_with_arena(const VanillaOption &i, const Market &m,
                                            int64_t n_paths, int64_t /*n_steps*/ = 1) {
    PricingResult r;
    if (n_paths == 0 || !(i.strike > 0.0) || i.expiry < 0.0 ||
        i.exercise != ExerciseStyle::European) {
        r.price = std::nan("");
        return r;
    }
    if (i.expiry <= 0.0) return detail::exact_intrinsic(i, m);
    return detail::exact_simulate(i, m, n_paths, OFG_EXACT_ARENA, 1);
}

// mc_european_qmc_with_seed / mc_european_qmc (mc_qmc.rs:29-144): the
// reference's Sobol points on the GPU, one path per point.
inline PricingResult mc_european_qmc_with_seed(const VanillaOption &i, const Market &m,
                                               int64_t n_paths, int64_t n_steps, uint64_t seed) {
    PricingResult r;
    if (i.exercise != ExerciseStyle::European) {
        r.price = std::nan("");
        return r;
    }
    Market eff = m;
    eff.dividend_yield = m.effective_dividend_yield(i.expiry);
    ofg_option o = detail::to_ofg(i, eff, n_paths, n_steps, seed);
    ofg_result res{};
    check(ofg_qmc_price(device_context(), &o, 1, &res));
    r.price = res.price;
    if (res.has_stderr) r.stderr_ = res.stderr_;
    if (!std::isnan(res.num_paths)) {
        r.diagnostics.insert("num_paths", res.num_paths);
        r.diagnostics.insert("num_steps", res.num_steps);
        r.diagnostics.insert("vol", res.vol);
    }
    r.kernel_ms = res.kernel_ms;
    return r;
}
inline PricingResult mc_european_qmc(const VanillaOption &i, const Market &m, int64_t n_paths,
                                     int64_t n_steps) {
    return mc_european_qmc_with_seed(i, m, n_paths, n_steps, 0);
}

// LongstaffSchwartzEngine (longstaff_schwartz.rs:66-735) on the GPU:
// VanillaOption (European / American / Bermudan exercise, GBM),
// BermudanOption under GBM, flat local vol or Heston with its exercise
// boundary, BarrierOption by plain Monte Carlo. Reference draws (the engine's
// own single stream) unless with_rng_mode says otherwise.
class LongstaffSchwartzEngine : public PricingEngine<VanillaOption>,
                                public PricingEngine<BermudanOption>,
                                public PricingEngine<BarrierOption> {
public:
    struct BermudanLsmOutput {
        PricingResult result;
        std::vector<ExerciseBoundaryPoint> exercise_boundary;
    };

    LongstaffSchwartzEngine(int64_t num_paths, int64_t num_steps, uint64_t seed)
        : num_paths(num_paths), num_steps(num_steps), seed(seed) {}
    LongstaffSchwartzEngine &with_local_vol_dynamics() {
        dynamics_ = OFG_LSM_LOCAL_VOL;
        return *this;
    }
    // The spot drift is the market's r - q; the model's mu is not used.
    LongstaffSchwartzEngine &with_heston_dynamics(const Heston &h) {
        dynamics_ = OFG_LSM_HESTON;
        heston_ = h;
        return *this;
    }
    LongstaffSchwartzEngine &with_rng_mode(RngMode m) {
        rng_ = m;
        return *this;
    }

    PricingResult price(const VanillaOption &i, const Market &m) const override {
        ofg_lsm_option o = base(i.option_type, i.expiry, m);
        o.instrument = OFG_LSM_VANILLA;
        o.exercise = i.exercise == ExerciseStyle::European   ? OFG_EUROPEAN
                     : i.exercise == ExerciseStyle::American ? OFG_AMERICAN
                                                             : OFG_BERMUDAN;
        o.strike = i.strike;
        o.vol = m.vol_for(i.strike, i.expiry);
        if (i.exercise == ExerciseStyle::Bermudan) {
            o.n_dates = (int32_t)i.bermudan_dates.size();
            o.dates = i.bermudan_dates.data();
        }
        return run(o, nullptr);
    }
    PricingResult price(const BermudanOption &i, const Market &m) const override {
        return price_bermudan_with_boundary(i, m).result;
    }
    PricingResult price(const BarrierOption &i, const Market &m) const override {
        ofg_lsm_option o = base(i.option_type, i.expiry, m);
        o.instrument = OFG_LSM_BARRIER;
        o.strike = i.strike;
        o.vol = m.vol_for(i.strike, i.expiry);
        o.barrier_dir = i.barrier.direction == BarrierDirection::Up ? OFG_UP : OFG_DOWN;
        o.barrier_style = i.barrier.style == BarrierStyle::In ? OFG_IN : OFG_OUT;
        o.barrier_level = i.barrier.level;
        o.rebate = i.barrier.rebate;
        return run(o, nullptr);
    }
    BermudanLsmOutput price_bermudan_with_boundary(const BermudanOption &i,
                                                   const Market &m) const {
        ofg_lsm_option o = base(i.option_type, i.expiry, m);
        o.instrument = OFG_LSM_BERMUDAN;
        o.dynamics = dynamics_;
        o.n_dates = (int32_t)i.exercise_dates.size();
        o.dates = i.exercise_dates.data();
        if (i.strike_schedule.size() == i.exercise_dates.size()) o.strikes = i.strike_schedule.data();
        o.strike = i.strike_schedule.empty() ? 0.0 : i.strike_schedule.back();
        o.vol = m.vol_for(o.strike, i.expiry);
        o.kappa = heston_.kappa;
        o.theta = heston_.theta;
        o.xi = heston_.xi;
        o.rho = heston_.rho;
        o.v0 = heston_.v0;
        BermudanLsmOutput out;
        out.result = run(o, &out.exercise_boundary);
        // exercise_dates: the effective schedule's length (bermudan.rs:85-103)
        if (out.result.diagnostics.get("num_paths")) {
            const size_t n = i.exercise_dates.size();
            const bool appended = n && i.expiry - i.exercise_dates.back() > 1.0e-12;
            out.result.diagnostics.insert("exercise_dates", (double)(n + (appended ? 1 : 0)));
        }
        return out;
    }

    int64_t num_paths, num_steps;
    uint64_t seed;

private:
    ofg_lsm_option base(OptionType t, double expiry, const Market &m) const {
        ofg_lsm_option o{};
        o.option_type = t == OptionType::Call ? OFG_CALL : OFG_PUT;
        o.spot = m.spot;
        o.expiry = expiry;
        o.rate = m.rate;
        o.div_yield = m.effective_dividend_yield(expiry);
        o.num_paths = num_paths;
        o.num_steps = num_steps;
        o.seed = seed;
        return o;
    }
    PricingResult run(const ofg_lsm_option &o, std::vector<ExerciseBoundaryPoint> *bnd) const {
        ofg_engine_cfg c{};
        c.rng_mode = (int32_t)rng_;
        ofg_result res{};
        std::vector<ofg_lsm_boundary> slots(bnd ? (size_t)std::max<int64_t>(num_steps, 1) : 0);
        int64_t len = 0;
        check(ofg_lsm_price(device_context(), &o, 1, &c, &res, bnd ? slots.data() : nullptr,
                            bnd ? &len : nullptr));
        PricingResult r;
        r.price = res.price;
        r.stderr_ = res.stderr_;
        if (!std::isnan(res.num_paths)) {
            r.diagnostics.insert("num_paths", res.num_paths);
            r.diagnostics.insert("num_steps", res.num_steps);
            if (!std::isnan(res.vol)) r.diagnostics.insert("vol", res.vol);
        }
        r.kernel_ms = res.kernel_ms;
        if (bnd) {
            for (int64_t k = 0; k < len; ++k) {
                const ofg_lsm_boundary &b = slots[(size_t)k];
                ExerciseBoundaryPoint p;
                p.time = b.time;
                p.strike = b.strike;
                if (b.has_boundary) p.boundary_spot = b.boundary_spot;
                p.itm_paths = b.itm_paths;
                p.exercised_paths = b.exercised_paths;
                bnd->push_back(p);
            }
        }
        return r;
    }
    int32_t dynamics_ = OFG_LSM_GBM;
    Heston heston_{0, 0, 0, 0, 0, 0};
    RngMode rng_ = RngMode::Reference;
};

// Direct MC helpers of pricing::asian (asian.rs:139-200) and pricing::barrier
// (barrier.rs:199-239): GBM { mu: r, sigma }, antithetic, discount e^{-rt},
// the Asian average over all steps + 1 path points, a zero payoff when the
// barrier option is inactive. Return (price, stderr).
struct AsianStrike {
    bool floating = false;
    double k = 0.0;
    static AsianStrike Fixed(double k) { return AsianStrike{false, k}; }
    static AsianStrike Floating() { return AsianStrike{true, 0.0}; }
};

namespace detail {
inline std::pair<double, double> legacy_price(ofg_option o) {
    ofg_engine_cfg c{};
    c.antithetic = 1;
    ofg_result r{};
    check(ofg_price(device_context(), &o, 1, &c, &r));
    return {r.price, r.stderr_};
}
inline std::pair<double, double> legacy_asian(OptionType t, AsianStrike strike, double s0, double r,
                                              double sigma, double T, int64_t steps, int64_t paths,
                                              uint64_t seed, bool geometric) {
    const Market m{s0, r, 0.0, sigma};
    ofg_option o = base(t, strike.floating ? s0 : strike.k, T, m, paths, steps, seed);
    o.vol = sigma;
    o.payoff = OFG_PAYOFF_ASIAN;
    o.averaging = geometric ? OFG_GEOMETRIC : OFG_ARITHMETIC;
    o.strike_type = strike.floating ? OFG_FLOATING : OFG_FIXED;
    std::vector<double> obs;
    for (int64_t j = 0; j <= steps; ++j) obs.push_back(T * ((double)j / (double)steps));
    o.n_obs = (int32_t)obs.size();
    o.obs_times = obs.data();
    return legacy_price(o);
}
}  // namespace detail

inline std::pair<double, double> arithmetic_asian_price_mc(OptionType t, AsianStrike strike,
                                                           double s0, double r, double sigma,
                                                           double T, int64_t steps,
                                                           int64_t paths, uint64_t seed) {
    return detail::legacy_asian(t, strike, s0, r, sigma, T, steps, paths, seed, false);
}
inline std::pair<double, double> geometric_asian_price_mc(OptionType t, AsianStrike strike,
                                                          double s0, double r, double sigma,
                                                          double T, int64_t steps, int64_t paths,
                                                          uint64_t seed) {
    return detail::legacy_asian(t, strike, s0, r, sigma, T, steps, paths, seed, true);
}
inline std::pair<double, double> barrier_price_mc(OptionType t, BarrierStyle style,
                                                  BarrierDirection dir, double s0, double k,
                                                  double barrier, double r, double sigma, double T,
                                                  int64_t steps, int64_t num_paths,
                                                  uint64_t seed) {
    const Market m{s0, r, 0.0, sigma};
    ofg_option o = detail::base(t, k, T, m, num_paths, steps, seed);
    o.vol = sigma;
    o.payoff = OFG_PAYOFF_BARRIER;
    o.barrier_dir = dir == BarrierDirection::Up ? OFG_UP : OFG_DOWN;
    o.barrier_style = style == BarrierStyle::In ? OFG_IN : OFG_OUT;
    o.barrier_level = barrier;
    o.rebate = 0.0;
    return detail::legacy_price(o);
}

// ------------------------------------------------------- multi-asset --

struct AutocallableSensitivities {   // pricing/autocallable.rs:25-33
    std::vector<double> delta;
    double vega = 0.0, cega = 0.0;
};

// instruments/autocallable.rs:18-157 (phoenix / memory select the Phoenix note)
struct Autocallable {
    std::vector<int32_t> underlyings;
    double notional = 100.0;
    std::vector<double> autocall_dates;
    double autocall_barrier = 1.0, coupon_rate = 0.0;
    double ki_barrier = 0.6, ki_strike = 1.0, maturity = 1.0;
    bool phoenix = false, memory = false;
    double coupon_barrier = 0.0;
};

namespace detail {
inline ofg_engine_cfg cfg_for(int32_t rng) {
    ofg_engine_cfg c{};
    c.rng_mode = rng;
    return c;
}
inline ofg_autocall autocall_struct(const Autocallable &a, const std::vector<double> &spots,
                                    const std::vector<double> &vols,
                                    const std::vector<double> &corr, double r, double q,
                                    int64_t paths, int64_t steps) {
    ofg_autocall c{};
    c.phoenix = a.phoenix;
    c.memory = a.memory;
    c.n_underlyings = (int32_t)a.underlyings.size();
    c.underlyings = a.underlyings.data();
    c.n_dates = (int32_t)a.autocall_dates.size();
    c.autocall_dates = a.autocall_dates.data();
    c.notional = a.notional;
    c.maturity = a.maturity;
    c.autocall_barrier = a.autocall_barrier;
    c.coupon_rate = a.coupon_rate;
    c.coupon_barrier = a.coupon_barrier;
    c.ki_barrier = a.ki_barrier;
    c.ki_strike = a.ki_strike;
    c.n_assets = (int32_t)spots.size();
    c.spots = spots.data();
    c.vols = vols.data();
    c.corr = corr.data();
    c.rate = r;
    c.div_yield = q;
    c.num_paths = paths;
    c.num_steps = steps;
    return c;
}
}  // namespace detail

// price_autocallable / price_phoenix_autocallable (pricing/autocallable.rs:52-190):
// price, stderr and bump-and-reprice Greeks (delta = sum of deltas, vega,
// rho = cega); fast draws (the reference's single stream cannot be split:
// a called path stops drawing). corr: row-major n x n over the market.
inline PricingResult price_autocallable(const Autocallable &a, const std::vector<double> &spots,
                                        const std::vector<double> &vols,
                                        const std::vector<double> &corr, double r, double q,
                                        int64_t n_paths, int64_t n_steps,
                                        int32_t rng_mode = OFG_RNG_FAST) {
    const ofg_autocall c = detail::autocall_struct(a, spots, vols, corr, r, q, n_paths, n_steps);
    const ofg_engine_cfg cfg = detail::cfg_for(rng_mode);
    ofg_result res{};
    ofg_autocall_sens sens{};
    check(ofg_autocall_price(device_context(), &c, 1, &cfg, 1, &res, &sens));
    PricingResult out;
    out.price = res.price;
    if (res.has_stderr) {
        out.stderr_ = res.stderr_;
        out.diagnostics.insert("num_paths", res.num_paths);
        out.diagnostics.insert("num_steps", res.num_steps);
        out.diagnostics.insert("observation_count", (double)a.autocall_dates.size());
    }
    if (res.has_greeks) out.greeks = Greeks{res.delta, 0.0, res.vega, 0.0, res.rho};
    return out;
}

// autocallable_sensitivities (pricing/autocallable.rs:101-133, 192-223)
inline AutocallableSensitivities autocallable_sensitivities(
    const Autocallable &a, const std::vector<double> &spots, const std::vector<double> &vols,
    const std::vector<double> &corr, double r, double q, int64_t n_paths, int64_t n_steps,
    int32_t rng_mode = OFG_RNG_FAST) {
    const ofg_autocall c = detail::autocall_struct(a, spots, vols, corr, r, q, n_paths, n_steps);
    const ofg_engine_cfg cfg = detail::cfg_for(rng_mode);
    ofg_result res{};
    ofg_autocall_sens sens{};
    check(ofg_autocall_price(device_context(), &c, 1, &cfg, 1, &res, &sens));
    if (!res.has_stderr || !sens.ok)
        throw PricingError(PricingError::Kind::InvalidInput, ofg_last_error());
    return AutocallableSensitivities{std::vector<double>(sens.delta, sens.delta + sens.n_delta),
                                     sens.vega, sens.cega};
}

enum class BasketType { Average = OFG_BASKET_AVERAGE, BestOf = OFG_BASKET_BEST_OF,
                        WorstOf = OFG_BASKET_WORST_OF };

struct BasketOption {   // instruments/basket.rs:25-34
    std::vector<double> weights;
    double strike = 0.0, maturity = 1.0;
    bool is_call = true;
    BasketType basket_type = BasketType::Average;
};

struct BasketSensitivities {   // pricing/basket.rs:57-65
    std::vector<double> delta;
    double vega = 0.0, cega = 0.0;
};

namespace detail {
inline ofg_basket basket_struct(const BasketOption &b, const std::vector<double> &spots,
                                const std::vector<double> &vols, const std::vector<double> &corr,
                                double r, const std::vector<double> &divs, int64_t paths) {
    ofg_basket o{};
    o.basket_type = (int32_t)b.basket_type;
    o.is_call = b.is_call;
    o.n_weights = (int32_t)b.weights.size();
    o.weights = b.weights.data();
    o.n_assets = (int32_t)spots.size();
    o.strike = b.strike;
    o.maturity = b.maturity;
    o.spots = spots.data();
    o.vols = vols.data();
    o.dividends = divs.data();
    o.corr = corr.data();
    o.rate = r;
    o.num_paths = paths;
    o.seed = 3148159;   // MC_SEED (pricing/basket.rs:27)
    return o;
}
}  // namespace detail

// price_basket_mc (pricing/basket.rs:99-160): the reference's draws; an
// invalid basket gives price NaN and no stderr, as the reference.
inline PricingResult price_basket_mc(const BasketOption &b, const std::vector<double> &spots,
                                     const std::vector<double> &vols,
                                     const std::vector<double> &corr, double r,
                                     const std::vector<double> &divs, int64_t n_paths) {
    const ofg_basket o = detail::basket_struct(b, spots, vols, corr, r, divs, n_paths);
    const ofg_engine_cfg cfg = detail::cfg_for(OFG_RNG_REFERENCE);
    ofg_result res{};
    ofg_basket_sens sens{};
    check(ofg_basket_price(device_context(), &o, 1, &cfg, 0, &res, &sens));
    PricingResult out;
    out.price = res.price;
    if (res.has_stderr) {
        out.stderr_ = res.stderr_;
        out.diagnostics.insert("num_paths", res.num_paths);
        out.diagnostics.insert("rho", sens.rho);
    }
    return out;
}

// BasketCopula (pricing/basket.rs:32-46): Gaussian, or StudentT with
// degrees_of_freedom >= 2.
struct BasketCopula {
    bool student_t = false;
    int32_t degrees_of_freedom = 0;
    static BasketCopula Gaussian() { return {}; }
    static BasketCopula StudentT(int32_t dof) { return {true, dof}; }
};

// price_basket_mc_with_copula (pricing/basket.rs:120-160), the reference's
// draws: under StudentT each path draws its asset normals, then dof more.
inline PricingResult price_basket_mc_with_copula(const BasketOption &b,
                                                 const std::vector<double> &spots,
                                                 const std::vector<double> &vols,
                                                 const std::vector<double> &corr, double r,
                                                 const std::vector<double> &divs, int64_t n_paths,
                                                 BasketCopula copula) {
    ofg_basket o = detail::basket_struct(b, spots, vols, corr, r, divs, n_paths);
    o.copula = copula.student_t ? OFG_COPULA_STUDENT_T : OFG_COPULA_GAUSSIAN;
    o.copula_dof = copula.degrees_of_freedom;
    const ofg_engine_cfg cfg = detail::cfg_for(OFG_RNG_REFERENCE);
    ofg_result res{};
    ofg_basket_sens sens{};
    check(ofg_basket_price(device_context(), &o, 1, &cfg, 0, &res, &sens));
    PricingResult out;
    out.price = res.price;
    if (res.has_stderr) {
        out.stderr_ = res.stderr_;
        out.diagnostics.insert("num_paths", res.num_paths);
        out.diagnostics.insert("rho", sens.rho);
    }
    return out;
}

// basket_sensitivities (pricing/basket.rs:456-539): every bump on the base
// price's draws, in one pass per path.
inline BasketSensitivities basket_sensitivities(const BasketOption &b,
                                                const std::vector<double> &spots,
                                                const std::vector<double> &vols,
                                                const std::vector<double> &corr, double r,
                                                const std::vector<double> &divs, int64_t n_paths) {
    const ofg_basket o = detail::basket_struct(b, spots, vols, corr, r, divs, n_paths);
    const ofg_engine_cfg cfg = detail::cfg_for(OFG_RNG_REFERENCE);
    ofg_result res{};
    ofg_basket_sens sens{};
    check(ofg_basket_price(device_context(), &o, 1, &cfg, 1, &res, &sens));
    if (!res.has_stderr || !sens.ok)
        throw PricingError(PricingError::Kind::InvalidInput, ofg_last_error());
    return BasketSensitivities{std::vector<double>(sens.delta, sens.delta + sens.n_delta),
                               sens.vega, sens.cega};
}

// stressed_correlation_matrix (pricing/basket.rs:447-454); corr row-major.
inline std::vector<double> stressed_correlation_matrix(const std::vector<double> &corr, int32_t n,
                                                       const std::vector<ofg_corr_stress> &sc) {
    std::vector<double> out((size_t)n * n);
    check(ofg_stress_correlation(corr.data(), n, sc.data(), (int32_t)sc.size(), out.data()));
    return out;
}

// -------------------------------------------------- product language --
namespace dsl {

// CompiledProduct (src/dsl/ir.rs:141-151), owned.
class CompiledProduct {
public:
    explicit CompiledProduct(ofg_dsl_product *p) : p_(p, &ofg_dsl_free) {
        ofg_dsl_info info{};
        check(ofg_dsl_get_info(p, &info));
        name = info.name;
        notional = info.notional;
        maturity = info.maturity;
        num_underlyings = info.num_underlyings;
        num_state_vars = info.n_state;
        num_schedules = info.n_schedules;
    }
    std::vector<double> schedule_dates(int32_t k) const {
        int32_t n = 0;
        check(ofg_dsl_schedule_dates(p_.get(), k, nullptr, 0, &n));
        std::vector<double> d((size_t)n);
        check(ofg_dsl_schedule_dates(p_.get(), k, d.data(), n, &n));
        return d;
    }
    const ofg_dsl_product *handle() const { return p_.get(); }
    std::string name;
    double notional = 0.0, maturity = 0.0;
    int32_t num_underlyings = 0, num_state_vars = 0, num_schedules = 0;

private:
    std::shared_ptr<ofg_dsl_product> p_;
};

// parse_and_compile (src/dsl/mod.rs:24-28): throws InvalidInput with the
// DslError text.
inline CompiledProduct parse_and_compile(const std::string &source) {
    ofg_dsl_product *p = nullptr;
    check(ofg_dsl_compile(source.c_str(), &p));
    return CompiledProduct(p);
}

// AssetMarketData (src/dsl/market.rs:4-30)
struct AssetMarketData {
    ofg_dsl_asset a{};
    static AssetMarketData equity(double spot, double vol, double dividend_yield) {
        AssetMarketData d;
        d.a.kind = OFG_DSL_EQUITY;
        d.a.spot = spot;
        d.a.vol = vol;
        d.a.dividend_yield = dividend_yield;
        return d;
    }
    static AssetMarketData fx(double spot, double vol, double domestic_rate, double foreign_rate) {
        AssetMarketData d;
        d.a.kind = OFG_DSL_FX;
        d.a.spot = spot;
        d.a.vol = vol;
        d.a.domestic_rate = domestic_rate;
        d.a.foreign_rate = foreign_rate;
        return d;
    }
    static AssetMarketData commodity(double spot, double vol, double convenience_yield, double kappa,
                                     double mu) {
        AssetMarketData d;
        d.a.kind = OFG_DSL_COMMODITY;
        d.a.spot = spot;
        d.a.vol = vol;
        d.a.convenience_yield = convenience_yield;
        d.a.kappa = kappa;
        d.a.mu = mu;
        return d;
    }
    static AssetMarketData rate(double initial_rate, double vol, double mean_reversion,
                                double long_run_mean) {
        AssetMarketData d;
        d.a.kind = OFG_DSL_RATE;
        d.a.spot = initial_rate;
        d.a.vol = vol;
        d.a.mean_reversion = mean_reversion;
        d.a.long_run_mean = long_run_mean;
        return d;
    }
    double initial_value() const { return a.spot; }
    AssetMarketData with_spot_bump(double x) const {
        AssetMarketData d = *this;
        d.a.spot += x;
        return d;
    }
    AssetMarketData with_vol_bump(double x) const {
        AssetMarketData d = *this;
        d.a.vol += x;
        return d;
    }
};

// MultiAssetMarket (src/dsl/market.rs:158-168); correlation row-major.
struct MultiAssetMarket {
    std::vector<AssetMarketData> assets;
    std::vector<double> correlation;
    double rate = 0.0;
    static MultiAssetMarket single(double spot, double vol, double rate, double q) {
        return MultiAssetMarket{{AssetMarketData::equity(spot, vol, q)}, {1.0}, rate};
    }
};

// DslMonteCarloEngine (src/dsl/engine.rs:237-837)
class DslMonteCarloEngine {
public:
    DslMonteCarloEngine(int64_t num_paths, int64_t num_steps, uint64_t seed)
        : num_paths(num_paths), num_steps(num_steps), seed(seed) {}
    DslMonteCarloEngine &with_rng_mode(int32_t mode) {
        rng_mode = mode;
        return *this;
    }

    // one product under several markets, one launch, the same draws
    std::vector<PricingResult> price_markets(const CompiledProduct &p,
                                             const std::vector<MultiAssetMarket> &markets) const {
        std::vector<std::vector<ofg_dsl_asset>> keep(markets.size());
        std::vector<ofg_dsl_market> ms(markets.size());
        for (size_t i = 0; i < markets.size(); ++i) {
            for (const auto &a : markets[i].assets) keep[i].push_back(a.a);
            ms[i].n_assets = (int32_t)keep[i].size();
            ms[i].assets = keep[i].data();
            ms[i].corr = markets[i].correlation.data();
            ms[i].rate = markets[i].rate;
        }
        const ofg_dsl_engine eng{num_paths, num_steps, seed};
        const ofg_engine_cfg cfg = detail::cfg_for(rng_mode);
        std::vector<ofg_result> out(markets.size());
        check(ofg_dsl_price(device_context(), p.handle(), ms.data(), ms.size(), &eng, &cfg,
                            out.data()));
        std::vector<PricingResult> res;
        for (const auto &r : out) {
            PricingResult x;
            x.price = r.price;
            x.stderr_ = r.stderr_;
            x.diagnostics.insert("num_paths", r.num_paths);
            x.diagnostics.insert("num_steps", r.num_steps);
            res.push_back(x);
        }
        return res;
    }

    PricingResult price_multi_asset(const CompiledProduct &p, const MultiAssetMarket &m) const {
        return price_markets(p, {m})[0];
    }

    // greeks_multi_asset (engine.rs:785-837): the five markets in one launch
    Greeks greeks_multi_asset(const CompiledProduct &p, const MultiAssetMarket &m,
                              size_t asset_index) const {
        if (asset_index >= m.assets.size())
            throw PricingError(PricingError::Kind::InvalidInput,
                               "asset_index " + std::to_string(asset_index) +
                                   " out of range (have " + std::to_string(m.assets.size()) +
                                   " assets)");
        const double h = 0.01 * m.assets[asset_index].initial_value();
        std::vector<MultiAssetMarket> s(5, m);
        s[1].assets[asset_index] = m.assets[asset_index].with_spot_bump(h);
        s[2].assets[asset_index] = m.assets[asset_index].with_spot_bump(-h);
        s[3].assets[asset_index] = m.assets[asset_index].with_vol_bump(0.01);
        s[4].rate = m.rate + 0.0001;
        const auto r = price_markets(p, s);
        const double base = r[0].price, up = r[1].price, dn = r[2].price;
        return Greeks{(up - dn) / (2.0 * h), (up - 2.0 * base + dn) / (h * h),
                      (r[3].price - base) / 0.01, 0.0, (r[4].price - base) / 0.0001};
    }

    int64_t num_paths, num_steps;
    uint64_t seed;
    int32_t rng_mode = OFG_RNG_REFERENCE;
};

}  // namespace dsl

// SpreadOption (src/instruments/spread.rs:15-78) and SpreadMonteCarloEngine
// (src/engines/monte_carlo/spread_mc.rs:21-175) over ofg_spread_price.
struct SpreadOption {
    double s1, s2, k, vol1, vol2, rho, q1, q2, r, t;

    // spread.rs:31-58 (the same comparisons: NaN fields pass where they pass there)
    void validate() const {
        auto bad = [](const char *m) { throw PricingError(PricingError::Kind::InvalidInput, m); };
        if (s1 <= 0.0 || s2 <= 0.0) bad("spread spots s1 and s2 must be > 0");
        if (k < 0.0) bad("spread strike k must be >= 0");
        if (vol1 < 0.0 || vol2 < 0.0) bad("spread volatilities vol1 and vol2 must be >= 0");
        if (!(rho >= -1.0 && rho <= 1.0)) bad("spread correlation rho must be in [-1, 1]");
        if (t < 0.0) bad("spread maturity t must be >= 0");
    }
    // spread.rs:61-72
    double effective_volatility() const {
        const double v = vol1 * vol1 - 2.0 * rho * vol1 * vol2 + vol2 * vol2;
        if (v < -1.0e-14)
            throw PricingError(PricingError::Kind::InvalidInput,
                               "spread effective variance is negative");
        return std::sqrt(std::fmax(v, 0.0));
    }
};

class SpreadMonteCarloEngine {
public:
    SpreadMonteCarloEngine(int64_t num_paths, uint64_t seed) : paths_(num_paths), seed_(seed) {}
    SpreadMonteCarloEngine &with_antithetic(bool a) { anti_ = a; return *this; }
    SpreadMonteCarloEngine &with_rng_kind(FastRngKind k) {
        kind_ = k;
        if (k == FastRngKind::ThreadRng) reproducible_ = false;
        return *this;
    }
    SpreadMonteCarloEngine &with_seed(uint64_t seed) {
        seed_ = seed;
        reproducible_ = true;
        return *this;
    }
    SpreadMonteCarloEngine &with_randomized_streams() { reproducible_ = false; return *this; }
    SpreadMonteCarloEngine &with_thread_rng() { return with_rng_kind(FastRngKind::ThreadRng); }
    // Throughput generator instead of the reference's kinds.
    SpreadMonteCarloEngine &with_fast_rng(bool fast) { fast_ = fast; return *this; }

    // spread_mc.rs:83-175 (the market is not used there either)
    PricingResult price(const SpreadOption &i) const {
        i.validate();
        if (paths_ == 0)
            throw PricingError(PricingError::Kind::InvalidInput, "spread MC num_paths must be > 0");
        const RngMode mode = fast_ ? RngMode::Fast : detail::kind_mode(kind_);
        ofg_spread_option o{i.s1, i.s2, i.k, i.vol1, i.vol2, i.rho, i.q1, i.q2, i.r, i.t,
                            paths_, reproducible_ ? seed_ : detail::random_base_seed()};
        ofg_result r{};
        check(ofg_spread_price(device_context(), &o, 1, anti_ ? 1 : 0, (int32_t)mode, &r,
                               nullptr));
        PricingResult out;
        out.price = r.price;
        if (r.has_stderr) out.stderr_ = r.stderr_;
        if (r.num_paths > 0) {   // simulated (t > 0)
            out.diagnostics.insert("num_paths", r.num_paths);
            out.diagnostics.insert("rho", i.rho);
        }
        out.kernel_ms = r.kernel_ms;
        return out;
    }

private:
    int64_t paths_;
    uint64_t seed_;
    bool anti_ = true;
    FastRngKind kind_ = FastRngKind::Xoshiro256PlusPlus;
    bool reproducible_ = true;
    bool fast_ = false;
};

}  // namespace openferric_amd
