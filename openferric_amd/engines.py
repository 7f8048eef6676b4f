"""GPU pricing engines behind the reference's PricingEngine contract.

`GpuMonteCarloPricingEngine` is the drop-in for MonteCarloPricingEngine
(src/engines/monte_carlo/mc_engine.rs:587-670, 871-988): same constructor
(num_paths, num_steps, seed), same builders, `price(instrument, market)`
returning PricingResult and `price_with_greeks_aad`. Every call goes through
the C ABI of libopenferric_amd.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .core import (AsianOption, Averaging, BarrierDirection, BarrierOption, BarrierStyle,
                   BermudanLsmOutput, BermudanOption, ConvergenceFailure, Diagnostics,
                   ExerciseBoundaryPoint, ExerciseStyle, Greeks, Heston, InvalidInput, Market,
                   MarketDataMissing, NumericalError, OptionType, PricingResult, StrikeType,
                   VanillaOption)


class VarianceReduction(enum.Enum):
    """mc_engine.rs:31-41"""
    None_ = 0
    Antithetic = 1
    ControlVariate = 2


class GreekMethod(enum.Enum):
    """How price_with_greeks_aad differentiates: the checkpointed reverse
    sweep (adjoint kernel), or the same pathwise estimator in one forward
    sweep (GBM running sums; Heston falls back to the adjoint)."""
    Adjoint = _lib.GREEKS_ADJOINT
    Forward = _lib.GREEKS_FORWARD


class RngMode(enum.Enum):
    Reference = _lib.RNG_REFERENCE
    External = _lib.RNG_EXTERNAL
    Fast = _lib.RNG_FAST
    Pcg64 = _lib.RNG_PCG64


class FastRngKind(enum.Enum):
    """fast_rng.rs:20-27: the reference's per-path stream generators."""
    Xoshiro256PlusPlus = 0
    Pcg64 = 1
    ThreadRng = 2
    StdRng = 3


def rng_kind_from_str(value: Optional[str]) -> FastRngKind:
    """python/src/engines.rs:92-108"""
    v = (value if value is not None else "xoshiro256plusplus").lower()
    if v in ("xoshiro256plusplus", "xoshiro256_plus_plus", "xoshiro"):
        return FastRngKind.Xoshiro256PlusPlus
    if v in ("pcg64", "pcg"):
        return FastRngKind.Pcg64
    if v in ("threadrng", "thread_rng", "thread"):
        return FastRngKind.ThreadRng
    if v in ("stdrng", "std_rng", "std"):
        return FastRngKind.StdRng
    raise ValueError(f"unsupported rng kind '{value}'")


def kind_rng_mode(kind: FastRngKind) -> RngMode:
    """The library's draws for a reference stream kind: xoshiro256++ or
    Pcg64 per path as the reference. ThreadRng streams are non-reproducible
    there; here they are xoshiro256++ streams on a fresh random base seed.
    StdRng (rand's ChaCha12) has no GPU stream."""
    if kind is FastRngKind.StdRng:
        raise InvalidInput("FastRngKind::StdRng streams are not available on the GPU")
    return RngMode.Pcg64 if kind is FastRngKind.Pcg64 else RngMode.Reference


class _StreamOptions:
    """rng_kind / reproducible builders shared by the engines whose
    reference counterparts carry them (MonteCarloPricingEngine,
    mc_engine.rs:626-656; MonteCarloGreeksEngine, mc_greeks.rs:44-75;
    SpreadMonteCarloEngine, spread_mc.rs:53-80). Randomized streams draw a
    fresh base seed per call (the reference a fresh seed per path: both are
    independent non-reproducible streams)."""

    rng_kind = FastRngKind.Xoshiro256PlusPlus
    reproducible = True

    def with_rng_kind(self, rng_kind: FastRngKind):
        self.rng_kind = rng_kind
        if rng_kind is FastRngKind.ThreadRng:
            self.reproducible = False
        if rng_kind is not FastRngKind.StdRng:
            self.rng_mode = kind_rng_mode(rng_kind)
        return self

    def with_randomized_streams(self):
        self.reproducible = False
        return self

    def with_thread_rng(self):
        return self.with_rng_kind(FastRngKind.ThreadRng)

    def _stream_mode(self) -> RngMode:
        if self.rng_kind is FastRngKind.StdRng and self.rng_mode in (RngMode.Reference,
                                                                     RngMode.Pcg64):
            kind_rng_mode(self.rng_kind)   # raises
        return self.rng_mode

    def _stream_seed(self) -> int:
        import secrets
        return self.seed if self.reproducible else secrets.randbits(64)


_ERRORS = {
    _lib.OFG_INVALID_INPUT: InvalidInput,
    _lib.OFG_CONVERGENCE_FAILURE: ConvergenceFailure,
    _lib.OFG_MARKET_DATA_MISSING: MarketDataMissing,
    _lib.OFG_NUMERICAL_ERROR: NumericalError,
    _lib.OFG_DEVICE_ERROR: _lib.DeviceError,
}


def _raise(rc: int):
    if rc != _lib.OFG_OK:
        raise _ERRORS.get(rc, NumericalError)(_lib.last_error())


class _Keep:
    """Keeps numpy arrays referenced by ctypes structs alive."""

    def __init__(self):
        self.refs = []

    def arr(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        self.refs.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double))


def option_struct(instrument, market: Market, num_paths: int, num_steps: int, seed: int,
                  keep: _Keep, heston: Optional[Heston] = None) -> _lib.OfgOption:
    """Flattens (instrument, market[, Heston]) into the C ABI's ofg_option."""
    o = _lib.OfgOption()
    o.model = _lib.MODEL_HESTON if heston is not None else _lib.MODEL_GBM
    o.option_type = _lib.CALL if instrument.option_type is OptionType.Call else _lib.PUT
    o.spot, o.strike, o.expiry = market.spot, instrument.strike, instrument.expiry
    o.rate = market.rate
    # drift mu = rate - dividend_yield (mc_engine.rs:916); a Heston model's own
    # mu is honoured by shifting the carry.
    o.div_yield = market.dividend_yield if heston is None else market.rate - heston.mu
    if isinstance(instrument, AsianOption):
        ref_strike = instrument.strike if instrument.asian.strike_type is StrikeType.Fixed \
            else market.spot
    else:
        ref_strike = instrument.strike
    o.vol = market.vol_for(max(ref_strike, 1e-12), instrument.expiry)
    if heston is not None:
        o.kappa, o.theta, o.xi, o.rho, o.v0 = heston.kappa, heston.theta, heston.xi, \
            heston.rho, heston.v0
    if isinstance(instrument, VanillaOption):
        o.payoff = _lib.PAYOFF_VANILLA
    elif isinstance(instrument, AsianOption):
        o.payoff = _lib.PAYOFF_ASIAN
        o.averaging = _lib.ARITHMETIC if instrument.asian.averaging is Averaging.Arithmetic \
            else _lib.GEOMETRIC
        o.strike_type = _lib.FIXED if instrument.asian.strike_type is StrikeType.Fixed \
            else _lib.FLOATING
        obs = list(instrument.asian.observation_times)
        o.n_obs = len(obs)
        o.obs_times = keep.arr(obs) if obs else None
    elif isinstance(instrument, BarrierOption):
        o.payoff = _lib.PAYOFF_BARRIER
        o.barrier_dir = _lib.UP if instrument.barrier.direction is BarrierDirection.Up \
            else _lib.DOWN
        o.barrier_style = _lib.IN if instrument.barrier.style is BarrierStyle.In else _lib.OUT
        o.barrier_level = instrument.barrier.level
        o.rebate = instrument.barrier.rebate
    else:
        raise InvalidInput(f"unsupported instrument {type(instrument).__name__}")
    o.num_paths, o.num_steps, o.seed = int(num_paths), int(num_steps), int(seed) & (2**64 - 1)
    o._keep = keep  # obs_times points into an array this keeps alive
    return o


def engine_cfg(antithetic=False, control_variate=False, rng_mode=RngMode.Reference,
               greek_mode=_lib.GREEKS_NONE, bump_rel=1e-4, normals1=None, normals2=None,
               keep: Optional[_Keep] = None, tune_block=0, tune_item_steps=0) -> _lib.OfgEngineCfg:
    c = _lib.OfgEngineCfg()
    c.antithetic = 1 if antithetic else 0
    c.control_variate = 1 if control_variate else 0
    c.rng_mode = rng_mode.value if isinstance(rng_mode, RngMode) else int(rng_mode)
    c.greek_mode = greek_mode
    c.bump_rel = bump_rel
    keep = keep or _Keep()
    c.normals1 = keep.arr(normals1) if normals1 is not None else None
    c.normals2 = keep.arr(normals2) if normals2 is not None else None
    c.tune_block, c.tune_item_steps = tune_block, tune_item_steps
    c._keep = keep  # the struct holds raw pointers into these arrays
    return c


def price_raw(opts: Sequence[_lib.OfgOption], cfg: _lib.OfgEngineCfg, device: int = 0,
              greeks: bool = False) -> List[_lib.OfgResult]:
    """Calls ofg_price (or ofg_price_greeks) on an array of ofg_option."""
    L = _lib.load()
    ctx = _lib.context(device)
    arr = (_lib.OfgOption * len(opts))(*opts)
    out = (_lib.OfgResult * len(opts))()
    fn = L.ofg_price_greeks if greeks else L.ofg_price
    _raise(fn(ctx, arr, len(opts), C.byref(cfg), out))
    return list(out)


def price_noalloc_raw(opts: Sequence[_lib.OfgOption], cfg: _lib.OfgEngineCfg,
                      device: int = 0) -> List[_lib.OfgResult]:
    """Calls ofg_price_noalloc: ofg_price with every device buffer in the
    library's fixed per-GPU region (no device allocation in the call)."""
    L = _lib.load()
    ctx = _lib.context(device)
    arr = (_lib.OfgOption * len(opts))(*opts)
    out = (_lib.OfgResult * len(opts))()
    _raise(L.ofg_price_noalloc(ctx, arr, len(opts), C.byref(cfg), out))
    return list(out)


def fill_normals(seed: int, n_streams: int, n_per_stream: int, first_stream: int = 0,
                 rng_mode: "RngMode" = None, step_major: bool = False,
                 device: int = 0) -> np.ndarray:
    """The reference's per-sample normal streams drawn on the GPU
    (ofg_fill_normals): row i holds the n_per_stream normals of the stream
    seeded stream_seed(seed, first_stream + i), as fill_standard_normals
    draws them (src/math/fast_rng.rs:228-240). Shape [n_streams,
    n_per_stream], or [n_per_stream, n_streams] with step_major."""
    L = _lib.load()
    ctx = _lib.context(device)
    mode = (rng_mode or RngMode.Reference).value
    shape = (n_per_stream, n_streams) if step_major else (n_streams, n_per_stream)
    out = np.zeros(shape)
    ms = C.c_double()
    _raise(L.ofg_fill_normals(ctx, mode, seed % (1 << 64), first_stream, n_streams, n_per_stream,
                              1 if step_major else 0,
                              out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms)))
    return out


def sample_payoffs(opt: _lib.OfgOption, cfg: _lib.OfgEngineCfg, antithetic: bool,
                   device: int = 0, want_y: bool = True) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """Per-sample (x, y) of one option via ofg_sample_payoffs."""
    L = _lib.load()
    ctx = _lib.context(device)
    samples = (opt.num_paths + 1) // 2 if antithetic else opt.num_paths
    xs = np.zeros(samples)
    ys = np.zeros(samples) if want_y else None
    _raise(L.ofg_sample_payoffs(ctx, C.byref(opt), C.byref(cfg),
                                xs.ctypes.data_as(C.POINTER(C.c_double)),
                                ys.ctypes.data_as(C.POINTER(C.c_double)) if want_y else None))
    return xs, ys


def _to_result(r: _lib.OfgResult, with_greeks: bool, heston: bool) -> PricingResult:
    diag = Diagnostics()
    if r.num_paths or r.num_steps:
        diag.insert("num_paths", r.num_paths)
        diag.insert("num_steps", r.num_steps)
        diag.insert("vol", r.vol)
    res = PricingResult(price=r.price, stderr=r.stderr_ if r.has_stderr else None,
                        diagnostics=diag, kernel_ms=r.kernel_ms)
    if with_greeks and r.has_greeks:
        res.greeks = Greeks(r.delta, r.gamma, r.vega, r.theta, r.rho)
        if heston:
            res.heston_gradient = dict(zip(["s0", "v0", "kappa", "theta", "xi", "rho"],
                                           list(r.heston_grad)))
    return res


class GpuMonteCarloPricingEngine(_StreamOptions):
    """Drop-in for MonteCarloPricingEngine (mc_engine.rs:587-670, 871-988)."""

    def __init__(self, num_paths: int, num_steps: int, seed: int):
        self.num_paths = num_paths
        self.num_steps = num_steps
        self.seed = seed
        self.variance_reduction = VarianceReduction.None_
        self.rng_mode = RngMode.Reference
        self.greek_method = GreekMethod.Adjoint
        self.device = 0

    def with_greek_method(self, method: GreekMethod) -> "GpuMonteCarloPricingEngine":
        self.greek_method = method
        return self

    def with_variance_reduction(self, vr: VarianceReduction) -> "GpuMonteCarloPricingEngine":
        self.variance_reduction = vr
        return self

    def with_seed(self, seed: int) -> "GpuMonteCarloPricingEngine":
        self.seed = seed
        self.reproducible = True
        return self

    def with_rng_mode(self, mode: RngMode) -> "GpuMonteCarloPricingEngine":
        self.rng_mode = mode
        return self

    def with_device(self, device: int) -> "GpuMonteCarloPricingEngine":
        self.device = device
        return self

    def _validate(self, instrument):
        instrument.validate()
        if isinstance(instrument, VanillaOption) and instrument.exercise is not ExerciseStyle.European:
            # the MC payoff ignores the exercise style like the reference's
            # MonteCarloInstrument impl (mc_engine.rs:441-456)
            pass

    def price(self, instrument, market: Market) -> PricingResult:
        self._validate(instrument)
        keep = _Keep()
        o = option_struct(instrument, market, self.num_paths, self.num_steps, self._stream_seed(),
                          keep)
        cfg = engine_cfg(antithetic=self.variance_reduction is VarianceReduction.Antithetic,
                         control_variate=self.variance_reduction is VarianceReduction.ControlVariate,
                         rng_mode=self._stream_mode(), keep=keep)
        r = price_raw([o], cfg, self.device)[0]
        return _to_result(r, False, False)

    def price_with_greeks_aad(self, instrument, market: Market) -> PricingResult:
        """Adjoint Greeks for any supported payoff (the reference only has
        them for VanillaOption, mc_engine.rs:977-987)."""
        self._validate(instrument)
        keep = _Keep()
        o = option_struct(instrument, market, self.num_paths, self.num_steps, self._stream_seed(),
                          keep)
        cfg = engine_cfg(antithetic=self.variance_reduction is VarianceReduction.Antithetic,
                         rng_mode=self._stream_mode(), greek_mode=self.greek_method.value,
                         keep=keep)
        r = price_raw([o], cfg, self.device)[0]
        return _to_result(r, True, False)


class GpuArithmeticAsianMC:
    """Drop-in for ArithmeticAsianMC (mc_engine.rs:672-869): antithetic always,
    geometric-Asian control variate on by default."""

    def __init__(self, paths: int, steps: int, seed: int):
        self.paths, self.steps, self.seed = paths, steps, seed
        self.control_variate = True
        self.rng_mode = RngMode.Reference

    def with_control_variate(self, cv: bool) -> "GpuArithmeticAsianMC":
        self.control_variate = cv
        return self

    def with_rng_mode(self, mode: RngMode) -> "GpuArithmeticAsianMC":
        self.rng_mode = mode
        return self

    def price(self, instrument: AsianOption, market: Market) -> PricingResult:
        instrument.validate()
        if instrument.asian.averaging is not Averaging.Arithmetic:
            raise InvalidInput("ArithmeticAsianMC requires Averaging::Arithmetic")
        if instrument.asian.strike_type is not StrikeType.Fixed:
            raise InvalidInput("ArithmeticAsianMC currently supports StrikeType::Fixed only")
        if self.paths == 0:
            raise InvalidInput("arithmetic asian mc paths must be > 0")
        if self.steps == 0:
            raise InvalidInput("arithmetic asian mc steps must be > 0")
        vol = market.vol_for(instrument.strike, instrument.expiry)
        if not (vol > 0.0) or vol != vol or vol == float("inf"):
            raise InvalidInput("market volatility must be finite and > 0")
        keep = _Keep()
        o = option_struct(instrument, market, self.paths, self.steps, self.seed, keep)
        cfg = engine_cfg(antithetic=True, control_variate=self.control_variate,
                         rng_mode=self.rng_mode, keep=keep)
        r = price_raw([o], cfg)[0]
        return _to_result(r, False, False)


class HestonMonteCarloPricingEngine:
    """PricingEngine for Heston paths: MonteCarloEngine::run with
    HestonPathGenerator (simulation.rs:79-117, 182-348) behind the engine
    contract. The Heston model's mu is the drift; market.rate discounts."""

    def __init__(self, model: Heston, num_paths: int, num_steps: int, seed: int):
        self.model = model
        self.num_paths, self.num_steps, self.seed = num_paths, num_steps, seed
        self.antithetic = False
        self.rng_mode = RngMode.Reference

    def with_antithetic(self, a: bool) -> "HestonMonteCarloPricingEngine":
        self.antithetic = a
        return self

    def with_rng_mode(self, mode: RngMode) -> "HestonMonteCarloPricingEngine":
        self.rng_mode = mode
        return self

    def price(self, instrument, market: Market) -> PricingResult:
        instrument.validate()
        keep = _Keep()
        o = option_struct(instrument, market, self.num_paths, self.num_steps, self.seed, keep,
                          heston=self.model)
        cfg = engine_cfg(antithetic=self.antithetic, rng_mode=self.rng_mode, keep=keep)
        return _to_result(price_raw([o], cfg)[0], False, True)

    def price_with_greeks_aad(self, instrument, market: Market) -> PricingResult:
        instrument.validate()
        keep = _Keep()
        o = option_struct(instrument, market, self.num_paths, self.num_steps, self.seed, keep,
                          heston=self.model)
        cfg = engine_cfg(antithetic=self.antithetic, rng_mode=self.rng_mode, keep=keep)
        return _to_result(price_raw([o], cfg, greeks=True)[0], True, True)


@dataclass
class HestonAadConfig:
    """mc_aad.rs:221-237"""
    num_paths: int
    num_steps: int
    seed: int


def heston_price_delta_aad(option_type: OptionType, strike: float, maturity: float, spot: float,
                           rate: float, model: Heston, config: HestonAadConfig) -> Tuple[float, float]:
    """GPU counterpart of heston_price_delta_aad (mc_aad.rs:279-342): price
    and dPrice/dS0 on the reference draws (no antithetic)."""
    if config.num_paths == 0:
        raise InvalidInput("num_paths must be > 0")
    if config.num_steps == 0:
        raise InvalidInput("num_steps must be > 0")
    if maturity <= 0.0:
        intrinsic = max(spot - strike, 0.0) if option_type is OptionType.Call \
            else max(strike - spot, 0.0)
        return intrinsic, 0.0
    if not model.validate():
        raise InvalidInput("invalid Heston parameters")
    eng = HestonMonteCarloPricingEngine(model, config.num_paths, config.num_steps, config.seed)
    mkt = Market(spot, rate, 0.0, 0.2)
    res = eng.price_with_greeks_aad(VanillaOption(option_type, strike, maturity), mkt)
    return res.price, res.greeks.delta


def mc_european_pathwise_aad(engine: GpuMonteCarloPricingEngine, instrument: VanillaOption,
                             market: Market) -> PricingResult:
    """GPU counterpart of mc_european_pathwise_aad (mc_aad.rs:75-219)."""
    instrument.validate()
    if instrument.exercise is not ExerciseStyle.European:
        raise InvalidInput("mc pathwise AAD supports European exercise only")
    if engine.num_paths == 0:
        raise InvalidInput("num_paths must be > 0")
    if engine.num_steps == 0:
        raise InvalidInput("num_steps must be > 0")
    if instrument.expiry <= 0.0:
        intrinsic = max(market.spot - instrument.strike, 0.0) \
            if instrument.option_type is OptionType.Call else max(instrument.strike - market.spot, 0.0)
        return PricingResult(intrinsic, 0.0, Greeks(0.0, 0.0, 0.0, 0.0, 0.0))
    vol = market.vol_for(instrument.strike, instrument.expiry)
    if not (vol > 0.0) or vol == float("inf"):
        raise InvalidInput("market volatility must be finite and > 0")
    # the drift uses the market's effective dividend yield (mc_aad.rs:138)
    eff = Market(market.spot, market.rate, market.effective_dividend_yield(instrument.expiry),
                 market.flat_vol)
    return engine.price_with_greeks_aad(instrument, eff)


class MonteCarloGreeksEngine(_StreamOptions):
    """Drop-in for MonteCarloGreeksEngine (mc_greeks.rs:20-129): pathwise and
    likelihood-ratio Greeks of European GBM vanillas from one exact-terminal
    normal per sample, computed by the GPU estimator kernel
    (ofg_mc_greeks_estimate)."""

    def __init__(self, num_paths: int, seed: int):
        self.num_paths = num_paths
        self.seed = seed
        self.antithetic = True
        self.spot_bump_rel = 1.0e-2
        self.rng_mode = RngMode.Reference
        self.device = 0

    def with_antithetic(self, antithetic: bool) -> "MonteCarloGreeksEngine":
        self.antithetic = antithetic
        return self

    def with_spot_bump_rel(self, spot_bump_rel: float) -> "MonteCarloGreeksEngine":
        self.spot_bump_rel = max(spot_bump_rel, 1.0e-6)
        return self

    def with_seed(self, seed: int) -> "MonteCarloGreeksEngine":
        self.seed = seed
        self.reproducible = True
        return self

    def with_rng_mode(self, mode: RngMode) -> "MonteCarloGreeksEngine":
        self.rng_mode = mode
        return self

    def estimates(self, instrument: VanillaOption, market: Market) -> dict:
        """All five discounted estimates (run_estimators, mc_greeks.rs:131-243)."""
        instrument.validate()
        if instrument.exercise is not ExerciseStyle.European:
            raise InvalidInput("MonteCarloGreeksEngine supports European exercise only")
        keep = _Keep()
        eff = Market(market.spot, market.rate, market.effective_dividend_yield(instrument.expiry),
                     market.flat_vol)
        o = option_struct(instrument, eff, self.num_paths, 1, self._stream_seed(), keep)
        cfg = engine_cfg(antithetic=self.antithetic, rng_mode=self._stream_mode(), keep=keep)
        L = _lib.load()
        ctx = _lib.context(self.device)
        out = _lib.OfgMcGreeks()
        _raise(L.ofg_mc_greeks_estimate(ctx, C.byref(o), 1, C.byref(cfg), self.spot_bump_rel,
                                        C.byref(out)))
        return {"pathwise_delta": out.pathwise_delta, "pathwise_gamma": out.pathwise_gamma,
                "lr_delta": out.lr_delta, "lr_gamma": out.lr_gamma, "lr_vega": out.lr_vega,
                "num_samples": out.num_samples, "kernel_ms": out.kernel_ms}

    def estimate_pathwise(self, instrument: VanillaOption, market: Market) -> Greeks:
        """Pathwise delta and CRN-difference gamma (mc_greeks.rs:93-108)."""
        e = self.estimates(instrument, market)
        return Greeks(e["pathwise_delta"], e["pathwise_gamma"], 0.0, 0.0, 0.0)

    def estimate_likelihood_ratio(self, instrument: VanillaOption, market: Market) -> Greeks:
        """Likelihood-ratio delta, gamma and vega (mc_greeks.rs:110-129)."""
        e = self.estimates(instrument, market)
        return Greeks(e["lr_delta"], e["lr_gamma"], e["lr_vega"], 0.0, 0.0)


_NAN_RESULT = float("nan")


def european_exact_raw(opts: Sequence[_lib.OfgOption], kind: int, n_chunks: int = 1,
                       want_payoffs: bool = False, device: int = 0):
    """ofg_european_exact: the reference's exact-terminal European pricers on
    its own single streams. Returns the results (and, with want_payoffs,
    every path's payoff, options concatenated)."""
    L = _lib.load()
    ctx = _lib.context(device)
    arr = (_lib.OfgOption * len(opts))(*opts)
    out = (_lib.OfgResult * len(opts))()
    pay = np.full(sum(int(o.num_paths) for o in opts), np.nan) if want_payoffs else None
    _raise(L.ofg_european_exact(ctx, arr, len(opts), kind, n_chunks, out,
                                pay.ctypes.data_as(C.POINTER(C.c_double)) if want_payoffs
                                else None))
    return (list(out), pay) if want_payoffs else list(out)


def _nan_result() -> PricingResult:
    return PricingResult(_NAN_RESULT, None)


def _reference_threads() -> int:
    """rayon::current_num_threads() of a default pool: the available
    parallelism of this process."""
    import os
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        return max(1, os.cpu_count() or 1)


def _exact(instrument: VanillaOption, market: Market, n_paths: int, kind: int, n_chunks: int,
           rng_mode: "RngMode", device: int) -> PricingResult:
    """The simulation of the three exact-terminal pricers once their early
    returns are past: expiry > 0 and a European instrument."""
    vol = market.vol_for(instrument.strike, instrument.expiry)
    keep = _Keep()
    # mu: rate - effective dividend yield (mc_parallel.rs:317, 397), rate -
    # dividend_yield for the arena (mc_engine.rs:230)
    q = market.dividend_yield if kind == _lib.EXACT_ARENA else \
        market.effective_dividend_yield(instrument.expiry)
    eff = Market(market.spot, market.rate, q, market.flat_vol)
    seed = 42 if kind == _lib.EXACT_ARENA else 0xDEADBEEFCAFEBABE
    o = option_struct(instrument, eff, n_paths, 1, seed, keep)
    if rng_mode is RngMode.Reference:
        r = european_exact_raw([o], kind, n_chunks, device=device)[0]
    else:
        # throughput streams: per-path generators instead of the single stream
        if not (vol > 0.0) or vol == float("inf"):
            return _nan_result()
        r = price_raw([o], engine_cfg(rng_mode=rng_mode, keep=keep), device)[0]
    if not r.has_stderr:
        return _nan_result()
    diag = Diagnostics()
    diag.insert("num_paths", float(n_paths))
    diag.insert("num_steps", 1.0)
    if kind == _lib.EXACT_PARALLEL:
        diag.insert("num_threads", float(n_chunks))
    diag.insert("vol", vol)
    return PricingResult(r.price, r.stderr_, diagnostics=diag, kernel_ms=r.kernel_ms)


def _intrinsic(instrument: VanillaOption, market: Market) -> PricingResult:
    call = instrument.option_type is OptionType.Call
    v = max(market.spot - instrument.strike, 0.0) if call else \
        max(instrument.strike - market.spot, 0.0)
    return PricingResult(v, 0.0)


def qmc_price_raw(opts: Sequence[_lib.OfgOption], device: int = 0) -> List[_lib.OfgResult]:
    """Calls ofg_qmc_price on an array of ofg_option (GBM vanillas)."""
    L = _lib.load()
    ctx = _lib.context(device)
    arr = (_lib.OfgOption * len(opts))(*opts)
    out = (_lib.OfgResult * len(opts))()
    _raise(L.ofg_qmc_price(ctx, arr, len(opts), out))
    return list(out)


def _qmc_result(r: _lib.OfgResult) -> PricingResult:
    diag = Diagnostics()
    if not math.isnan(r.num_paths):
        diag.insert("num_paths", r.num_paths)
        diag.insert("num_steps", r.num_steps)
        diag.insert("vol", r.vol)
    return PricingResult(r.price, r.stderr_ if r.has_stderr else None, diagnostics=diag,
                         kernel_ms=r.kernel_ms)


def mc_european_qmc_with_seed(instrument: VanillaOption, market: Market, n_paths: int,
                              n_steps: int, seed: int, device: int = 0) -> PricingResult:
    """mc_european_qmc_with_seed (engines/monte_carlo/mc_qmc.rs:40-144) on the
    GPU: the reference's Sobol points, one path per point. Non-European
    exercise gives a NaN price with no stderr, as the reference."""
    if instrument.exercise is not ExerciseStyle.European:
        return PricingResult(_NAN_RESULT, None)
    keep = _Keep()
    q = market.effective_dividend_yield(instrument.expiry)
    eff = Market(market.spot, market.rate, q, market.flat_vol)
    o = option_struct(instrument, eff, n_paths, n_steps, seed, keep)
    return _qmc_result(qmc_price_raw([o], device)[0])


def mc_european_qmc(instrument: VanillaOption, market: Market, n_paths: int,
                    n_steps: int) -> PricingResult:
    """mc_european_qmc (mc_qmc.rs:29-37): seed 0."""
    return mc_european_qmc_with_seed(instrument, market, n_paths, n_steps, 0)


def mc_european_parallel(instrument: VanillaOption, market: Market, n_paths: int,
                         n_steps: int = 1, n_chunks: Optional[int] = None,
                         rng_mode: "RngMode" = None, device: int = 0) -> PricingResult:
    """mc_european_parallel (mc_parallel.rs:264-362) on the GPU. n_steps is
    ignored, as in the reference. The reference splits the paths into
    n_chunks = rayon::current_num_threads() chunks, each its own stream
    (default: this process's available parallelism, as rayon's default pool);
    with the reference's thread count the draws are the reference's, path by
    path. rng_mode=RngMode.Fast uses per-path throughput streams instead."""
    rng_mode = RngMode.Reference if rng_mode is None else rng_mode
    if n_paths == 0 or instrument.exercise is not ExerciseStyle.European:
        return _nan_result()
    if instrument.expiry <= 0.0:
        return _intrinsic(instrument, market)
    chunks = _reference_threads() if n_chunks is None else int(n_chunks)
    return _exact(instrument, market, n_paths, _lib.EXACT_PARALLEL, chunks, rng_mode, device)


def mc_european_sequential(instrument: VanillaOption, market: Market, n_paths: int,
                           n_steps: int = 1, rng_mode: "RngMode" = None,
                           device: int = 0) -> PricingResult:
    """mc_european_sequential (mc_parallel.rs:365-434) on the GPU: the
    reference's single stream (seed 0xDEAD_BEEF_CAFE_BABE), path by path."""
    rng_mode = RngMode.Reference if rng_mode is None else rng_mode
    if n_paths == 0:
        return _nan_result()
    if instrument.expiry <= 0.0:
        return _intrinsic(instrument, market)
    if instrument.exercise is not ExerciseStyle.European:
        return _nan_result()
    return _exact(instrument, market, n_paths, _lib.EXACT_SEQUENTIAL, 1, rng_mode, device)


def mc_european_with_arena(instrument: VanillaOption, market: Market, n_paths: int,
                           n_steps: int = 1, arena=None, rng_mode: "RngMode" = None,
                           device: int = 0) -> PricingResult:
    """mc_european_with_arena (mc_engine.rs:186-343) on the GPU: the
    reference's single stream (seed 42), path by path, mu = r - q. The arena
    (a host payoff buffer in the reference) is accepted and unused: payoffs
    are reduced on the device. An instrument that fails validation gives NaN."""
    rng_mode = RngMode.Reference if rng_mode is None else rng_mode
    if n_paths == 0:
        return _nan_result()
    try:
        instrument.validate()
    except InvalidInput:
        return _nan_result()
    if instrument.exercise is not ExerciseStyle.European:
        return _nan_result()
    if instrument.expiry <= 0.0:
        return _intrinsic(instrument, market)
    return _exact(instrument, market, n_paths, _lib.EXACT_ARENA, 1, rng_mode, device)


class BatchPricer:
    """Many contracts priced in one persistent-kernel launch, replayed through
    a captured hipGraph (ofg_batch_*)."""

    def __init__(self, opts: Sequence[_lib.OfgOption], cfg: _lib.OfgEngineCfg, use_graph=True,
                 device: int = 0):
        self._L = _lib.load()
        self._ctx = _lib.context(device)
        self.n = len(opts)
        self._arr = (_lib.OfgOption * self.n)(*opts)
        self._cfg = cfg
        self._b = self._L.ofg_batch_create(self._ctx, self._arr, self.n, C.byref(cfg),
                                           1 if use_graph else 0)
        if not self._b:
            raise InvalidInput(_lib.last_error())

    def run(self) -> float:
        ms = C.c_double()
        _raise(self._L.ofg_batch_run(self._b, C.byref(ms)))
        return ms.value

    def pricing_kernel_ms(self) -> float:
        return self._L.ofg_batch_pricing_kernel_ms(self._b)

    def results(self) -> List[_lib.OfgResult]:
        out = (_lib.OfgResult * self.n)()
        _raise(self._L.ofg_batch_results(self._b, out))
        return list(out)

    def close(self):
        if self._b:
            self._L.ofg_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(opts: Sequence[_lib.OfgOption], cfg: _lib.OfgEngineCfg, cache_path: Optional[str] = None,
             device: int = 0) -> Tuple[int, int]:
    L = _lib.load()
    ctx = _lib.context(device)
    arr = (_lib.OfgOption * len(opts))(*opts)
    b, it = C.c_int32(), C.c_int32()
    _raise(L.ofg_autotune(ctx, arr, len(opts), C.byref(cfg),
                          cache_path.encode() if cache_path else None, C.byref(b), C.byref(it)))
    return b.value, it.value


# ------------------------------------------------------- Longstaff-Schwartz ---
@dataclass(frozen=True)
class LsmDynamics:
    """LsmDynamics (engines/lsm/longstaff_schwartz.rs:25-39). `kind` is one of
    "gbm", "local_vol" (the market's flat vol as the surface), "heston"."""
    kind: str = "gbm"
    kappa: float = 0.0
    theta: float = 0.0
    xi: float = 0.0
    rho: float = 0.0
    v0: float = 0.0


_EXERCISE = {ExerciseStyle.European: _lib.EUROPEAN, ExerciseStyle.American: _lib.AMERICAN,
             ExerciseStyle.Bermudan: _lib.BERMUDAN}
_DYNAMICS = {"gbm": _lib.LSM_GBM, "local_vol": _lib.LSM_LOCAL_VOL, "heston": _lib.LSM_HESTON}


def lsm_option_struct(instrument, market: Market, num_paths: int, num_steps: int, seed: int,
                      keep: _Keep, dynamics: LsmDynamics = LsmDynamics()) -> _lib.OfgLsmOption:
    """Flattens (instrument, market, engine) into the C ABI's ofg_lsm_option."""
    o = _lib.OfgLsmOption()
    o.option_type = _lib.CALL if instrument.option_type is OptionType.Call else _lib.PUT
    o.spot, o.expiry, o.rate = market.spot, instrument.expiry, market.rate
    o.div_yield = market.effective_dividend_yield(instrument.expiry)
    if isinstance(instrument, VanillaOption):
        o.instrument = _lib.LSM_VANILLA
        o.exercise = _EXERCISE[instrument.exercise]
        o.strike = instrument.strike
        o.vol = market.vol_for(instrument.strike, instrument.expiry)
        dates = list(instrument.bermudan_dates) if instrument.exercise is ExerciseStyle.Bermudan \
            else []
        o.n_dates = len(dates)
        o.dates = keep.arr(dates) if dates else None
    elif isinstance(instrument, BermudanOption):
        o.instrument = _lib.LSM_BERMUDAN
        o.dynamics = _DYNAMICS[dynamics.kind]
        dates = list(instrument.exercise_dates)
        strikes = list(instrument.strike_schedule)
        o.n_dates = len(dates)
        o.dates = keep.arr(dates) if dates else None
        o.strikes = keep.arr(strikes) if strikes else None
        o.strike = strikes[-1] if strikes else 0.0
        o.vol = market.vol_for(o.strike, instrument.expiry)
        o.kappa, o.theta, o.xi, o.rho, o.v0 = dynamics.kappa, dynamics.theta, dynamics.xi, \
            dynamics.rho, dynamics.v0
    elif isinstance(instrument, BarrierOption):
        o.instrument = _lib.LSM_BARRIER
        o.strike = instrument.strike
        o.vol = market.vol_for(instrument.strike, instrument.expiry)
        o.barrier_dir = _lib.UP if instrument.barrier.direction is BarrierDirection.Up \
            else _lib.DOWN
        o.barrier_style = _lib.IN if instrument.barrier.style is BarrierStyle.In else _lib.OUT
        o.barrier_level = instrument.barrier.level
        o.rebate = instrument.barrier.rebate
    else:
        raise InvalidInput(f"unsupported instrument {type(instrument).__name__}")
    o.num_paths, o.num_steps, o.seed = int(num_paths), int(num_steps), int(seed) & (2**64 - 1)
    o._keep = keep
    return o


def lsm_price_raw(opts: Sequence[_lib.OfgLsmOption], cfg: _lib.OfgEngineCfg, device: int = 0,
                  want_boundary: bool = False):
    """Calls ofg_lsm_price. Returns the results, and with want_boundary also
    one list of boundary structs per option (empty unless BermudanOption)."""
    L = _lib.load()
    ctx = _lib.context(device)
    n = len(opts)
    arr = (_lib.OfgLsmOption * n)(*opts)
    out = (_lib.OfgResult * n)()
    bnd = blen = None
    if want_boundary:
        slots = sum(max(int(o.num_steps), 0) for o in opts)
        bnd = (_lib.OfgLsmBoundary * max(slots, 1))()
        blen = (C.c_int64 * max(n, 1))()
    _raise(L.ofg_lsm_price(ctx, arr, n, C.byref(cfg), out, bnd, blen))
    if not want_boundary:
        return list(out)
    per, slot = [], 0
    for i, o in enumerate(opts):
        per.append([bnd[slot + k] for k in range(blen[i])])
        slot += max(int(o.num_steps), 0)
    return list(out), per


def _lsm_result(r: _lib.OfgResult, instrument, dynamics: LsmDynamics) -> PricingResult:
    diag = Diagnostics()
    if not math.isnan(r.num_paths):
        diag.insert("num_paths", r.num_paths)
        diag.insert("num_steps", r.num_steps)
        if isinstance(instrument, BermudanOption):
            diag.insert("exercise_dates", float(len(instrument.effective_schedule())))
        if not math.isnan(r.vol):
            diag.insert("vol", r.vol)
    return PricingResult(price=r.price, stderr=r.stderr_, diagnostics=diag, kernel_ms=r.kernel_ms)


class LongstaffSchwartzEngine:
    """LongstaffSchwartzEngine (engines/lsm/longstaff_schwartz.rs:66-112)
    on the GPU: VanillaOption (European / American / Bermudan exercise),
    BermudanOption with GBM, flat local-vol or Heston dynamics and its
    exercise boundary, and BarrierOption by plain Monte Carlo. The default
    Reference RNG mode draws the reference engine's own single stream."""

    def __init__(self, num_paths: int, num_steps: int, seed: int):
        self.num_paths = num_paths
        self.num_steps = num_steps
        self.seed = seed
        self.dynamics = LsmDynamics()
        self.rng_mode = RngMode.Reference
        self.device = 0

    @staticmethod
    def new(num_paths: int, num_steps: int, seed: int) -> "LongstaffSchwartzEngine":
        return LongstaffSchwartzEngine(num_paths, num_steps, seed)

    def with_local_vol_dynamics(self) -> "LongstaffSchwartzEngine":
        self.dynamics = LsmDynamics("local_vol")
        return self

    def with_heston_dynamics(self, model: Heston) -> "LongstaffSchwartzEngine":
        """The spot drift is the market's r - q; the model's mu is ignored
        (longstaff_schwartz.rs:98-111)."""
        self.dynamics = LsmDynamics("heston", model.kappa, model.theta, model.xi, model.rho,
                                    model.v0)
        return self

    def with_rng_mode(self, mode: RngMode) -> "LongstaffSchwartzEngine":
        self.rng_mode = mode
        return self

    def with_device(self, device: int) -> "LongstaffSchwartzEngine":
        self.device = device
        return self

    def _structs(self, instruments, markets, keep):
        return [lsm_option_struct(ins, mk, self.num_paths, self.num_steps, self.seed, keep,
                                  self.dynamics) for ins, mk in zip(instruments, markets)]

    def price(self, instrument, market: Market) -> PricingResult:
        return self.price_batch([instrument], market)[0]

    def price_batch(self, instruments, market) -> List[PricingResult]:
        """Prices several instruments in one batched GPU sequence (one market
        for all, or one per instrument)."""
        markets = market if isinstance(market, (list, tuple)) else [market] * len(instruments)
        keep = _Keep()
        opts = self._structs(instruments, markets, keep)
        res = lsm_price_raw(opts, engine_cfg(rng_mode=self.rng_mode), self.device)
        return [_lsm_result(r, ins, self.dynamics) for r, ins in zip(res, instruments)]

    def price_bermudan_with_boundary(self, instrument: BermudanOption,
                                     market: Market) -> BermudanLsmOutput:
        keep = _Keep()
        opts = self._structs([instrument], [market], keep)
        res, bnd = lsm_price_raw(opts, engine_cfg(rng_mode=self.rng_mode), self.device,
                                 want_boundary=True)
        pts = [ExerciseBoundaryPoint(time=b.time, strike=b.strike,
                                     boundary_spot=b.boundary_spot if b.has_boundary else None,
                                     itm_paths=int(b.itm_paths),
                                     exercised_paths=int(b.exercised_paths)) for b in bnd[0]]
        return BermudanLsmOutput(result=_lsm_result(res[0], instrument, self.dynamics),
                                 exercise_boundary=pts)


def lsm_last_timing(device: int = 0) -> dict:
    """Device time of the last lsm_price_raw call on this device: path
    simulation (seed + path kernels) and backward rounds."""
    L = _lib.load()
    a, b, r = C.c_double(), C.c_double(), C.c_int32()
    _raise(L.ofg_lsm_last_timing(_lib.context(device), C.byref(a), C.byref(b), C.byref(r)))
    return {"paths_ms": a.value, "rounds_ms": b.value, "rounds": r.value}


def stream_jump(state: Sequence[int], k: int) -> List[int]:
    """ofg_stream_jump: the xoshiro256++ state k draws ahead (host-only)."""
    L = _lib.load()
    src = (C.c_uint64 * 4)(*[int(x) & (2**64 - 1) for x in state])
    dst = (C.c_uint64 * 4)()
    _raise(L.ofg_stream_jump(src, int(k) & (2**64 - 1), dst))
    return list(dst)
