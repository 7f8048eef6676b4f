"""Multi-asset contracts on the GPU: worst-of autocallables, Phoenix notes and
basket options.

Mirrors src/instruments/autocallable.rs (Autocallable, PhoenixAutocallable)
and src/pricing/autocallable.rs (price_autocallable,
autocallable_sensitivities, price_phoenix_autocallable,
phoenix_autocallable_sensitivities) over the C ABI's ofg_autocall_price, and
src/instruments/basket.rs (BasketOption, OutperformanceBasketOption,
QuantoBasketOption) with src/pricing/basket.rs (price_basket_mc,
basket_sensitivities, price_outperformance_basket_mc,
price_quanto_basket_mc) over ofg_basket_price.
Draws: the reference walks one xoshiro256++ stream whose position depends on
when earlier paths were called, which no parallel simulator can reproduce;
`rng_mode` is Fast (per-path streams from the reference's seed 94210) or
External (caller normals, [path][step][underlying] per contract). Baskets
draw a fixed count per path, so they also take the reference's own stream
(Reference, the default), split across paths by jump-ahead.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .core import Diagnostics, Greeks, InvalidInput, PricingResult
from .engines import RngMode, _Keep, _raise


@dataclass
class Autocallable:
    """instruments/autocallable.rs:18-34"""
    underlyings: Sequence[int]
    notional: float
    autocall_dates: Sequence[float]
    autocall_barrier: float
    coupon_rate: float
    ki_barrier: float
    ki_strike: float
    maturity: float

    def validate(self) -> None:
        if not self.underlyings:
            raise InvalidInput("autocallable underlyings cannot be empty")
        if self.notional <= 0.0:
            raise InvalidInput("autocallable notional must be > 0")
        if self.maturity <= 0.0:
            raise InvalidInput("autocallable maturity must be > 0")
        if not self.autocall_dates:
            raise InvalidInput("autocallable autocall_dates cannot be empty")
        if self.coupon_rate < 0.0:
            raise InvalidInput("autocallable coupon_rate must be >= 0")
        if self.autocall_barrier <= 0.0:
            raise InvalidInput("autocallable autocall_barrier must be > 0")
        if self.ki_barrier <= 0.0:
            raise InvalidInput("autocallable ki_barrier must be > 0")
        if self.ki_strike <= 0.0:
            raise InvalidInput("autocallable ki_strike must be > 0")
        if any(t <= 0.0 or t > self.maturity for t in self.autocall_dates):
            raise InvalidInput("autocallable dates must lie in (0, maturity]")
        if any(b <= a for a, b in zip(self.autocall_dates, self.autocall_dates[1:])):
            raise InvalidInput("autocallable dates must be strictly increasing")
        if len(set(self.underlyings)) != len(self.underlyings):
            raise InvalidInput("autocallable underlyings must be unique")

    def instrument_type(self) -> str:
        return "Autocallable"


@dataclass
class PhoenixAutocallable:
    """instruments/autocallable.rs:113-157"""
    underlyings: Sequence[int]
    notional: float
    autocall_dates: Sequence[float]
    autocall_barrier: float
    coupon_barrier: float
    coupon_rate: float
    memory: bool
    ki_barrier: float
    ki_strike: float
    maturity: float

    def validate(self) -> None:
        Autocallable(self.underlyings, self.notional, self.autocall_dates,
                     self.autocall_barrier, self.coupon_rate, self.ki_barrier, self.ki_strike,
                     self.maturity).validate()
        if self.coupon_barrier <= 0.0:
            raise InvalidInput("phoenix coupon_barrier must be > 0")

    def instrument_type(self) -> str:
        return "PhoenixAutocallable"


@dataclass
class AutocallableSensitivities:
    """pricing/autocallable.rs:25-33"""
    delta: List[float]
    vega: float
    cega: float


def autocall_struct(contract, spots, vols, corr_matrix, r, q, n_paths, n_steps,
                    keep: _Keep) -> _lib.OfgAutocall:
    """Flattens a contract and its market inputs into the C ABI's ofg_autocall."""
    a = _lib.OfgAutocall()
    phoenix = isinstance(contract, PhoenixAutocallable)
    a.phoenix = int(phoenix)
    a.memory = int(bool(getattr(contract, "memory", False)))
    und = np.ascontiguousarray(list(contract.underlyings), dtype=np.int32)
    keep.refs.append(und)
    a.underlyings = und.ctypes.data_as(C.POINTER(C.c_int32))
    a.n_underlyings = len(und)
    a.autocall_dates = keep.arr(list(contract.autocall_dates)) if contract.autocall_dates else None
    a.n_dates = len(contract.autocall_dates)
    a.notional, a.maturity = contract.notional, contract.maturity
    a.autocall_barrier, a.coupon_rate = contract.autocall_barrier, contract.coupon_rate
    a.coupon_barrier = contract.coupon_barrier if phoenix else 0.0
    a.ki_barrier, a.ki_strike = contract.ki_barrier, contract.ki_strike
    a.n_assets = len(spots)
    a.spots = keep.arr(list(spots)) if len(spots) else None
    if len(vols) != len(spots):
        raise InvalidInput("autocallable spots and vols lengths must match")
    a.vols = keep.arr(list(vols)) if len(vols) else None
    cm = np.asarray(corr_matrix, dtype=np.float64)
    if cm.shape != (len(spots), len(spots)):
        raise InvalidInput("autocallable correlation matrix dimensions must match assets")
    a.corr = keep.arr(cm.ravel()) if len(spots) else None
    a.rate, a.div_yield = r, q
    a.num_paths, a.num_steps = int(n_paths), int(n_steps)
    a._keep = keep
    return a


def autocall_price_raw(structs, rng_mode=RngMode.Fast, normals=None, with_greeks=False,
                       device: int = 0):
    """Calls ofg_autocall_price. normals: per contract [path][step][underlying]
    arrays (External mode). Returns (results, sensitivities)."""
    from .engines import engine_cfg
    L = _lib.load()
    ctx = _lib.context(device)
    n = len(structs)
    z = None
    if normals is not None:
        # the C ABI's layout: [step][underlying][path] per contract
        z = np.concatenate([np.ascontiguousarray(np.transpose(np.asarray(x), (1, 2, 0))).ravel()
                            for x in normals])
    cfg = engine_cfg(rng_mode=rng_mode, normals1=z)
    arr = (_lib.OfgAutocall * n)(*structs)
    out = (_lib.OfgResult * n)()
    sens = (_lib.OfgAutocallSens * n)()
    _raise(L.ofg_autocall_price(ctx, arr, n, C.byref(cfg), 1 if with_greeks else 0, out, sens))
    return list(out), list(sens)


def _result(r: _lib.OfgResult, n_dates: int, greeks: bool) -> PricingResult:
    diag = Diagnostics()
    if r.has_stderr:
        diag.insert("num_paths", r.num_paths)
        diag.insert("num_steps", r.num_steps)
        diag.insert("observation_count", float(n_dates))
    res = PricingResult(price=r.price, stderr=r.stderr_ if r.has_stderr else None,
                        diagnostics=diag, kernel_ms=r.kernel_ms)
    if greeks and r.has_greeks:
        res.greeks = Greeks(delta=r.delta, gamma=0.0, vega=r.vega, theta=0.0, rho=r.rho)
    return res


def price_autocallable(autocall: Autocallable, spots, vols, corr_matrix, r, q, n_paths, n_steps,
                       rng_mode=RngMode.Fast, normals=None) -> PricingResult:
    """price_autocallable (pricing/autocallable.rs:52-99): price, stderr and
    the bump-and-reprice Greeks (delta = sum of deltas, vega, rho = cega)."""
    keep = _Keep()
    s = autocall_struct(autocall, spots, vols, corr_matrix, r, q, n_paths, n_steps, keep)
    res, _ = autocall_price_raw([s], rng_mode, None if normals is None else [normals], True)
    return _result(res[0], len(autocall.autocall_dates), True)


def price_phoenix_autocallable(phoenix: PhoenixAutocallable, spots, vols, corr_matrix, r, q,
                               n_paths, n_steps, rng_mode=RngMode.Fast,
                               normals=None) -> PricingResult:
    """price_phoenix_autocallable (pricing/autocallable.rs:135-190)."""
    return price_autocallable(phoenix, spots, vols, corr_matrix, r, q, n_paths, n_steps,
                              rng_mode, normals)


def autocallable_sensitivities(autocall, spots, vols, corr_matrix, r, q, n_paths, n_steps,
                               rng_mode=RngMode.Fast, normals=None) -> AutocallableSensitivities:
    """autocallable_sensitivities / phoenix_autocallable_sensitivities
    (pricing/autocallable.rs:101-133, 192-223): all bumps in one launch."""
    autocall.validate()
    keep = _Keep()
    s = autocall_struct(autocall, spots, vols, corr_matrix, r, q, n_paths, n_steps, keep)
    res, sens = autocall_price_raw([s], rng_mode, None if normals is None else [normals], True)
    if not res[0].has_stderr or not sens[0].ok:
        raise InvalidInput(_lib.last_error() or "autocallable sensitivities failed")
    sv = sens[0]
    return AutocallableSensitivities([sv.delta[k] for k in range(sv.n_delta)], sv.vega, sv.cega)


phoenix_autocallable_sensitivities = autocallable_sensitivities


# ---------------------------------------------------------------------------
# Basket options (src/instruments/basket.rs, src/pricing/basket.rs)
# ---------------------------------------------------------------------------

class BasketType:
    """instruments/basket.rs:14-23"""
    Average = _lib.BASKET_AVERAGE
    BestOf = _lib.BASKET_BEST_OF
    WorstOf = _lib.BASKET_WORST_OF


BASKET_MC_SEED = 3_148_159   # pricing/basket.rs:27


@dataclass
class BasketOption:
    """instruments/basket.rs:25-60"""
    weights: Sequence[float]
    strike: float
    maturity: float
    is_call: bool
    basket_type: int = BasketType.Average

    def validate(self) -> None:
        if self.strike < 0.0:
            raise InvalidInput("basket strike must be >= 0")
        if self.maturity < 0.0:
            raise InvalidInput("basket maturity must be >= 0")
        if any(not np.isfinite(w) for w in self.weights):
            raise InvalidInput("basket weights must be finite")
        if self.basket_type == BasketType.Average and len(self.weights) == 0:
            raise InvalidInput("average basket requires non-empty weights")

    def instrument_type(self) -> str:
        return "BasketOption"


@dataclass
class BasketSensitivities:
    """pricing/basket.rs:57-65"""
    delta: List[float]
    vega: float
    cega: float


def basket_struct(basket: BasketOption, spots, vols, corr_matrix, r, dividends, n_paths,
                  keep: _Keep, seed: int = BASKET_MC_SEED) -> _lib.OfgBasket:
    """Flattens a basket and its market inputs into the C ABI's ofg_basket."""
    b = _lib.OfgBasket()
    b.basket_type = int(basket.basket_type)
    b.is_call = int(bool(basket.is_call))
    b.n_weights = len(basket.weights)
    b.weights = keep.arr(list(basket.weights)) if len(basket.weights) else None
    b.strike, b.maturity = basket.strike, basket.maturity
    n = len(spots)
    if len(vols) != n or len(dividends) != n:
        raise InvalidInput("basket spots/vols/dividends lengths must match")
    cm = np.asarray(corr_matrix, dtype=np.float64)
    if cm.shape != (n, n):
        raise InvalidInput("correlation matrix dimensions must match asset count")
    b.n_assets = n
    b.spots = keep.arr(list(spots)) if n else None
    b.vols = keep.arr(list(vols)) if n else None
    b.dividends = keep.arr(list(dividends)) if n else None
    b.corr = keep.arr(cm.ravel()) if n else None
    b.rate = r
    b.num_paths = int(n_paths)
    b.seed = int(seed)
    b._keep = keep
    return b


def basket_price_raw(structs, rng_mode=RngMode.Reference, normals=None, with_sens=False,
                     device: int = 0):
    """Calls ofg_basket_price. normals: per basket [path][draw] arrays
    (External mode; draws = assets, or factors then assets under a factor
    model). Returns (results, sensitivities)."""
    from .engines import engine_cfg
    L = _lib.load()
    ctx = _lib.context(device)
    n = len(structs)
    z = None
    if normals is not None:
        # the C ABI's layout: [asset][path] per basket
        z = np.concatenate([np.ascontiguousarray(np.asarray(x, dtype=np.float64).T).ravel()
                            for x in normals])
    cfg = engine_cfg(rng_mode=rng_mode, normals1=z)
    arr = (_lib.OfgBasket * n)(*structs)
    out = (_lib.OfgResult * n)()
    sens = (_lib.OfgBasketSens * n)()
    _raise(L.ofg_basket_price(ctx, arr, n, C.byref(cfg), 1 if with_sens else 0, out, sens))
    return list(out), list(sens)


def price_basket_mc(basket: BasketOption, spots, vols, corr_matrix, r, dividends, n_paths,
                    rng_mode=RngMode.Reference, normals=None) -> PricingResult:
    """price_basket_mc (pricing/basket.rs:99-160): Gaussian copula, one exact
    terminal step, the reference's seed 3148159 and, by default, its draws."""
    keep = _Keep()
    s = basket_struct(basket, spots, vols, corr_matrix, r, dividends, n_paths, keep)
    res, sens = basket_price_raw([s], rng_mode, None if normals is None else [normals])
    r0 = res[0]
    diag = Diagnostics()
    if r0.has_stderr:
        diag.insert("num_paths", r0.num_paths)
        diag.insert("rho", sens[0].rho)
    return PricingResult(price=r0.price, stderr=r0.stderr_ if r0.has_stderr else None,
                         diagnostics=diag, kernel_ms=r0.kernel_ms)


def basket_sensitivities(basket: BasketOption, spots, vols, corr_matrix, r, dividends, n_paths,
                         rng_mode=RngMode.Reference, normals=None) -> BasketSensitivities:
    """basket_sensitivities (pricing/basket.rs:456-539): every bump on the
    base price's draws, all in one launch."""
    keep = _Keep()
    s = basket_struct(basket, spots, vols, corr_matrix, r, dividends, n_paths, keep)
    res, sens = basket_price_raw([s], rng_mode, None if normals is None else [normals], True)
    if not res[0].has_stderr or not sens[0].ok:
        raise InvalidInput(_lib.last_error() or "basket sensitivities failed")
    sv = sens[0]
    return BasketSensitivities([sv.delta[k] for k in range(sv.n_delta)], sv.vega, sv.cega)


@dataclass
class FactorCorrelationModel:
    """math/correlation.rs:62-66: OneFactor (loadings = one beta per asset)
    or MultiFactor (loadings = one row of factor loadings per asset)."""
    loadings: Sequence
    one_factor: bool = True

    @staticmethod
    def OneFactor(loadings) -> "FactorCorrelationModel":
        return FactorCorrelationModel(list(loadings), True)

    @staticmethod
    def MultiFactor(loadings) -> "FactorCorrelationModel":
        return FactorCorrelationModel([list(r) for r in loadings], False)

    def n_assets(self) -> int:
        return len(self.loadings)

    def correlation_matrix(self):
        """correlation_matrix (correlation.rs:123-161): implied correlations."""
        L = np.asarray(self.loadings, dtype=np.float64)
        if self.one_factor:
            L = L.reshape(-1, 1)
        c = np.clip(L @ L.T, -1.0, 1.0)
        np.fill_diagonal(c, 1.0)
        return c.tolist()


@dataclass(frozen=True)
class _StudentTCopula:
    degrees_of_freedom: int


class BasketCopula:
    """pricing/basket.rs:32-46: BasketCopula.Gaussian or
    BasketCopula.StudentT(degrees_of_freedom) (at least 2, checked when the
    first path is drawn, as math/correlation.rs:559-562)."""
    Gaussian = "gaussian"

    @staticmethod
    def StudentT(degrees_of_freedom: int) -> _StudentTCopula:
        return _StudentTCopula(int(degrees_of_freedom))


def _apply_copula(b, copula) -> None:
    if copula == BasketCopula.Gaussian:
        b.copula, b.copula_dof = _lib.COPULA_GAUSSIAN, 0
    elif isinstance(copula, _StudentTCopula):
        b.copula, b.copula_dof = _lib.COPULA_STUDENT_T, copula.degrees_of_freedom
    else:
        raise InvalidInput("copula must be BasketCopula.Gaussian or BasketCopula.StudentT(dof)")


def price_basket_mc_with_copula(basket: BasketOption, spots, vols, corr_matrix, r, dividends,
                                n_paths, copula=BasketCopula.Gaussian,
                                rng_mode=RngMode.Reference, normals=None) -> PricingResult:
    """price_basket_mc_with_copula (pricing/basket.rs:120-160). Under
    StudentT(dof) each path draws its asset normals, then dof more for the
    chi-square (normals, External mode: [path][asset + dof])."""
    keep = _Keep()
    s = basket_struct(basket, spots, vols, corr_matrix, r, dividends, n_paths, keep)
    _apply_copula(s, copula)
    res, sens = basket_price_raw([s], rng_mode, None if normals is None else [normals])
    r0 = res[0]
    diag = Diagnostics()
    if r0.has_stderr:
        diag.insert("num_paths", r0.num_paths)
        diag.insert("rho", sens[0].rho)
    return PricingResult(price=r0.price, stderr=r0.stderr_ if r0.has_stderr else None,
                         diagnostics=diag, kernel_ms=r0.kernel_ms)


def price_basket_mc_with_factor_model(basket: BasketOption, spots, vols,
                                      factor_model: FactorCorrelationModel, r, dividends, n_paths,
                                      copula=BasketCopula.Gaussian, rng_mode=RngMode.Reference,
                                      normals=None) -> PricingResult:
    """price_basket_mc_with_factor_model (pricing/basket.rs:162-205).
    normals (External): [path][draw], draws = factors, one per asset, then
    the Student-t copula's dof."""
    keep = _Keep()
    n = len(spots)
    b = basket_struct(basket, spots, vols, np.eye(n) if n else [], r, dividends, n_paths, keep)
    b.corr = None
    rows = np.asarray(factor_model.loadings, dtype=np.float64)
    k = 1 if factor_model.one_factor else (rows.shape[1] if rows.ndim == 2 and rows.size else 0)
    b.factor_kind = 1 if factor_model.one_factor else 2
    b.n_loading_rows, b.n_factors = len(factor_model.loadings), k
    b.loadings = keep.arr(rows.ravel()) if rows.size else None
    _apply_copula(b, copula)
    res, sens = basket_price_raw([b], rng_mode, None if normals is None else [normals])
    r0 = res[0]
    diag = Diagnostics()
    if r0.has_stderr:
        diag.insert("num_paths", r0.num_paths)
        diag.insert("rho", 0.0)
    return PricingResult(price=r0.price, stderr=r0.stderr_ if r0.has_stderr else None,
                         diagnostics=diag, kernel_ms=r0.kernel_ms)


@dataclass
class OutperformanceBasketOption:
    """instruments/basket.rs:70-118"""
    leader_index: int
    lagger_weights: Sequence[float]
    strike: float
    maturity: float
    is_call: bool = True   # option_type

    def instrument_type(self) -> str:
        return "OutperformanceBasketOption"


@dataclass
class QuantoBasketOption:
    """instruments/basket.rs:120-165"""
    basket: BasketOption
    fx_rate: float
    fx_vol: float
    asset_fx_corr: Sequence[float]
    domestic_rate: float
    foreign_rate: float

    def instrument_type(self) -> str:
        return "QuantoBasketOption"


def _variant_result(r) -> PricingResult:
    # the reference's outperformance / quanto results carry no diagnostics
    return PricingResult(price=r.price, stderr=r.stderr_ if r.has_stderr else None,
                         diagnostics=Diagnostics(), kernel_ms=r.kernel_ms)


def price_outperformance_basket_mc(option: OutperformanceBasketOption, spots, vols, corr_matrix,
                                   r, dividends, n_paths, rng_mode=RngMode.Reference,
                                   normals=None) -> PricingResult:
    """price_outperformance_basket_mc (pricing/basket.rs:288-366)."""
    keep = _Keep()
    b = basket_struct(BasketOption(list(option.lagger_weights), option.strike, option.maturity,
                                   option.is_call, BasketType.Average), spots, vols, corr_matrix,
                      r, dividends, n_paths, keep)
    b.product, b.leader_index = 1, int(option.leader_index)
    res, _ = basket_price_raw([b], rng_mode, None if normals is None else [normals])
    return _variant_result(res[0])


def price_quanto_basket_mc(option: QuantoBasketOption, spots, vols, corr_matrix, dividends,
                           n_paths, rng_mode=RngMode.Reference, normals=None) -> PricingResult:
    """price_quanto_basket_mc (pricing/basket.rs:368-445)."""
    keep = _Keep()
    b = basket_struct(option.basket, spots, vols, corr_matrix, 0.0, dividends, n_paths, keep)
    b.product = 2
    b.fx_rate, b.fx_vol = option.fx_rate, option.fx_vol
    b.domestic_rate, b.foreign_rate = option.domestic_rate, option.foreign_rate
    b.n_asset_fx_corr = len(option.asset_fx_corr)
    b.asset_fx_corr = keep.arr(list(option.asset_fx_corr)) if len(option.asset_fx_corr) else None
    res, _ = basket_price_raw([b], rng_mode, None if normals is None else [normals])
    return _variant_result(res[0])


def validate_or_repair_correlation_matrix(corr_matrix):
    """validate_or_repair_correlation_matrix (math/correlation.rs:340-358)
    with the default PsdProjectionConfig: (matrix, was_repaired)."""
    L = _lib.load()
    c = np.ascontiguousarray(np.asarray(corr_matrix, dtype=np.float64))
    n = c.shape[0]
    if c.shape != (n, n):
        raise InvalidInput("correlation matrix dimensions must match asset count")
    outm = np.zeros_like(c)
    rep = C.c_int32(0)
    _raise(L.ofg_repair_correlation(c.ctypes.data_as(C.POINTER(C.c_double)), n,
                                    outm.ctypes.data_as(C.POINTER(C.c_double)), C.byref(rep)))
    return outm, bool(rep.value)


@dataclass(frozen=True)
class CorrelationStressScenario:
    """CorrelationStressScenario (math/correlation.rs:49-60); build with the
    constructors below."""
    kind: int
    value: float
    i: int = 0
    j: int = 0

    @staticmethod
    def ScaleOffDiagonal(factor: float) -> "CorrelationStressScenario":
        return CorrelationStressScenario(_lib.STRESS_SCALE_OFF_DIAGONAL, float(factor))

    @staticmethod
    def AdditiveShift(shift: float) -> "CorrelationStressScenario":
        return CorrelationStressScenario(_lib.STRESS_ADDITIVE_SHIFT, float(shift))

    @staticmethod
    def FloorOffDiagonal(floor: float) -> "CorrelationStressScenario":
        return CorrelationStressScenario(_lib.STRESS_FLOOR_OFF_DIAGONAL, float(floor))

    @staticmethod
    def CapOffDiagonal(cap: float) -> "CorrelationStressScenario":
        return CorrelationStressScenario(_lib.STRESS_CAP_OFF_DIAGONAL, float(cap))

    @staticmethod
    def OverridePair(i: int, j: int, value: float) -> "CorrelationStressScenario":
        return CorrelationStressScenario(_lib.STRESS_OVERRIDE_PAIR, float(value), int(i), int(j))


def stressed_correlation_matrix(corr_matrix, scenarios):
    """stressed_correlation_matrix (pricing/basket.rs:447-454): the scenarios
    applied in order, then the Higham repair when the result is not PSD."""
    L = _lib.load()
    c = np.ascontiguousarray(np.asarray(corr_matrix, dtype=np.float64))
    n = c.shape[0] if c.ndim == 2 else 0
    if c.shape != (n, n) or n == 0:
        raise InvalidInput("correlation matrix dimensions must match asset count")
    arr = (_lib.OfgCorrStress * max(len(scenarios), 1))()
    for k, sc in enumerate(scenarios):
        arr[k].kind, arr[k].i, arr[k].j, arr[k].value = sc.kind, sc.i, sc.j, sc.value
    out = np.zeros_like(c)
    _raise(L.ofg_stress_correlation(c.ctypes.data_as(C.POINTER(C.c_double)), n, arr,
                                    len(scenarios), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out
