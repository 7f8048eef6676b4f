"""Structured-product language on the GPU.

Mirrors src/dsl of rosssaunders/openferric: `parse_and_compile`
(mod.rs:24-28), `CompiledProduct` (ir.rs:141-151), `DslProduct`
(engine.rs:217-234), `AssetMarketData` / `MultiAssetMarket` (market.rs) and
`DslMonteCarloEngine` (engine.rs:237-1000) with `price_multi_asset`,
`greeks_multi_asset`, `extended_greeks_multi_asset` and
`cross_greeks_multi_asset`. Compilation runs in the library
(ofg_dsl_compile); pricing runs the compiled product per path on the GPU
(ofg_dsl_price), every bump of a Greeks call being one market of a single
launch on the base price's draws, as the reference reprices each bump with
the same seed.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .core import Diagnostics, Greeks, InvalidInput, PricingResult
from .engines import RngMode, _Keep, _raise, engine_cfg


class CompiledProduct:
    """A compiled product (ir.rs:141-151), owned by the library."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        info = _lib.OfgDslInfo()
        _raise(_lib.load().ofg_dsl_get_info(self._h, C.byref(info)))
        self.name = info.name.decode()
        self.notional = info.notional
        self.maturity = info.maturity
        self.num_underlyings = info.num_underlyings
        self.num_state_vars = info.n_state
        self.num_schedules = info.n_schedules

    def schedule_dates(self, k: int) -> List[float]:
        L = _lib.load()
        n = C.c_int32(0)
        _raise(L.ofg_dsl_schedule_dates(self._h, k, None, 0, C.byref(n)))
        buf = (C.c_double * max(1, n.value))()
        _raise(L.ofg_dsl_schedule_dates(self._h, k, buf, n.value, C.byref(n)))
        return list(buf[:n.value])

    def jit_compile(self, n_assets: Optional[int] = None, device: int = 0) -> float:
        """Compiles this product's programs into a native evaluation kernel
        for markets of n_assets assets (default: the product's underlyings)
        on `device`'s context: later prices of the product there run it, with
        the interpreter's results bit for bit (ofg_dsl_jit_compile; the
        counterpart of JitCompiledProgram::compile, jit.rs:168-200). Returns
        the compile time in ms (0 when already compiled)."""
        L = _lib.load()
        ms = C.c_double(0.0)
        n = self.num_underlyings if n_assets is None else int(n_assets)
        _raise(L.ofg_dsl_jit_compile(_lib.context(device), self._h, n, C.byref(ms)))
        return ms.value

    def jit_source(self, n_assets: Optional[int] = None) -> str:
        """The HIP source of this product's native evaluation kernel
        (ofg_dsl_jit_source)."""
        L = _lib.load()
        n = self.num_underlyings if n_assets is None else int(n_assets)
        size = C.c_size_t(0)
        _raise(L.ofg_dsl_jit_source(self._h, max(1, n), None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value + 1)
        _raise(L.ofg_dsl_jit_source(self._h, max(1, n), buf, size.value + 1, C.byref(size)))
        return buf.value.decode()

    def max_local_slots(self) -> int:
        info = _lib.OfgDslInfo()
        _raise(_lib.load().ofg_dsl_get_info(self._h, C.byref(info)))
        return info.n_locals

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().ofg_dsl_free(h)
            except Exception:
                pass
            self._h = C.c_void_p(0)


def jit_stats(device: int = 0) -> Tuple[int, int]:
    """(native evaluation kernels attached to `device`'s context, launches of
    them so far)."""
    k, n = C.c_int64(0), C.c_int64(0)
    _raise(_lib.load().ofg_dsl_jit_stats(_lib.context(device), C.byref(k), C.byref(n)))
    return k.value, n.value


def parse_and_compile(source: str) -> CompiledProduct:
    """parse_and_compile (src/dsl/mod.rs:24-28); raises InvalidInput with the
    DslError text."""
    L = _lib.load()
    h = C.c_void_p()
    _raise(L.ofg_dsl_compile(source.encode(), C.byref(h)))
    return CompiledProduct(h.value)


@dataclass
class DslProduct:
    """engine.rs:217-234"""
    product: CompiledProduct

    def instrument_type(self) -> str:
        return "DslProduct"


# ---- market data (market.rs) ------------------------------------------------

@dataclass
class Equity:
    spot: float
    vol: float
    dividend_yield: float = 0.0
    kind = _lib.DSL_EQUITY

    def initial_value(self) -> float:
        return self.spot


@dataclass
class Fx:
    spot: float
    vol: float
    domestic_rate: float
    foreign_rate: float
    kind = _lib.DSL_FX

    def initial_value(self) -> float:
        return self.spot


@dataclass
class Commodity:
    spot: float
    vol: float
    convenience_yield: float
    kappa: float
    mu: float
    kind = _lib.DSL_COMMODITY

    def initial_value(self) -> float:
        return self.spot


@dataclass
class Rate:
    initial_rate: float
    vol: float
    mean_reversion: float
    long_run_mean: float
    kind = _lib.DSL_RATE

    def initial_value(self) -> float:
        return self.initial_rate


def with_spot_bump(a, amount: float):
    """AssetMarketData::with_spot_bump (market.rs:54-98)"""
    if isinstance(a, Rate):
        return replace(a, initial_rate=a.initial_rate + amount)
    return replace(a, spot=a.spot + amount)


def with_vol_bump(a, amount: float):
    """AssetMarketData::with_vol_bump (market.rs:100-144)"""
    return replace(a, vol=a.vol + amount)


class AssetMarketData:
    """The reference's enum variants (market.rs:4-30)."""
    Equity = Equity
    Fx = Fx
    Commodity = Commodity
    Rate = Rate


@dataclass
class MultiAssetMarket:
    """market.rs:158-168"""
    assets: List
    correlation: List[List[float]]
    rate: float

    @staticmethod
    def single(spot: float, vol: float, rate: float, dividend_yield: float) -> "MultiAssetMarket":
        return MultiAssetMarket([Equity(spot, vol, dividend_yield)], [[1.0]], rate)

    def initial_spots(self) -> List[float]:
        return [a.initial_value() for a in self.assets]


def _market_struct(m: MultiAssetMarket, keep: _Keep) -> _lib.OfgDslMarket:
    n = len(m.assets)
    arr = (_lib.OfgDslAsset * max(1, n))()
    for i, a in enumerate(m.assets):
        s = arr[i]
        s.kind = a.kind
        s.vol = a.vol
        if isinstance(a, Rate):
            s.spot, s.mean_reversion, s.long_run_mean = a.initial_rate, a.mean_reversion, \
                a.long_run_mean
        else:
            s.spot = a.spot
        if isinstance(a, Equity):
            s.dividend_yield = a.dividend_yield
        if isinstance(a, Fx):
            s.domestic_rate, s.foreign_rate = a.domestic_rate, a.foreign_rate
        if isinstance(a, Commodity):
            s.convenience_yield, s.kappa, s.mu = a.convenience_yield, a.kappa, a.mu
    keep.refs.append(arr)
    cm = np.asarray(m.correlation, dtype=np.float64)
    if n and cm.shape != (n, n):
        rows = len(m.correlation)
        if rows != n:
            raise InvalidInput(f"correlation matrix rows ({rows}) must match number of assets "
                               f"({n})")
        for i, row in enumerate(m.correlation):
            if len(row) != n:
                raise InvalidInput(f"correlation matrix row {i} has {len(row)} columns, "
                                   f"expected {n}")
    st = _lib.OfgDslMarket()
    st.n_assets = n
    st.assets = C.cast(arr, C.POINTER(_lib.OfgDslAsset))
    st.corr = keep.arr(cm.ravel()) if n else None
    st.rate = m.rate
    return st


@dataclass
class ExtendedGreeks:
    """engine.rs:1034-1044"""
    delta: float
    gamma: float
    vega: float
    theta: float
    rho: float
    vanna: float
    volga: float


@dataclass
class CrossGreeks:
    """engine.rs:1046-1051"""
    cross_gamma: float
    corr_sens: float


@dataclass
class DslMonteCarloEngine:
    """engine.rs:237-253. rng_mode Reference reproduces the default build's
    single stream; Fast uses per-path streams; External takes normals
    [step][asset][path]."""
    num_paths: int
    num_steps: int
    seed: int
    rng_mode: int = RngMode.Reference
    normals: Optional[np.ndarray] = None
    device: int = 0
    last_kernel_ms: float = field(default=0.0, repr=False)

    def price_markets(self, product: CompiledProduct,
                      markets: Sequence[MultiAssetMarket]) -> List[PricingResult]:
        """Prices one product under several markets in one launch, all on the
        same draws."""
        L = _lib.load()
        ctx = _lib.context(self.device)
        keep = _Keep()
        ms = (_lib.OfgDslMarket * len(markets))(*[_market_struct(m, keep) for m in markets])
        eng = _lib.OfgDslEngine(int(self.num_paths), int(self.num_steps),
                                int(self.seed) & (2 ** 64 - 1))
        z = None
        if self.normals is not None:
            z = np.ascontiguousarray(self.normals, dtype=np.float64).ravel()
        cfg = engine_cfg(rng_mode=self.rng_mode, normals1=z)
        out = (_lib.OfgResult * len(markets))()
        _raise(L.ofg_dsl_price(ctx, product._h, ms, len(markets), C.byref(eng), C.byref(cfg), out))
        res = []
        for r in out:
            d = Diagnostics()
            d.insert("num_paths", r.num_paths)
            d.insert("num_steps", r.num_steps)
            res.append(PricingResult(price=r.price, stderr=r.stderr_, diagnostics=d,
                                     kernel_ms=r.kernel_ms))
        self.last_kernel_ms = out[0].kernel_ms if len(markets) else 0.0
        return res

    def price_multi_asset(self, product: CompiledProduct,
                          market: MultiAssetMarket) -> PricingResult:
        """price_multi_asset (engine.rs:255-353)"""
        return self.price_markets(product, [market])[0]

    def price(self, instrument: DslProduct, market: MultiAssetMarket) -> PricingResult:
        return self.price_multi_asset(instrument.product, market)

    @staticmethod
    def _check_index(market, i):
        n = len(market.assets)
        if i >= n:
            raise InvalidInput(f"asset_index {i} out of range (have {n} assets)")

    @staticmethod
    def _bumped(market, idx, spot=0.0, vol=0.0):
        assets = list(market.assets)
        a = assets[idx]
        if spot:
            a = with_spot_bump(a, spot)
        if vol:
            a = with_vol_bump(a, vol)
        assets[idx] = a
        return MultiAssetMarket(assets, market.correlation, market.rate)

    def greeks_multi_asset(self, product: CompiledProduct, market: MultiAssetMarket,
                           asset_index: int) -> Greeks:
        """greeks_multi_asset (engine.rs:785-837): base, spot +-1 %, vol +0.01
        and rate +1bp in one launch."""
        self._check_index(market, asset_index)
        h = 0.01 * market.assets[asset_index].initial_value()
        scen = [market, self._bumped(market, asset_index, spot=h),
                self._bumped(market, asset_index, spot=-h),
                self._bumped(market, asset_index, vol=0.01),
                MultiAssetMarket(market.assets, market.correlation, market.rate + 0.0001)]
        base, up, dn, vg, rr = [r.price for r in self.price_markets(product, scen)]
        return Greeks(delta=(up - dn) / (2.0 * h), gamma=(up - 2.0 * base + dn) / (h * h),
                      vega=(vg - base) / 0.01, theta=0.0, rho=(rr - base) / 0.0001)

    def extended_greeks_multi_asset(self, product: CompiledProduct, market: MultiAssetMarket,
                                    asset_index: int, base_price: float) -> ExtendedGreeks:
        """extended_greeks_multi_asset (engine.rs:839-924)"""
        self._check_index(market, asset_index)
        h = 0.01 * market.assets[asset_index].initial_value()
        v = 0.01
        b = self._bumped
        scen = [b(market, asset_index, spot=h), b(market, asset_index, spot=-h),
                b(market, asset_index, vol=v), b(market, asset_index, vol=-v),
                b(market, asset_index, spot=h, vol=v), b(market, asset_index, spot=h, vol=-v),
                b(market, asset_index, spot=-h, vol=v), b(market, asset_index, spot=-h, vol=-v),
                MultiAssetMarket(market.assets, market.correlation, market.rate + 0.0001)]
        su, sd, vu, vd, uu, ud, du, dd, rr = [r.price for r in self.price_markets(product, scen)]
        return ExtendedGreeks(
            delta=(su - sd) / (2.0 * h), gamma=(su - 2.0 * base_price + sd) / (h * h),
            vega=(vu - vd) / (2.0 * v), theta=0.0, rho=(rr - base_price) / 0.0001,
            vanna=(uu - ud - du + dd) / (4.0 * h * v), volga=(vu - 2.0 * base_price + vd) / (v * v))

    def cross_greeks_multi_asset(self, product: CompiledProduct, market: MultiAssetMarket,
                                 asset_i: int, asset_j: int, base_price: float) -> CrossGreeks:
        """cross_greeks_multi_asset (engine.rs:926-1000)"""
        n = len(market.assets)
        if asset_i >= n or asset_j >= n:
            raise InvalidInput(f"asset indices ({asset_i}, {asset_j}) out of range (have {n} "
                               "assets)")
        bi = 0.01 * market.assets[asset_i].initial_value()
        bj = 0.01 * market.assets[asset_j].initial_value()

        def two(si, sj):
            assets = list(market.assets)
            assets[asset_i] = with_spot_bump(market.assets[asset_i], si)
            assets[asset_j] = with_spot_bump(market.assets[asset_j], sj)
            return MultiAssetMarket(assets, market.correlation, market.rate)

        rho = market.correlation[asset_i][asset_j]
        r_up, r_dn = min(rho + 0.01, 1.0), max(rho - 0.01, -1.0)

        def corr_with(v):
            c = [list(r) for r in market.correlation]
            c[asset_i][asset_j] = v
            c[asset_j][asset_i] = v
            return MultiAssetMarket(market.assets, c, market.rate)

        scen = [two(bi, bj), two(bi, -bj), two(-bi, bj), two(-bi, -bj), corr_with(r_up),
                corr_with(r_dn)]
        pp, pm, mp, mm, cu, cd = [r.price for r in self.price_markets(product, scen)]
        return CrossGreeks(cross_gamma=(pp - pm - mp + mm) / (4.0 * bi * bj),
                           corr_sens=(cu - cd) / (r_up - r_dn))
