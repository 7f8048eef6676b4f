"""Spread options on the GPU.

Mirrors SpreadOption (src/instruments/spread.rs:15-78) and
SpreadMonteCarloEngine (src/engines/monte_carlo/spread_mc.rs:21-175) over
the C ABI's ofg_spread_price, and the Python binding McEngine.spread_price
(python/src/engines.rs:1392-1430). Sample i draws z1 then z2 from its own
stream seeded with seed + 7919 i: xoshiro256++ (the default) or Pcg64 per
path as the reference. Randomized streams (with_randomized_streams,
ThreadRng) take a fresh random base seed per call instead of a fresh random
seed per sample: both are non-reproducible draws of independent streams.
FastRngKind.StdRng (rand's ChaCha12) has no GPU stream and is refused.
FastRngKind and rng_kind_from_str are re-exported from engines.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .core import Diagnostics, InvalidInput, Market, PricingResult
from .engines import FastRngKind, RngMode, _StreamOptions, _raise, rng_kind_from_str


@dataclass
class SpreadOption:
    """Two-asset spread option input bundle (spread.rs:15-26)."""
    s1: float
    s2: float
    k: float
    vol1: float
    vol2: float
    rho: float
    q1: float
    q2: float
    r: float
    t: float

    def validate(self) -> None:
        """spread.rs:31-58 (NaN fields pass where they pass there)."""
        if self.s1 <= 0.0 or self.s2 <= 0.0:
            raise InvalidInput("spread spots s1 and s2 must be > 0")
        if self.k < 0.0:
            raise InvalidInput("spread strike k must be >= 0")
        if self.vol1 < 0.0 or self.vol2 < 0.0:
            raise InvalidInput("spread volatilities vol1 and vol2 must be >= 0")
        if not (-1.0 <= self.rho <= 1.0):
            raise InvalidInput("spread correlation rho must be in [-1, 1]")
        if self.t < 0.0:
            raise InvalidInput("spread maturity t must be >= 0")

    def effective_volatility(self) -> float:
        """spread.rs:61-72"""
        variance = (self.vol1 * self.vol1 - 2.0 * self.rho * self.vol1 * self.vol2
                    + self.vol2 * self.vol2)
        if variance < -1.0e-14:
            raise InvalidInput("spread effective variance is negative")
        return max(variance, 0.0) ** 0.5


def spread_struct(option: SpreadOption, num_paths: int, seed: int) -> _lib.OfgSpreadOption:
    o = _lib.OfgSpreadOption()
    for f in ("s1", "s2", "k", "vol1", "vol2", "rho", "q1", "q2", "r", "t"):
        setattr(o, f, float(getattr(option, f)))
    o.num_paths, o.seed = int(num_paths), int(seed) & ((1 << 64) - 1)
    return o


def spread_price_raw(structs: Sequence[_lib.OfgSpreadOption], antithetic: bool = True,
                     rng_mode: int = _lib.RNG_REFERENCE, want_payoffs: bool = False,
                     device: int = 0):
    """ofg_spread_price over a batch: the results, and with want_payoffs
    each option's sample payoffs (None for options priced without
    simulation)."""
    n = len(structs)
    arr = (_lib.OfgSpreadOption * max(n, 1))(*structs)
    out = (_lib.OfgResult * max(n, 1))()
    sizes = [((s.num_paths + 1) // 2 if antithetic else s.num_paths) if s.num_paths > 0 else 0
             for s in structs]
    pay = np.full(max(sum(sizes), 1), np.nan) if want_payoffs else None
    ptr = pay.ctypes.data_as(C.POINTER(C.c_double)) if want_payoffs else None
    _raise(_lib.load().ofg_spread_price(_lib.context(device), arr, n, 1 if antithetic else 0,
                                        rng_mode, out, ptr))
    res = [out[i] for i in range(n)]
    if not want_payoffs:
        return res
    per, first = [], 0
    for s, m, r in zip(structs, sizes, res):
        per.append(pay[first:first + m].copy() if r.num_paths > 0 else None)
        first += m
    return res, per


class SpreadMonteCarloEngine(_StreamOptions):
    """Drop-in for SpreadMonteCarloEngine (spread_mc.rs:21-80); prices
    through ofg_spread_price. with_rng_mode(RngMode.Fast) selects the
    throughput generator instead of the reference's kinds."""

    def __init__(self, num_paths: int, seed: int):
        self.num_paths = num_paths
        self.seed = seed
        self.antithetic = True
        self.rng_kind = FastRngKind.Xoshiro256PlusPlus
        self.reproducible = True
        self.rng_mode = RngMode.Reference
        self.device = 0

    @staticmethod
    def new(num_paths: int, seed: int) -> "SpreadMonteCarloEngine":
        return SpreadMonteCarloEngine(num_paths, seed)

    def with_antithetic(self, antithetic: bool) -> "SpreadMonteCarloEngine":
        self.antithetic = antithetic
        return self

    def with_seed(self, seed: int) -> "SpreadMonteCarloEngine":
        self.seed = seed
        self.reproducible = True
        return self

    def with_rng_mode(self, mode: RngMode) -> "SpreadMonteCarloEngine":
        if mode is RngMode.External:
            raise InvalidInput("the spread engine draws its own streams")
        self.rng_mode = mode
        return self

    def with_device(self, device: int) -> "SpreadMonteCarloEngine":
        self.device = device
        return self

    def price(self, instrument: SpreadOption, market: Optional[Market] = None) -> PricingResult:
        """spread_mc.rs:83-175; the market is unused, as there."""
        return self.price_batch([instrument])[0]

    def price_batch(self, instruments: Sequence[SpreadOption]) -> List[PricingResult]:
        """Every instrument in one launch (one error fails the batch)."""
        for ins in instruments:
            ins.validate()
        if self.num_paths == 0:
            raise InvalidInput("spread MC num_paths must be > 0")
        mode = self._stream_mode()
        seed = self._stream_seed()
        structs = [spread_struct(ins, self.num_paths, seed) for ins in instruments]
        res = spread_price_raw(structs, self.antithetic, mode.value, device=self.device)
        return [_result(r, ins) for r, ins in zip(res, instruments)]


def _result(r: _lib.OfgResult, ins: SpreadOption) -> PricingResult:
    d = Diagnostics()
    if r.num_paths > 0:   # simulated (t > 0): spread_mc.rs:165-167
        d.insert("num_paths", r.num_paths)
        d.insert("rho", ins.rho)
    return PricingResult(price=r.price, stderr=r.stderr_ if r.has_stderr else None, greeks=None,
                         diagnostics=d, kernel_ms=r.kernel_ms)


def spread_price(s1: float, s2: float, k: float, vol1: float, vol2: float, rho: float, q1: float,
                 q2: float, rate: float, expiry: float, num_paths: int, seed: int,
                 antithetic: Optional[bool] = None, rng_kind: Optional[str] = None,
                 reproducible: Optional[bool] = None) -> PricingResult:
    """McEngine.spread_price of the Python binding (python/src/engines.rs:1392-1430)."""
    option = SpreadOption(s1, s2, k, vol1, vol2, rho, q1, q2, rate, expiry)
    eng = (SpreadMonteCarloEngine(num_paths, seed)
           .with_antithetic(True if antithetic is None else antithetic)
           .with_rng_kind(rng_kind_from_str(rng_kind)))
    if reproducible is False:
        eng = eng.with_randomized_streams()
    return eng.price(option)
