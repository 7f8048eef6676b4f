"""The reference's direct Monte Carlo pricing helpers, priced on the GPU:
arithmetic and geometric Asian (src/pricing/asian.rs:139-200) and barrier
(src/pricing/barrier.rs:199-239).

Each helper is MonteCarloEngine::new(paths, seed).with_antithetic(true) run
on GbmPathGenerator { mu: r, sigma } with discount exp(-r t); here that is
one ofg_price call on the reference draws, antithetic, no control variate:
  * Asian: the average is over all steps + 1 path points including S0
    (observation times j t / steps, j = 0..steps, which the engine's index
    map sends to path index j);
  * barrier: inclusive monitoring of every path point including S0 and a
    zero payoff when the option is not active (rebate 0).
Returns (price, stderr) like the reference.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

from . import _lib
from .core import BarrierDirection, BarrierStyle, InvalidInput, OptionType
from .engines import RngMode, _Keep, engine_cfg, price_raw


@dataclass(frozen=True)
class AsianStrike:
    """pricing/asian.rs:18-21: Fixed(k) or Floating."""
    fixed: Optional[float] = None

    @staticmethod
    def Fixed(k: float) -> "AsianStrike":
        return AsianStrike(fixed=float(k))

    @staticmethod
    def Floating() -> "AsianStrike":
        return AsianStrike(fixed=None)

    @property
    def is_floating(self) -> bool:
        return self.fixed is None


def _gbm_option(option_type: OptionType, s0: float, strike: float, r: float, sigma: float,
                t: float, steps: int, paths: int, seed: int) -> _lib.OfgOption:
    if paths <= 0:
        raise InvalidInput("num_paths must be > 0")
    if steps <= 0:
        raise InvalidInput("num_steps must be > 0")
    o = _lib.OfgOption()
    o.model = _lib.MODEL_GBM
    o.option_type = _lib.CALL if option_type is OptionType.Call else _lib.PUT
    o.spot, o.strike, o.expiry, o.rate, o.div_yield, o.vol = s0, strike, t, r, 0.0, sigma
    o.num_paths, o.num_steps, o.seed = int(paths), int(steps), int(seed) & (2**64 - 1)
    return o


def _asian(option_type: OptionType, strike: AsianStrike, s0: float, r: float, sigma: float,
           t: float, steps: int, paths: int, seed: int, geometric: bool) -> Tuple[float, float]:
    # the floating-strike payoff ignores the strike; S0 fills the field
    k = s0 if strike.is_floating else strike.fixed
    o = _gbm_option(option_type, s0, k, r, sigma, t, steps, paths, seed)
    keep = _Keep()
    o.payoff = _lib.PAYOFF_ASIAN
    o.averaging = _lib.GEOMETRIC if geometric else _lib.ARITHMETIC
    o.strike_type = _lib.FLOATING if strike.is_floating else _lib.FIXED
    obs = [t * (j / steps) for j in range(steps + 1)]
    o.n_obs = len(obs)
    o.obs_times = keep.arr(obs)
    r_ = price_raw([o], engine_cfg(antithetic=True, rng_mode=RngMode.Reference, keep=keep))[0]
    return r_.price, r_.stderr_


def arithmetic_asian_price_mc(option_type: OptionType, strike: AsianStrike, s0: float, r: float,
                              sigma: float, t: float, steps: int, paths: int,
                              seed: int) -> Tuple[float, float]:
    """pricing/asian.rs:139-168 on the GPU."""
    return _asian(option_type, strike, s0, r, sigma, t, steps, paths, seed, False)


def geometric_asian_price_mc(option_type: OptionType, strike: AsianStrike, s0: float, r: float,
                             sigma: float, t: float, steps: int, paths: int,
                             seed: int) -> Tuple[float, float]:
    """pricing/asian.rs:170-199 on the GPU."""
    return _asian(option_type, strike, s0, r, sigma, t, steps, paths, seed, True)


def barrier_price_mc(option_type: OptionType, style: BarrierStyle, direction: BarrierDirection,
                     s0: float, k: float, barrier: float, r: float, sigma: float, t: float,
                     steps: int, num_paths: int, seed: int) -> Tuple[float, float]:
    """pricing/barrier.rs:199-239 on the GPU."""
    o = _gbm_option(option_type, s0, k, r, sigma, t, steps, num_paths, seed)
    o.payoff = _lib.PAYOFF_BARRIER
    o.barrier_dir = _lib.UP if direction is BarrierDirection.Up else _lib.DOWN
    o.barrier_style = _lib.IN if style is BarrierStyle.In else _lib.OUT
    o.barrier_level, o.rebate = barrier, 0.0
    r_ = price_raw([o], engine_cfg(antithetic=True, rng_mode=RngMode.Reference))[0]
    return r_.price, r_.stderr_
