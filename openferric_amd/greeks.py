"""Finite-difference Greeks behind the reference's harness
(src/greeks/bsm.rs:16-55, 130-212).

`finite_difference_greeks(pricer, s, k, r, sigma, t, ds, dr, dv, dt)` and
`bump_and_reprice(pricer, s, k, r, sigma, t, bump_s, bump_r, bump_sigma,
bump_t)` take a GenericPricer: any callable or object with
`price(s, k, r, sigma, t)`. With a `McPricer` -- the GPU Monte Carlo engine
as a GenericPricer -- every revaluation is priced in one batched launch on
the option's own seed (common random numbers) through the C ABI
(`ofg_finite_difference_greeks`, `ofg_bump_and_reprice`); any other pricer is
evaluated point by point exactly as the reference does.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass
from typing import List, Sequence, Tuple

from . import _lib
from .core import AsianOption, Market
from .engines import (GpuMonteCarloPricingEngine, VarianceReduction, _Keep, _raise, engine_cfg,
                      option_struct)


@dataclass
class FiniteDifferenceGreeks:
    """FiniteDifferenceGreeks (bsm.rs:16-24)."""
    delta: float
    gamma: float
    vega: float
    theta: float
    rho: float
    vanna: float
    volga: float


def _call(pricer, s, k, r, sigma, t) -> float:
    return pricer.price(s, k, r, sigma, t) if hasattr(pricer, "price") else pricer(s, k, r, sigma, t)


class McPricer:
    """The GPU Monte Carlo engine as a GenericPricer (bsm.rs:44-55):
    price(s, k, r, sigma, t) prices `instrument` with strike k and expiry t
    (an Asian option's fixings scaled by t / its expiry) in a market with
    spot s, rate r, flat vol sigma and `dividend_yield`, on the engine's
    paths, steps, seed, variance reduction and RNG mode. GBM only."""

    def __init__(self, engine: GpuMonteCarloPricingEngine, instrument, dividend_yield: float = 0.0):
        self.engine = engine
        self.instrument = instrument
        self.dividend_yield = dividend_yield

    def _instrument(self, k, t):
        ins = self.instrument
        if isinstance(ins, AsianOption) and ins.expiry > 0.0 and t != ins.expiry:
            obs = [min(x * (t / ins.expiry), t) for x in ins.asian.observation_times]
            ins = dataclasses.replace(ins, asian=dataclasses.replace(ins.asian,
                                                                     observation_times=obs))
        return dataclasses.replace(ins, strike=k, expiry=t)

    def struct(self, s, k, r, sigma, t, keep: _Keep) -> _lib.OfgOption:
        m = Market(s, r, self.dividend_yield, sigma)
        e = self.engine
        return option_struct(self._instrument(k, t), m, e.num_paths, e.num_steps, e.seed, keep)

    def cfg(self, keep: _Keep) -> _lib.OfgEngineCfg:
        vr = self.engine.variance_reduction
        return engine_cfg(antithetic=vr is VarianceReduction.Antithetic,
                          control_variate=vr is VarianceReduction.ControlVariate,
                          rng_mode=self.engine.rng_mode, keep=keep)

    def price(self, s, k, r, sigma, t) -> float:
        return self.engine.price(self._instrument(k, t), Market(s, r, self.dividend_yield,
                                                                 sigma)).price


def fd_greeks_raw(opts: Sequence[_lib.OfgOption], cfg: _lib.OfgEngineCfg,
                  bumps: Sequence[Tuple[float, float, float, float]],
                  device: int = 0) -> List[_lib.OfgFdGreeks]:
    """ofg_finite_difference_greeks over many options in one batched launch;
    bumps[i] = (ds, dr, dv, dt)."""
    L = _lib.load()
    ctx = _lib.context(device)
    n = len(opts)
    arr = (_lib.OfgOption * n)(*opts)
    b = (_lib.OfgFdBumps * n)(*[_lib.OfgFdBumps(*x) for x in bumps])
    out = (_lib.OfgFdGreeks * n)()
    _raise(L.ofg_finite_difference_greeks(ctx, arr, n, C.byref(cfg), b, out))
    return list(out)


def bump_and_reprice_raw(opts: Sequence[_lib.OfgOption], cfg: _lib.OfgEngineCfg,
                         bumps: Sequence[Tuple[float, float, float, float]],
                         device: int = 0) -> List[Tuple[float, ...]]:
    """ofg_bump_and_reprice over many options in one batched launch."""
    L = _lib.load()
    ctx = _lib.context(device)
    n = len(opts)
    arr = (_lib.OfgOption * n)(*opts)
    b = (_lib.OfgFdBumps * n)(*[_lib.OfgFdBumps(*x) for x in bumps])
    out = (C.c_double * (6 * n))()
    _raise(L.ofg_bump_and_reprice(ctx, arr, n, C.byref(cfg), b, out))
    return [tuple(out[6 * i:6 * i + 6]) for i in range(n)]


def finite_difference_greeks(pricer, s, k, r, sigma, t, ds, dr, dv, dt) -> FiniteDifferenceGreeks:
    """finite_difference_greeks (bsm.rs:165-212)."""
    if isinstance(pricer, McPricer):
        keep = _Keep()
        g = fd_greeks_raw([pricer.struct(s, k, r, sigma, t, keep)], pricer.cfg(keep),
                          [(ds, dr, dv, dt)], pricer.engine.device)[0]
        return FiniteDifferenceGreeks(g.delta, g.gamma, g.vega, g.theta, g.rho, g.vanna, g.volga)
    s_dn = max(s - ds, 1e-8)
    v_dn = max(sigma - dv, 1e-8)
    p0 = _call(pricer, s, k, r, sigma, t)
    p_s_up = _call(pricer, s + ds, k, r, sigma, t)
    p_s_dn = _call(pricer, s_dn, k, r, sigma, t)
    p_r_up = _call(pricer, s, k, r + dr, sigma, t)
    p_r_dn = _call(pricer, s, k, r - dr, sigma, t)
    p_v_up = _call(pricer, s, k, r, sigma + dv, t)
    p_v_dn = _call(pricer, s, k, r, v_dn, t)
    p_t_up = _call(pricer, s, k, r, sigma, t + dt)
    p_t_dn = _call(pricer, s, k, r, sigma, max(t - dt, 1e-8))
    p_uu = _call(pricer, s + ds, k, r, sigma + dv, t)
    p_ud = _call(pricer, s + ds, k, r, v_dn, t)
    p_du = _call(pricer, s_dn, k, r, sigma + dv, t)
    p_dd = _call(pricer, s_dn, k, r, v_dn, t)
    return FiniteDifferenceGreeks(
        delta=(p_s_up - p_s_dn) / (2.0 * ds),
        gamma=(p_s_up - 2.0 * p0 + p_s_dn) / (ds * ds),
        vega=(p_v_up - p_v_dn) / (2.0 * dv),
        # market convention: theta is dV/d(calendar time), -dV/d(maturity)
        theta=-(p_t_up - p_t_dn) / (2.0 * dt),
        rho=(p_r_up - p_r_dn) / (2.0 * dr),
        vanna=(p_uu - p_ud - p_du + p_dd) / (4.0 * ds * dv),
        volga=(p_v_up - 2.0 * p0 + p_v_dn) / (dv * dv))


def bump_and_reprice(pricer, s, k, r, sigma, t, bump_s, bump_r, bump_sigma,
                     bump_t) -> Tuple[float, float, float, float, float, float]:
    """bump_and_reprice (bsm.rs:130-163): (base, s_up, r_up, v_up, t_up,
    s_dn) with t_up at max(t + bump_t, 1e-8) and s_dn unfloored."""
    if isinstance(pricer, McPricer):
        keep = _Keep()
        return bump_and_reprice_raw([pricer.struct(s, k, r, sigma, t, keep)], pricer.cfg(keep),
                                    [(bump_s, bump_r, bump_sigma, bump_t)],
                                    pricer.engine.device)[0]
    base = _call(pricer, s, k, r, sigma, t)
    s_up = _call(pricer, s + bump_s, k, r, sigma, t)
    r_up = _call(pricer, s, k, r + bump_r, sigma, t)
    v_up = _call(pricer, s, k, r, sigma + bump_sigma, t)
    t_up = _call(pricer, s, k, r, sigma, max(t + bump_t, 1e-8))
    return (base, s_up, r_up, v_up, t_up, _call(pricer, s - bump_s, k, r, sigma, t))
