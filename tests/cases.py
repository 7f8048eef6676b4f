"""Matched contract descriptions for the oracle and the HIP library."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def make_pair(model="gbm", payoff="vanilla", option_type="call", spot=100.0, strike=100.0,
              expiry=1.0, rate=0.05, q=0.0, vol=0.2, kappa=0.0, theta=0.0, xi=0.0, rho=0.0, v0=0.0,
              averaging="arith", strike_type="fixed", barrier_dir="up", barrier_style="out",
              barrier_level=0.0, rebate=0.0, obs_times=None, num_paths=1000, num_steps=16, seed=42):
    """Returns (oracle OrOption, ofg OfgOption, keepalive)."""
    from openferric_amd import _lib
    from openferric_amd.engines import _Keep

    m = {"gbm": (O.GBM, _lib.MODEL_GBM), "heston": (O.HESTON, _lib.MODEL_HESTON)}[model]
    p = {"vanilla": (O.VANILLA, _lib.PAYOFF_VANILLA), "asian": (O.ASIAN, _lib.PAYOFF_ASIAN),
         "barrier": (O.BARRIER, _lib.PAYOFF_BARRIER)}[payoff]
    t = {"call": (O.CALL, _lib.CALL), "put": (O.PUT, _lib.PUT)}[option_type]
    av = {"arith": (O.ARITH, _lib.ARITHMETIC), "geo": (O.GEO, _lib.GEOMETRIC)}[averaging]
    st = {"fixed": (O.FIXED, _lib.FIXED), "floating": (O.FLOATING, _lib.FLOATING)}[strike_type]
    bd = {"up": (O.UP, _lib.UP), "down": (O.DOWN, _lib.DOWN)}[barrier_dir]
    bs = {"in": (O.IN, _lib.IN), "out": (O.OUT, _lib.OUT)}[barrier_style]
    oo = O.make_option(model=m[0], payoff=p[0], option_type=t[0], spot=spot, strike=strike,
                       expiry=expiry, rate=rate, div_yield=q, vol=vol, kappa=kappa, theta=theta,
                       xi=xi, rho=rho, v0=v0, averaging=av[0], strike_type=st[0],
                       barrier_dir=bd[0], barrier_style=bs[0], barrier_level=barrier_level,
                       rebate=rebate, obs_times=obs_times)
    keep = _Keep()
    g = _lib.OfgOption()
    g.model, g.payoff, g.option_type = m[1], p[1], t[1]
    g.averaging, g.strike_type, g.barrier_dir, g.barrier_style = av[1], st[1], bd[1], bs[1]
    g.spot, g.strike, g.expiry, g.rate, g.div_yield, g.vol = spot, strike, expiry, rate, q, vol
    g.kappa, g.theta, g.xi, g.rho, g.v0 = kappa, theta, xi, rho, v0
    g.barrier_level, g.rebate = barrier_level, rebate
    if obs_times is not None:
        g.obs_times = keep.arr(obs_times)
        g.n_obs = len(obs_times)
    g.num_paths, g.num_steps, g.seed = num_paths, num_steps, seed
    return oo, g, keep


def daily_obs(expiry, steps):
    """Observation at every step t_j = j T / n, j = 1..n."""
    return [expiry * (j / steps) for j in range(1, steps + 1)]


def step_major(z):
    """[sample][step] -> [step][sample], contiguous."""
    return np.ascontiguousarray(np.asarray(z).T)
