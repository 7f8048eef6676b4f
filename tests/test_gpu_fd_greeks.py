"""finite_difference_greeks / bump_and_reprice (src/greeks/bsm.rs:130-212) on
the GPU through the C ABI against the oracle's restatement of the harness
with the oracle's MonteCarloEngine::run as the pricer, on the same draws.

Each Greek is a fixed linear combination of revaluations, so with every
revaluation within 2e-13 relative of the oracle's price (the pricing parity
bar is 1e-11; measured ~1e-15) a Greek agrees to 2e-13 x max|price| x the
sum of its coefficients' magnitudes (delta 1/ds, gamma 4/ds^2, vanna
1/(ds dv), volga 4/dv^2, ...). bump_and_reprice's six prices agree to
2e-13 relative. Cases: vanilla calls and puts, a daily-fixed Asian, an up-and-
out barrier, antithetic on and off, and the floors at 1e-8 (spot, vol and
maturity bumped below zero). The Python mirror's batched path equals its
point-by-point path bit for bit.
"""
import numpy as np
import pytest

from tests.cases import daily_obs

pytestmark = pytest.mark.gpu

EPS = 2e-13
PATHS, STEPS = 20_000, 48

CASES = [
    dict(payoff="vanilla", option_type="call", spot=100.0, strike=100.0, expiry=1.0, rate=0.05,
         q=0.0, vol=0.2, bumps=(1.0, 1e-4, 1e-3, 1e-3)),
    dict(payoff="vanilla", option_type="put", spot=100.0, strike=95.0, expiry=0.8, rate=0.03,
         q=0.01, vol=0.22, bumps=(0.5, 1e-4, 1e-3, 1e-3)),
    dict(payoff="asian", option_type="call", spot=100.0, strike=98.0, expiry=1.5, rate=0.04,
         q=0.0, vol=0.3, bumps=(1.0, 1e-4, 1e-3, 1e-3)),
    dict(payoff="barrier", option_type="call", spot=100.0, strike=95.0, expiry=1.0, rate=0.03,
         q=0.0, vol=0.25, barrier_level=140.0, bumps=(1.0, 1e-4, 1e-3, 1e-3)),
    # floors: spot - ds, vol - dv and T - dt all below 1e-8
    dict(payoff="vanilla", option_type="call", spot=0.5, strike=0.45, expiry=0.5, rate=0.02,
         q=0.0, vol=0.2, bumps=(0.8, 1e-4, 0.3, 0.7)),
]
COEF = dict(delta=lambda b: 1 / b[0], gamma=lambda b: 4 / b[0] ** 2, rho=lambda b: 1 / b[1],
            vega=lambda b: 1 / b[2], theta=lambda b: 1 / b[3],
            vanna=lambda b: 1 / (b[0] * b[2]), volga=lambda b: 4 / b[2] ** 2)


def _struct(d, seed, keep):
    from openferric_amd import _lib
    o = _lib.OfgOption()
    o.model = _lib.MODEL_GBM
    o.payoff = {"asian": _lib.PAYOFF_ASIAN, "barrier": _lib.PAYOFF_BARRIER}.get(
        d["payoff"], _lib.PAYOFF_VANILLA)
    o.option_type = _lib.CALL if d["option_type"] == "call" else _lib.PUT
    o.spot, o.strike, o.expiry = d["spot"], d["strike"], d["expiry"]
    o.rate, o.div_yield, o.vol = d["rate"], d["q"], d["vol"]
    if d["payoff"] == "asian":
        o.obs_times = keep.arr(daily_obs(d["expiry"], STEPS))
        o.n_obs = STEPS
    if d["payoff"] == "barrier":
        o.barrier_dir, o.barrier_style, o.barrier_level = _lib.UP, _lib.OUT, d["barrier_level"]
    o.num_paths, o.num_steps, o.seed = PATHS, STEPS, seed
    return o


def _oracle_pricer(oracle, d, seed, antithetic):
    """price(s, k, r, sigma, t) of the oracle's engine on the option's seed
    (Asian fixings scaled by t / expiry, as the library's pricer)."""
    def price(s, k, r, v, t):
        obs = None
        if d["payoff"] == "asian":
            obs = [min(x * (t / d["expiry"]), t) for x in daily_obs(d["expiry"], STEPS)]
        o = oracle.make_option(payoff={"asian": oracle.ASIAN, "barrier": oracle.BARRIER}.get(
            d["payoff"], oracle.VANILLA), option_type=oracle.CALL if d["option_type"] == "call"
            else oracle.PUT, spot=s, strike=k, expiry=t, rate=r, div_yield=d["q"], vol=v,
            barrier_dir=oracle.UP, barrier_style=oracle.OUT,
            barrier_level=d.get("barrier_level", 0.0), obs_times=obs)
        return oracle.mc_run(o, PATHS, STEPS, seed, antithetic=antithetic, nthreads=8)[0]
    return price


@pytest.mark.parametrize("antithetic", [False, True])
def test_fd_greeks_match_oracle_harness(oracle, gpu, antithetic):
    from openferric_amd import RngMode, engine_cfg, fd_greeks_raw, bump_and_reprice_raw
    from openferric_amd.engines import _Keep
    keep = _Keep()
    seeds = [11 + i for i in range(len(CASES))]
    structs = [_struct(d, s, keep) for d, s in zip(CASES, seeds)]
    cfg = engine_cfg(antithetic=antithetic, rng_mode=RngMode.Reference)
    got = fd_greeks_raw(structs, cfg, [d["bumps"] for d in CASES])
    bar = bump_and_reprice_raw(structs, cfg, [d["bumps"] for d in CASES])
    for i, d in enumerate(CASES):
        pr = _oracle_pricer(oracle, d, seeds[i], antithetic)
        b = d["bumps"]
        args = (d["spot"], d["strike"], d["rate"], d["vol"], d["expiry"])
        want = oracle.finite_difference_greeks(pr, *args, *b)
        pmax = max(abs(x) for x in oracle.bump_and_reprice(pr, *args, *b)) + 1e-300
        for name in oracle.FD_FIELDS:
            tol = EPS * pmax * COEF[name](b) + 1e-300
            assert abs(getattr(got[i], name) - want[name]) <= tol, (i, name, getattr(got[i], name),
                                                                    want[name], tol)
        want6 = oracle.bump_and_reprice(pr, *args, *b)
        for g, w in zip(bar[i], want6):
            assert abs(g - w) <= EPS * max(abs(w), 1e-12), (i, g, w)


def test_mirror_batched_equals_point_by_point(gpu):
    from openferric_amd import (AsianOption, AsianSpec, Averaging, GpuMonteCarloPricingEngine,
                                McPricer, OptionType, RngMode, StrikeType, VanillaOption,
                                bump_and_reprice, finite_difference_greeks)
    eng = GpuMonteCarloPricingEngine(PATHS, STEPS, 5).with_rng_mode(RngMode.Reference)
    for ins in (VanillaOption.european_call(100.0, 1.0),
                AsianOption(OptionType.Put, 100.0, 1.0,
                            AsianSpec(Averaging.Arithmetic, StrikeType.Fixed,
                                      daily_obs(1.0, STEPS)))):
        mp = McPricer(eng, ins, dividend_yield=0.01)
        batched = finite_difference_greeks(mp, 100.0, 100.0, 0.03, 0.25, 1.0, 1.0, 1e-4, 1e-3,
                                           1e-3)
        one_by_one = finite_difference_greeks(lambda *a: mp.price(*a), 100.0, 100.0, 0.03, 0.25,
                                              1.0, 1.0, 1e-4, 1e-3, 1e-3)
        assert batched == one_by_one
        assert bump_and_reprice(mp, 100.0, 100.0, 0.03, 0.25, 1.0, 1.0, 1e-4, 1e-3, 1e-3) == \
            bump_and_reprice(lambda *a: mp.price(*a), 100.0, 100.0, 0.03, 0.25, 1.0, 1.0, 1e-4,
                             1e-3, 1e-3)


def test_heston_and_bad_bumps_refused(gpu):
    from openferric_amd import InvalidInput, engine_cfg, fd_greeks_raw
    from tests.cases import make_pair
    _, g, keep = make_pair(model="heston", kappa=1.5, theta=0.04, xi=0.5, rho=-0.5, v0=0.04)
    with pytest.raises(InvalidInput, match="GBM"):
        fd_greeks_raw([g], engine_cfg(), [(1.0, 1e-4, 1e-3, 1e-3)])
    _, g, keep = make_pair()
    with pytest.raises(InvalidInput, match="finite"):
        fd_greeks_raw([g], engine_cfg(), [(float("nan"), 1e-4, 1e-3, 1e-3)])
