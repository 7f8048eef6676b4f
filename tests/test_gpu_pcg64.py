"""The reference's Pcg64 stream kind (FastRngKind::Pcg64, fast_rng.rs:93-137)
through the library's pricing entry points, against the oracle drawing the
same per-sample Pcg64 streams (its Pcg64 pinned by numpy's PCG64):

* ofg_sample_payoffs: every sample's payoff to 1e-12 of max(|x|, S0) for
  GBM Europeans (antithetic on and off), daily-fixed Asians, up-and-out
  barriers and Heston Europeans (two normals per step);
* ofg_price: batch prices to 1e-11 relative;
* ofg_finite_difference_greeks over Pcg64 revaluations against the oracle
  harness on the same draws (the FD bar of the xoshiro test);
* ofg_mc_greeks_estimate: the five estimators to 1e-11;
* the engine mirrors' with_rng_kind; pathwise Greeks refuse the kind.
"""
import numpy as np
import pytest

from tests.cases import daily_obs, make_pair

pytestmark = pytest.mark.gpu


def _cfg(antithetic=False, greek_mode=0):
    from openferric_amd import RngMode, engine_cfg
    return engine_cfg(antithetic=antithetic, rng_mode=RngMode.Pcg64, greek_mode=greek_mode)


CASES = [
    dict(model="gbm", payoff="vanilla", strike=105.0, vol=0.25, num_paths=4001, num_steps=64,
         seed=7),
    dict(model="gbm", payoff="asian", strike=98.0, vol=0.3, expiry=1.5, num_paths=2001,
         num_steps=48, seed=8, obs_times=daily_obs(1.5, 48)),
    dict(model="gbm", payoff="barrier", barrier_dir="up", barrier_style="out",
         barrier_level=150.0, strike=95.0, vol=0.25, num_paths=2001, num_steps=100, seed=9),
    dict(model="heston", payoff="vanilla", strike=100.0, rate=0.03, kappa=1.5, theta=0.04,
         xi=0.5, rho=-0.6, v0=0.05, num_paths=2001, num_steps=64, seed=10),
]


@pytest.mark.parametrize("antithetic", [False, True])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_sample_payoffs_on_pcg64_streams(oracle, gpu, case, antithetic):
    from openferric_amd import sample_payoffs
    c = CASES[case]
    oo, g, keep = make_pair(**c)
    with oracle.rng_kind(oracle.RNG_PCG64):
        ref = oracle.mc_run(oo, c["num_paths"], c["num_steps"], c["seed"], antithetic=antithetic,
                            want_samples=True)
    xs, _ = sample_payoffs(g, _cfg(antithetic), antithetic, want_y=False)
    err = np.abs(xs - ref[2]) / np.maximum(np.abs(ref[2]), 100.0)
    assert err.max() <= 1e-12, err.max()


def test_batch_prices_on_pcg64_streams(oracle, gpu):
    from openferric_amd import price_raw
    pairs = [make_pair(**c) for c in CASES]
    res = price_raw([p[1] for p in pairs], _cfg(True))
    with oracle.rng_kind(oracle.RNG_PCG64):
        for c, (oo, g, keep), r in zip(CASES, pairs, res):
            p, se = oracle.mc_run(oo, c["num_paths"], c["num_steps"], c["seed"], antithetic=True)
            assert abs(r.price - p) <= 1e-11 * abs(p), (c, r.price, p)
            assert abs(r.stderr_ - se) <= 1e-9 * se


def test_fd_greeks_on_pcg64_streams(oracle, gpu):
    from openferric_amd import fd_greeks_raw
    oo, g, keep = make_pair(spot=100.0, strike=100.0, rate=0.05, vol=0.2, expiry=1.0,
                            num_paths=20_000, num_steps=16, seed=3)
    b = (1.0, 1e-4, 1e-3, 1e-3)
    got = fd_greeks_raw([g], _cfg(), [b])[0]

    def pricer(s, k, r, v, t):
        o = oracle.make_option(spot=s, strike=k, rate=r, vol=v, expiry=t)
        with oracle.rng_kind(oracle.RNG_PCG64):
            return oracle.mc_run(o, 20_000, 16, 3, nthreads=8)[0]

    want = oracle.finite_difference_greeks(pricer, 100.0, 100.0, 0.05, 0.2, 1.0, *b)
    pmax = max(abs(x) for x in oracle.bump_and_reprice(pricer, 100.0, 100.0, 0.05, 0.2, 1.0, *b))
    coef = dict(delta=1 / b[0], gamma=4 / b[0] ** 2, rho=1 / b[1], vega=1 / b[2], theta=1 / b[3],
                vanna=1 / (b[0] * b[2]), volga=4 / b[2] ** 2)
    for name in oracle.FD_FIELDS:
        assert abs(getattr(got, name) - want[name]) <= 2e-13 * pmax * coef[name] + 1e-300, name


def test_mc_greeks_estimators_on_pcg64_streams(oracle, gpu):
    from openferric_amd import MonteCarloGreeksEngine, FastRngKind, Market, VanillaOption
    eng = MonteCarloGreeksEngine(50_001, 17).with_rng_kind(FastRngKind.Pcg64)
    m = Market.builder().spot(100.0).rate(0.03).dividend_yield(0.01).flat_vol(0.3).build()
    got = eng.estimates(VanillaOption.european_put(105.0, 0.75), m)
    oo, _, _ = make_pair(option_type="put", spot=100.0, strike=105.0, rate=0.03, q=0.01,
                         vol=0.3, expiry=0.75, num_steps=1)
    with oracle.rng_kind(oracle.RNG_PCG64):
        ref = oracle.mc_greeks_estimate(oo, 50_001, 17, True, 1e-2)
    for k, ok in zip(["pathwise_delta", "pathwise_gamma", "lr_delta", "lr_gamma", "lr_vega"],
                     ["pw_delta", "pw_gamma", "lr_delta", "lr_gamma", "lr_vega"]):
        assert abs(got[k] - ref[ok]) <= 1e-11 * abs(ref[ok]) + 1e-14, k


def test_engine_mirror_rng_kind(oracle, gpu):
    from openferric_amd import (FastRngKind, GpuMonteCarloPricingEngine, InvalidInput, Market,
                                VanillaOption)
    m = Market.builder().spot(100.0).rate(0.05).dividend_yield(0.0).flat_vol(0.2).build()
    call = VanillaOption.european_call(100.0, 1.0)
    eng = GpuMonteCarloPricingEngine(10_000, 32, 5).with_rng_kind(FastRngKind.Pcg64)
    r = eng.price(call, m)
    oo, _, _ = make_pair(num_paths=10_000, num_steps=32, seed=5)
    with oracle.rng_kind(oracle.RNG_PCG64):
        p, _ = oracle.mc_run(oo, 10_000, 32, 5)
    assert abs(r.price - p) <= 1e-11 * p
    with pytest.raises(InvalidInput, match="pathwise Greeks"):
        eng.price_with_greeks_aad(call, m)
    with pytest.raises(InvalidInput, match="StdRng"):
        GpuMonteCarloPricingEngine(100, 2, 1).with_rng_kind(FastRngKind.StdRng).price(call, m)
