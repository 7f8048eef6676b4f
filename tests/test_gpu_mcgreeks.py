"""GPU tests of the exact-terminal European estimators and the direct MC
helpers (C ABI calls):

* MonteCarloGreeksEngine (mc_greeks.rs): pathwise and likelihood-ratio Greeks
  vs the oracle on the reference draws and external normals, and the fast
  mode against the Black-Scholes closed forms;
* mc_european_parallel / _sequential (mc_parallel.rs:264-434) and
  mc_european_with_arena (mc_engine.rs:186-343);
* arithmetic/geometric Asian and barrier helpers (pricing/asian.rs:139-200,
  pricing/barrier.rs:199-239) vs the oracle's engine on the same options.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ["pathwise_delta", "pathwise_gamma", "lr_delta", "lr_gamma", "lr_vega"]
OKEYS = ["pw_delta", "pw_gamma", "lr_delta", "lr_gamma", "lr_vega"]


def _bs(S, K, r, q, vol, T, call=True):
    from scipy.stats import norm
    d1 = (math.log(S / K) + (r - q + 0.5 * vol * vol) * T) / (vol * math.sqrt(T))
    delta = math.exp(-q * T) * (norm.cdf(d1) if call else norm.cdf(d1) - 1.0)
    gamma = math.exp(-q * T) * norm.pdf(d1) / (S * vol * math.sqrt(T))
    vega = S * math.exp(-q * T) * norm.pdf(d1) * math.sqrt(T)
    return delta, gamma, vega


def _raw_estimate(opts, rng="reference", antithetic=True, rel=1e-2, normals=None):
    import ctypes as C
    from openferric_amd import RngMode, _lib, engine_cfg
    from openferric_amd.engines import _raise
    mode = {"reference": RngMode.Reference, "fast": RngMode.Fast,
            "external": RngMode.External}[rng]
    cfg = engine_cfg(antithetic=antithetic, rng_mode=mode, normals1=normals)
    arr = (_lib.OfgOption * len(opts))(*opts)
    out = (_lib.OfgMcGreeks * len(opts))()
    _raise(_lib.load().ofg_mc_greeks_estimate(_lib.context(0), arr, len(opts), C.byref(cfg),
                                               rel, out))
    return [{k: getattr(o, k) for k in KEYS + ["num_samples"]} for o in out]


CASES = [
    dict(option_type="call", strike=100.0, rate=0.05, vol=0.2, expiry=1.0),
    dict(option_type="put", strike=110.0, rate=0.03, q=0.01, vol=0.35, expiry=0.5),
    dict(option_type="call", strike=80.0, rate=0.0, vol=0.15, expiry=2.0),
    dict(option_type="put", strike=95.0, rate=0.07, q=0.02, vol=0.5, expiry=3.0),
]


@pytest.mark.parametrize("antithetic", [True, False])
def test_mc_greeks_estimators_match_oracle_on_reference_draws(oracle, gpu, antithetic):
    """All five estimates, four options in one launch, against the oracle's
    sequential fold on the same draws: sums differ only by summation order
    and 1-ulp exp differences."""
    from tests.cases import make_pair
    pairs = [make_pair(num_paths=60_001, num_steps=1, seed=42 + i, **c)
             for i, c in enumerate(CASES)]
    got = _raw_estimate([p[1] for p in pairs], antithetic=antithetic, rel=0.015)
    for (oo, g, keep), r in zip(pairs, got):
        ref = oracle.mc_greeks_estimate(oo, 60_001, g.seed, antithetic, 0.015)
        assert r["num_samples"] == (30_001 if antithetic else 60_001)
        for k, ok in zip(KEYS, OKEYS):
            assert abs(r[k] - ref[ok]) <= 1e-11 * abs(ref[ok]) + 1e-14, (k, r[k], ref[ok])


def test_mc_greeks_external_normals_and_batch_independence(oracle, gpu):
    """External normals (the reference streams' draws) reproduce the
    reference-mode estimates; an option priced alone equals the same option
    inside a batch."""
    from tests.cases import make_pair
    n, seed = 20_000, 9
    oo, g, keep = make_pair(num_paths=n, num_steps=1, seed=seed, **CASES[1])
    z = np.array([oracle.sample_normals(seed, i, 1, 1)[0][0] for i in range(n)])
    ext = _raw_estimate([g], rng="external", antithetic=False, normals=z)[0]
    ref = _raw_estimate([g], antithetic=False)[0]
    for k in KEYS:
        assert abs(ext[k] - ref[k]) <= 1e-12 * abs(ref[k]) + 1e-15, k
    others = [make_pair(num_paths=5000 + 37 * i, num_steps=1, seed=i, **CASES[i % 4])[1]
              for i in range(7)]
    batch = _raw_estimate(others[:3] + [g] + others[3:], antithetic=False)
    for k in KEYS:
        assert batch[3][k] == ref[k], k


@pytest.mark.parametrize("case", range(len(CASES)))
def test_mc_greeks_fast_mode_against_black_scholes(gpu, case):
    """Fast draws: every estimator within 3.5 SE of its exact expectation
    (BS delta, gamma, vega; the CRN difference of BS deltas for the pathwise
    gamma), 12 independent seeds x 2^20 paths."""
    from tests.cases import make_pair
    c = dict(CASES[case])
    q = c.pop("q", 0.0)
    call = c["option_type"] == "call"
    opts = [make_pair(num_paths=1 << 20, num_steps=1, seed=1000 + b, q=q, **c)[1]
            for b in range(12)]
    est = _raw_estimate(opts, rng="fast", rel=0.01)
    S, K, r, vol, T = 100.0, c["strike"], c["rate"], c["vol"], c["expiry"]
    delta, gamma, vega = _bs(S, K, r, q, vol, T, call)
    h = S * 0.01
    dg = (_bs(S + h, K, r, q, vol, T, call)[0] - _bs(S - h, K, r, q, vol, T, call)[0]) / (2 * h)
    for k, e in zip(KEYS, [delta, dg, delta, gamma, vega]):
        v = np.array([x[k] for x in est])
        se = v.std(ddof=1) / math.sqrt(len(v))
        assert abs(v.mean() - e) <= 3.5 * se + 1e-12, (k, v.mean(), e, se)


def test_mc_greeks_engine_mirror_and_reference_values(gpu):
    """The Python mirror reproduces the reference's unit tests
    (mc_greeks.rs:330-413): delta 0.6368 (2 %), gamma 0.01876 (20 %), vega
    37.524 (5 %); validation errors and the zero-expiry default."""
    import openferric_amd as ofa
    m = ofa.Market.builder().spot(100.0).rate(0.05).dividend_yield(0.0).flat_vol(0.2).build()
    opt = ofa.VanillaOption.european_call(100.0, 1.0)
    eng = ofa.MonteCarloGreeksEngine(100_000, 42)
    pw = eng.estimate_pathwise(opt, m)
    lr = eng.estimate_likelihood_ratio(opt, m)
    assert abs(pw.delta - 0.6368) / 0.6368 < 0.02
    assert abs(pw.gamma - 0.01876) / 0.01876 < 0.20
    assert abs(lr.vega - 37.524) / 37.524 < 0.05
    z = eng.estimate_pathwise(ofa.VanillaOption.european_call(100.0, 0.0), m)
    assert (z.delta, z.gamma) == (0.0, 0.0)
    with pytest.raises(ofa.InvalidInput, match="num_paths must be > 0"):
        ofa.MonteCarloGreeksEngine(0, 1).estimate_pathwise(opt, m)
    bad = ofa.MonteCarloGreeksEngine(10, 1)
    bad.spot_bump_rel = 0.0
    with pytest.raises(ofa.InvalidInput, match="spot_bump_rel must be finite and > 0"):
        bad.estimate_pathwise(opt, m)
    assert ofa.MonteCarloGreeksEngine(10, 1).with_spot_bump_rel(0.0).spot_bump_rel == 1e-6


def test_mc_european_parallel_and_sequential(gpu):
    """Exact-terminal European: within 3 SE of Black-Scholes; NaN for zero
    paths, intrinsic value at zero expiry, num_steps diagnostic 1."""
    from scipy.stats import norm
    import openferric_amd as ofa
    m = ofa.Market.builder().spot(100.0).rate(0.05).dividend_yield(0.02).flat_vol(0.25).build()
    for opt in [ofa.VanillaOption.european_call(105.0, 1.5),
                ofa.VanillaOption.european_put(95.0, 0.75)]:
        S, K, r, q, v, T = 100.0, opt.strike, 0.05, 0.02, 0.25, opt.expiry
        d1 = (math.log(S / K) + (r - q + 0.5 * v * v) * T) / (v * math.sqrt(T))
        d2 = d1 - v * math.sqrt(T)
        call = opt.option_type is ofa.OptionType.Call
        bs = (S * math.exp(-q * T) * norm.cdf(d1) - K * math.exp(-r * T) * norm.cdf(d2)) if call \
            else (K * math.exp(-r * T) * norm.cdf(-d2) - S * math.exp(-q * T) * norm.cdf(-d1))
        for fn in (ofa.mc_european_parallel, ofa.mc_european_sequential,
                   ofa.mc_european_with_arena):
            res = fn(opt, m, 2_000_000, 252)
            assert abs(res.price - bs) <= 3 * res.stderr, (res.price, bs, res.stderr)
            assert res.diagnostics.get("num_steps") == 1.0
            assert res.diagnostics.get("num_paths") == 2_000_000
    assert math.isnan(ofa.mc_european_parallel(ofa.VanillaOption.european_call(100.0, 1.0), m,
                                                0, 1).price)
    r0 = ofa.mc_european_sequential(ofa.VanillaOption.european_put(110.0, 0.0), m, 10, 1)
    assert r0.price == 10.0 and r0.stderr == 0.0
    # mc_engine.rs:1170-1194: 200,000 paths within 1 % of BS; invalid -> NaN
    m0 = ofa.Market.builder().spot(100.0).rate(0.05).dividend_yield(0.0).flat_vol(0.2).build()
    ra = ofa.mc_european_with_arena(ofa.VanillaOption.european_call(100.0, 1.0), m0, 200_000, 1)
    assert abs(ra.price - 10.450583572185565) / 10.450583572185565 <= 0.01
    assert math.isnan(ofa.mc_european_with_arena(ofa.VanillaOption.european_call(-1.0, 1.0), m0,
                                                 100, 1).price)


def _legacy_oracle(oracle, kind, option_type, s0, k, r, sigma, t, steps, paths, seed, **kw):
    from tests.cases import make_pair
    extra = {}
    if kind in ("arith", "geo"):
        extra = dict(payoff="asian", averaging=kind, strike_type=kw.get("strike_type", "fixed"),
                     obs_times=[t * (j / steps) for j in range(steps + 1)])
    else:
        extra = dict(payoff="barrier", barrier_dir=kw["direction"], barrier_style=kw["style"],
                     barrier_level=kw["barrier"], rebate=0.0)
    oo, g, keep = make_pair(option_type=option_type, spot=s0, strike=k, rate=r, vol=sigma,
                            expiry=t, num_paths=paths, num_steps=steps, seed=seed, **extra)
    return oracle.mc_run(oo, paths, steps, seed, antithetic=True)


def test_legacy_asian_and_barrier_helpers_match_oracle(oracle, gpu):
    """Direct helpers on the GPU equal the oracle's MonteCarloEngine::run on
    the same options (average over all path points incl. S0; barrier pays 0
    when inactive) to the price bar."""
    from openferric_amd import BarrierDirection, BarrierStyle, OptionType
    from openferric_amd.pricing import (AsianStrike, arithmetic_asian_price_mc,
                                        barrier_price_mc, geometric_asian_price_mc)
    args = (100.0, 0.05, 0.2, 1.0, 64, 20_001, 44)
    for fn, kind in [(arithmetic_asian_price_mc, "arith"), (geometric_asian_price_mc, "geo")]:
        for ot, ots in [(OptionType.Call, "call"), (OptionType.Put, "put")]:
            for strike, st in [(AsianStrike.Fixed(100.0), "fixed"), (AsianStrike.Floating(), "floating")]:
                p, se = fn(ot, strike, *args)
                k = 100.0
                pr, ser = _legacy_oracle(oracle, kind, ots, 100.0, k, 0.05, 0.2, 1.0, 64, 20_001,
                                         44, strike_type=st)
                assert abs(p - pr) <= 1e-11 * abs(pr) + 1e-13, (kind, ots, st, p, pr)
                assert abs(se - ser) <= 1e-8 * ser + 1e-13
    for style, ss in [(BarrierStyle.Out, "out"), (BarrierStyle.In, "in")]:
        for d, ds, h in [(BarrierDirection.Up, "up", 115.0), (BarrierDirection.Down, "down", 88.0)]:
            p, se = barrier_price_mc(OptionType.Call, style, d, 100.0, 100.0, h, 0.01, 0.2, 1.0,
                                     128, 20_001, 10)
            pr, ser = _legacy_oracle(oracle, "barrier", "call", 100.0, 100.0, 0.01, 0.2, 1.0, 128,
                                     20_001, 10, direction=ds, style=ss, barrier=h)
            assert abs(p - pr) <= 1e-11 * max(abs(pr), 1e-3), (ss, ds, p, pr)


def test_legacy_helpers_reference_assertions(gpu, oracle):
    """pricing/asian.rs and pricing/barrier.rs tests: geometric MC within
    2 SE + 0.02 of the continuous closed form (and within 3 SE of the exact
    discrete expectation), arithmetic >= geometric - 0.05, barrier in + out
    = vanilla within 2.5 (SE_in + SE_out)."""
    from scipy.stats import norm
    from openferric_amd import BarrierDirection, BarrierStyle, OptionType
    from openferric_amd.pricing import (AsianStrike, arithmetic_asian_price_mc,
                                        barrier_price_mc, geometric_asian_price_mc)
    s0, k, r, sigma, t = 100.0, 100.0, 0.05, 0.2, 1.0
    m = math.log(s0) + (r - 0.5 * sigma * sigma) * t / 2.0
    v = sigma * sigma * t / 3.0
    d1 = (m - math.log(k) + v) / math.sqrt(v)
    closed = math.exp(-r * t) * (math.exp(m + 0.5 * v) * norm.cdf(d1)
                                 - k * norm.cdf(d1 - math.sqrt(v)))
    mc, se = geometric_asian_price_mc(OptionType.Call, AsianStrike.Fixed(k), s0, r, sigma, t, 252,
                                      80_000, 44)
    assert abs(mc - closed) <= 2.0 * se + 2e-2
    obs = np.array([t * (j / 252) for j in range(253)])
    discrete = oracle.geometric_asian_discrete(oracle.CALL, s0, k, r, 0.0, sigma, obs)
    assert abs(mc - discrete) <= 3.0 * se
    a, _ = arithmetic_asian_price_mc(OptionType.Call, AsianStrike.Fixed(100.0), 100.0, 0.02, 0.2,
                                     1.0, 128, 30_000, 11)
    g, _ = geometric_asian_price_mc(OptionType.Call, AsianStrike.Fixed(100.0), 100.0, 0.02, 0.2,
                                    1.0, 128, 30_000, 11)
    assert a >= g - 0.05
    out, e_out = barrier_price_mc(OptionType.Call, BarrierStyle.Out, BarrierDirection.Up, 100.0,
                                  100.0, 110.0, 0.01, 0.2, 1.0, 252, 30_000, 10)
    inn, e_in = barrier_price_mc(OptionType.Call, BarrierStyle.In, BarrierDirection.Up, 100.0,
                                 100.0, 110.0, 0.01, 0.2, 1.0, 252, 30_000, 10)
    d1 = (math.log(1.0) + (0.01 + 0.02) * 1.0) / 0.2
    vanilla = 100.0 * norm.cdf(d1) - 100.0 * math.exp(-0.01) * norm.cdf(d1 - 0.2)
    assert abs(out + inn - vanilla) <= 2.5 * (e_out + e_in)
