"""Size-independent properties at the benchmark's full sizes (C ABI calls):

* the 4096-option GBM batch (fast draws, one seed per option, one persistent
  launch replayed from a hipGraph): the European half's z-scores against
  Black-Scholes behave as independent standard normals (mean, variance,
  extremes), and a graph replay equals the first run bit for bit;
* Heston up barriers at 2^18 paths x 365 steps, fast and reference draws:
  knock-in + knock-out payoffs equal the vanilla payoff sample by sample,
  exactly (in + out = vanilla is exact per path for inclusive barriers);
* GBM down barriers likewise.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bs(call, S, K, r, q, v, T):
    from scipy.stats import norm
    d1 = (math.log(S / K) + (r - q + 0.5 * v * v) * T) / (v * math.sqrt(T))
    d2 = d1 - v * math.sqrt(T)
    if call:
        return S * math.exp(-q * T) * norm.cdf(d1) - K * math.exp(-r * T) * norm.cdf(d2)
    return K * math.exp(-r * T) * norm.cdf(-d2) - S * math.exp(-q * T) * norm.cdf(-d1)


def test_batched_gbm_4096_european_z_scores(gpu):
    import bench
    from openferric_amd import BatchPricer, RngMode, engine_cfg
    opts = bench.make_workload()
    # one seed per option so the options' errors are independent (the bench
    # itself gives every option seed 42: common random numbers across the book)
    structs, keeps = zip(*[bench.to_ofg(d, seed=7000 + i) for i, d in enumerate(opts)])
    b = BatchPricer(list(structs), engine_cfg(rng_mode=RngMode.Fast), use_graph=True)
    b.run()
    first = [r.price for r in b.results()]
    b.run()
    again = [r.price for r in b.results()]
    res = b.results()
    b.close()
    assert first == again
    z, zero = [], 0
    for d, r in zip(opts, res):
        if d["payoff"] != "vanilla":
            continue
        bs = _bs(d["option_type"] == "call", 100.0, d["strike"], d["rate"], 0.0, d["vol"],
                 d["expiry"])
        if r.stderr_ == 0.0:
            # every sampled payoff zero: only for options worth almost nothing
            assert r.price == 0.0 and bs < 1e-3, (d, bs)
            zero += 1
            continue
        z.append((r.price - bs) / r.stderr_)
    z = np.array(z)
    n = len(z)
    assert n + zero == 2048 and zero < 20
    assert abs(z.mean()) < 4.0 / math.sqrt(n), z.mean()
    assert abs(z.var() - 1.0) < 4.0 * math.sqrt(2.0 / n), z.var()
    assert np.abs(z).max() < 5.0, np.abs(z).max()


def _payoffs(kw, rng):
    from openferric_amd import RngMode, engine_cfg, sample_payoffs
    from tests.cases import make_pair
    _, g, keep = make_pair(**kw)
    mode = {"fast": RngMode.Fast, "reference": RngMode.Reference}[rng]
    xs, _ = sample_payoffs(g, engine_cfg(rng_mode=mode), False, want_y=False)
    return xs


@pytest.mark.parametrize("rng", ["fast", "reference"])
def test_heston_barrier_in_plus_out_equals_vanilla_full_size(gpu, rng):
    base = dict(model="heston", spot=100.0, strike=100.0, expiry=1.0, rate=0.04, kappa=2.5,
                theta=0.05, xi=0.8, rho=-0.7, v0=0.04, num_paths=1 << 18, num_steps=365,
                seed=77)
    van = _payoffs(base, rng)
    for level in (112.0, 135.0):
        bar = dict(base, payoff="barrier", barrier_dir="up", barrier_level=level, rebate=0.0)
        xin = _payoffs(dict(bar, barrier_style="in"), rng)
        xout = _payoffs(dict(bar, barrier_style="out"), rng)
        assert np.array_equal(xin + xout, van)
        assert 0 < np.count_nonzero(xin) < np.count_nonzero(van)


@pytest.mark.parametrize("rng", ["fast", "reference"])
def test_gbm_down_barrier_in_plus_out_equals_vanilla_full_size(gpu, rng):
    base = dict(option_type="put", spot=100.0, strike=95.0, expiry=2.0, rate=0.03, vol=0.35,
                num_paths=1 << 18, num_steps=504, seed=5)
    van = _payoffs(base, rng)
    bar = dict(base, payoff="barrier", barrier_dir="down", barrier_level=80.0, rebate=0.0)
    xin = _payoffs(dict(bar, barrier_style="in"), rng)
    xout = _payoffs(dict(bar, barrier_style="out"), rng)
    assert np.array_equal(xin + xout, van)
