"""Quasi-Monte Carlo on the GPU (ofg_qmc_price through the C ABI) against the
oracle's restatement of mc_european_qmc_with_seed.

* the GPU forms each path's Sobol point from its index (Gray-code identity)
  and prices the reference sequence's points: one batched call over calls
  and puts with 1..300 dimensions (more than one 256-dimension segment),
  ragged point counts, seeds 0 / 7 / 2^63+5, and tiles with several indices
  per thread -- prices within 1e-11 relative of the oracle, stderr 1e-9;
* the reference's QMC assertions (0.1 % of Black-Scholes at 10k points,
  stderr competitive with PRNG Monte Carlo) on the GPU;
* the reference's degenerate results (NaN, no stderr; payoff at the spot for
  expiry 0), the dimension limit, and the Python mirror.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(oracle, **kw):
    from tests.cases import make_pair
    kw.setdefault("model", "gbm")
    kw.setdefault("payoff", "vanilla")
    return make_pair(**kw)


CASES = [
    dict(option_type="call", spot=100.0, strike=100.0, expiry=1.0, rate=0.05, vol=0.2,
         num_paths=10_000, num_steps=1, seed=11),
    dict(option_type="put", spot=100.0, strike=105.0, expiry=2.0, rate=0.03, q=0.01,
         vol=0.25, num_paths=12_345, num_steps=8, seed=3),
    dict(option_type="call", spot=90.0, strike=100.0, expiry=0.5, rate=0.01, vol=0.4,
         num_paths=65_536, num_steps=64, seed=0),
    dict(option_type="put", spot=110.0, strike=100.0, expiry=1.5, rate=0.02, vol=0.3,
         num_paths=3_001, num_steps=300, seed=2**63 + 5),
    dict(option_type="call", spot=100.0, strike=120.0, expiry=3.0, rate=0.04, vol=0.15,
         num_paths=(1 << 20) + 17, num_steps=4, seed=7),
]


def test_batch_matches_oracle(oracle, gpu):
    from openferric_amd import qmc_price_raw
    pairs = [_pair(oracle, **kw) for kw in CASES]
    res = qmc_price_raw([g for _, g, _ in pairs])
    for kw, (oo, _, _), r in zip(CASES, pairs, res):
        p, se = oracle.mc_european_qmc(oo, kw["num_paths"], kw["num_steps"], kw["seed"])
        assert abs(r.price - p) <= 1e-11 * abs(p), (kw, r.price, p)
        assert r.has_stderr and abs(r.stderr_ - se) <= 1e-9 * se, (kw, r.stderr_, se)
        assert r.num_paths == kw["num_paths"] and r.num_steps == kw["num_steps"]
    # batching does not change a price
    alone = qmc_price_raw([pairs[2][1]])[0]
    assert alone.price == res[2].price and alone.stderr_ == res[2].stderr_


def test_reference_assertions(oracle, gpu):
    from openferric_amd import qmc_price_raw
    _, g, _ = _pair(oracle, **CASES[0])
    r = qmc_price_raw([g])[0]
    bs = oracle.bs_price(oracle.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0, exact=True)
    assert abs(r.price - bs) / bs <= 0.001
    oo, g50, _ = _pair(oracle, **dict(CASES[0], num_paths=50_000, seed=7))
    q = qmc_price_raw([g50])[0]
    _, mse = oracle.mc_run(oo, 50_000, 1, 7)
    assert q.stderr_ < 1.2 * mse


def test_convergence_at_scale(oracle, gpu):
    from openferric_amd import qmc_price_raw
    kw = dict(option_type="put", spot=100.0, strike=100.0, expiry=1.0, rate=0.03, vol=0.2,
              num_paths=1 << 22, num_steps=16, seed=1)
    _, g, _ = _pair(oracle, **kw)
    r = qmc_price_raw([g])[0]
    bs = oracle.bs_price(oracle.PUT, 100.0, 100.0, 0.03, 0.0, 0.2, 1.0, exact=True)
    # a 16-step discretisation of GBM is exact in law: the error is QMC error
    assert abs(r.price - bs) < 2e-3, (r.price, bs)
    assert qmc_price_raw([g])[0].price == r.price


def test_degenerate_results_and_limits(oracle, gpu):
    from openferric_amd import InvalidInput, qmc_price_raw
    _, g0, _ = _pair(oracle, **dict(CASES[0], num_paths=0))
    _, gT, _ = _pair(oracle, **dict(CASES[0], expiry=0.0, strike=90.0))
    _, gv, _ = _pair(oracle, **dict(CASES[0], vol=0.0))
    r0, rT, rv = qmc_price_raw([g0, gT, gv])
    assert math.isnan(r0.price) and not r0.has_stderr
    assert rT.price == 10.0 and rT.has_stderr and rT.stderr_ == 0.0
    assert math.isnan(rv.price) and not rv.has_stderr
    _, gbig, _ = _pair(oracle, **dict(CASES[0], num_steps=21_202, num_paths=10))
    with pytest.raises(InvalidInput, match="Sobol dimensions"):
        qmc_price_raw([gbig])
    _, gasian, _ = _pair(oracle, payoff="asian", option_type="call", spot=100.0,
                         strike=100.0, expiry=1.0, rate=0.05, vol=0.2, num_paths=100,
                         num_steps=4, seed=1)
    with pytest.raises(InvalidInput):
        qmc_price_raw([gasian])


def test_python_mirror(gpu):
    from openferric_amd import Market, VanillaOption, mc_european_qmc, mc_european_qmc_with_seed
    m = Market.builder().spot(100.0).rate(0.05).flat_vol(0.2).build()
    r = mc_european_qmc_with_seed(VanillaOption.european_call(100.0, 1.0), m, 10_000, 1, 11)
    assert abs(r.price - 10.450583572185565) / 10.450583572185565 <= 0.001
    assert r.diagnostics.get("num_paths") == 10_000.0 and r.diagnostics.get("vol") == 0.2
    am = mc_european_qmc(VanillaOption.american_put(100.0, 1.0), m, 1000, 4)
    assert math.isnan(am.price) and am.stderr is None
