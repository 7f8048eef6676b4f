"""Worst-of autocallables and Phoenix notes on the GPU (ofg_autocall_price
through the C ABI) against the oracle's restatement of
simulate_autocallable_paths.

* external normals: one batched call over standard / Phoenix (memory and
  not) contracts with 1, 2, 3, 4 and 8 underlyings (selected out of a larger
  market), knock-ins, early calls and ragged path counts -- prices within
  1e-10 relative of the oracle fed the same per-path normals, stderr 1e-8;
* the bump-and-reprice sensitivities: each bumped scenario's price equals
  the oracle's price of the bumped inputs on the same normals, so delta,
  vega and cega agree to 1e-7;
* fast draws: the reference's tests (lower barrier lowers the value, higher
  coupon raises it, Phoenix coupons raise it, a single underlying selected
  out of a two-asset market prices exactly as the one-asset market) and the
  certain-knock-in Margrabe closed form within 4 standard errors;
* invalid contracts give NaN and the reference's message; reference draws
  are refused with the reason.
"""
import math

import numpy as np
import pytest
from scipy.stats import norm

pytestmark = pytest.mark.gpu

DATES = [0.25, 0.5, 0.75, 1.0]

MARKET = dict(
    spots=[100.0, 80.0, 120.0, 95.0, 60.0, 140.0, 110.0, 70.0, 105.0],
    vols=[0.25, 0.3, 0.2, 0.35, 0.4, 0.15, 0.22, 0.28, 0.18],
)


def _corr(n, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n + 2))
    c = a @ a.T
    d = np.sqrt(np.diag(c))
    c = c / np.outer(d, d)
    np.fill_diagonal(c, 1.0)
    return (c + c.T) / 2


CORR = _corr(9)

CONTRACTS = [
    dict(phoenix=False, underlyings=[0, 1], autocall_barrier=1.0, coupon_rate=0.08,
         ki_barrier=0.6, ki_strike=1.0, num_paths=5_000, num_steps=64),
    dict(phoenix=True, memory=True, underlyings=[2, 0, 5], autocall_barrier=1.05,
         coupon_barrier=0.8, coupon_rate=0.12, ki_barrier=0.7, ki_strike=1.0, num_paths=3_001,
         num_steps=80),
    dict(phoenix=True, memory=False, underlyings=[4], autocall_barrier=1.1, coupon_barrier=0.9,
         coupon_rate=0.1, ki_barrier=0.75, ki_strike=0.9, num_paths=4_096, num_steps=48),
    dict(phoenix=False, underlyings=[1, 3, 4, 7], autocall_barrier=1.02, coupon_rate=0.09,
         ki_barrier=0.8, ki_strike=1.0, num_paths=2_500, num_steps=40),
    dict(phoenix=True, memory=True, underlyings=[8, 7, 6, 5, 4, 3, 2, 1], autocall_barrier=0.95,
         coupon_barrier=0.7, coupon_rate=0.15, ki_barrier=0.5, ki_strike=1.0, num_paths=1_500,
         num_steps=32),
]


def _contract(kw):
    from openferric_amd import Autocallable, PhoenixAutocallable
    common = dict(underlyings=kw["underlyings"], notional=100.0, autocall_dates=DATES,
                  autocall_barrier=kw["autocall_barrier"], coupon_rate=kw["coupon_rate"],
                  ki_barrier=kw["ki_barrier"], ki_strike=kw["ki_strike"], maturity=1.0)
    if kw["phoenix"]:
        return PhoenixAutocallable(coupon_barrier=kw["coupon_barrier"], memory=kw["memory"],
                                   **common)
    return Autocallable(**common)


def _oracle(oracle, kw, spots, vols, corr, normals, r=0.02, q=0.01):
    u = kw["underlyings"]
    c = np.asarray(corr)[np.ix_(u, u)]
    a = oracle.make_autocall(spots=[spots[i] for i in u], vols=[vols[i] for i in u], corr=c,
                             dates=DATES, notional=100.0, maturity=1.0,
                             autocall_barrier=kw["autocall_barrier"],
                             coupon_rate=kw["coupon_rate"],
                             coupon_barrier=kw.get("coupon_barrier", 0.0),
                             ki_barrier=kw["ki_barrier"], ki_strike=kw["ki_strike"], rate=r,
                             div_yield=q, phoenix=kw["phoenix"], memory=kw.get("memory", False),
                             num_paths=kw["num_paths"], num_steps=kw["num_steps"])
    return oracle.autocall_price(a, normals)


def _normals(kw, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((kw["num_paths"], kw["num_steps"], len(kw["underlyings"])))


def test_external_normals_batch_matches_oracle_with_sensitivities(oracle, gpu):
    from openferric_amd import RngMode, autocall_price_raw, autocall_struct
    from openferric_amd.engines import _Keep
    keep = _Keep()
    structs = [autocall_struct(_contract(kw), MARKET["spots"], MARKET["vols"], CORR, 0.02, 0.01,
                               kw["num_paths"], kw["num_steps"], keep) for kw in CONTRACTS]
    zs = [_normals(kw, 100 + i) for i, kw in enumerate(CONTRACTS)]
    res, sens = autocall_price_raw(structs, RngMode.External, zs, with_greeks=True)
    spots, vols = MARKET["spots"], MARKET["vols"]
    for kw, z, r, s in zip(CONTRACTS, zs, res, sens):
        p, se = _oracle(oracle, kw, spots, vols, CORR, z)
        assert abs(r.price - p) <= 1e-10 * abs(p), (kw, r.price, p)
        assert abs(r.stderr_ - se) <= 1e-8 * max(se, 1e-9), (kw, r.stderr_, se)
        assert s.ok and s.n_delta == len(kw["underlyings"])
        # spot bumps, one underlying at a time
        for k, idx in enumerate(kw["underlyings"]):
            b = max(abs(spots[idx]) * 0.01, 1e-4)
            up = list(spots)
            dn = list(spots)
            up[idx] += b
            dn[idx] = max(dn[idx] - b, 1e-8)
            d = (_oracle(oracle, kw, up, vols, CORR, z)[0] -
                 _oracle(oracle, kw, dn, vols, CORR, z)[0]) / (2 * b)
            assert abs(s.delta[k] - d) <= 1e-7, (kw, k, s.delta[k], d)
        vu = [v + 0.01 if i in kw["underlyings"] else v for i, v in enumerate(vols)]
        vd = [max(v - 0.01, 1e-6) if i in kw["underlyings"] else v for i, v in enumerate(vols)]
        vega = (_oracle(oracle, kw, spots, vu, CORR, z)[0] -
                _oracle(oracle, kw, spots, vd, CORR, z)[0]) / 0.02
        assert abs(s.vega - vega) <= 1e-7 * max(1.0, abs(vega)), (kw, s.vega, vega)
        n = len(spots)
        cu = np.zeros(n * n)
        cd = np.zeros(n * n)
        oracle.lib().or_bump_corr_matrix(oracle._ptr(np.ascontiguousarray(CORR).ravel()), n, 0.01,
                                         oracle._ptr(cu))
        oracle.lib().or_bump_corr_matrix(oracle._ptr(np.ascontiguousarray(CORR).ravel()), n, -0.01,
                                         oracle._ptr(cd))
        cega = (_oracle(oracle, kw, spots, vols, cu.reshape(n, n), z)[0] -
                _oracle(oracle, kw, spots, vols, cd.reshape(n, n), z)[0]) / 0.02
        assert abs(s.cega - cega) <= 1e-7 * max(1.0, abs(cega)), (kw, s.cega, cega)
        assert r.has_greeks and abs(r.delta - sum(s.delta[:s.n_delta])) < 1e-12
        assert r.rho == s.cega


def test_reference_tests_in_fast_mode(gpu):
    from openferric_amd import (Autocallable, PhoenixAutocallable, price_autocallable,
                                price_phoenix_autocallable)

    def note(c, b, ki=0.2):
        return Autocallable([0, 1], 100.0, DATES, b, c, ki, 1.0, 1.0)

    s2, c4 = [100.0, 100.0], [[1.0, 0.4], [0.4, 1.0]]
    hi = price_autocallable(note(0.08, 1.0), s2, [0.12, 0.12], c4, 0.01, 0.0, 16_000, 64).price
    lo = price_autocallable(note(0.08, 0.8), s2, [0.12, 0.12], c4, 0.01, 0.0, 16_000, 64).price
    assert lo < hi
    c3 = [[1.0, 0.3], [0.3, 1.0]]
    a = price_autocallable(note(0.02, 1.0), s2, [0.2, 0.2], c3, 0.01, 0.0, 16_000, 64).price
    b = price_autocallable(note(0.10, 1.0), s2, [0.2, 0.2], c3, 0.01, 0.0, 16_000, 64).price
    assert b > a
    one = Autocallable([0], 100.0, [0.5, 1.0], 1.0, 0.07, 0.6, 1.0, 1.0)
    shifted = Autocallable([1], 100.0, [0.5, 1.0], 1.0, 0.07, 0.6, 1.0, 1.0)
    single = price_autocallable(one, [100.0], [0.22], [[1.0]], 0.01, 0.0, 12_000, 64).price
    indexed = price_autocallable(shifted, [80.0, 100.0], [0.35, 0.22], [[1.0, -0.2], [-0.2, 1.0]],
                                 0.01, 0.0, 12_000, 64).price
    assert abs(single - indexed) < 1e-10
    std = Autocallable([0, 1], 100.0, DATES, 1.05, 0.12, 0.9, 1.0, 1.0)
    ph = PhoenixAutocallable([0, 1], 100.0, DATES, 1.05, 0.7, 0.12, True, 0.9, 1.0, 1.0)
    sp = price_autocallable(std, s2, [0.35, 0.35], [[1, 0.2], [0.2, 1]], 0.01, 0.0, 20_000, 80)
    pp = price_phoenix_autocallable(ph, s2, [0.35, 0.35], [[1, 0.2], [0.2, 1]], 0.01, 0.0,
                                    20_000, 80)
    assert pp.price > sp.price
    assert sp.diagnostics.get("observation_count") == 4.0 and sp.greeks is not None


@pytest.mark.parametrize("rho", [-0.5, 0.6])
def test_fast_mode_certain_knock_in_matches_margrabe(gpu, rho):
    from openferric_amd import Autocallable, price_autocallable
    r, T, s1, s2 = 0.02, 1.0, 0.3, 0.2
    c = Autocallable([0, 1], 100.0, DATES, 1e9, 0.0, 1e9, 1.0, T)
    res = price_autocallable(c, [100.0, 50.0], [s1, s2], [[1, rho], [rho, 1]], r, 0.0,
                             1 << 20, 16)
    sig = math.sqrt(s1 * s1 + s2 * s2 - 2 * rho * s1 * s2)
    want = math.exp(-r * T) * 100.0 * 2 * math.exp(r * T) * (1 - norm.cdf(sig * math.sqrt(T) / 2))
    assert abs(res.price - want) < 4 * res.stderr, (res.price, want, res.stderr)


def test_invalid_contract_and_reference_draws(gpu):
    from openferric_amd import (Autocallable, InvalidInput, RngMode, _lib, autocall_price_raw,
                                autocall_struct, price_autocallable)
    from openferric_amd.engines import _Keep
    bad = Autocallable([0, 0], 100.0, DATES, 1.0, 0.08, 0.6, 1.0, 1.0)
    r = price_autocallable(bad, [100.0], [0.2], [[1.0]], 0.01, 0.0, 1000, 16)
    assert math.isnan(r.price) and r.stderr is None
    assert "unique" in _lib.last_error()
    good = Autocallable([0], 100.0, DATES, 1.0, 0.08, 0.6, 1.0, 1.0)
    s = autocall_struct(good, [100.0], [0.2], [[1.0]], 0.01, 0.0, 100, 16, _Keep())
    with pytest.raises(InvalidInput, match="single stream"):
        autocall_price_raw([s], RngMode.Reference)


@pytest.mark.parametrize("mode", ["external", "fast"])
def test_fused_sensitivities_match_scenario_by_scenario(gpu, monkeypatch, mode):
    """The fused pass (every bump of a contract on each path's shared draws)
    prices each scenario as pricing it on its own does."""
    from openferric_amd import RngMode, autocall_price_raw, autocall_struct
    from openferric_amd.engines import _Keep
    rng_mode = RngMode.External if mode == "external" else RngMode.Fast
    zs = [_normals(kw, 300 + i) for i, kw in enumerate(CONTRACTS)] if mode == "external" else None

    def run():
        keep = _Keep()
        structs = [autocall_struct(_contract(kw), MARKET["spots"], MARKET["vols"], CORR, 0.02, 0.01,
                                   kw["num_paths"], kw["num_steps"], keep) for kw in CONTRACTS]
        return autocall_price_raw(structs, rng_mode, zs, with_greeks=True)

    fused, fs = run()
    monkeypatch.setenv("OFG_AUTOCALL_FUSED", "0")
    alone, as_ = run()
    for f, a, x, y in zip(fused, alone, fs, as_):
        assert abs(f.price - a.price) <= 1e-12 * abs(a.price)
        assert abs(f.stderr_ - a.stderr_) <= 1e-9 * a.stderr_
        assert x.ok and y.ok and x.n_delta == y.n_delta
        for k in range(x.n_delta):
            assert abs(x.delta[k] - y.delta[k]) <= 1e-9 * max(1.0, abs(y.delta[k]))
        assert abs(x.vega - y.vega) <= 1e-9 * max(1.0, abs(y.vega))
        assert abs(x.cega - y.cega) <= 1e-9 * max(1.0, abs(y.cega))
