"""Structured-product engine on the GPU (ofg_dsl_price through the C ABI)
against the oracle's restatement (oracle/dsl_oracle.py).

* caller normals: products using every statement and operator (nested
  if / else if / else, let, set, pay, redeem, skip, two schedules, state of
  both types, worst_of / best_of of lists and of performances, price, min,
  max, abs, exp, log, ==, !=, and / or / not) on markets mixing equity, FX,
  commodity (Schwartz) and rate (Vasicek) assets -- prices within 1e-10
  relative of the oracle fed the same normals, stderr 1e-8;
* the reference's single stream split across paths by jump-ahead: the
  same bar against the oracle walking the stream;
* several markets in one call price as separate calls do (bit for bit), and
  greeks_multi_asset / extended / cross Greeks equal the finite differences
  of oracle prices of the bumped markets;
* out-of-range price() is the reference's NumericalError (but not in a
  branch no path takes), market / engine validation messages;
* fast draws: a call against Black-Scholes within 4 standard errors.
"""
import math
import re

import numpy as np
import pytest

from oracle import dsl_oracle as D
from tests import dsl_products as P
from tests.basket_numpy import bs_price

pytestmark = pytest.mark.gpu


def _mk(spec):
    """(library market, oracle assets, corr, rate) from a spec dict."""
    from openferric_amd import AssetMarketData as A, MultiAssetMarket
    lib_assets, ora = [], []
    for a in spec["assets"]:
        k = a["kind"]
        if k == "equity":
            lib_assets.append(A.Equity(a["spot"], a["vol"], a.get("dividend_yield", 0.0)))
        elif k == "fx":
            lib_assets.append(A.Fx(a["spot"], a["vol"], a["domestic_rate"], a["foreign_rate"]))
        elif k == "commodity":
            lib_assets.append(A.Commodity(a["spot"], a["vol"], 0.0, a["kappa"], a["mu"]))
        else:
            lib_assets.append(A.Rate(a["spot"], a["vol"], a["mean_reversion"], a["long_run_mean"]))
        ora.append(dict(a))
    return MultiAssetMarket(lib_assets, spec["corr"], spec["rate"]), ora, spec["corr"], spec["rate"]


def _corr(n, seed, rho=0.4):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n + 2)) + rho * rng.normal(size=(1, n + 2))
    c = a @ a.T
    d = np.sqrt(np.diag(c))
    c = c / np.outer(d, d)
    np.fill_diagonal(c, 1.0)
    return ((c + c.T) / 2).tolist()


EQ = lambda s, v, q=0.0: dict(kind="equity", spot=s, vol=v, dividend_yield=q)
MIXED = dict(assets=[EQ(100.0, 0.22, 0.01),
                     dict(kind="fx", spot=1.1, vol=0.1, domestic_rate=0.03, foreign_rate=0.01),
                     dict(kind="commodity", spot=80.0, vol=0.35, kappa=1.5, mu=math.log(75.0)),
                     dict(kind="rate", spot=0.03, vol=0.01, mean_reversion=0.5,
                          long_run_mean=0.04)],
             corr=_corr(4, 1), rate=0.025)
THREE = dict(assets=[EQ(100.0, 0.2, 0.02), EQ(95.0, 0.25, 0.01), EQ(105.0, 0.3, 0.0)],
             corr=_corr(3, 2), rate=0.03)
TWO = dict(assets=[EQ(100.0, 0.3), EQ(90.0, 0.25, 0.02)], corr=[[1.0, 0.3], [0.3, 1.0]],
           rate=0.02)
ONE = dict(assets=[EQ(100.0, 0.2, 0.01)], corr=[[1.0]], rate=0.03)
# four assets that all step as GBM (three equities and an FX rate)
FOUR = dict(assets=[EQ(100.0, 0.2, 0.02), EQ(95.0, 0.25, 0.01), EQ(105.0, 0.3, 0.0),
                    dict(kind="fx", spot=1.2, vol=0.12, domestic_rate=0.03, foreign_rate=0.015)],
            corr=_corr(4, 4), rate=0.03)

CASES = [
    (P.KITCHEN_SINK, MIXED, 1_500, 48),
    (P.RANGE_ACCRUAL, THREE, 2_000, 24),
    (P.standard_autocall(3, 1.02, 0.09, 1.5), THREE, 2_500, 30),
    (P.phoenix_memory(2, 1.05, 0.75, 0.1, 2.0), TWO, 2_001, 16),
    (P.WORST_DIGITAL, TWO, 3_333, 5),
    (P.CALL, ONE, 4_097, 3),
    (P.standard_autocall(4, 1.0, 0.07, 1.0), FOUR, 2_222, 20),
]


def _engine(paths, steps, mode, normals=None, seed=42):
    from openferric_amd import DslMonteCarloEngine
    return DslMonteCarloEngine(paths, steps, seed, rng_mode=mode, normals=normals)


def _close(got, want, rel=1e-10, absol=1e-9):
    return abs(got - want) <= rel * abs(want) + absol


@pytest.mark.parametrize("case", range(len(CASES)))
def test_external_normals_match_oracle(gpu, case):
    from openferric_amd import RngMode, parse_and_compile
    src, spec, paths, steps = CASES[case]
    mkt, ora, corr, rate = _mk(spec)
    n = len(ora)
    z = np.random.default_rng(100 + case).standard_normal((paths, steps, n))
    prod = parse_and_compile(src)
    eng = _engine(paths, steps, RngMode.External, np.transpose(z, (1, 2, 0)))
    r = eng.price_multi_asset(prod, mkt)
    p, se = D.price(src, ora, corr, rate, paths, steps, z)
    print(case, r.price, p, flush=True)
    assert _close(r.price, p), (r.price, p)
    assert abs(r.stderr - se) <= 1e-8 * se + 1e-12, (r.stderr, se)
    assert r.diagnostics.get("num_paths") == paths and r.diagnostics.get("num_steps") == steps


# MIXED steps Schwartz and Vasicek assets too; the others are GBM-only
# markets of 3, 2, 1 and 4 assets
@pytest.mark.parametrize("case", [0, 2, 3, 5, 6])
def test_reference_draws_match_oracle(gpu, case):
    from openferric_amd import RngMode, parse_and_compile
    src, spec, paths, steps = CASES[case]
    mkt, ora, corr, rate = _mk(spec)
    r = _engine(paths, steps, RngMode.Reference, seed=7).price_multi_asset(parse_and_compile(src),
                                                                          mkt)
    p, se = D.price(src, ora, corr, rate, paths, steps, None, seed=7)
    assert _close(r.price, p), (r.price, p)
    assert abs(r.stderr - se) <= 1e-8 * se + 1e-12


def test_scenarios_share_draws_and_greeks_are_bumped_oracle_prices(gpu):
    from openferric_amd import MultiAssetMarket, RngMode, parse_and_compile
    from openferric_amd.dsl import with_spot_bump, with_vol_bump
    src, spec, paths, steps = CASES[2]
    mkt, ora, corr, rate = _mk(spec)
    prod = parse_and_compile(src)
    eng = _engine(paths, steps, RngMode.Reference, seed=11)
    alone = [eng.price_multi_asset(prod, m).price for m in (mkt, MultiAssetMarket(
        mkt.assets, mkt.correlation, mkt.rate + 0.01))]
    both = [r.price for r in eng.price_markets(prod, [mkt, MultiAssetMarket(
        mkt.assets, mkt.correlation, mkt.rate + 0.01)])]
    assert alone == both
    g = eng.greeks_multi_asset(prod, mkt, 1)

    def oprice(assets=None, rate_=None):
        return D.price(src, assets or ora, corr, rate if rate_ is None else rate_, paths, steps,
                       None, seed=11)[0]

    h = 0.01 * ora[1]["spot"]
    up = [dict(a) for a in ora]
    dn = [dict(a) for a in ora]
    vu = [dict(a) for a in ora]
    up[1]["spot"] += h
    dn[1]["spot"] -= h
    vu[1]["vol"] += 0.01
    base = oprice()
    pu, pd, pv, pr = oprice(up), oprice(dn), oprice(vu), oprice(rate_=rate + 0.0001)
    assert abs(g.delta - (pu - pd) / (2 * h)) <= 1e-7 * max(1.0, abs(g.delta))
    assert abs(g.gamma - (pu - 2 * base + pd) / (h * h)) <= 1e-5 * max(1.0, abs(g.gamma))
    assert abs(g.vega - (pv - base) / 0.01) <= 1e-7 * max(1.0, abs(g.vega))
    assert abs(g.rho - (pr - base) / 0.0001) <= 1e-5 * max(1.0, abs(g.rho))
    x = eng.extended_greeks_multi_asset(prod, mkt, 0, base)
    assert math.isfinite(x.vanna) and math.isfinite(x.volga) and x.theta == 0.0
    c = eng.cross_greeks_multi_asset(prod, mkt, 0, 2, base)
    assert math.isfinite(c.cross_gamma) and math.isfinite(c.corr_sens)


def test_errors_and_masked_index(gpu):
    from openferric_amd import InvalidInput, NumericalError, RngMode, parse_and_compile
    from openferric_amd import MultiAssetMarket, AssetMarketData as A
    eng = _engine(1000, 4, RngMode.Reference)
    one = MultiAssetMarket.single(100.0, 0.2, 0.01, 0.0)
    with pytest.raises(NumericalError, match=r"eval error: asset index 5 out of range \(have 1 "
                                             r"assets\)"):
        eng.price_multi_asset(parse_and_compile(P.OUT_OF_RANGE_PRICE), one)
    r = eng.price_multi_asset(parse_and_compile(P.MASKED_BAD_INDEX), one)
    assert math.isfinite(r.price)
    with pytest.raises(NumericalError, match="standalone"):
        eng.price_multi_asset(parse_and_compile(P.CALL.replace("max(price(0) - 100, 0)",
                                                               "performances()")), one)
    with pytest.raises(InvalidInput, match="asset vol must be finite and > 0"):
        eng.price_multi_asset(parse_and_compile(P.CALL), MultiAssetMarket.single(100, 0.0, 0, 0))
    with pytest.raises(InvalidInput, match="rate asset vol"):
        eng.price_multi_asset(parse_and_compile(P.CALL),
                              MultiAssetMarket([A.Rate(0.01, 0.0, 0.1, 0.02)], [[1.0]], 0.0))
    with pytest.raises(InvalidInput, match="num_paths must be > 0"):
        _engine(0, 4, RngMode.Reference).price_multi_asset(parse_and_compile(P.CALL), one)
    with pytest.raises(InvalidInput, match="correlation matrix diagonal must be 1"):
        eng.price_multi_asset(parse_and_compile(P.CALL), MultiAssetMarket(one.assets, [[0.9]], 0.0))
    deep = "(1 + " * 40 + "1" + ")" * 40   # 41 stack slots: beyond the device's 32 registers
    with pytest.raises(InvalidInput, match="exceeds the device limits"):
        eng.price_multi_asset(parse_and_compile(P.CALL.replace("max(price(0) - 100, 0)", deep)),
                              one)


def test_fast_draws_call_matches_black_scholes(gpu):
    from openferric_amd import RngMode, parse_and_compile
    mkt, *_ = _mk(ONE)
    r = _engine(1 << 20, 4, RngMode.Fast).price_multi_asset(parse_and_compile(P.CALL), mkt)
    want = bs_price(True, 100.0, 100.0, 0.03, 0.01, 0.2, 1.0)
    assert abs(r.price - want) < 4 * r.stderr, (r.price, want, r.stderr)


def _forward_sum(n):
    names = "\n".join(f"        A{i} = asset({i})" for i in range(n))
    total = " + ".join(f"price({i})" for i in range(n))
    return (f'product "Forward sum"\n    notional: 1\n    maturity: 1.5\n    underlyings\n'
            f'{names}\n    schedule 1.5 from 1.5 to 1.5\n        redeem {total}\n')


@pytest.mark.parametrize("spec", [ONE, TWO, THREE, FOUR], ids=["1", "2", "3", "4"])
def test_fast_draws_forwards_match_closed_form(gpu, spec):
    """Fast draws on GBM-only markets of 1 to 4 assets: the discounted sum of
    the terminal prices equals the sum of the forwards discounted at the
    market rate, within 4 standard errors."""
    from openferric_amd import RngMode, parse_and_compile
    mkt, ora, corr, rate = _mk(spec)
    T = 1.5
    want = 0.0
    for a in ora:
        mu = (a["domestic_rate"] - a["foreign_rate"]) if a["kind"] == "fx" else \
            rate - a.get("dividend_yield", 0.0)
        want += a["spot"] * math.exp((mu - rate) * T)
    r = _engine(1 << 20, 9, RngMode.Fast, seed=11).price_multi_asset(
        parse_and_compile(_forward_sum(len(ora))), mkt)
    assert abs(r.price - want) < 4 * r.stderr, (r.price, want, r.stderr)


def test_many_paths_per_thread(gpu):
    """300,001 paths: several paths per thread and a ragged last tile."""
    from openferric_amd import RngMode, parse_and_compile
    mkt, ora, corr, rate = _mk(ONE)
    paths, steps = 300_001, 2
    z = np.random.default_rng(5).standard_normal((paths, steps, 1))
    r = _engine(paths, steps, RngMode.External, np.transpose(z, (1, 2, 0))).price_multi_asset(
        parse_and_compile(P.CALL), mkt)
    p, se = D.price(P.CALL, ora, corr, rate, paths, steps, z)
    assert _close(r.price, p) and abs(r.stderr - se) <= 1e-8 * se


def test_reference_draws_several_paths_per_thread(gpu):
    """300,001 paths on the reference's stream: two paths per simulation
    thread, consecutive in the stream, and a ragged last tile (the seed
    kernel's one state per thread); one market and two markets simulated
    together, both equal to the oracle's stream order."""
    from openferric_amd import MultiAssetMarket, RngMode, parse_and_compile
    mkt, ora, corr, rate = _mk(TWO)
    paths, steps = 300_001, 2
    eng = _engine(paths, steps, RngMode.Reference, seed=21)
    prod = parse_and_compile(P.WORST_DIGITAL)
    r = eng.price_multi_asset(prod, mkt)
    p, se = D.price(P.WORST_DIGITAL, ora, corr, rate, paths, steps, None, seed=21)
    assert _close(r.price, p) and abs(r.stderr - se) <= 1e-8 * se, (r.price, p)
    both = eng.price_markets(prod, [mkt, MultiAssetMarket(mkt.assets, mkt.correlation,
                                                          mkt.rate + 0.01)])
    assert _close(both[0].price, p), (both[0].price, p)


def test_fused_form_matches_split_form(gpu, monkeypatch):
    """The single-kernel form (used when all snapshots exceed the buffer
    budget) prices as the two-kernel form."""
    from openferric_amd import RngMode, parse_and_compile
    src, spec, paths, steps = CASES[0]
    mkt, *_ = _mk(spec)
    prod = parse_and_compile(src)
    split = _engine(paths, steps, RngMode.Reference, seed=3).price_multi_asset(prod, mkt)
    monkeypatch.setenv("OFG_DSL_SPLIT_BUDGET", "0")
    fused = _engine(paths, steps, RngMode.Reference, seed=3).price_multi_asset(prod, mkt)
    assert abs(fused.price - split.price) <= 1e-12 * abs(split.price)
    assert abs(fused.stderr - split.stderr) <= 1e-9 * split.stderr


@pytest.mark.parametrize("case", [0, 2, 3, 6])
def test_markets_fused_simulation_matches_separate(gpu, monkeypatch, case):
    """Greeks bumps simulated in one pass (shared draws and factors) price as
    markets simulated one by one -- equity, FX, Schwartz and Vasicek assets,
    spot / vol / rate / correlation bumps."""
    from openferric_amd import MultiAssetMarket, RngMode, parse_and_compile
    from openferric_amd.dsl import with_spot_bump, with_vol_bump
    src, spec, paths, steps = CASES[case]
    mkt, *_ = _mk(spec)
    prod = parse_and_compile(src)
    n = len(mkt.assets)
    corr_up = [list(r) for r in mkt.correlation]
    corr_up[0][1] = corr_up[1][0] = min(corr_up[0][1] + 0.01, 1.0)
    bumps = [mkt]
    for i in range(n):
        a = list(mkt.assets)
        a[i] = with_spot_bump(a[i], 0.01 * a[i].initial_value())
        bumps.append(MultiAssetMarket(a, mkt.correlation, mkt.rate))
    a = list(mkt.assets)
    a[0] = with_vol_bump(a[0], 0.01)
    bumps.append(MultiAssetMarket(a, mkt.correlation, mkt.rate))
    bumps.append(MultiAssetMarket(mkt.assets, mkt.correlation, mkt.rate + 0.0001))
    bumps.append(MultiAssetMarket(mkt.assets, corr_up, mkt.rate))
    bumps = bumps[:9]
    for mode in (RngMode.Reference, RngMode.Fast):
        eng = _engine(paths, steps, mode, seed=5)
        fused = [r.price for r in eng.price_markets(prod, bumps)]
        monkeypatch.setenv("OFG_DSL_FUSE", "0")
        apart = [r.price for r in eng.price_markets(prod, bumps)]
        monkeypatch.delenv("OFG_DSL_FUSE")
        for f, s_ in zip(fused, apart):
            assert abs(f - s_) <= 1e-12 * abs(s_) + 1e-12, (f, s_)


# ---- native evaluation kernels (run-time compiled product programs) ----

def _all_products():
    """Every product of the test corpus that the library compiles, with a
    market of its underlyings' count."""
    from openferric_amd import parse_and_compile
    out = []
    for name, src in sorted(vars(P).items()):
        if not (isinstance(src, str) and src.startswith("product")):
            continue
        try:
            prod = parse_and_compile(src)
        except Exception:
            continue
        n = max(1, prod.num_underlyings)
        spec = {1: ONE, 2: TWO, 3: THREE, 4: MIXED}.get(n)
        if spec is not None:
            out.append((name, src, spec))
    for n, spec in ((1, ONE), (2, TWO), (3, THREE), (4, FOUR)):
        out.append((f"standard_autocall_{n}", P.standard_autocall(n, 1.0, 0.07, 1.0), spec))
        out.append((f"phoenix_memory_{n}", P.phoenix_memory(n, 1.05, 0.75, 0.1, 2.0), spec))
        out.append((f"forward_sum_{n}", _forward_sum(n), spec))
    import bench
    out.append(("bench_worst_of_autocall", bench.DSL_PRODUCT, THREE))
    return out


@pytest.mark.parametrize("mode", ["reference", "fast", "external"])
def test_native_kernel_prices_equal_interpreter(gpu, mode):
    """Every corpus product, priced by the interpreter and then by its native
    kernel, on the same draws: prices and standard errors bit for bit; each
    native price is one launch of the native kernel."""
    from openferric_amd import RngMode, parse_and_compile
    from openferric_amd.dsl import jit_stats
    rng = {"reference": RngMode.Reference, "fast": RngMode.Fast, "external": RngMode.External}[mode]
    checked = 0
    for name, src, spec in _all_products():
        mkt, ora, corr, rate = _mk(spec)
        paths, steps = 3_001, 12
        z = None
        if rng == RngMode.External:
            z = np.transpose(np.random.default_rng(9).standard_normal((paths, steps, len(ora))),
                             (1, 2, 0))
        eng = _engine(paths, steps, rng, normals=z, seed=3)
        prod = parse_and_compile(src)
        try:
            want = eng.price_multi_asset(prod, mkt)
        except Exception as e:   # the product's own error: the native kernel must give it too
            try:
                prod.jit_compile(len(ora))
            except type(e):
                continue   # raised when building the program, before any kernel
            with pytest.raises(type(e), match=re.escape(str(e))):
                eng.price_multi_asset(prod, mkt)
            continue
        ms = prod.jit_compile(len(ora))
        assert ms > 0.0
        assert prod.jit_compile(len(ora)) == 0.0   # attached once per context
        before = jit_stats()[1]
        got = eng.price_multi_asset(prod, mkt)
        assert jit_stats()[1] == before + 1, name
        assert got.price == want.price and got.stderr == want.stderr, (name, got.price, want.price)
        checked += 1
    assert checked >= 15, checked


def test_native_kernel_greeks_markets_and_errors(gpu):
    """The native kernel under several markets in one call (the fused
    simulation) and the interpreter's error for an out-of-range price()
    index, with the masked index of a branch no path takes still allowed."""
    from openferric_amd import MultiAssetMarket, NumericalError, RngMode, parse_and_compile
    src, spec, paths, steps = CASES[2]
    mkt, *_ = _mk(spec)
    bumps = [mkt, MultiAssetMarket(mkt.assets, mkt.correlation, mkt.rate + 0.01)]
    eng = _engine(paths, steps, RngMode.Reference, seed=13)
    prod = parse_and_compile(src)
    want = [r.price for r in eng.price_markets(prod, bumps)]
    g_want = eng.greeks_multi_asset(prod, mkt, 1)
    prod.jit_compile(3)
    assert [r.price for r in eng.price_markets(prod, bumps)] == want
    g = eng.greeks_multi_asset(prod, mkt, 1)
    assert (g.delta, g.gamma, g.vega, g.rho) == (g_want.delta, g_want.gamma, g_want.vega,
                                                 g_want.rho)
    one = MultiAssetMarket.single(100.0, 0.2, 0.01, 0.0)
    e1 = _engine(1000, 4, RngMode.Reference)
    bad = parse_and_compile(P.OUT_OF_RANGE_PRICE)
    bad.jit_compile(1)
    with pytest.raises(NumericalError, match=r"eval error: asset index 5 out of range \(have 1 "
                                             r"assets\)"):
        e1.price_multi_asset(bad, one)
    masked = parse_and_compile(P.MASKED_BAD_INDEX)
    want_m = e1.price_multi_asset(masked, one).price
    masked.jit_compile(1)
    assert e1.price_multi_asset(masked, one).price == want_m


def test_native_kernel_invalid_arguments(gpu):
    from openferric_amd import InvalidInput, parse_and_compile
    prod = parse_and_compile(P.CALL)
    with pytest.raises(InvalidInput, match="n_assets must be in 1..16"):
        prod.jit_compile(0)
    with pytest.raises(InvalidInput, match="n_assets must be in 1..16"):
        prod.jit_compile(17)
