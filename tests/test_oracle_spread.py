"""Pins the oracle's Pcg64 stream and spread-option engine
(oracle/oracle_spread.c: fast_rng.rs:93-137, spread_mc.rs:83-200):

* Pcg64: numpy's PCG64 is PCG XSL-RR 128/64 with the same multiplier but
  outputs the state AFTER each step, so from the reference's seeded state
  (before its discarded draw) numpy's outputs are the reference's draws
  after the discard: exact u64 equality, 4,000 draws x 8 seeds (every
  rotation count occurs).
* Spread engine: Margrabe's closed form at k = 0 within 4 SE (both kinds,
  antithetic on and off, rho of both signs); with vol2 = 0 the spread is a
  Black-Scholes call on S1 struck at S2's forward plus k, with vol1 = 0 a put
  on S2 struck at S1's forward minus k (4 SE); the reference's own test
  (Kirk within 1 % at 300,000 paths, seed 13); each sample's payoff equals
  the payoff of the pinned streams' draws, computed here; the statistics;
  the early returns and validation messages in the reference's order.
* The analytic restatements used above reproduce the reference's tests
  (src/engines/analytic/spread.rs:155-232).
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

XO, PCG = O.RNG_XOSHIRO, O.RNG_PCG64


def _splitmix3(seed):
    L = O.lib()
    sm = C.c_uint64(seed)
    return [L.or_splitmix64_next(C.byref(sm)) for _ in range(3)]


@pytest.mark.parametrize("seed", [0, 1, 13, 42, 7919, 0xDEADBEEFCAFEBABE, 2**63, 2**64 - 1])
def test_pcg64_equals_numpy_pcg64(seed):
    hi, lo, stream = _splitmix3(seed)
    g = np.random.PCG64()
    g.state = {"bit_generator": "PCG64", "state": {"state": (hi << 64) | lo,
                                                   "inc": (stream << 1) | 1},
               "has_uint32": 0, "uinteger": 0}
    want = [int(x) for x in g.random_raw(4000)]
    assert O.pcg64_stream(seed, 4000) == want


def test_pcg64_next_f64_is_top_53_bits():
    st = (C.c_uint64 * 4)()
    st2 = (C.c_uint64 * 4)()
    O.lib().or_pcg64_seed(99, st)
    O.lib().or_pcg64_seed(99, st2)
    for _ in range(100):
        u = O.lib().or_pcg64_next_f64(st)
        assert u == (O.lib().or_pcg64_next_u64(st2) >> 11) * 2.0 ** -53


def _bs(call, s, k, r, q, v, t):
    vt = v * math.sqrt(t)
    d1 = (math.log(s / k) + (r - q + 0.5 * v * v) * t) / vt
    d2 = d1 - vt
    N = O._ncdf
    if call:
        return s * math.exp(-q * t) * N(d1) - k * math.exp(-r * t) * N(d2)
    return k * math.exp(-r * t) * N(-d2) - s * math.exp(-q * t) * N(-d1)


def _opt(**kw):
    o = dict(s1=100.0, s2=96.0, k=3.0, vol1=0.2, vol2=0.15, rho=0.5, q1=0.0, q2=0.0, r=0.05,
             t=0.5)
    o.update(kw)
    return o


MARGRABE = [_opt(s1=100.0, s2=105.0, k=0.0, q1=0.04, q2=0.06, t=1.0),
            _opt(s1=110.0, s2=100.0, k=0.0, vol1=0.22, vol2=0.18, rho=-0.6, q1=0.01, q2=0.02,
                 r=0.03, t=1.2)]


@pytest.mark.parametrize("kind", [XO, PCG])
@pytest.mark.parametrize("anti", [True, False])
@pytest.mark.parametrize("case", range(len(MARGRABE)))
def test_zero_strike_is_margrabe(kind, anti, case):
    o = MARGRABE[case]
    p, se = O.spread_mc(o, 200_000, 5 + case, antithetic=anti, rng_kind=kind)
    m = O.margrabe_exchange_price(o)
    assert abs(p - m) <= 4 * se, (p, m, se)


@pytest.mark.parametrize("kind", [XO, PCG])
def test_zero_vol_legs_are_black_scholes(kind):
    o = _opt(vol2=0.0, q1=0.01, q2=0.03, k=4.0)
    p, se = O.spread_mc(o, 200_000, 21, rng_kind=kind)
    strike = o["s2"] * math.exp((o["r"] - o["q2"]) * o["t"]) + o["k"]
    want = _bs(True, o["s1"], strike, o["r"], o["q1"], o["vol1"], o["t"])
    assert abs(p - want) <= 4 * se, (p, want, se)
    o = _opt(vol1=0.0, vol2=0.3, s1=110.0, q1=0.02, q2=0.01, k=2.0)
    p, se = O.spread_mc(o, 200_000, 22, rng_kind=kind)
    strike = o["s1"] * math.exp((o["r"] - o["q1"]) * o["t"]) - o["k"]
    want = _bs(False, o["s2"], strike, o["r"], o["q2"], o["vol2"], o["t"])
    assert abs(p - want) <= 4 * se, (p, want, se)


def test_reference_test_kirk_within_one_percent():
    """spread_mc.rs:206-241"""
    o = _opt()
    mc, _ = O.spread_mc(o, 300_000, 13)
    kirk = O.kirk_spread_price(o)
    assert abs((mc - kirk) / kirk) <= 0.01


def test_analytic_restatements_reproduce_reference_tests():
    # spread.rs:157-191, 193-209, 211-231
    o = _opt(s1=100.0, s2=105.0, k=0.0, q1=0.04, q2=0.06, t=1.0)
    m = O.margrabe_exchange_price(o)
    sigma = math.sqrt(0.2 ** 2 - 2 * 0.5 * 0.2 * 0.15 + 0.15 ** 2)
    assert abs(sigma - 0.1802775638) <= 1e-10
    ratio = o["s2"] * math.exp(-o["q1"] * o["t"]) * _bs(True, o["s1"] / o["s2"], 1.0,
                                                        o["q2"] - o["q1"], 0.0, sigma, o["t"])
    assert abs(m - ratio) <= 1e-10 and abs(m - 5.68713039) <= 2e-5
    assert abs(O.kirk_spread_price(_opt()) - 5.57702191) <= 2e-5
    o = _opt(s1=110.0, s2=100.0, k=0.0, vol1=0.22, vol2=0.18, rho=0.4, q1=0.01, q2=0.02,
             r=0.03, t=1.2)
    assert abs(O.kirk_spread_price(o) - O.margrabe_exchange_price(o)) <= 1e-10


def _stream_normals(kind, seed, i):
    L = O.lib()
    st = (C.c_uint64 * 4)()
    s = O.lib().or_stream_seed(seed, i)
    if kind == PCG:
        L.or_pcg64_seed(s, st)
        nxt = L.or_pcg64_next_f64
    else:
        L.or_xoshiro_seed(s, st)
        nxt = L.or_xoshiro_next_f64
    z1 = L.or_acklam_inv_cdf(L.or_uniform_open01(nxt(st)))
    z2 = L.or_acklam_inv_cdf(L.or_uniform_open01(nxt(st)))
    return z1, z2


def _payoff(o, z1, z2):
    ct = math.sqrt(max(1.0 - o["rho"] ** 2, 0.0))
    st = math.sqrt(o["t"])
    w2 = o["rho"] * z1 + ct * z2
    a = o["s1"] * math.exp(o["vol1"] * st * z1 + (o["r"] - o["q1"] - 0.5 * o["vol1"] ** 2) * o["t"])
    b = o["s2"] * math.exp(o["vol2"] * st * w2 + (o["r"] - o["q2"] - 0.5 * o["vol2"] ** 2) * o["t"])
    return max(a - b - o["k"], 0.0)


@pytest.mark.parametrize("kind", [XO, PCG])
@pytest.mark.parametrize("anti", [True, False])
def test_sample_payoffs_are_the_streams(kind, anti):
    o = _opt(k=1.0, rho=-0.3)
    n = 401
    p, se, pay = O.spread_mc(o, n, 77, antithetic=anti, rng_kind=kind, want_payoffs=True)
    samples = (n + 1) // 2 if anti else n
    assert len(pay) == samples
    want = np.empty(samples)
    for i in range(samples):
        z1, z2 = _stream_normals(kind, 77, i)
        x = _payoff(o, z1, z2)
        if anti:
            x = 0.5 * (x + _payoff(o, -z1, -z2))
        want[i] = x
    assert np.allclose(pay, want, rtol=1e-13, atol=1e-12)
    disc = math.exp(-o["r"] * o["t"])
    assert abs(p - disc * want.mean()) <= 1e-12 * abs(p)
    assert abs(se - disc * want.std(ddof=1) / math.sqrt(samples)) <= 1e-9 * se


def test_early_returns_and_messages():
    p, se = O.spread_mc(_opt(t=0.0, s1=110.0), 10, 1)
    assert p == 11.0 and se == 0.0
    p, se = O.spread_mc(_opt(), 1, 1)   # one sample: variance 0
    assert se == 0.0 and p >= 0.0
    cases = [(dict(s1=-1.0, k=-1.0), "spread spots s1 and s2 must be > 0"),
             (dict(s2=0.0), "spread spots s1 and s2 must be > 0"),
             (dict(k=-0.5, vol1=-1.0), "spread strike k must be >= 0"),
             (dict(vol2=-0.1, rho=2.0), "spread volatilities vol1 and vol2 must be >= 0"),
             (dict(rho=1.5, t=-1.0), "spread correlation rho must be in [-1, 1]"),
             (dict(rho=float("nan")), "spread correlation rho must be in [-1, 1]"),
             (dict(t=-1.0), "spread maturity t must be >= 0")]
    for kw, msg in cases:
        with pytest.raises(ValueError, match=r"^" + msg.replace("[", r"\[").replace("]", r"\]")):
            O.spread_mc(_opt(**kw), 10, 1)
    with pytest.raises(ValueError, match="num_paths must be > 0"):
        O.spread_mc(_opt(), 0, 1)
    # validation comes before the path-count check, and NaN spots pass it
    with pytest.raises(ValueError, match="spots"):
        O.spread_mc(_opt(s1=-1.0), 0, 1)
    p, _ = O.spread_mc(_opt(s1=float("nan")), 10, 1)
    assert p == 0.0   # fmax(NaN - ..., 0) is 0, as Rust's f64::max


def test_engine_streams_of_the_pcg64_kind():
    """MonteCarloEngine's per-sample streams of the Pcg64 kind (simulation.rs
    :215-225 with FastRngKind::Pcg64): the oracle's sample normals are the
    Acklam inverses of the pinned Pcg64 draws, z1 and z2 interleaved, and a
    GBM European priced on them lands within 4 SE of Black-Scholes."""
    L = O.lib()
    with O.rng_kind(PCG):
        z1, z2 = O.sample_normals(19, 7, 5, 2)
        o = O.make_option(spot=100.0, strike=105.0, expiry=1.0, rate=0.03, vol=0.25)
        p, se = O.mc_run(o, 100_000, 4, 11)
    assert L.or_rng_kind() == XO
    st = (C.c_uint64 * 4)()
    L.or_pcg64_seed(L.or_stream_seed(19, 7), st)
    want = [L.or_acklam_inv_cdf(L.or_uniform_open01(L.or_pcg64_next_f64(st))) for _ in range(10)]
    assert list(z1) == want[0::2] and list(z2) == want[1::2]
    bs = _bs(True, 100.0, 105.0, 0.03, 0.0, 0.25, 1.0)
    assert abs(p - bs) <= 4 * se
    p_x, _ = O.mc_run(o, 100_000, 4, 11)
    assert p_x != p   # the kinds draw different streams
