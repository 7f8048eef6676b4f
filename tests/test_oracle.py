"""Pins the CPU oracle (CPU only, no GPU).

Each check is chosen so a plausible restatement mistake (wrong constant,
shift, sign, index, dropped term, wrong operand) fails: known-answer vectors,
hand-derived unit-state transitions, high-precision mathematics, closed forms,
exact small sums, and the reference's own test assertions and fixtures.
"""
import ctypes as C
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fixtures():
    with open(os.path.join(GOLD, "reference_fixtures.json")) as f:
        return json.load(f)


# ----------------------------------------------------------------- RNG ----

def test_splitmix64_known_answer(oracle):
    # Published SplitMix64 outputs for seed 0 (Steele, Lea, Flood 2014;
    # Vigna's splitmix64.c reference implementation).
    s = (C.c_uint64 * 1)(0)
    got = [oracle.lib().or_splitmix64_next(s) for _ in range(4)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                   0xF88BB8A8724C81EC]


def _state(*v):
    return (C.c_uint64 * 4)(*v)


def test_xoshiro_output_function_on_unit_states(oracle):
    # result = rotl(s0 + s3, 23) + s0 (Blackman & Vigna 2018, xoshiro256++)
    L = oracle.lib()
    assert L.or_xoshiro_next_u64(_state(1, 0, 0, 0)) == (1 << 23) + 1
    assert L.or_xoshiro_next_u64(_state(0, 0, 0, 1)) == 1 << 23
    assert L.or_xoshiro_next_u64(_state(0, 5, 7, 0)) == 0
    # rotation wraps: s0 + s3 = 2^63 -> rotl 23 = 2^22
    assert L.or_xoshiro_next_u64(_state(0, 0, 0, 1 << 63)) == 1 << 22


def test_xoshiro_state_transition_on_unit_states(oracle):
    # Hand-derived from the update t = s1 << 17; s2 ^= s0; s3 ^= s1; s1 ^= s2;
    # s0 ^= s3; s2 ^= t; s3 = rotl(s3, 45).
    L = oracle.lib()
    s = _state(0, 1, 0, 0)
    L.or_xoshiro_next_u64(s)
    assert list(s) == [1, 1, 1 << 17, 1 << 45]
    s = _state(1, 0, 0, 0)
    L.or_xoshiro_next_u64(s)
    assert list(s) == [1, 1, 1, 0]
    s = _state(0, 0, 1, 0)
    L.or_xoshiro_next_u64(s)
    assert list(s) == [0, 1, 1, 0]
    s = _state(0, 0, 0, 1)
    L.or_xoshiro_next_u64(s)
    assert list(s) == [1, 0, 0, 1 << 45]


def test_xoshiro_transition_is_gf2_linear(oracle):
    rng = np.random.default_rng(5)
    L = oracle.lib()
    for _ in range(50):
        a = [int(x) for x in rng.integers(0, 2**63, 4, dtype=np.uint64)]
        b = [int(x) for x in rng.integers(0, 2**63, 4, dtype=np.uint64)]
        sa, sb, sab = _state(*a), _state(*b), _state(*[x ^ y for x, y in zip(a, b)])
        for s in (sa, sb, sab):
            L.or_xoshiro_next_u64(s)
        assert [x ^ y for x, y in zip(sa, sb)] == list(sab)


def test_xoshiro_seed_uses_splitmix_and_reproduces(oracle):
    # fast_rng.rs:250-277: same seed reproduces; next_f64 in [0, 1)
    L = oracle.lib()
    s = (C.c_uint64 * 4)()
    L.or_xoshiro_seed(0, s)
    assert list(s) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                       0xF88BB8A8724C81EC]
    a, b = (C.c_uint64 * 4)(), (C.c_uint64 * 4)()
    L.or_xoshiro_seed(42, a)
    L.or_xoshiro_seed(42, b)
    assert [L.or_xoshiro_next_u64(a) for _ in range(128)] == \
        [L.or_xoshiro_next_u64(b) for _ in range(128)]
    L.or_xoshiro_seed(1, a)
    for _ in range(1000):
        u = L.or_xoshiro_next_f64(a)
        assert 0.0 <= u < 1.0


def test_stream_seed_wraps(oracle):
    L = oracle.lib()
    assert L.or_stream_seed(42, 3) == 42 + 3 * 7919
    assert L.or_stream_seed(2**64 - 1, 1) == 7918


def test_uniform_open01_clamp(oracle):
    L = oracle.lib()
    eps = 2.220446049250313e-16
    assert L.or_uniform_open01(0.0) == eps
    assert L.or_uniform_open01(1.0 - 1e-17) == 1.0 - eps
    assert L.or_uniform_open01(0.5) == 0.5


# ------------------------------------------------------------- normals ----

def test_acklam_against_high_precision_inverse(oracle):
    with open(os.path.join(GOLD, "inverse_normal_mpmath.json")) as f:
        grid = json.load(f)
    L = oracle.lib()
    worst = 0.0
    for p, z in grid:
        got = L.or_acklam_inv_cdf(p)
        rel = abs(got - z) / max(abs(z), 1e-300) if z != 0 else abs(got)
        worst = max(worst, rel)
    # Acklam's published relative error bound is 1.15e-9
    assert worst <= 1.2e-9, worst


def test_acklam_branch_points_and_guards(oracle):
    L = oracle.lib()
    assert L.or_acklam_inv_cdf(0.5) == 0.0
    assert math.isinf(L.or_acklam_inv_cdf(0.0)) and L.or_acklam_inv_cdf(0.0) < 0
    assert math.isinf(L.or_acklam_inv_cdf(1.0)) and L.or_acklam_inv_cdf(1.0) > 0
    assert math.isnan(L.or_acklam_inv_cdf(1.5))
    assert math.isnan(L.or_acklam_inv_cdf(float("nan")))
    # symmetry across the tails: z(1-p) = -z(p) to rounding of 1 - p
    for p in (1e-6, 1e-3, 0.02, 0.0242, 0.3):
        assert abs(L.or_acklam_inv_cdf(1 - p) + L.or_acklam_inv_cdf(p)) < 1e-9
    # fast_norm.rs:209-219
    assert abs(L.or_acklam_inv_cdf(0.841344746068543) - 1.0) < 1e-6
    assert abs(L.or_acklam_inv_cdf(0.9772498680518208) - 2.0) < 1e-6


def test_hart_cdf_table(oracle):
    # fast_norm.rs:156-193 (A&S table 26.1, tol 1e-7; symmetry 1e-12)
    table = [(-8.0, 6.22096057427178e-16), (-5.0, 2.866515718791939e-7),
             (-4.0, 3.167124183311998e-5), (-3.0, 0.0013498980316300946),
             (-2.0, 0.02275013194817921), (-1.0, 0.15865525393145702),
             (-0.5, 0.3085375387259869), (0.0, 0.5), (0.5, 0.6914624612740131),
             (1.0, 0.8413447460685429), (2.0, 0.9772498680518208), (3.0, 0.99865010196837),
             (4.0, 0.9999683287581669), (5.0, 0.9999997133484281), (8.0, 1.0 - 6.22096057427178e-16)]
    L = oracle.lib()
    for x, e in table:
        assert abs(L.or_hart_norm_cdf(x) - e) < 1e-7
    for i in range(81):
        x = i / 10.0
        assert abs(L.or_hart_norm_cdf(x) + L.or_hart_norm_cdf(-x) - 1.0) < 1e-12
    for i in range(1, 1000):
        p = i / 1000.0
        assert abs(L.or_hart_norm_cdf(L.or_acklam_inv_cdf(p)) - p) < 2e-7


# ---------------------------------------------------------- generators ----

def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_gbm_generator_closed_forms(oracle):
    L = oracle.lib()
    # stochastic.rs:108-116: one unit step, zero noise -> 100 e^{0.03}
    z = np.zeros(1)
    out = np.zeros(2)
    L.or_gbm_generate(0.05, 0.2, 100.0, 1.0, 1, _ptr(z), _ptr(out))
    assert abs(out[1] - 103.0454533953517) <= 1e-12 * 103.0454533953517
    # zero noise: S_n = S0 exp(n drift)
    n = 252
    z = np.zeros(n)
    out = np.zeros(n + 1)
    mu, sig, T = 0.07, 0.3, 2.0
    L.or_gbm_generate(mu, sig, 50.0, T, n, _ptr(z), _ptr(out))
    dt = T / n
    drift = (mu - 0.5 * sig * sig) * dt
    j = np.arange(n + 1)
    np.testing.assert_allclose(out, 50.0 * np.exp(j * drift), rtol=1e-13)
    # log-increments: ln(S_n/S_0) = n drift + sigma sqrt(dt) sum z
    z = np.random.default_rng(1).standard_normal(n)
    L.or_gbm_generate(mu, sig, 50.0, T, n, _ptr(z), _ptr(out))
    lhs = np.log(out[-1] / 50.0)
    rhs = n * drift + sig * math.sqrt(dt) * z.sum()
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs))
    assert out[0] == 50.0


def test_heston_generator_special_cases(oracle):
    L = oracle.lib()
    n = 100
    T = 1.0
    dt = T / n
    kappa, theta, v0, rho = 2.0, 0.09, 0.04, -0.7
    rng = np.random.default_rng(2)
    z1, z2 = rng.standard_normal(n), rng.standard_normal(n)
    out = np.zeros(n + 1)
    # xi = 0: deterministic Euler variance v_j = theta + (v0 - theta)(1 - kappa dt)^j,
    # and log S increments (mu - v/2) dt + sqrt(v dt) zs with zs = rho z1 + sqrt(1-rho^2) z2
    mu = 0.03
    L.or_heston_generate(mu, kappa, theta, 0.0, rho, v0, 100.0, T, n, _ptr(z1), _ptr(z2),
                         _ptr(out))
    v = theta + (v0 - theta) * (1 - kappa * dt) ** np.arange(n)
    zs = rho * z1 + math.sqrt(1 - rho * rho) * z2
    logs = np.log(100.0) + np.concatenate([[0.0], np.cumsum((mu - 0.5 * v) * dt +
                                                             np.sqrt(v * dt) * zs)])
    np.testing.assert_allclose(np.log(out), logs, rtol=0, atol=1e-12)
    # xi = 0 and v0 = theta: a GBM with sigma = sqrt(theta) driven by zs
    L.or_heston_generate(mu, kappa, theta, 0.0, rho, theta, 100.0, T, n, _ptr(z1), _ptr(z2),
                         _ptr(out))
    g = np.zeros(n + 1)
    L.or_gbm_generate(mu, math.sqrt(theta), 100.0, T, n, _ptr(np.ascontiguousarray(zs)), _ptr(g))
    np.testing.assert_allclose(out, g, rtol=1e-12)
    # clamps: v >= 0 (stochastic.rs:118-131) and S >= 1e-12
    s_next, v_next = C.c_double(), C.c_double()
    L.or_heston_step_euler(0.03, 2.0, 0.04, 0.7, -0.6, 100.0, 0.001, 1.0 / 252, -15.0, 1.2,
                           C.byref(s_next), C.byref(v_next))
    assert v_next.value >= 0.0
    zz = np.full(n, -40.0)
    L.or_heston_generate(mu, kappa, 4.0, 0.0, 0.0, 4.0, 100.0, T, n, _ptr(zz), _ptr(zz), _ptr(out))
    assert out.min() >= 1e-12


def test_heston_step_by_brute_force_long_double(oracle):
    # Recompute one Euler step in exact rational arithmetic on representable
    # inputs and compare with the oracle to rounding (no sqrt/exp involved
    # in the variance update for v = 0.04 = 0.2^2, sqrt exact enough).
    L = oracle.lib()
    s_next, v_next = C.c_double(), C.c_double()
    kappa, theta, xi, rho, v, dt, z1, z2 = 1.5, 0.0625, 0.5, -0.75, 0.25, 0.25, 0.5, -1.0
    L.or_heston_step_euler(0.0, kappa, theta, xi, rho, 100.0, v, dt, z1, z2, C.byref(s_next),
                           C.byref(v_next))
    F = Fraction
    vn = F(v) + F(kappa) * (F(theta) - F(v)) * F(dt) + F(xi) * F(1, 2) * F(1, 2) * F(z1)
    assert v_next.value == float(vn)
    zs = rho * z1 + math.sqrt(1 - rho * rho) * z2
    e = (0.0 - 0.5 * v) * dt + 0.5 * 0.5 * zs
    assert abs(s_next.value - 100.0 * math.exp(e)) <= 1e-13 * 100


# ------------------------------------------------------------- payoffs ----

def test_payoffs_on_hand_paths(oracle):
    L = oracle.lib()
    path = np.array([100.0, 110.0, 95.0, 120.0, 105.0])
    # vanilla: terminal only
    o = oracle.make_option(payoff=oracle.VANILLA, strike=100.0)
    assert L.or_payoff(C.byref(o), _ptr(path), 5) == 5.0
    o = oracle.make_option(payoff=oracle.VANILLA, option_type=oracle.PUT, strike=110.0)
    assert L.or_payoff(C.byref(o), _ptr(path), 5) == 5.0
    # barrier inclusive, includes path[0]
    for lvl, hit_up, hit_dn in [(120.0, True, True), (120.5, False, True), (95.0, True, True),
                                (94.9, True, False), (100.0, True, True)]:
        for d, hit in ((oracle.UP, hit_up), (oracle.DOWN, hit_dn)):
            oi = oracle.make_option(payoff=oracle.BARRIER, strike=100.0, barrier_dir=d,
                                    barrier_style=oracle.IN, barrier_level=lvl, rebate=0.0)
            oo = oracle.make_option(payoff=oracle.BARRIER, strike=100.0, barrier_dir=d,
                                    barrier_style=oracle.OUT, barrier_level=lvl, rebate=0.0)
            xi_, xo = L.or_payoff(C.byref(oi), _ptr(path), 5), L.or_payoff(C.byref(oo), _ptr(path), 5)
            assert xi_ + xo == 5.0  # in + out = vanilla per path
            assert (xi_ == 5.0) == hit
    # rebate paid when inactive
    o = oracle.make_option(payoff=oracle.BARRIER, strike=100.0, barrier_dir=oracle.UP,
                           barrier_style=oracle.OUT, barrier_level=115.0, rebate=2.5)
    assert L.or_payoff(C.byref(o), _ptr(path), 5) == 2.5
    # Asian index map round((t/T) n): T=1, n=4 -> t=0.3 -> idx round(1.2)=1,
    # t=0.375 -> round(1.5)=2 (half away from zero), t=1 -> 4
    o = oracle.make_option(payoff=oracle.ASIAN, strike=100.0, expiry=1.0,
                           obs_times=[0.3, 0.375, 1.0])
    assert L.or_payoff(C.byref(o), _ptr(path), 5) == (110.0 + 95.0 + 105.0) / 3 - 100.0
    o = oracle.make_option(payoff=oracle.ASIAN, strike=100.0, expiry=1.0, averaging=oracle.GEO,
                           obs_times=[0.0, 0.75])
    assert abs(L.or_payoff(C.byref(o), _ptr(path), 5) - (math.sqrt(100.0 * 120.0) - 100.0)) < 1e-12
    o = oracle.make_option(payoff=oracle.ASIAN, strike=100.0, expiry=1.0,
                           strike_type=oracle.FLOATING, option_type=oracle.CALL, obs_times=[0.25, 0.5])
    # floating strike: vanilla(type, S_T, avg) (mc_engine.rs:540-543)
    assert L.or_payoff(C.byref(o), _ptr(path), 5) == 105.0 - (110.0 + 95.0) / 2
    o.option_type = oracle.PUT
    assert L.or_payoff(C.byref(o), _ptr(path), 5) == 0.0


# --------------------------------------------------------- statistics -----

def test_run_close_matches_exact_rational_sums(oracle):
    o = oracle.make_option(strike=95.0)
    for anti in (False, True):
        price, se, xs, ys = oracle.mc_run(o, 7, 3, 11, antithetic=anti, want_samples=True)
        n = len(xs)
        F = [Fraction(x) for x in xs]
        mean = sum(F) / n
        var = (sum(f * f for f in F) - sum(F) ** 2 / n) / (n - 1)
        df = math.exp(-0.05)
        assert abs(price - df * float(mean)) <= 1e-14 * abs(price)
        assert abs(se - df * math.sqrt(float(var) / n)) <= 1e-12 * se


def test_cv_close_equals_regression_estimator(oracle):
    obs = [m / 11.0 for m in range(12)]
    o = oracle.make_option(payoff=oracle.ASIAN, obs_times=obs)
    e, kind = oracle.builtin_cv_expected(o)
    assert kind == 1
    price, se, xs, ys = oracle.mc_run(o, 4000, 24, 3, control_variate=True, want_samples=True)
    beta = np.cov(xs, ys, ddof=1)[0, 1] / np.var(ys, ddof=1)
    adj = xs + beta * (e - ys)
    df = math.exp(-0.05)
    assert abs(price - df * adj.mean()) <= 1e-10 * abs(price)
    assert abs(se - df * adj.std(ddof=1) / math.sqrt(len(adj))) <= 1e-8 * se


def test_threaded_fold_equals_serial(oracle):
    o = oracle.make_option()
    p1, s1 = oracle.mc_run(o, 20000, 64, 123, nthreads=1)
    p8, s8 = oracle.mc_run(o, 20000, 64, 123, nthreads=8)
    # mc_engine.rs:1135-1167: same seed, tolerance 1e-12 (summation order)
    assert abs(p1 - p8) < 1e-12 and abs(s1 - s8) < 1e-12


# ------------------------------------------- reference assertions (MC) ----

BS_CALL = 10.450583572185565


def test_bs_against_haug_fixture(oracle):
    for c in _fixtures()["european_haug"]:
        t = oracle.CALL if c["option_type"] == "Call" else oracle.PUT
        got = oracle.bs_price(t, c["spot"], c["strike"], c["rate"], c["q"], c["vol"], c["expiry"],
                              exact=True)
        assert abs(got - c["expected"]) <= c["tol"] + 1e-4 * 0, (c, got)
        hart = oracle.bs_price(t, c["spot"], c["strike"], c["rate"], c["q"], c["vol"], c["expiry"])
        assert abs(hart - got) <= 2e-6 * max(c["spot"], c["strike"])


def test_reference_engine_assertions(oracle):
    o = oracle.make_option()
    bs = oracle.bs_price(oracle.CALL, 100, 100, 0.05, 0.0, 0.2, 1.0)
    assert abs(bs - BS_CALL) < 1e-5
    # mc_engine.rs:998-1021 (1%)
    p, se = oracle.mc_run(o, 100_000, 252, 42)
    assert abs(p - bs) / bs <= 0.01
    # :1024-1046 antithetic SE < plain SE
    pa, sea = oracle.mc_run(o, 100_000, 252, 42, antithetic=True)
    assert sea < se
    # :1049-1073 CV within 0.5%
    pc, _ = oracle.mc_run(o, 100_000, 252, 42, control_variate=True)
    assert abs(pc - bs) / bs <= 0.005
    # simulation.rs:397-421: 60k antithetic x 252, within 2 SE + 2e-2
    p6, se6 = oracle.mc_run(o, 60_000, 252, 42, antithetic=True)
    assert abs(p6 - bs) <= 2 * se6 + 2e-2


def test_asian_reference_assertions(oracle):
    obs = [m / 11.0 for m in range(12)]
    # mc_engine.rs:1076-1094 Kemna-Vorst ~5.31
    g = oracle.geometric_asian_discrete(oracle.CALL, 100, 100, 0.05, 0.0, 0.2, obs)
    assert abs(g - 5.31) <= 0.15
    # asian_quantlib.rs:109-144: 5.3425606635 (1e-4)
    g2 = oracle.geometric_asian_discrete(oracle.CALL, 100, 100, 0.06, 0.03, 0.2,
                                         [i / 10.0 for i in range(1, 11)])
    assert abs(g2 - 5.3425606635) <= 1e-4
    # mc_engine.rs:1097-1132 ArithmeticAsianMC 100k x 252 vs 5.73 (2%)
    o = oracle.make_option(payoff=oracle.ASIAN, obs_times=obs)
    p, se = oracle.mc_run(o, 100_000, 252, 42, antithetic=True, control_variate=True)
    assert abs(p - 5.73) / 5.73 <= 0.02


def test_quantlib_asian_case4(oracle):
    # asian_quantlib.rs:185-256: cases4 row first=1/12, 12 fixings, 80k x 128
    # via MonteCarloPricingEngine (no variance reduction) within 2 SE.
    c = [r for r in _fixtures()["asian_arithmetic_cases4"]
         if abs(r["first"] - 1 / 12) < 1e-15 and r["fixings"] == 12][0]
    dt = c["length"] / (c["fixings"] - 1)
    obs = [c["first"] + i * dt for i in range(c["fixings"])]
    o = oracle.make_option(payoff=oracle.ASIAN, option_type=oracle.PUT, spot=c["underlying"],
                           strike=c["strike"], expiry=obs[-1], rate=c["rate"],
                           div_yield=c["dividend_yield"], vol=c["vol"], obs_times=obs)
    p, se = oracle.mc_run(o, 80_000, 128, 42)
    assert abs(p - c["expected"]) <= 2 * se
    pc, sec = oracle.mc_run(o, 20_000, 128, 42)
    assert abs(p - c["expected"]) <= abs(pc - c["expected"]) + 5e-3


def test_pathwise_aad_reference_assertions(oracle):
    # mc_aad.rs:353-409: 120k x 64 antithetic vs BS Greeks
    o = oracle.make_option(rate=0.03, div_yield=0.01)
    g = oracle.mc_european_pathwise_aad(o, 120_000, 64, 42, True)
    ref = oracle.bs_greeks(oracle.CALL, 100, 100, 0.03, 0.01, 0.2, 1.0)
    assert abs(g["delta"] - ref["delta"]) < 6e-3
    assert abs(g["vega"] - ref["vega"]) < 8e-2
    assert abs(g["rho"] - ref["rho"]) < 7e-2
    assert abs(g["theta"] - ref["theta"]) < 9e-2


def test_heston_delta_aad_reference_assertion(oracle):
    # mc_aad.rs:438-480: AAD delta vs CRN central bump h = 1e-3 (1e-4 abs)
    kw = dict(model=oracle.HESTON, strike=70.0, rate=0.03, kappa=1.8, theta=0.04, xi=0.6,
              rho=-0.5, v0=0.04)
    o = oracle.make_option(spot=100.0, **kw)
    _, d = oracle.heston_price_delta_aad(o, 60_000, 64, 7)
    up, _ = oracle.heston_price_delta_aad(oracle.make_option(spot=100.001, **kw), 60_000, 64, 7)
    dn, _ = oracle.heston_price_delta_aad(oracle.make_option(spot=99.999, **kw), 60_000, 64, 7)
    assert abs(d - (up - dn) / 2e-3) < 1e-4


def test_heston_semi_analytic_against_lewis_and_fang_oosterlee():
    from oracle.heston_analytic import heston_call, heston_put
    fx = _fixtures()
    for c in fx["heston_lewis_inline"]:
        f = heston_call if c["option_type"] == "Call" else heston_put
        got = f(100.0, c["strike"], 1.0, 0.01, 0.02, 4.0, 0.25, 1.0, -0.5, 0.04)
        assert abs(got - c["expected"]) <= 1e-8
    p = fx["fang_oosterlee"]["params"]
    for c in fx["fang_oosterlee"]["cases"]:
        got = heston_call(p["spot"], c["strike"], c["expiry"], p["rate"], p["q"], p["kappa"],
                          p["theta"], p["xi"], p["rho"], p["v0"])
        assert abs(got - c["call"]) <= 1e-7


@pytest.mark.slow
def test_heston_oracle_mc_against_semi_analytic(oracle):
    # The oracle's Heston Euler MC (reference discretisation) within 3 SE plus
    # a small Euler-bias allowance of the semi-analytic price, on parameters
    # that satisfy the Feller condition (2 kappa theta > xi^2) so the bias of
    # the full-truncation scheme is well below the SE at 252 steps.
    from oracle.heston_analytic import heston_call
    o = oracle.make_option(model=oracle.HESTON, spot=100, strike=100, expiry=1.0, rate=0.02,
                           kappa=2.0, theta=0.04, xi=0.3, rho=-0.7, v0=0.04)
    p, se = oracle.mc_run(o, 400_000, 252, 9, nthreads=8)
    ref = heston_call(100, 100, 1.0, 0.02, 0.0, 2.0, 0.04, 0.3, -0.7, 0.04)
    assert abs(p - ref) <= 3 * se + 0.01


def test_bs_greeks_quantlib_atm_and_finite_differences(oracle):
    # greeks/bsm.rs:246-257: QuantLib ATM call Greeks
    g = oracle.bs_greeks(oracle.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0)
    assert abs(g["delta"] - 0.6368) < 5e-4
    assert abs(g["gamma"] - 0.01876) < 5e-5
    assert abs(g["theta"] - -6.414) < 5e-3
    assert abs(g["vega"] - 37.524) < 5e-3
    assert abs(g["rho"] - 51.522) < 2.0
    # closed-form Greeks are the derivatives of the closed-form price
    # (central differences of the exact-CDF price, both option types, q != 0)
    for t in (oracle.CALL, oracle.PUT):
        args = dict(spot=95.0, strike=100.0, rate=0.03, q=0.01, vol=0.25, expiry=0.8)
        g = oracle.bs_greeks(t, *args.values())

        def px(**kw):
            a = dict(args, **kw)
            return oracle.bs_price(t, a["spot"], a["strike"], a["rate"], a["q"], a["vol"],
                                   a["expiry"], exact=True)

        h = 1e-4
        fd = {"delta": (px(spot=95.0 + h) - px(spot=95.0 - h)) / (2 * h),
              "gamma": (px(spot=95.0 + h) - 2 * px() + px(spot=95.0 - h)) / h ** 2,
              "vega": (px(vol=0.25 + h) - px(vol=0.25 - h)) / (2 * h),
              "rho": (px(rate=0.03 + h) - px(rate=0.03 - h)) / (2 * h),
              "theta": -(px(expiry=0.8 + h) - px(expiry=0.8 - h)) / (2 * h)}
        for k, v in fd.items():
            assert abs(g[k] - v) <= 1e-5 * max(1.0, abs(v)), (t, k, g[k], v)


def test_mc_greeks_engine_reference_assertions(oracle):
    """mc_greeks.rs:330-413: pathwise delta within 2 % and LR vega within 5 %
    of the closed forms; the QuantLib reference values 0.6368, 0.01876,
    37.524 (S=K=100, r=5 %, sigma=20 %, T=1, 100,000 antithetic paths, seed 42)."""
    from scipy.stats import norm
    o = oracle.make_option()
    g = oracle.mc_greeks_estimate(o, 100_000, 42)
    d1 = (math.log(100.0 / 100.0) + (0.05 + 0.5 * 0.04) * 1.0) / 0.2
    assert abs(g["pw_delta"] - norm.cdf(d1)) / norm.cdf(d1) <= 0.02
    vega = 100.0 * norm.pdf(d1)
    assert abs(g["lr_vega"] - vega) / vega <= 0.05
    assert abs(g["pw_delta"] - 0.6368) / 0.6368 < 0.02
    assert abs(g["pw_gamma"] - 0.01876) / 0.01876 < 0.20
    assert abs(g["lr_vega"] - 37.524) / 37.524 < 0.05


@pytest.mark.parametrize("opt_type", ["call", "put"])
def test_mc_greeks_engine_estimators_unbiased(oracle, opt_type):
    """Each estimator against its exact expectation at 3 standard errors of
    20 independent batches: pathwise and LR delta -> BS delta, LR gamma -> BS
    gamma, LR vega -> BS vega, and the CRN difference of pathwise deltas ->
    (Delta_BS(S+h) - Delta_BS(S-h)) / (2h) exactly (pathwise delta at a bumped
    spot is unbiased for the BS delta there). A dropped term or wrong sign in
    any score weight moves its mean by many standard errors."""
    from scipy.stats import norm
    S, K, r, q, vol, T, rel = 100.0, 105.0, 0.03, 0.01, 0.25, 0.75, 0.02
    t = oracle.CALL if opt_type == "call" else oracle.PUT
    o = oracle.make_option(option_type=t, strike=K, rate=r, div_yield=q, vol=vol, expiry=T)
    batches = np.array([[*oracle.mc_greeks_estimate(o, 100_000, 1000 + b, True, rel).values()]
                        for b in range(20)])
    mean, se = batches.mean(axis=0), batches.std(axis=0, ddof=1) / np.sqrt(len(batches))

    def bs_delta(s):
        d1 = (math.log(s / K) + (r - q + 0.5 * vol * vol) * T) / (vol * math.sqrt(T))
        return math.exp(-q * T) * (norm.cdf(d1) if t == oracle.CALL else norm.cdf(d1) - 1.0)

    d1 = (math.log(S / K) + (r - q + 0.5 * vol * vol) * T) / (vol * math.sqrt(T))
    gamma = math.exp(-q * T) * norm.pdf(d1) / (S * vol * math.sqrt(T))
    vega = S * math.exp(-q * T) * norm.pdf(d1) * math.sqrt(T)
    h = S * rel
    expected = [bs_delta(S), (bs_delta(S + h) - bs_delta(S - h)) / (2 * h), bs_delta(S), gamma,
                vega]
    for k, e in enumerate(expected):
        assert abs(mean[k] - e) <= 3.5 * se[k] + 1e-12, (k, mean[k], e, se[k])


def test_mc_greeks_engine_edge_cases(oracle):
    o = oracle.make_option(expiry=0.0)
    assert list(oracle.mc_greeks_estimate(o, 1000, 1).values()) == [0.0] * 5
    with pytest.raises(ValueError):
        oracle.mc_greeks_estimate(oracle.make_option(), 0, 1)
    with pytest.raises(ValueError):
        oracle.mc_greeks_estimate(oracle.make_option(), 10, 1, True, 0.0)
    with pytest.raises(ValueError):
        oracle.mc_greeks_estimate(oracle.make_option(vol=0.0), 10, 1)
    # non-antithetic uses ceil-free sample count and the plain estimator
    a = oracle.mc_greeks_estimate(oracle.make_option(), 1, 7, False)
    b = oracle.mc_greeks_estimate(oracle.make_option(), 2, 7, True)
    assert a["lr_delta"] != b["lr_delta"]
