"""Pins of the oracle's Longstaff-Schwartz restatement (CPU only).

* nalgebra-LU 3x3 solve: against numpy.linalg.solve on random systems, with
  pivoting forced, and the singular cases that make the reference fall back
  to a zero regression;
* Longstaff-Schwartz (2001) Table 1, the 20 American puts that the
  reference's tests/lsm_reference.rs checks (100k paths, 50 steps, seed 42,
  tolerance 0.10), plus its other assertions (American >= European,
  diagnostics, American call on a non-dividend stock ~ European);
* QuantLib FD Bermudan put prices of tests/bermudan_quantlib.rs within 0.5 %
  (450k paths, 120 steps, seed 7) and its step-down strike ordering;
* an independent numpy least-squares restatement (tests/lsm_numpy.py, lstsq
  regression) on small GBM / local-vol / Heston / Bermudan cases: prices to
  1e-9 and per-date in-the-money and exercise counts exactly;
* European exercise = discounted mean payoff, against Black-Scholes;
* a dense Bermudan against a CRR binomial American price (the reference's
  convergence test);
* barrier pricer: knock-in + knock-out = vanilla on the same stream.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from tests import lsm_numpy as LN

TABLE1 = [
    # spot, vol, expiry, expected (Longstaff & Schwartz 2001, Table 1)
    (36.0, 0.2, 1.0, 4.474), (36.0, 0.2, 2.0, 4.831), (36.0, 0.4, 1.0, 7.081),
    (36.0, 0.4, 2.0, 8.480), (38.0, 0.2, 1.0, 3.242), (38.0, 0.2, 2.0, 3.746),
    (38.0, 0.4, 1.0, 6.141), (38.0, 0.4, 2.0, 7.647), (40.0, 0.2, 1.0, 2.319),
    (40.0, 0.2, 2.0, 2.875), (40.0, 0.4, 1.0, 5.318), (40.0, 0.4, 2.0, 6.912),
    (42.0, 0.2, 1.0, 1.615), (42.0, 0.2, 2.0, 2.216), (42.0, 0.4, 1.0, 4.586),
    (42.0, 0.4, 2.0, 6.241), (44.0, 0.2, 1.0, 1.100), (44.0, 0.2, 2.0, 1.678),
    (44.0, 0.4, 1.0, 3.942), (44.0, 0.4, 2.0, 5.642),
]

# tests/bermudan_quantlib.rs: puts (spot, strike, rate, dividend, vol, expiry,
# dates, QuantLib 1.41 FD price)
D4_1Y = [0.2493150685, 0.4986301370, 0.7506849315, 1.0]
D6_15Y = [0.2493150685, 0.5013698630, 0.7506849315, 1.0, 1.2520547945, 1.5013698630]
D8_2Y = [0.2493150685, 0.4986301370, 0.7506849315, 1.0, 1.2493150685, 1.5013698630,
         1.7506849315, 2.0]
D3_075Y = [0.2493150685, 0.5013698630, 0.7506849315]
D12_3Y = [0.2493150685, 0.4986301370, 0.7506849315, 1.0, 1.2493150685, 1.5013698630,
          1.7506849315, 2.0, 2.2493150685, 2.4986301370, 2.7506849315, 3.0]
QL_PUTS = [
    (100.0, 100.0, 0.03, 0.00, 0.20, 1.0, D4_1Y, 6.6581251992),
    (95.0, 100.0, 0.05, 0.01, 0.25, 1.5013698630, D6_15Y, 11.7181754029),
    (110.0, 100.0, 0.02, 0.00, 0.30, 2.0, D8_2Y, 11.4282877043),
    (80.0, 90.0, 0.06, 0.00, 0.35, 0.7506849315, D3_075Y, 13.8596661589),
    (150.0, 140.0, 0.025, 0.005, 0.28, 3.0, D12_3Y, 19.2975666673),
]


def _eff_q(spot, r, q, T):
    """Market::effective_dividend_yield with no discrete dividends."""
    fwd = spot * math.exp((r - q) * T)
    return r - math.log(max(fwd / spot, 1e-16)) / T


def _pool_map(fn, items):
    with ThreadPoolExecutor(max_workers=8) as ex:
        return list(ex.map(fn, items))


# ------------------------------------------------------------- LU solve ---
def test_lu3_solve_matches_numpy_with_pivoting():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        x = O.lu3_solve(a, b)
        ref = np.linalg.solve(a, b)
        assert np.allclose(x, ref, rtol=1e-10, atol=1e-12)
    # pivot rows: a zero leading entry needs a row swap
    a = np.array([[0.0, 2.0, 1.0], [1.0, 0.0, 0.0], [0.0, 0.0, 3.0]])
    assert np.allclose(O.lu3_solve(a, [1.0, 2.0, 3.0]), np.linalg.solve(a, [1.0, 2.0, 3.0]))


def test_lu3_solve_singular_is_none():
    assert O.lu3_solve(np.zeros((3, 3)), [1.0, 1.0, 1.0]) is None
    # all ITM paths at one spot: rank-1 normal equations
    s = 37.5
    xtx = np.array([[5.0, 5 * s, 5 * s * s], [5 * s, 5 * s * s, 5 * s ** 3],
                    [5 * s * s, 5 * s ** 3, 5 * s ** 4]])
    x = O.lu3_solve(xtx, [1.0, s, s * s])
    # rounding may leave a tiny non-zero pivot; either None or a solution
    assert x is None or np.all(np.isfinite(x))
    a = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 0.0, 0.0]])
    assert O.lu3_solve(a, [1.0, 2.0, 0.0]) is None


# -------------------------------------------------------------- Table 1 ---
def test_table1_american_puts_within_reference_tolerance():
    def run(case):
        spot, vol, T, _ = case
        o = O.make_lsm_option(spot=spot, strike=40.0, vol=vol, expiry=T, rate=0.06,
                              num_paths=100_000, num_steps=50, seed=42)
        return O.lsm_price(o)

    got = _pool_map(run, TABLE1)
    for (spot, vol, T, expected), (p, se) in zip(TABLE1, got):
        assert abs(p - expected) <= 0.10, (spot, vol, T, p, expected)
        assert 0.0 < se < 0.05


def test_american_put_above_european_and_seed_reproducible():
    kw = dict(spot=40.0, strike=40.0, vol=0.2, expiry=1.0, rate=0.06, num_paths=20_000,
              num_steps=50, seed=42)
    am = O.lsm_price(O.make_lsm_option(exercise=O.AMERICAN, **kw))
    eu = O.lsm_price(O.make_lsm_option(exercise=O.EUROPEAN, **kw))
    assert am[0] > eu[0]
    assert O.lsm_price(O.make_lsm_option(exercise=O.AMERICAN, **kw)) == am


def test_american_call_without_dividends_matches_european():
    kw = dict(option_type=O.CALL, spot=40.0, strike=40.0, vol=0.2, expiry=1.0, rate=0.06,
              num_paths=100_000, num_steps=50, seed=42)
    am, _ = O.lsm_price(O.make_lsm_option(exercise=O.AMERICAN, **kw))
    eu, _ = O.lsm_price(O.make_lsm_option(exercise=O.EUROPEAN, **kw))
    bs = O.bs_price(O.CALL, 40.0, 40.0, 0.06, 0.0, 0.2, 1.0, exact=True)
    assert abs(am - eu) < 0.50
    assert abs(eu - bs) < 0.10


def test_european_exercise_is_discounted_mean_payoff():
    n, steps = 4000, 12
    o = O.make_lsm_option(exercise=O.EUROPEAN, spot=100.0, strike=105.0, vol=0.3,
                          expiry=1.5, rate=0.02, div_yield=0.01, num_paths=n,
                          num_steps=steps, seed=9)
    p, se = O.lsm_price(o)
    z = O.lsm_stream_normals(9, n * steps).reshape(n, steps)
    paths = LN.gbm_paths(100.0, 0.02, 0.01, 0.3, 1.5, steps, z)
    pv = np.maximum(105.0 - paths[:, -1], 0.0) * math.exp(-0.02 * 1.5 / steps) ** steps
    assert abs(p - pv.mean()) < 1e-9 * pv.mean()
    bs = O.bs_price(O.PUT, 100.0, 105.0, 0.02, 0.01, 0.3, 1.5, exact=True)
    assert abs(p - bs) < 4 * se


# ------------------------------------------------------ QuantLib Bermudan ---
def test_bermudan_puts_match_quantlib_fd_within_half_percent():
    def run(case):
        spot, k, r, q, vol, T, dates, _ = case
        o = O.make_lsm_option(instrument=O.LSM_BERMUDAN, spot=spot, strike=k, rate=r,
                              div_yield=_eff_q(spot, r, q, T), vol=vol, expiry=T,
                              dates=dates, strikes=[k] * len(dates), num_paths=450_000,
                              num_steps=120, seed=7)
        return O.lsm_price(o, want_boundary=True)

    got = _pool_map(run, QL_PUTS)
    for case, (p, se, pts) in zip(QL_PUTS, got):
        ref = case[-1]
        assert abs(p - ref) / ref <= 0.005, (case[:6], p, ref, se)
        assert any(pt["boundary_spot"] is not None for pt in pts)
        # terminal point last, chronological order
        assert pts[-1]["time"] == case[5] and pts[-1]["boundary_spot"] == case[1]
        assert all(a["time"] < b["time"] for a, b in zip(pts, pts[1:]))


def test_bermudan_step_down_strike_prices_below_constant_strike():
    dates = [0.25, 0.5, 0.75, 1.0]
    kw = dict(instrument=O.LSM_BERMUDAN, spot=100.0, rate=0.03, vol=0.25, expiry=1.0,
              dates=dates, num_paths=100_000, num_steps=100, seed=42)
    step = O.lsm_price(O.make_lsm_option(strikes=[100.0, 97.5, 95.0, 92.5], **kw))[0]
    const = O.lsm_price(O.make_lsm_option(strikes=[100.0] * 4, **kw))[0]
    assert step < const


def test_dense_bermudan_close_to_binomial_american():
    def crr_american_put(s, k, r, vol, T, n):
        dt = T / n
        u = math.exp(vol * math.sqrt(dt))
        d = 1 / u
        p = (math.exp(r * dt) - d) / (u - d)
        disc = math.exp(-r * dt)
        j = np.arange(n + 1)
        st = s * u ** (n - 2 * j).astype(float)
        v = np.maximum(k - st, 0.0)
        for i in range(n - 1, -1, -1):
            st = s * u ** (i - 2 * np.arange(i + 1)).astype(float)
            v = np.maximum(disc * (p * v[:-1] + (1 - p) * v[1:]), k - st)
        return float(v[0])

    am = crr_american_put(100.0, 100.0, 0.05, 0.2, 1.0, 2500)
    assert abs(am - 6.09) < 0.01     # textbook value of this put
    dates = [i / 52 for i in range(1, 53)]
    o = O.make_lsm_option(instrument=O.LSM_BERMUDAN, spot=100.0, rate=0.05, vol=0.2,
                          expiry=1.0, dates=dates, strikes=[100.0] * 52, num_paths=100_000,
                          num_steps=120, seed=11)
    p, se = O.lsm_price(o)
    assert p <= am + 0.20
    assert abs(p - am) / am <= 0.03


# --------------------------------------------- independent restatement ---
def _check_against_numpy(o, paths, put, strikes_by_step, exercisable, terminal_strike,
                         bermudan_times=None):
    p, se, pts = O.lsm_price(o, want_boundary=True)
    q, qse, qpts = LN.lsm(paths, put, strikes_by_step, exercisable, o.rate, o.expiry,
                          terminal_strike)
    assert abs(p - q) <= 1e-9 * max(abs(q), 1.0), (p, q)
    assert abs(se - qse) <= 1e-7 * max(qse, 1e-9)
    if bermudan_times is not None:
        # oracle boundary is chronological with the terminal point last
        inner = pts[:-1]
        assert len(inner) == len(qpts)
        for b, (t, m, cnt, extreme) in zip(inner, reversed(qpts)):
            assert b["itm_paths"] == m and b["exercised_paths"] == cnt
            if extreme is None:
                assert b["boundary_spot"] is None
            else:
                assert abs(b["boundary_spot"] - extreme) <= 1e-12 * extreme


@pytest.mark.parametrize("put,exercise", [(True, "american"), (False, "american"),
                                          (True, "bermudan")])
def test_vanilla_lsm_matches_numpy_lstsq(put, exercise):
    n, steps, seed = 3000, 16, 123
    dates = [0.3, 0.55, 0.8] if exercise == "bermudan" else []
    o = O.make_lsm_option(option_type=O.PUT if put else O.CALL,
                          exercise=O.AMERICAN if exercise == "american" else O.BERMUDAN,
                          spot=100.0, strike=100.0 if put else 95.0, vol=0.25, expiry=1.0,
                          rate=0.04, div_yield=0.03, dates=dates, num_paths=n,
                          num_steps=steps, seed=seed)
    z = O.lsm_stream_normals(seed, n * steps).reshape(n, steps)
    paths = LN.gbm_paths(100.0, 0.04, 0.03, 0.25, 1.0, steps, z)
    if exercise == "american":
        ex = set(range(1, steps))
    else:
        ex = {min(int(round(d / 1.0 * steps)), steps) for d in dates}
    k = 100.0 if put else 95.0
    _check_against_numpy(o, paths, put, {t: k for t in range(steps + 1)}, ex, k)


@pytest.mark.parametrize("dynamics", ["gbm", "local_vol", "heston"])
def test_bermudan_lsm_matches_numpy_lstsq(dynamics):
    n, steps, seed, T = 3000, 20, 77, 1.25
    dates = [0.2, 0.45, 0.7, 1.0]
    strikes = [104.0, 102.0, 100.0, 98.0]
    dyn = {"gbm": O.LSM_GBM, "local_vol": O.LSM_LOCAL_VOL, "heston": O.LSM_HESTON}[dynamics]
    hp = dict(kappa=2.0, theta=0.04, xi=0.6, rho=-0.5, v0=0.04)
    o = O.make_lsm_option(instrument=O.LSM_BERMUDAN, dynamics=dyn, spot=100.0, vol=0.22,
                          expiry=T, rate=0.02, div_yield=0.01, dates=dates, strikes=strikes,
                          num_paths=n, num_steps=steps, seed=seed, **hp)
    if dynamics == "heston":
        z = O.lsm_stream_normals(seed, 2 * n * steps).reshape(n, steps, 2)
        paths = LN.heston_paths(100.0, 0.02, 0.01, hp["kappa"], hp["theta"], hp["xi"],
                                hp["rho"], hp["v0"], T, steps, z[:, :, 0], z[:, :, 1])
    else:
        z = O.lsm_stream_normals(seed, n * steps).reshape(n, steps)
        paths = LN.gbm_paths(100.0, 0.02, 0.01, 0.22, T, steps, z)
    sched = {}
    for d, k in zip(dates, strikes):
        sched[min(max(int(round(d / T * steps)), 1), steps)] = k
    ex = {t for t in sched if t < steps}
    _check_against_numpy(o, paths, True, sched, ex, strikes[-1], bermudan_times=dates)


def test_external_normals_reproduce_the_stream():
    n, steps, seed = 500, 10, 5
    o = O.make_lsm_option(num_paths=n, num_steps=steps, seed=seed)
    z = O.lsm_stream_normals(seed, n * steps)
    assert O.lsm_price(o) == O.lsm_price(o, normals=z)


# ---------------------------------------------------------------- barrier ---
def test_barrier_in_plus_out_equals_vanilla():
    kw = dict(instrument=O.LSM_BARRIER, option_type=O.CALL, spot=100.0, strike=100.0,
              vol=0.25, expiry=1.0, rate=0.03, barrier_dir=O.UP, barrier_level=125.0,
              num_paths=20_000, num_steps=64, seed=17)
    pin = O.lsm_price(O.make_lsm_option(barrier_style=O.IN, **kw))[0]
    pout = O.lsm_price(O.make_lsm_option(barrier_style=O.OUT, **kw))[0]
    van = O.lsm_price(O.make_lsm_option(barrier_style=O.OUT,
                                        **dict(kw, barrier_level=1e300)))[0]
    assert abs(pin + pout - van) < 1e-12 * van
    assert 0.0 < pin < van and 0.0 < pout < van
