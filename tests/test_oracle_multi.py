"""Pins of the oracle's autocallable restatement (CPU only).

* zero volatility: every path is the deterministic forward, so the present
  value follows in closed form -- call date and coupon of a standard note,
  Phoenix coupons with and without memory on rising and falling paths,
  knock-in redemption;
* never called / no knock-in: e^{-rT} N (1 + cT) exactly, any volatility;
  always called at the first date: e^{-r t1} N (1 + c t1);
* certain knock-in, never called: e^{-rT} N E[min_i S_i(T)/S_i(0)] / K_ki,
  with E[S(T)/S(0)] = e^{(r-q)T} for one asset and, for two correlated
  assets, E[min] = 2 e^{(r-q)T} (1 - N(sigma sqrt(T) / 2)),
  sigma^2 = s1^2 + s2^2 - 2 rho s1 s2 (Margrabe's exchange option), within
  4 standard errors -- a transposed Cholesky factor or a wrong correlation
  sign fails it;
* the reference's own tests (pricing/autocallable.rs:716-923): lower
  autocall barrier lowers the value, higher coupon raises it, Phoenix coupons
  raise it;
* the single stream equals its own draws fed as [path][step][asset] normals
  when no path is called early.
"""
import math

import numpy as np
import pytest
from scipy.stats import norm

from oracle import oracle as O

DATES = [0.25, 0.5, 0.75, 1.0]


def _pv(**kw):
    return O.autocall_price(O.make_autocall(**kw))


def test_zero_vol_standard_note_called_on_first_date_above_barrier():
    # forward ratio e^{(r-q)t}: 1.0253, 1.0513, 1.0779, 1.1052 with r = 0.1
    r = 0.1
    for barrier, t_call in ((1.04, 0.5), (1.07, 0.75), (1.10, 1.0)):
        p, se = _pv(spots=[100.0, 120.0], vols=[0.0, 0.0], corr=[[1, 0.3], [0.3, 1]],
                    dates=DATES, autocall_barrier=barrier, coupon_rate=0.08, rate=r,
                    num_paths=64, num_steps=100)
        assert se < 1e-10
        assert abs(p - math.exp(-r * t_call) * 100.0 * (1 + 0.08 * t_call)) < 1e-9, (barrier, p)


@pytest.mark.parametrize("memory", [False, True])
def test_zero_vol_phoenix_memory_on_rising_path(memory):
    r, c, N = 0.1, 0.12, 100.0
    # coupon barrier 1.06: missed at 0.25 and 0.5, met at 0.75; call at 1.0
    p, _ = _pv(spots=[100.0], vols=[0.0], corr=[[1.0]], dates=DATES, phoenix=True, memory=memory,
               coupon_barrier=1.06, autocall_barrier=1.09, coupon_rate=c, rate=r, num_paths=8,
               num_steps=100)
    due = N * c * (0.75 if memory else 0.25)
    want = math.exp(-r * 0.75) * due + math.exp(-r * 1.0) * (N * c * 0.25 + N)
    assert abs(p - want) < 1e-9, (p, want)


@pytest.mark.parametrize("memory", [False, True])
def test_zero_vol_phoenix_on_falling_path_with_knock_in(memory):
    r, q, c, N = 0.01, 0.11, 0.12, 100.0
    # ratio e^{-0.1 t}: 0.9753, 0.9512, 0.9277, 0.9048; coupon barrier 0.94
    for ki, ki_hit in ((0.5, False), (0.91, True)):
        p, _ = _pv(spots=[100.0], vols=[0.0], corr=[[1.0]], dates=DATES, phoenix=True,
                   memory=memory, coupon_barrier=0.94, autocall_barrier=1.2, coupon_rate=c,
                   ki_barrier=ki, ki_strike=1.0, rate=r, div_yield=q, num_paths=8,
                   num_steps=100)
        coupons = math.exp(-r * 0.25) * N * c * 0.25 + math.exp(-r * 0.5) * N * c * 0.25
        ratio_t = math.exp(-0.1)
        red = N * ratio_t if ki_hit else N
        want = coupons + math.exp(-r) * red
        assert abs(p - want) < 1e-6 * N, (ki, p, want)


def test_never_called_no_knock_in_is_deterministic():
    p, se = _pv(spots=[100.0, 90.0], vols=[0.3, 0.25], corr=[[1, -0.2], [-0.2, 1]], dates=DATES,
                autocall_barrier=1e9, ki_barrier=1e-9, coupon_rate=0.05, rate=0.02,
                num_paths=500, num_steps=50)
    assert abs(p - math.exp(-0.02) * 100.0 * 1.05) < 1e-12 and se < 1e-12


def test_always_called_at_first_date():
    p, se = _pv(spots=[100.0, 90.0], vols=[0.3, 0.25], corr=[[1, 0.5], [0.5, 1]], dates=DATES,
                autocall_barrier=1e-9, coupon_rate=0.05, rate=0.02, num_paths=500, num_steps=50)
    assert abs(p - math.exp(-0.02 * 0.25) * 100.0 * (1 + 0.05 * 0.25)) < 1e-12


def test_certain_knock_in_one_asset_matches_forward():
    r, q, T = 0.03, 0.01, 1.0
    p, se = _pv(spots=[100.0], vols=[0.25], corr=[[1.0]], dates=DATES, autocall_barrier=1e9,
                ki_barrier=1e9, ki_strike=0.8, rate=r, div_yield=q, num_paths=200_000,
                num_steps=16, seed=5)
    want = 100.0 * math.exp(-q * T) / 0.8
    assert abs(p - want) < 4 * se, (p, want, se)


@pytest.mark.parametrize("rho", [-0.6, 0.0, 0.7])
def test_certain_knock_in_two_assets_matches_margrabe(rho):
    r, q, T, s1, s2 = 0.02, 0.0, 1.0, 0.3, 0.2
    p, se = _pv(spots=[100.0, 50.0], vols=[s1, s2], corr=[[1, rho], [rho, 1]], dates=DATES,
                autocall_barrier=1e9, ki_barrier=1e9, ki_strike=1.0, rate=r, div_yield=q,
                num_paths=200_000, num_steps=8, seed=11)
    sig = math.sqrt(s1 * s1 + s2 * s2 - 2 * rho * s1 * s2)
    e_min = 2 * math.exp((r - q) * T) * (1 - norm.cdf(sig * math.sqrt(T) / 2))
    want = math.exp(-r * T) * 100.0 * e_min
    assert abs(p - want) < 4 * se, (rho, p, want, se)


def test_reference_monotonicity():
    spots, corr4 = [100.0, 100.0], [[1.0, 0.4], [0.4, 1.0]]
    hi = _pv(spots=spots, vols=[0.12, 0.12], corr=corr4, dates=DATES, coupon_rate=0.08,
             autocall_barrier=1.0, num_paths=16_000, num_steps=64)[0]
    lo = _pv(spots=spots, vols=[0.12, 0.12], corr=corr4, dates=DATES, coupon_rate=0.08,
             autocall_barrier=0.8, num_paths=16_000, num_steps=64)[0]
    assert lo < hi
    corr3 = [[1.0, 0.3], [0.3, 1.0]]
    c_lo = _pv(spots=spots, vols=[0.2, 0.2], corr=corr3, dates=DATES, coupon_rate=0.02,
               num_paths=16_000, num_steps=64)[0]
    c_hi = _pv(spots=spots, vols=[0.2, 0.2], corr=corr3, dates=DATES, coupon_rate=0.10,
               num_paths=16_000, num_steps=64)[0]
    assert c_hi > c_lo
    kw = dict(spots=spots, vols=[0.35, 0.35], corr=[[1.0, 0.2], [0.2, 1.0]], dates=DATES,
              autocall_barrier=1.05, coupon_rate=0.12, ki_barrier=0.9, ki_strike=1.0,
              num_paths=20_000, num_steps=80)
    std = _pv(**kw)[0]
    ph = _pv(phoenix=True, memory=True, coupon_barrier=0.7, **kw)[0]
    assert ph > std


def test_stream_equals_its_draws_as_external_normals_when_never_called():
    n, steps = 300, 12
    kw = dict(spots=[100.0, 80.0, 120.0], vols=[0.2, 0.3, 0.25],
              corr=[[1, 0.2, 0.1], [0.2, 1, -0.3], [0.1, -0.3, 1]], dates=DATES,
              autocall_barrier=1e9, ki_barrier=0.7, num_paths=n, num_steps=steps, seed=9)
    z = O.lsm_stream_normals(9, n * steps * 3).reshape(n, steps, 3)
    assert O.autocall_price(O.make_autocall(**kw)) == O.autocall_price(O.make_autocall(**kw), z)


def test_cholesky_reproduces_matrix_and_rejects_indefinite():
    c = np.array([[1, 0.3, -0.2], [0.3, 1, 0.5], [-0.2, 0.5, 1]])
    l = O.cholesky_lower(c)
    assert np.allclose(l @ l.T, c, atol=1e-14) and np.allclose(np.triu(l, 1), 0)
    assert O.cholesky_lower(np.array([[1, 0.95, 0.95], [0.95, 1, -0.95], [0.95, -0.95, 1]])) is None
