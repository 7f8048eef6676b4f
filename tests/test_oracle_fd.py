"""Pins the oracle's restatement of the finite-difference Greeks harness
(finite_difference_greeks / bump_and_reprice, src/greeks/bsm.rs:130-212) and
of the closed-form BSM Greeks it is checked against:

* closed-form vanna and volga (and the other five) equal mpmath's numerical
  derivatives of the Black-Scholes-Merton price to 1e-10 relative;
* the reference's own test (bsm.rs:226-244): the harness over the
  Black-Scholes pricer is within its stated tolerances of the closed forms,
  and at smaller bumps within the O(h^2) error of central differences;
* the harness calls the pricer at exactly the reference's points, in its
  order, with its floors: a recording pricer;
* bump_and_reprice returns the pricer's base value (bsm.rs:260-270) and its
  six points in order, the spot-down point unfloored.
"""
import math

import mpmath as mp
import pytest

from oracle import oracle as O


def _bs(call, s, k, r, q, sigma, t):
    if t <= 0.0:
        return max(s - k, 0.0) if call else max(k - s, 0.0)
    vt = sigma * math.sqrt(t)
    d1 = (math.log(s / k) + (r - q + 0.5 * sigma * sigma) * t) / vt
    d2 = d1 - vt
    N = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))
    if call:
        return s * math.exp(-q * t) * N(d1) - k * math.exp(-r * t) * N(d2)
    return k * math.exp(-r * t) * N(-d2) - s * math.exp(-q * t) * N(-d1)


def _bs_mp(call, s, k, r, q, sigma, t):
    s, k, r, q, sigma, t = map(mp.mpf, (s, k, r, q, sigma, t))
    vt = sigma * mp.sqrt(t)
    d1 = (mp.log(s / k) + (r - q + sigma ** 2 / 2) * t) / vt
    d2 = d1 - vt
    if call:
        return s * mp.exp(-q * t) * mp.ncdf(d1) - k * mp.exp(-r * t) * mp.ncdf(d2)
    return k * mp.exp(-r * t) * mp.ncdf(-d2) - s * mp.exp(-q * t) * mp.ncdf(-d1)


POINTS = [(True, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0), (False, 100.0, 95.0, 0.03, 0.01, 0.22, 0.8),
          (True, 80.0, 110.0, 0.02, 0.03, 0.45, 2.5), (False, 130.0, 100.0, 0.07, 0.0, 0.15, 0.3)]


@pytest.mark.parametrize("pt", POINTS)
def test_closed_form_greeks_equal_mpmath_derivatives(pt):
    call, s, k, r, q, v, t = pt
    mp.mp.dps = 40
    f = lambda S, V, R, T: _bs_mp(call, S, k, R, q, V, T)
    want = dict(
        delta=mp.diff(lambda x: f(x, v, r, t), s),
        gamma=mp.diff(lambda x: f(x, v, r, t), s, 2),
        vega=mp.diff(lambda x: f(s, x, r, t), v),
        theta=-mp.diff(lambda x: f(s, v, r, x), t),
        rho=mp.diff(lambda x: f(s, v, x, t), r),
        vanna=mp.diff(lambda a, b: f(a, b, r, t), (s, v), (1, 1)),
        volga=mp.diff(lambda x: f(s, x, r, t), v, 2),
    )
    got = O.bsm_greeks_closed_form(call, s, k, r, q, v, t)
    for name in O.FD_FIELDS:
        assert abs(got[name] - float(want[name])) <= 1e-10 * max(abs(float(want[name])), 1e-3), name


def test_reference_fd_test_against_closed_form():
    """bsm.rs:226-244 on the Black-Scholes pricer, the reference's bumps and
    tolerances."""
    pricer = lambda s, k, r, v, t: _bs(True, s, k, r, 0.0, v, t)
    fd = O.finite_difference_greeks(pricer, 100.0, 100.0, 0.05, 0.2, 1.0, 1e-3, 1e-4, 1e-4, 1e-4)
    cf = O.bsm_greeks_closed_form(True, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0)
    tol = dict(delta=2e-4, gamma=2e-4, vega=4e-3, theta=1e-2, rho=5e-3, vanna=4e-3, volga=1e-2)
    for name in O.FD_FIELDS:
        assert abs(fd[name] - cf[name]) < tol[name], name


@pytest.mark.parametrize("pt", POINTS)
def test_harness_converges_to_closed_form(pt):
    call, s, k, r, q, v, t = pt
    pricer = lambda S, K, R, V, T: _bs(call, S, K, R, q, V, T)
    fd = O.finite_difference_greeks(pricer, s, k, r, v, t, 1e-3 * s, 1e-4, 1e-4, 1e-4)
    cf = O.bsm_greeks_closed_form(call, s, k, r, q, v, t)
    for name in O.FD_FIELDS:
        scale = max(abs(cf[name]), 1e-2 * max(abs(cf["vega"]), 1.0))
        assert abs(fd[name] - cf[name]) <= 5e-4 * scale, (name, fd[name], cf[name])


def test_harness_calls_and_floors_recorded():
    calls = []

    def rec(s, k, r, v, t):
        calls.append((s, k, r, v, t))
        return 0.0

    O.finite_difference_greeks(rec, 1.0, 2.0, 0.03, 0.1, 0.5, 3.0, 0.01, 0.2, 0.7)
    s, k, r, v, t = 1.0, 2.0, 0.03, 0.1, 0.5
    e = 1e-8
    assert calls == [(s, k, r, v, t), (s + 3.0, k, r, v, t), (e, k, r, v, t),
                     (s, k, r + 0.01, v, t), (s, k, r - 0.01, v, t), (s, k, r, v + 0.2, t),
                     (s, k, r, e, t), (s, k, r, v, t + 0.7), (s, k, r, v, e),
                     (s + 3.0, k, r, v + 0.2, t), (s + 3.0, k, r, e, t), (e, k, r, v + 0.2, t),
                     (e, k, r, e, t)]
    calls.clear()
    out = O.bump_and_reprice(rec, s, k, r, v, t, 3.0, 0.01, 0.2, -0.9)
    assert calls == [(s, k, r, v, t), (s + 3.0, k, r, v, t), (s, k, r + 0.01, v, t),
                     (s, k, r, v + 0.2, t), (s, k, r, v, 1e-8), (s - 3.0, k, r, v, t)]
    assert len(out) == 6


def test_bump_and_reprice_base():
    """bsm.rs:260-270."""
    pricer = lambda s, k, r, v, t: _bs(False, s, k, r, 0.0, v, t)
    base, s_up, r_up, v_up, t_up, s_dn = O.bump_and_reprice(pricer, 100.0, 95.0, 0.03, 0.22, 0.8,
                                                            0.1, 1e-4, 1e-4, 1e-4)
    assert abs(base - _bs(False, 100.0, 95.0, 0.03, 0.0, 0.22, 0.8)) < 1e-12
    assert s_up < base < s_dn  # a put falls with the spot
