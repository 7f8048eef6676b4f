"""Heston semi-analytic price -- TEST INFRASTRUCTURE ONLY.

Lewis (2001) single-integral formula over the Heston characteristic function
in Gatheral's numerically stable form (the form of HestonCharFn,
src/engines/fft/char_fn.rs:85-163 of the reference), integrated with adaptive
quadrature instead of the reference's Carr-Madan FFT. Pinned by the QuantLib
Lewis 12-digit prices that the reference's own fixtures hold
(tests/golden/reference_fixtures.json, "heston_lewis").
"""
from __future__ import annotations

import cmath
import math

from scipy import integrate


def heston_cf(u: complex, T: float, kappa: float, theta: float, xi: float, rho: float,
              v0: float) -> complex:
    """E[exp(i u X_T)], X_T = ln(S_T/S_0) - (r-q)T (Gatheral, 'The Volatility
    Surface' (2006) eqs. 2.8-2.10)."""
    alpha = -0.5 * u * u - 0.5j * u
    beta = kappa - rho * xi * 1j * u
    gamma = 0.5 * xi * xi
    d = cmath.sqrt(beta * beta - 4.0 * alpha * gamma)
    rm = (beta - d) / (xi * xi)
    rp = (beta + d) / (xi * xi)
    g = rm / rp
    edt = cmath.exp(-d * T)
    D = rm * (1.0 - edt) / (1.0 - g * edt)
    C = kappa * (rm * T - (2.0 / (xi * xi)) * cmath.log((1.0 - g * edt) / (1.0 - g)))
    return cmath.exp(C * theta + D * v0)


def heston_call(spot, strike, T, r, q, kappa, theta, xi, rho, v0) -> float:
    """Lewis formula: C = S e^{-qT} - sqrt(S K) e^{-(r+q)T/2} / pi *
    int_0^inf Re[e^{i u x} phi(u - i/2)] / (u^2 + 1/4) du, x = ln(S/K) + (r-q)T."""
    x = math.log(spot / strike) + (r - q) * T

    def f(u):
        return (cmath.exp(1j * u * x) * heston_cf(u - 0.5j, T, kappa, theta, xi, rho, v0)).real \
            / (u * u + 0.25)

    val, _ = integrate.quad(f, 0.0, math.inf, limit=2000, epsabs=1e-14, epsrel=1e-13)
    return spot * math.exp(-q * T) - math.sqrt(spot * strike) * math.exp(-0.5 * (r + q) * T) \
        / math.pi * val


def heston_put(spot, strike, T, r, q, kappa, theta, xi, rho, v0) -> float:
    c = heston_call(spot, strike, T, r, q, kappa, theta, xi, rho, v0)
    return c - spot * math.exp(-q * T) + strike * math.exp(-r * T)
