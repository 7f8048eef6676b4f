"""Independent numpy / closed-form checks for basket options (test helpers).

* higham_nearest: Higham (2002) alternating projections with Dykstra's
  correction, written from the paper with numpy's eigh (the reference uses
  nalgebra's SymmetricEigen, the library a Jacobi solver), with the
  reference's defaults tol 1e-10 and 100 iterations and its final PSD pass,
  unit diagonal and clipping (src/math/correlation.rs:271-335);
* moment_matching: Levy's two-moment and the Gentle-style three-moment
  lognormal fits of a weighted arithmetic basket (Levy 1992; the moments are
  those of correlated lognormals, E[S_i S_j] = S_i S_j e^{(2r-q_i-q_j +
  rho_ij s_i s_j) T}, and the third from the multivariate lognormal mgf);
* bs_price and margrabe: the textbook formulas.
"""
import itertools
import math

import numpy as np
from scipy.stats import norm


def _sym(m):
    return 0.5 * (m + m.T)


def _proj_psd(m):
    w, v = np.linalg.eigh(_sym(m))
    return _sym((v * np.maximum(w, 0.0)) @ v.T)


def higham_nearest(c, tol=1e-10, max_iter=100):
    c = np.asarray(c, dtype=np.float64)
    n = c.shape[0]
    y = _sym(c.copy())
    np.fill_diagonal(y, 1.0)
    ds = np.zeros_like(y)
    for _ in range(max_iter):
        r = _sym(y - ds)
        x = _proj_psd(r)
        ds = x - r
        yn = x.copy()
        np.fill_diagonal(yn, 1.0)
        yn = _sym(yn)
        diff = np.linalg.norm(yn - y)
        y = yn
        if diff < tol:
            break
    y = _proj_psd(_sym(y))
    np.fill_diagonal(y, 1.0)
    y = _sym(y)
    iu = np.triu_indices(n, 1)
    vals = np.clip(y[iu], -1.0, 1.0)
    y[iu] = vals
    y[(iu[1], iu[0])] = vals
    np.fill_diagonal(y, 1.0)
    return y


def repair_if_needed(c):
    c = np.asarray(c, dtype=np.float64)
    if np.linalg.eigvalsh(c).min() >= -1e-10:
        return c, False
    return higham_nearest(c), True


def bs_price(is_call, s, k, r, q, vol, t):
    if vol <= 0 or t <= 0:
        f = s * math.exp((r - q) * t)
        return math.exp(-r * t) * max((f - k) if is_call else (k - f), 0.0)
    d1 = (math.log(s / k) + (r - q + 0.5 * vol * vol) * t) / (vol * math.sqrt(t))
    d2 = d1 - vol * math.sqrt(t)
    if is_call:
        return s * math.exp(-q * t) * norm.cdf(d1) - k * math.exp(-r * t) * norm.cdf(d2)
    return k * math.exp(-r * t) * norm.cdf(-d2) - s * math.exp(-q * t) * norm.cdf(-d1)


def margrabe(a, b, sig, t):
    """Present value of max(A_T - B_T, 0) for lognormal A, B with present
    values a, b (forward prices discounted) and spread volatility sig."""
    if sig <= 0:
        return max(a - b, 0.0)
    d1 = (math.log(a / b) + 0.5 * sig * sig * t) / (sig * math.sqrt(t))
    return a * norm.cdf(d1) - b * norm.cdf(d1 - sig * math.sqrt(t))


def basket_moments(spots, weights, vols, corr, r, divs, t):
    s, w, v, q = (np.asarray(x, dtype=np.float64) for x in (spots, weights, vols, divs))
    c = np.asarray(corr, dtype=np.float64)
    n = len(s)
    f = s * np.exp((r - q) * t)
    m1 = float(w @ f)
    cov = c * np.outer(v, v) * t
    m2 = float(w @ (np.outer(f, f) * np.exp(cov)) @ w)
    m3 = 0.0
    for i, j, k in itertools.product(range(n), repeat=3):
        m3 += w[i] * w[j] * w[k] * f[i] * f[j] * f[k] * math.exp(cov[i, j] + cov[i, k] + cov[j, k])
    return m1, m2, m3


def moment_matching(method, is_call, strike, spots, weights, vols, corr, r, divs, t):
    m1, m2, m3 = basket_moments(spots, weights, vols, corr, r, divs, t)
    s2_levy = math.log(max(m2 / (m1 * m1), 1 + 1e-14))
    if method == "levy":
        s2 = s2_levy
    else:
        s2 = 0.5 * max(s2_levy + math.log(max(m3 / m1 ** 3, 1 + 1e-14)) / 3.0, 1e-14)
    sig = math.sqrt(s2)
    d1 = (math.log(m1 / strike) + 0.5 * s2) / sig
    d2 = d1 - sig
    disc = math.exp(-r * t)
    if is_call:
        return disc * (m1 * norm.cdf(d1) - strike * norm.cdf(d2))
    return disc * (strike * norm.cdf(-d2) - m1 * norm.cdf(-d1))
