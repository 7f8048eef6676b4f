"""Independent numpy restatement of least-squares Monte Carlo for small
cases (test helper).

It follows Longstaff & Schwartz (2001) with the reference engine's conventions
(longstaff_schwartz.rs: regression of the discounted realised values on
{1, S, S^2} over in-the-money paths when there are at least 3, exercise when
intrinsic > continuation, values discounted one step at a time) but shares no
code with the oracle: paths are built with numpy's vectorised exp and
cumulative products, and the regression is a least-squares fit by numpy's
SVD-based lstsq instead of LU-solved normal equations. A dropped term, a
transposed basis or an off-by-one exercise step in the oracle therefore shows
up as a price difference far above the rounding noise.
"""
from __future__ import annotations

import numpy as np


def gbm_paths(spot, r, q, vol, expiry, steps, z):
    """z: [paths][steps]. Returns [paths][steps+1]."""
    dt = expiry / steps
    drift = (r - q - 0.5 * vol * vol) * dt
    step_vol = vol * np.sqrt(dt)
    growth = np.exp(step_vol * z + drift)
    out = np.empty((z.shape[0], steps + 1))
    out[:, 0] = spot
    out[:, 1:] = spot * np.cumprod(growth, axis=1)
    return out


def heston_paths(spot, r, q, kappa, theta, xi, rho, v0, expiry, steps, z1, z2):
    dt = expiry / steps
    n = z1.shape[0]
    s = np.full(n, float(spot))
    v = np.full(n, float(v0))
    out = np.empty((n, steps + 1))
    out[:, 0] = s
    for t in range(steps):
        vp = np.maximum(v, 0.0)
        zs = rho * z1[:, t] + np.sqrt(1.0 - rho * rho) * z2[:, t]
        vn = v + kappa * (theta - vp) * dt + xi * np.sqrt(vp) * np.sqrt(dt) * z1[:, t]
        sn = s * np.exp((r - q - 0.5 * vp) * dt + np.sqrt(vp) * np.sqrt(dt) * zs)
        s = np.maximum(sn, 1e-12)
        v = np.maximum(vn, 0.0)
        out[:, t + 1] = s
    return out


def lsm(paths, put, strikes_by_step, exercisable, rate, expiry, terminal_strike):
    """paths [n][steps+1]; exercisable: set of steps in [1, steps-1];
    strikes_by_step: dict step -> strike. Returns (price, stderr, points)
    where points lists (step, itm, exercised, extreme) in reverse time."""
    n, cols = paths.shape
    steps = cols - 1
    disc = np.exp(-rate * expiry / steps)

    def intrinsic(s, k):
        return np.maximum(k - s, 0.0) if put else np.maximum(s - k, 0.0)

    values = intrinsic(paths[:, steps], terminal_strike)
    points = []
    for t in range(steps - 1, 0, -1):
        values = values * disc
        if t not in exercisable:
            continue
        k = strikes_by_step[t]
        s = paths[:, t]
        ex = intrinsic(s, k)
        itm = ex > 0.0
        m = int(itm.sum())
        if m < 3:
            points.append((t, m, 0, None))
            continue
        basis = np.stack([np.ones(m), s[itm], s[itm] ** 2], axis=1)
        beta, *_ = np.linalg.lstsq(basis, values[itm], rcond=None)
        cont = beta[0] + beta[1] * s + beta[2] * s * s
        do = itm & (ex > cont)
        values = np.where(do, ex, values)
        cnt = int(do.sum())
        extreme = None
        if cnt:
            extreme = float(s[do].max() if put else s[do].min())
        points.append((t, m, cnt, extreme))
    values = values * disc
    mean = values.mean()
    se = values.std(ddof=1) / np.sqrt(n)
    return float(mean), float(se), points
