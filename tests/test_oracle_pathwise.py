"""Pins the oracle's pathwise sensitivities (oracle/oracle_pathwise.c).

* The reference's own estimators on the same draws: GBM Europeans equal
  or_mc_european_pathwise_aad (mc_aad.rs:16-219, the tape AAD restated) and
  the Heston S0 slot equals or_heston_price_delta_aad (mc_aad.rs:239-342).
* Every sample's pv equals the discounted payoff or_mc_run computes for that
  sample (bit for bit: the path values follow the reference's operations).
* Every slot, payoff and model against per-path central differences of
  or_mc_run's payoffs with common random numbers (the same stream per
  sample): all but the few paths whose payoff kink, barrier status or
  variance clamp lies between the two bumped revaluations agree to 2e-5.
* Each planted mistake (a dropped term of the tangent recursion) fails that
  per-path check, so the check has the power to catch it.
"""
import numpy as np
import pytest

from oracle import oracle as O

SLOTS = {"spot": 0, "vol": 1, "v0": 1, "rate": 2, "expiry": 3, "kappa": 4, "theta": 5,
         "xi": 6, "rho": 7}


def _opt(**kw):
    return O.make_option(**kw)


def _kw(base, **over):
    d = dict(base)
    d.update(over)
    return d


def _per_path_fd(kw, field, paths, steps, seed, antithetic, rel=1e-5):
    """Per-sample central difference of the discounted payoff in `field`."""
    x0 = kw[field]
    h = rel * max(abs(x0), 0.05)
    out = []
    for sgn in (1.0, -1.0):
        d = dict(kw)
        d[field] = x0 + sgn * h
        if field == "expiry" and d.get("obs_times") is not None:
            # schedule held fixed in index space: scale the fixings with T
            d["obs_times"] = [t * d["expiry"] / x0 for t in kw["obs_times"]]
        o = _opt(**d)
        _, _, xs, _ = O.mc_run(o, paths, steps, seed, antithetic=antithetic, want_samples=True)
        out.append(np.exp(-d["rate"] * d["expiry"]) * xs)
    return (out[0] - out[1]) / (2 * h)


GBM = dict(model=O.GBM, spot=100.0, strike=95.0, expiry=1.25, rate=0.04, div_yield=0.01,
           vol=0.25)
HES = dict(model=O.HESTON, spot=100.0, strike=100.0, expiry=1.0, rate=0.03, div_yield=0.01,
           kappa=1.5, theta=0.05, xi=0.5, rho=-0.6, v0=0.045)
# Feller condition violated (2 kappa theta < xi^2): the variance clamp is hit
HES_CLAMP = dict(HES, kappa=0.8, theta=0.02, xi=0.9, v0=0.02)

STEPS = 24


def _obs(T, n=12):
    return [T * k / n for k in range(1, n + 1)]


CASES = [
    ("gbm_call", GBM, ["spot", "vol", "rate", "expiry"]),
    ("gbm_put", _kw(GBM, option_type=O.PUT), ["spot", "vol", "rate", "expiry"]),
    ("gbm_asian_arith", _kw(GBM, payoff=O.ASIAN, obs_times=_obs(1.25)),
     ["spot", "vol", "rate", "expiry"]),
    ("gbm_asian_geo_put", _kw(GBM, payoff=O.ASIAN, obs_times=_obs(1.25), averaging=O.GEO,
                              option_type=O.PUT), ["spot", "vol", "rate", "expiry"]),
    ("gbm_asian_floating", _kw(GBM, payoff=O.ASIAN, obs_times=_obs(1.25),
                               strike_type=O.FLOATING), ["spot", "vol", "rate", "expiry"]),
    ("gbm_up_out", _kw(GBM, payoff=O.BARRIER, barrier_level=140.0),
     ["spot", "vol", "rate", "expiry"]),
    ("gbm_down_in_put", _kw(GBM, payoff=O.BARRIER, barrier_dir=O.DOWN, barrier_style=O.IN,
                            barrier_level=85.0, option_type=O.PUT, rebate=1.5),
     ["spot", "vol", "rate", "expiry"]),
    ("heston_call", HES, ["spot", "v0", "rate", "expiry", "kappa", "theta", "xi", "rho"]),
    ("heston_clamped_put", _kw(HES_CLAMP, option_type=O.PUT),
     ["spot", "v0", "rate", "expiry", "kappa", "theta", "xi", "rho"]),
    ("heston_up_out", _kw(HES, payoff=O.BARRIER, barrier_level=130.0),
     ["spot", "v0", "rate", "kappa", "theta", "xi", "rho"]),
    ("heston_asian", _kw(HES, payoff=O.ASIAN, obs_times=_obs(1.0)),
     ["spot", "v0", "rate", "expiry", "kappa", "theta", "xi", "rho"]),
]


def _mismatch(fd, g):
    scale = np.mean(np.abs(g)) + np.mean(np.abs(fd)) + 1e-12
    bad = np.abs(fd - g) > 2e-5 * (np.abs(g) + scale)
    return bad


@pytest.mark.parametrize("name,kw,fields", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("antithetic", [False, True])
def test_per_path_central_differences(name, kw, fields, antithetic):
    paths, seed = 3000, 17
    r = O.pathwise_greeks(_opt(**kw), paths, STEPS, seed, antithetic=antithetic,
                          want_samples=True)
    per = r["samples"]
    for field in fields:
        fd = _per_path_fd(kw, field, paths, STEPS, seed, antithetic)
        g = per[:, 1 + SLOTS[field]]
        bad = _mismatch(fd, g)
        # kink / barrier / clamp crossings between the bumps: a handful of paths
        assert bad.mean() <= 0.01, (name, field, bad.mean())
        good = ~bad
        assert np.all(np.isfinite(g))
        # the mean over the agreeing paths matches too
        assert abs(np.mean(g[good]) - np.mean(fd[good])) <= 1e-6 * (np.mean(np.abs(g)) + 1e-9)


@pytest.mark.parametrize("name,kw,fields", CASES, ids=[c[0] for c in CASES])
def test_sample_pv_equals_mc_run_payoff(name, kw, fields):
    o = _opt(**kw)
    r = O.pathwise_greeks(o, 1000, STEPS, 5, antithetic=True, want_samples=True)
    _, _, xs, _ = O.mc_run(o, 1000, STEPS, 5, antithetic=True, want_samples=True)
    df = np.exp(-kw["rate"] * kw["expiry"])
    assert np.array_equal(r["samples"][:, 0], df * xs)
    p, se = O.mc_run(o, 1000, STEPS, 5, antithetic=True)
    assert abs(r["price"] - p) <= 1e-13 * abs(p)
    assert abs(r["stderr"] - se) <= 1e-9 * se


@pytest.mark.parametrize("antithetic", [False, True])
def test_gbm_european_equals_tape_aad(antithetic):
    for kw in (_kw(GBM, div_yield=0.0), _kw(GBM, option_type=O.PUT)):
        o = _opt(**kw)
        ref = O.mc_european_pathwise_aad(o, 20_000, 48, 42, antithetic)
        r = O.pathwise_greeks(o, 20_000, 48, 42, antithetic=antithetic)
        g = r["g"]
        assert abs(r["price"] - ref["price"]) <= 1e-13 * ref["price"]
        for got, want in ((g[0], ref["delta"]), (g[1], ref["vega"]), (g[2], ref["rho"]),
                          (-g[3], ref["theta"])):
            assert abs(got - want) <= 1e-11 * max(abs(want), 1.0), (got, want)


def test_heston_delta_equals_forward_dual():
    for kw in (HES, _kw(HES_CLAMP, option_type=O.PUT, strike=90.0)):
        o = _opt(**kw)
        p, d = O.heston_price_delta_aad(o, 8000, 64, 7)
        r = O.pathwise_greeks(o, 8000, 64, 7)
        assert abs(r["price"] - p) <= 1e-13 * p
        assert abs(r["g"][0] - d) <= 1e-12 * abs(d)


def test_external_normals_equal_reference_draws():
    o = _opt(**HES)
    z1, z2 = O.normals_matrix(9, 300, STEPS, 2)
    a = O.pathwise_greeks(o, 300, STEPS, 9)
    b = O.pathwise_greeks(o, 300, STEPS, 0, normals1=z1, normals2=z2)
    assert a["price"] == b["price"] and np.array_equal(a["g"], b["g"])


def test_threads_agree():
    o = _opt(**_kw(HES, payoff=O.ASIAN, obs_times=_obs(1.0)))
    a = O.pathwise_greeks(o, 4000, STEPS, 3, antithetic=True, nthreads=1)
    b = O.pathwise_greeks(o, 4000, STEPS, 3, antithetic=True, nthreads=4)
    assert abs(a["price"] - b["price"]) <= 1e-13 * a["price"]
    assert np.allclose(a["g"], b["g"], rtol=1e-12, atol=1e-14)


# (mutation, case, field) -- each planted mistake must fail the per-path check
MUTATIONS = [
    (1, "gbm_call", "vol"),
    (2, "heston_call", "kappa"),
    (3, "heston_call", "v0"),
    (4, "heston_call", "rho"),
    (5, "gbm_asian_arith", "spot"),
    (5, "heston_call", "spot"),
    (6, "gbm_call", "expiry"),
    (6, "heston_call", "expiry"),
]


@pytest.mark.parametrize("mutate,case,field", MUTATIONS)
def test_planted_mistakes_are_caught(mutate, case, field):
    kw = {c[0]: c[1] for c in CASES}[case]
    paths, seed = 1500, 23
    r = O.pathwise_greeks(_opt(**kw), paths, STEPS, seed, mutate=mutate, want_samples=True)
    fd = _per_path_fd(kw, field, paths, STEPS, seed, False)
    bad = _mismatch(fd, r["samples"][:, 1 + SLOTS[field]])
    assert bad.mean() > 0.3, (mutate, case, field, bad.mean())


@pytest.mark.parametrize("name,kw,fields", CASES, ids=[c[0] for c in CASES])
def test_maturity_slot_from_scaling_identity(name, kw, fields):
    """The Euler path depends on T only through kappa dt, theta dt, xi dt,
    mu dt and v0 dt (GBM: mu dt and sigma sqrt(dt)) in the per-step variance
    w = v dt, so path by path
      dP/dT = (kappa P_kappa + theta P_theta + xi P_xi + mu P_mu + v0 P_v0) / T - r pv
    (GBM: (mu P_mu + sigma P_sigma / 2) / T - r pv), P_mu = P_r + T pv.
    The oracle's directly propagated T slot satisfies it to rounding."""
    o = _opt(**kw)
    r = O.pathwise_greeks(o, 2000, STEPS, 31, antithetic=True, want_samples=True)
    pv, g = r["samples"][:, 0], r["samples"][:, 1:]
    T, mu, rate = kw["expiry"], kw["rate"] - kw["div_yield"], kw["rate"]
    g_mu = g[:, 2] + T * pv
    if kw["model"] == O.HESTON:
        ident = (kw["kappa"] * g[:, 4] + kw["theta"] * g[:, 5] + kw["xi"] * g[:, 6] +
                 mu * g_mu + kw["v0"] * g[:, 1]) / T - rate * pv
    else:
        ident = (mu * g_mu + 0.5 * kw["vol"] * g[:, 1]) / T - rate * pv
    scale = np.abs(g[:, 3]).mean() + np.abs(ident).mean()
    assert np.max(np.abs(ident - g[:, 3])) <= 1e-11 * scale
