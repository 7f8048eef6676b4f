"""Adjoint Greeks on the GPU (C ABI calls).

Checks:
  * GBM European on the reference draws against the oracle's restatement of
    mc_european_pathwise_aad (same draws): price and Greeks to 1e-10 relative;
  * the reference's own assertion (mc_aad.rs:353-409) against BS Greeks;
  * Heston delta against the oracle's heston_price_delta_aad on the same draws;
  * every adjoint gradient (GBM S0/sigma/r/T, Heston S0/v0/kappa/theta/xi/rho)
    against central finite differences with common random numbers priced on
    the GPU (same seeds, so the same draws), for European, Asian and barrier
    payoffs, in both RNG modes.
"""
import math

import numpy as np
import pytest

from tests.cases import make_pair

pytestmark = pytest.mark.gpu


def _cfg(rng="reference", antithetic=False, greek=0, bump=1e-4):
    from openferric_amd import engine_cfg, RngMode
    mode = {"reference": RngMode.Reference, "fast": RngMode.Fast}[rng]
    return engine_cfg(antithetic=antithetic, rng_mode=mode, greek_mode=greek, bump_rel=bump)


def _adjoint(g, rng="reference", antithetic=False):
    from openferric_amd import price_raw
    return price_raw([g], _cfg(rng, antithetic), greeks=True)[0]


def test_gbm_european_matches_oracle_tape_aad(oracle, gpu):
    case = dict(rate=0.03, q=0.01, vol=0.2, num_paths=40_000, num_steps=64, seed=42)
    oo, g, keep = make_pair(**case)
    for anti in (True, False):
        ref = oracle.mc_european_pathwise_aad(oo, 40_000, 64, 42, anti)
        r = _adjoint(g, antithetic=anti)
        assert r.has_greeks
        assert abs(r.price - ref["price"]) <= 1e-12 * ref["price"]
        for name, got in (("delta", r.delta), ("vega", r.vega), ("rho", r.rho),
                          ("theta", r.theta)):
            assert abs(got - ref[name]) <= 1e-10 * max(1.0, abs(ref[name])), (name, got, ref[name])


def test_gbm_european_reference_assertion(oracle, gpu):
    """mc_aad.rs:353-409: 120k antithetic paths x 64 steps vs BS Greeks."""
    oo, g, keep = make_pair(rate=0.03, q=0.01, vol=0.2, num_paths=120_000, num_steps=64, seed=42)
    r = _adjoint(g, antithetic=True)
    ref = oracle.bs_greeks(oracle.CALL, 100, 100, 0.03, 0.01, 0.2, 1.0)
    assert abs(r.delta - ref["delta"]) < 6e-3
    assert abs(r.vega - ref["vega"]) < 8e-2
    assert abs(r.rho - ref["rho"]) < 7e-2
    assert abs(r.theta - ref["theta"]) < 9e-2


HESTON = dict(model="heston", spot=100.0, strike=70.0, expiry=1.0, rate=0.03, kappa=1.8,
              theta=0.04, xi=0.6, rho=-0.5, v0=0.04, num_paths=60_000, num_steps=64, seed=7)


def test_heston_delta_matches_oracle_forward_dual(oracle, gpu):
    """heston_price_delta_aad semantics (mc_aad.rs:239-342) on the same draws."""
    oo, g, keep = make_pair(**HESTON)
    p_ref, d_ref = oracle.heston_price_delta_aad(oo, 60_000, 64, 7)
    r = _adjoint(g)
    assert abs(r.price - p_ref) <= 1e-11 * p_ref
    assert abs(r.delta - d_ref) <= 1e-10 * abs(d_ref)
    assert abs(r.heston_grad[0] - d_ref) <= 1e-10 * abs(d_ref)


def _full(case):
    import inspect
    defaults = {k: v.default for k, v in inspect.signature(make_pair).parameters.items()}
    return {**defaults, **case}


def _fd(case, field, h, rng, antithetic):
    """CRN central difference of the GPU price in one ofg_option field."""
    from openferric_amd import price_raw
    case = _full(case)
    up = dict(case)
    dn = dict(case)
    up[field] = case[field] + h
    dn[field] = case[field] - h
    _, gu, k1 = make_pair(**up)
    _, gd, k2 = make_pair(**dn)
    ru, rd = price_raw([gu, gd], _cfg(rng, antithetic))
    return (ru.price - rd.price) / (2 * h)


GBM_FIELDS = [("spot", "delta"), ("vol", "vega"), ("rate", "rho"), ("expiry", "theta")]
HESTON_FIELDS = [("spot", 0), ("v0", 1), ("kappa", 2), ("theta", 3), ("xi", 4), ("rho", 5)]


def _check_fd(case, rng, antithetic, rel_tol, abs_tol):
    case = _full(case)
    oo, g, keep = make_pair(**case)
    r = _adjoint(g, rng, antithetic)
    heston = case.get("model") == "heston"
    fields = HESTON_FIELDS if heston else GBM_FIELDS
    for field, slot in fields:
        # small bumps: with common random numbers, a larger bump lets paths
        # cross the payoff kink between the two revaluations and the FD picks
        # up that sampling noise; the allowance covers FD round-off.
        h = 1e-7 * max(abs(case.get(field, 0.0)), 1e-2)
        fd = _fd(case, field, h, rng, antithetic)
        tol = rel_tol * abs(fd) + abs_tol + 20 * 2.2e-16 * abs(r.price) / h
        if heston:
            got = r.heston_grad[slot]
        else:
            got = getattr(r, slot)
            if slot == "theta":
                fd = -fd
        assert abs(got - fd) <= tol, (field, got, fd)


@pytest.mark.parametrize("rng", ["reference", "fast"])
def test_gbm_greeks_vs_crn_fd(gpu, rng):
    for payoff, extra in [("vanilla", {}),
                          ("asian", dict(obs_times=[m / 12.0 for m in range(1, 13)])),
                          ("asian", dict(obs_times=[m / 12.0 for m in range(1, 13)],
                                         averaging="geo", option_type="put")),
                          ("barrier", dict(barrier_dir="up", barrier_style="out",
                                           barrier_level=250.0))]:
        case = dict(payoff=payoff, strike=95.0, rate=0.04, q=0.01, vol=0.25, expiry=1.5,
                    num_paths=20_000, num_steps=48, seed=11, **extra)
        _check_fd(case, rng, True, 2e-4, 1e-6)


@pytest.mark.parametrize("rng", ["reference", "fast"])
def test_heston_greeks_vs_crn_fd(gpu, rng):
    # Feller satisfied (2 kappa theta > xi^2) so the variance floor is never hit
    case = dict(HESTON, strike=100.0, kappa=2.0, theta=0.05, xi=0.3, rho=-0.6, v0=0.045,
                num_paths=20_000, num_steps=50)
    _check_fd(case, rng, False, 5e-4, 1e-6)
    case_b = dict(case, payoff="barrier", barrier_dir="up", barrier_style="out",
                  barrier_level=200.0)
    _check_fd(case_b, rng, True, 5e-4, 1e-6)


def test_adjoint_price_equals_pricing_kernel(gpu):
    """The adjoint's forward sweep is the pricing path: same plain price."""
    from openferric_amd import price_raw
    for case in [dict(num_paths=10_000, num_steps=37, seed=5),
                 dict(HESTON, num_paths=10_000, num_steps=37)]:
        for rng in ("reference", "fast"):
            oo, g, keep = make_pair(**case)
            p = price_raw([g], _cfg(rng, True))[0]
            a = _adjoint(g, rng, True)
            assert abs(a.price - p.price) <= 1e-13 * abs(p.price)


def test_fd_greek_mode_matches_adjoint(gpu):
    from openferric_amd import price_raw, _lib
    oo, g, keep = make_pair(num_paths=40_000, num_steps=32, seed=3)
    fd = price_raw([g], _cfg("fast", True, greek=_lib.GREEKS_FD_CRN, bump=1e-5))[0]
    ad = _adjoint(g, "fast", True)
    for name in ("delta", "vega", "rho", "theta"):
        a, f = getattr(ad, name), getattr(fd, name)
        assert abs(a - f) <= 1e-3 * abs(f) + 1e-5, (name, a, f)


FWD_CASES = [
    dict(payoff="vanilla"),
    dict(payoff="vanilla", option_type="put", q=0.02),
    dict(payoff="asian", obs="daily"),
    dict(payoff="asian", obs="sparse", option_type="put"),
    dict(payoff="asian", obs="daily", averaging="geo"),
    dict(payoff="asian", obs="sparse", strike_type="floating"),
    dict(payoff="barrier", barrier_dir="up", barrier_style="out", barrier_level=135.0, rebate=1.0),
    dict(payoff="barrier", barrier_dir="down", barrier_style="in", barrier_level=85.0),
]


@pytest.mark.parametrize("rng", ["reference", "fast"])
@pytest.mark.parametrize("antithetic", [False, True])
def test_forward_pathwise_equals_adjoint(gpu, rng, antithetic):
    """OFG_GREEKS_FORWARD (GBM running sums, one forward sweep) is the same
    pathwise estimator as the adjoint kernel on the same draws: prices
    identical to 1e-14, every Greek within 1e-10 of the adjoint (relative to
    max(|Greek|, 1)), for every payoff kind, in one batched call."""
    from openferric_amd import _lib, price_raw
    from tests.cases import daily_obs
    structs, keeps = [], []
    for i, c in enumerate(FWD_CASES):
        c = dict(c)
        obs = c.pop("obs", None)
        steps = 48
        if obs == "daily":
            c["obs_times"] = daily_obs(1.0, steps)
        elif obs == "sparse":
            c["obs_times"] = [k / 7.0 for k in range(8)]
        _, g, keep = make_pair(vol=0.25, rate=0.04, num_paths=6_001, num_steps=steps, seed=100 + i,
                               **c)
        structs.append(g)
        keeps.append(keep)
    adj = price_raw(structs, _cfg(rng, antithetic, greek=_lib.GREEKS_ADJOINT))
    fwd = price_raw(structs, _cfg(rng, antithetic, greek=_lib.GREEKS_FORWARD))
    for c, a, f in zip(FWD_CASES, adj, fwd):
        assert f.has_greeks
        assert abs(f.price - a.price) <= 1e-14 * max(abs(a.price), 1.0), (c, f.price, a.price)
        for k in ("delta", "vega", "rho", "theta"):
            fa, ff = getattr(a, k), getattr(f, k)
            assert abs(ff - fa) <= 1e-10 * max(abs(fa), 1.0), (c, k, ff, fa)


def test_forward_pathwise_matches_oracle_tape_aad(oracle, gpu):
    """Forward mode against the oracle's mc_european_pathwise_aad on the
    reference draws (the reference's own estimator), antithetic on and off;
    a Heston option in the same call (forward tangent) agrees with the
    adjoint kernel."""
    from openferric_amd import _lib, price_raw
    case = dict(rate=0.03, q=0.01, vol=0.2, num_paths=40_000, num_steps=64, seed=42)
    oo, g, keep = make_pair(**case)
    _, h, hk = make_pair(**HESTON)
    for anti in (True, False):
        ref = oracle.mc_european_pathwise_aad(oo, 40_000, 64, 42, anti)
        r, rh = price_raw([g, h], _cfg("reference", anti, greek=_lib.GREEKS_FORWARD))
        assert abs(r.price - ref["price"]) <= 1e-12 * ref["price"]
        for name in ("delta", "vega", "rho", "theta"):
            got = getattr(r, name)
            assert abs(got - ref[name]) <= 1e-10 * max(1.0, abs(ref[name])), (name, got, ref[name])
        ha = price_raw([h], _cfg("reference", anti, greek=_lib.GREEKS_ADJOINT))[0]
        assert rh.price == ha.price
        for a_, f_ in zip(ha.heston_grad, rh.heston_grad):
            assert abs(f_ - a_) <= 1e-10 * max(abs(a_), 1.0), (f_, a_)


HESTON_FWD_CASES = [
    dict(payoff="vanilla"),
    dict(payoff="vanilla", option_type="put", xi=1.2, theta=0.02, v0=0.01),  # v hits 0 often
    dict(payoff="barrier", barrier_dir="up", barrier_style="out", barrier_level=125.0, rebate=0.5),
    dict(payoff="barrier", barrier_dir="down", barrier_style="in", barrier_level=88.0),
    dict(payoff="asian", obs="daily"),
    dict(payoff="asian", obs="sparse", averaging="geo"),
]


@pytest.mark.parametrize("rng", ["reference", "fast"])
@pytest.mark.parametrize("antithetic", [False, True])
def test_heston_forward_tangent_equals_adjoint(gpu, rng, antithetic):
    """OFG_GREEKS_FORWARD on Heston options (forward tangent of the Euler
    step, one sweep) against the checkpointed adjoint on the same draws:
    identical prices, all six Heston sensitivities and rho/theta within 1e-10
    (relative to max(|g|, 1)), every payoff kind, Feller condition violated in
    one case."""
    from openferric_amd import _lib, price_raw
    from tests.cases import daily_obs
    structs, keeps = [], []
    for i, c in enumerate(HESTON_FWD_CASES):
        c = dict(HESTON, **c)
        obs = c.pop("obs", None)
        c.update(num_paths=3_001, num_steps=40, seed=300 + i, strike=100.0)
        if obs == "daily":
            c["obs_times"] = daily_obs(1.0, 40)
        elif obs == "sparse":
            c["obs_times"] = [k / 5.0 for k in range(6)]
        _, g, keep = make_pair(**c)
        structs.append(g)
        keeps.append(keep)
    adj = price_raw(structs, _cfg(rng, antithetic, greek=_lib.GREEKS_ADJOINT))
    fwd = price_raw(structs, _cfg(rng, antithetic, greek=_lib.GREEKS_FORWARD))
    for c, a, f in zip(HESTON_FWD_CASES, adj, fwd):
        assert f.price == a.price, c
        for k in range(6):
            ga, gf = a.heston_grad[k], f.heston_grad[k]
            assert abs(gf - ga) <= 1e-10 * max(abs(ga), 1.0), (c, k, gf, ga)
        for k in ("rho", "theta", "vega"):
            ga, gf = getattr(a, k), getattr(f, k)
            assert abs(gf - ga) <= 1e-10 * max(abs(ga), 1.0), (c, k, gf, ga)
