"""GPU parity against the oracle on identical draws (C ABI calls).

Bar: per-sample payoffs within 1e-12 of the oracle relative to the payoff
scale max(|x_ref|, S0) (S_T's own rounding sets that scale near the payoff
kink), prices within 1e-11 relative. Knife-edge barrier paths (a spot within
1e-9 relative of the barrier, where 1-ulp exp differences can flip the
inclusive test) are counted and excluded.
"""
import numpy as np
import pytest

from tests.cases import daily_obs, make_pair, step_major

pytestmark = pytest.mark.gpu

PER_PATH_TOL = 1e-12
PRICE_TOL = 1e-11


def _cfg(rng="reference", antithetic=False, cv=False, n1=None, n2=None):
    from openferric_amd import engine_cfg, RngMode
    mode = {"reference": RngMode.Reference, "external": RngMode.External, "fast": RngMode.Fast}[rng]
    return engine_cfg(antithetic=antithetic, control_variate=cv, rng_mode=mode, normals1=n1,
                      normals2=n2)


def _per_path(oracle, case_kwargs, antithetic=False, cv=False, rng="reference"):
    from openferric_amd import sample_payoffs
    oo, g, keep = make_pair(**case_kwargs)
    n = case_kwargs["num_paths"]
    steps = case_kwargs["num_steps"]
    samples = (n + 1) // 2 if antithetic else n
    n1 = n2 = None
    on1 = on2 = None
    if rng == "external":
        streams = 2 if case_kwargs.get("model") == "heston" else 1
        z1, z2 = oracle.normals_matrix(case_kwargs["seed"], samples, steps, streams)
        on1, on2 = z1, (z2 if streams == 2 else None)
        n1 = step_major(z1)
        n2 = step_major(z2) if streams == 2 else None
    ref = oracle.mc_run(oo, n, steps, case_kwargs["seed"], antithetic=antithetic,
                        control_variate=cv, want_samples=True, normals1=on1, normals2=on2)
    cfg = _cfg(rng, antithetic, cv, n1, n2)
    xs, ys = sample_payoffs(g, cfg, antithetic)
    return ref, xs, ys, case_kwargs


def _check(ref, xs, ys, spot, mask=None):
    _, _, rx, ry = ref
    ex = np.abs(xs - rx) / np.maximum(np.abs(rx), spot)
    ey = np.abs(ys - ry) / np.maximum(np.abs(ry), spot)
    if mask is not None:
        ex, ey = ex[mask], ey[mask]
    assert ex.max() <= PER_PATH_TOL, f"max per-path x error {ex.max():.3e}"
    assert ey.max() <= PER_PATH_TOL, f"max per-path y error {ey.max():.3e}"


GBM_EURO = dict(model="gbm", payoff="vanilla", spot=100.0, strike=100.0, expiry=1.0, rate=0.05,
                vol=0.2, num_paths=4096, num_steps=252, seed=42)


@pytest.mark.parametrize("antithetic", [False, True])
def test_gbm_european_reference_draws_per_path(oracle, gpu, antithetic):
    ref, xs, ys, _ = _per_path(oracle, GBM_EURO, antithetic=antithetic)
    _check(ref, xs, ys, 100.0)


def test_gbm_european_external_normals_per_path(oracle, gpu):
    ref, xs, ys, _ = _per_path(oracle, dict(GBM_EURO, num_paths=2048), antithetic=True,
                               rng="external")
    _check(ref, xs, ys, 100.0)


def test_gbm_european_cv_per_path(oracle, gpu):
    ref, xs, ys, _ = _per_path(oracle, dict(GBM_EURO, option_type="put", q=0.02, num_paths=3000),
                               cv=True)
    _check(ref, xs, ys, 100.0)


@pytest.mark.parametrize("averaging,strike_type", [("arith", "fixed"), ("geo", "fixed"),
                                                   ("arith", "floating")])
def test_gbm_asian_per_path(oracle, gpu, averaging, strike_type):
    case = dict(GBM_EURO, payoff="asian", option_type="put", strike=98.0, num_paths=2048,
                num_steps=128, obs_times=[m / 11.0 for m in range(12)], averaging=averaging,
                strike_type=strike_type)
    ref, xs, ys, _ = _per_path(oracle, case, antithetic=True, cv=(averaging == "arith"
                                                                   and strike_type == "fixed"))
    _check(ref, xs, ys, 100.0)


def test_gbm_asian_daily_observation_per_path(oracle, gpu):
    case = dict(GBM_EURO, payoff="asian", num_paths=2048, num_steps=252,
                obs_times=daily_obs(1.0, 252))
    ref, xs, ys, _ = _per_path(oracle, case, cv=True)
    _check(ref, xs, ys, 100.0)


def _barrier_mask(oracle, case, antithetic):
    """True for samples whose legs never come within 1e-9 of the barrier."""
    from oracle.oracle import _ptr
    import numpy as np
    n = case["num_paths"]
    samples = (n + 1) // 2 if antithetic else n
    steps = case["num_steps"]
    H = case["barrier_level"]
    mask = np.ones(samples, dtype=bool)
    path = np.zeros(steps + 1)
    heston = case.get("model") == "heston"
    for i in range(samples):
        z1, z2 = oracle.sample_normals(case["seed"], i, steps, 2 if heston else 1)
        for sgn in ((1.0, -1.0) if antithetic else (1.0,)):
            a, b = sgn * z1, sgn * z2
            mu = case.get("rate", 0.05) - case.get("q", 0.0)
            if heston:
                oracle.lib().or_heston_generate(mu, case["kappa"], case["theta"], case["xi"],
                                                case["rho"], case["v0"], case["spot"],
                                                case["expiry"], steps, _ptr(a), _ptr(b), _ptr(path))
            else:
                oracle.lib().or_gbm_generate(mu, case["vol"], case["spot"], case["expiry"], steps,
                                             _ptr(a), _ptr(path))
            if np.min(np.abs(path - H)) / H < 1e-9:
                mask[i] = False
    return mask


@pytest.mark.parametrize("direction,style", [("up", "out"), ("down", "in")])
def test_gbm_barrier_per_path(oracle, gpu, direction, style):
    case = dict(GBM_EURO, payoff="barrier", num_paths=2048, num_steps=64,
                barrier_dir=direction, barrier_style=style,
                barrier_level=120.0 if direction == "up" else 85.0, rebate=1.5)
    ref, xs, ys, _ = _per_path(oracle, case, antithetic=True)
    mask = _barrier_mask(oracle, case, True)
    _check(ref, xs, ys, 100.0, mask)


HESTON = dict(model="heston", payoff="vanilla", spot=100.0, strike=100.0, expiry=1.0, rate=0.03,
              kappa=1.8, theta=0.04, xi=0.6, rho=-0.5, v0=0.04, num_paths=2048, num_steps=64,
              seed=7)


@pytest.mark.parametrize("antithetic", [False, True])
def test_heston_european_reference_draws_per_path(oracle, gpu, antithetic):
    ref, xs, ys, _ = _per_path(oracle, HESTON, antithetic=antithetic)
    _check(ref, xs, ys, 100.0)


def test_heston_external_normals_per_path(oracle, gpu):
    ref, xs, ys, _ = _per_path(oracle, dict(HESTON, num_paths=1024), rng="external")
    _check(ref, xs, ys, 100.0)


def test_heston_up_and_out_per_path(oracle, gpu):
    case = dict(HESTON, payoff="barrier", barrier_dir="up", barrier_style="out",
                barrier_level=125.0, num_paths=1024, num_steps=365, xi=0.9, v0=0.06, theta=0.06)
    ref, xs, ys, _ = _per_path(oracle, case)
    mask = _barrier_mask(oracle, case, False)
    _check(ref, xs, ys, 100.0, mask)


def test_config1_price_parity(oracle, gpu):
    """Config 1: European call, 100,000 antithetic paths x 252 steps, seed 42."""
    from openferric_amd import price_raw
    oo, g, keep = make_pair(**dict(GBM_EURO, num_paths=100_000))
    p_ref, se_ref = oracle.mc_run(oo, 100_000, 252, 42, antithetic=True)
    r = price_raw([g], _cfg("reference", antithetic=True))[0]
    assert abs(r.price - p_ref) / abs(p_ref) <= PRICE_TOL
    assert abs(r.stderr_ - se_ref) / se_ref <= 1e-8


def test_price_parity_cv_and_asian(oracle, gpu):
    from openferric_amd import price_raw
    cases = [
        (dict(GBM_EURO, num_paths=20_000, num_steps=64, seed=123), False, True),
        (dict(GBM_EURO, payoff="asian", num_paths=20_000, num_steps=128,
              obs_times=[m / 11.0 for m in range(12)]), True, True),
        (dict(HESTON, num_paths=8000), True, False),
    ]
    for case, anti, cv in cases:
        oo, g, keep = make_pair(**case)
        p_ref, se_ref = oracle.mc_run(oo, case["num_paths"], case["num_steps"], case["seed"],
                                      antithetic=anti, control_variate=cv)
        r = price_raw([g], _cfg("reference", antithetic=anti, cv=cv))[0]
        assert abs(r.price - p_ref) / abs(p_ref) <= PRICE_TOL, (case, r.price, p_ref)


def test_fast_mode_within_3se_of_bs(oracle, gpu):
    """Fast RNG mode (Box-Muller draws): config 1 within 3 SE of BS and of
    the oracle's reference-draw price."""
    from openferric_amd import price_raw
    oo, g, keep = make_pair(**dict(GBM_EURO, num_paths=1_000_000))
    r = price_raw([g], _cfg("fast", antithetic=True))[0]
    bs = 10.450583572185565
    assert abs(r.price - bs) <= 3 * r.stderr_
    p_ref, se_ref = oracle.mc_run(oo, 100_000, 252, 42, antithetic=True)
    assert abs(r.price - p_ref) <= 3 * np.hypot(r.stderr_, se_ref)


@pytest.mark.parametrize("steps", [1, 7, 64, 253])
def test_fast_mode_gbm_paths_equal_host_build(gpu, shim, steps):
    """Fast-mode GBM paths on the GPU equal, sample by sample, the host build
    of the same device math (tests/cpu_shim): xoroshiro128+ stream, Box-Muller
    from the LDS tables, 2^(y/256) steps. Only the GPU's reciprocal-sqrt seed
    and rounding of fused operations may move the last bits. A call struck at
    1e-9 makes the payoff S_T - 1e-9."""
    from openferric_amd import sample_payoffs
    n, seed, s0, r, vol, T = 3000, 17, 100.0, 0.04, 0.3, 1.5
    _, g, keep = make_pair(strike=1e-9, spot=s0, rate=r, vol=vol, expiry=T, num_paths=n,
                           num_steps=steps, seed=seed)
    xs, _ = sample_payoffs(g, _cfg("fast"), False, want_y=False)
    host = np.array([shim.shim_fast_gbm_terminal(seed, i, s0, r, vol, T, steps)
                     for i in range(n)]) - 1e-9
    np.testing.assert_allclose(xs, host, rtol=1e-13, atol=0)


@pytest.mark.parametrize("steps", [1, 6, 101])
def test_fast_mode_heston_paths_equal_host_build(gpu, shim, steps):
    """As above for the fast-mode Heston step (re-associated Euler step,
    2^(y/256) growth, generator clamps), including a path set whose variance
    hits zero often (Feller condition violated)."""
    from openferric_amd import sample_payoffs
    n, seed, s0, r, T = 2000, 23, 100.0, 0.03, 2.0
    for kappa, theta, xi, rho, v0 in [(1.5, 0.04, 0.5, -0.6, 0.05), (0.5, 0.02, 1.0, -0.9, 0.01)]:
        _, g, keep = make_pair(model="heston", strike=1e-9, spot=s0, rate=r, expiry=T,
                               kappa=kappa, theta=theta, xi=xi, rho=rho, v0=v0, num_paths=n,
                               num_steps=steps, seed=seed)
        xs, _ = sample_payoffs(g, _cfg("fast"), False, want_y=False)
        host = np.array([shim.shim_fast_heston_terminal(seed, i, s0, r, kappa, theta, xi, rho,
                                                         v0, T, steps)
                         for i in range(n)])
        host = np.maximum(host - 1e-9, 0.0)
        np.testing.assert_allclose(xs, host, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("model", ["gbm", "heston"])
@pytest.mark.parametrize("num_paths", [1, 2, 3, 7, 65, 130])
def test_ragged_waves_reference_draws_per_path(oracle, gpu, model, num_paths):
    """A sample range that leaves a wave mostly idle (one to seven busy
    lanes, or a last chunk of one or two): the tail values of a draw block
    are inverted by the busy lanes only, and with 400 steps a block often
    holds more tail values than busy lanes."""
    kw = dict(GBM_EURO, num_paths=num_paths, num_steps=400, seed=1234 + num_paths)
    if model == "heston":
        kw.update(model="heston", kappa=1.5, theta=0.04, xi=0.5, rho=-0.7, v0=0.04)
    ref, xs, ys, _ = _per_path(oracle, kw)
    _check(ref, xs, ys, 100.0)
