"""The oracle's restatement of the reference's basket Monte Carlo
(simulate_basket_once_matrix, src/pricing/basket.rs:599-683) pinned by
closed forms and the reference's own tests; the library's host-side
correlation repair (ofg_repair_correlation) against an independent numpy
Higham projection and Higham's published example. CPU only."""
import math

import numpy as np
import pytest

from tests.basket_numpy import bs_price, higham_nearest, margrabe, moment_matching

AVG, BEST, WORST = 0, 1, 2


def _z(n_paths, n_assets, seed):
    return np.random.default_rng(seed).standard_normal((n_paths, n_assets))


def test_zero_vol_is_discounted_forward_intrinsic(oracle):
    spots, divs, w = [100.0, 80.0, 120.0], [0.01, 0.03, 0.0], [0.5, 0.3, 0.2]
    corr = [[1, 0.3, 0.1], [0.3, 1, 0.2], [0.1, 0.2, 1]]
    r, t = 0.04, 1.5
    fwd = sum(wi * s * math.exp((r - q) * t) for wi, s, q in zip(w, spots, divs))
    for is_call, k in ((True, 90.0), (False, 110.0), (True, 130.0)):
        b = oracle.make_basket(spots, [0.0] * 3, divs, corr, w, AVG, is_call, k, t, r, 1000)
        p, se = oracle.basket_price(b, _z(1000, 3, 1))
        want = math.exp(-r * t) * max((fwd - k) if is_call else (k - fwd), 0.0)
        assert abs(p - want) < 1e-11 * max(1.0, want) and se < 1e-10


def test_one_asset_is_black_scholes(oracle):
    for is_call, k in ((True, 100.0), (False, 95.0)):
        b = oracle.make_basket([100.0], [0.25], [0.02], [[1.0]], [1.0], AVG, is_call, k, 1.0, 0.03,
                               400_000)
        p, se = oracle.basket_price(b)
        want = bs_price(is_call, 100.0, k, 0.03, 0.02, 0.25, 1.0)
        assert abs(p - want) < 4 * se, (p, want, se)


def test_zero_weight_asset_equals_one_asset_on_first_column(oracle):
    z = _z(20_000, 2, 5)
    corr = [[1.0, -0.4], [-0.4, 1.0]]
    two = oracle.make_basket([100.0, 70.0], [0.3, 0.2], [0.0, 0.01], corr, [1.0, 0.0], AVG, True,
                             100.0, 1.0, 0.02, 20_000)
    one = oracle.make_basket([100.0], [0.3], [0.0], [[1.0]], [1.0], AVG, True, 100.0, 1.0, 0.02,
                             20_000)
    p2, s2 = oracle.basket_price(two, z)
    p1, s1 = oracle.basket_price(one, z[:, :1])
    assert p1 == p2 and s1 == s2


@pytest.mark.parametrize("rho", [-0.5, 0.6])
def test_best_and_worst_of_two_match_margrabe(oracle, rho):
    s1, s2, q1, q2, r, t = 0.3, 0.2, 0.01, 0.03, 0.02, 1.0
    corr = [[1.0, rho], [rho, 1.0]]
    sig = math.sqrt(s1 * s1 + s2 * s2 - 2 * rho * s1 * s2)
    a, b = math.exp(-q1 * t), math.exp(-q2 * t)   # discounted forwards of S_i(T)/S_i(0)
    best_want = a + margrabe(b, a, sig, t)         # max(X1, X2) = X1 + (X2 - X1)^+
    worst_want = a - margrabe(a, b, sig, t)        # min(X1, X2) = X1 - (X1 - X2)^+
    for kind, want in ((BEST, best_want), (WORST, worst_want)):
        bk = oracle.make_basket([100.0, 50.0], [s1, s2], [q1, q2], corr, None, kind, True, 0.0, t,
                                r, 400_000)
        p, se = oracle.basket_price(bk)
        assert abs(p - want) < 4 * se, (kind, p, want, se)


def test_single_stream_is_path_major(oracle):
    n, m = 3_000, 4
    corr = np.full((m, m), 0.25)
    np.fill_diagonal(corr, 1.0)
    b = oracle.make_basket([100, 90, 110, 105], [0.2, 0.3, 0.25, 0.15], [0.0] * 4, corr,
                           [0.25] * 4, AVG, False, 100.0, 2.0, 0.01, n, seed=77)
    z = oracle.lsm_stream_normals(77, n * m).reshape(n, m)
    assert oracle.basket_price(b) == oracle.basket_price(b, z)


def test_zero_maturity_prices_the_spots(oracle):
    b = oracle.make_basket([100.0, 120.0], [0.2, 0.2], [0.0, 0.0], [[1, 0.2], [0.2, 1]],
                           None, BEST, True, 1.1, 0.0, 0.05, 100)
    assert oracle.basket_price(b) == (0.0, 0.0)   # max ratio is 1 < 1.1
    b = oracle.make_basket([100.0, 120.0], [0.2, 0.2], [0.0, 0.0], [[1, 0.2], [0.2, 1]],
                           [0.5, 0.5], AVG, True, 100.0, 0.0, 0.05, 100)
    assert oracle.basket_price(b) == (10.0, 0.0)   # no discounting at zero maturity


FIVE = dict(spots=[100.0, 98.0, 102.0, 95.0, 105.0], vols=[0.2, 0.21, 0.19, 0.23, 0.18],
            divs=[0.01] * 5, weights=[0.25, 0.2, 0.2, 0.2, 0.15],
            corr=[[1.0, 0.35, 0.30, 0.20, 0.15], [0.35, 1.0, 0.25, 0.30, 0.10],
                  [0.30, 0.25, 1.0, 0.28, 0.22], [0.20, 0.30, 0.28, 1.0, 0.26],
                  [0.15, 0.10, 0.22, 0.26, 1.0]])


def test_reference_five_asset_matches_moment_matching(oracle):
    """five_asset_mc_matches_moment_matching_within_two_percent (basket.rs:1155-1195)."""
    f = FIVE
    b = oracle.make_basket(f["spots"], f["vols"], f["divs"], f["corr"], f["weights"], AVG, True,
                           100.0, 1.0, 0.02, 250_000)
    mc, se = oracle.basket_price(b)
    mm = moment_matching("gentle", True, 100.0, f["spots"], f["weights"], f["vols"], f["corr"],
                         0.02, f["divs"], 1.0)
    assert abs(mc - mm) / mm <= 0.02, (mc, mm)
    levy = moment_matching("levy", True, 100.0, f["spots"], f["weights"], f["vols"], f["corr"],
                           0.02, f["divs"], 1.0)
    assert abs(mc - levy) / levy <= 0.02 and se < 0.05


def test_reference_one_asset_and_three_asset_tests(oracle):
    b = oracle.make_basket([100.0], [0.2], [0.0], [[1.0]], [1.0], AVG, True, 100.0, 1.0, 0.03,
                           120_000)
    p, se = oracle.basket_price(b)
    assert abs(p - bs_price(True, 100, 100, 0.03, 0.0, 0.2, 1.0)) <= 3 * se + 0.15
    corr = [[1.0, 0.4, 0.2], [0.4, 1.0, 0.3], [0.2, 0.3, 1.0]]
    prices = {}
    for kind in (WORST, BEST):
        b = oracle.make_basket([100.0, 98.0, 102.0], [0.25, 0.22, 0.2], [0.0] * 3, corr, None,
                               kind, True, 1.0, 1.0, 0.01, 120_000)
        prices[kind] = oracle.basket_price(b)[0]
        assert math.isfinite(prices[kind]) and prices[kind] >= 0.0
    assert prices[BEST] > prices[WORST]


def test_non_psd_matrix_is_refused(oracle):
    bad = [[1.0, 0.95, 0.95], [0.95, 1.0, -0.95], [0.95, -0.95, 1.0]]
    b = oracle.make_basket([100.0] * 3, [0.2] * 3, [0.0] * 3, bad, [1 / 3] * 3, AVG)
    with pytest.raises(ValueError):
        oracle.basket_price(b)


# ---- the library's host-side correlation repair ---------------------------

def test_higham_published_example():
    """Higham (2002), section 4: A = [[1,1,0],[1,1,1],[0,1,1]] has nearest
    correlation matrix with off-diagonals 0.7607 and 0.1573."""
    from openferric_amd import validate_or_repair_correlation_matrix
    a = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 1.0], [0.0, 1.0, 1.0]])
    m, rep = validate_or_repair_correlation_matrix(a)
    assert rep
    assert abs(m[0, 1] - 0.7607) < 5e-5 and abs(m[1, 2] - 0.7607) < 5e-5
    assert abs(m[0, 2] - 0.1573) < 5e-5
    assert np.abs(m - higham_nearest(a)).max() < 1e-9


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_repair_matches_numpy_projection(seed):
    from openferric_amd import validate_or_repair_correlation_matrix
    rng = np.random.default_rng(seed)
    n = 4 + 3 * seed
    a = rng.uniform(-0.9, 0.9, size=(n, n))
    a = (a + a.T) / 2
    np.fill_diagonal(a, 1.0)
    assert np.linalg.eigvalsh(a).min() < -1e-3
    m, rep = validate_or_repair_correlation_matrix(a)
    assert rep
    want = higham_nearest(a)
    assert np.abs(m - want).max() < 1e-8
    assert np.linalg.eigvalsh(m).min() >= -1e-8 and np.all(np.diag(m) == 1.0)
    assert np.array_equal(m, m.T)


def test_reference_repair_case_and_psd_passthrough():
    """psd_repair_validation_produces_valid_matrix (basket.rs:1197-1210)."""
    from openferric_amd import InvalidInput, validate_or_repair_correlation_matrix
    bad = [[1.0, 0.95, 0.95], [0.95, 1.0, -0.95], [0.95, -0.95, 1.0]]
    m, rep = validate_or_repair_correlation_matrix(bad)
    assert rep and np.linalg.eigvalsh(m).min() >= -1e-8
    good = np.array([[1.0, 0.3], [0.3, 1.0]])
    m, rep = validate_or_repair_correlation_matrix(good)
    assert not rep and np.array_equal(m, good)
    with pytest.raises(InvalidInput, match="symmetric"):
        validate_or_repair_correlation_matrix([[1.0, 0.3], [0.2, 1.0]])
    with pytest.raises(InvalidInput, match="diagonal"):
        validate_or_repair_correlation_matrix([[1.0, 0.3], [0.3, 0.9]])


# ---- outperformance and quanto variants (basket.rs:288-445) ----------------

def _black(f, k, sig, t):
    d1 = (math.log(f / k) + 0.5 * sig * sig * t) / (sig * math.sqrt(t))
    from scipy.stats import norm
    return f * norm.cdf(d1) - k * norm.cdf(d1 - sig * math.sqrt(t))


@pytest.mark.parametrize("strike", [0.0, 1.2])
def test_outperformance_ratio_matches_lognormal_closed_forms(oracle, strike):
    s1, s2, q1, q2, r, rho, t = 0.3, 0.2, 0.01, 0.03, 0.02, 0.4, 1.0
    b = oracle.make_basket([100.0, 50.0], [s1, s2], [q1, q2], [[1, rho], [rho, 1]], [0.0, 1.0],
                           AVG, True, strike, t, r, 400_000)
    p, se = oracle.outperformance_price(b, 0)
    f = 2.0 * math.exp((q2 - q1 + s2 * s2 - rho * s1 * s2) * t)   # E[S1(T)/S2(T)]
    sig = math.sqrt(s1 * s1 + s2 * s2 - 2 * rho * s1 * s2)
    want = math.exp(-r * t) * (f if strike == 0.0 else _black(f, strike, sig, t))
    assert abs(p - want) < 4 * se, (p, want, se)


def test_quanto_call_matches_adjusted_black_scholes(oracle):
    s, v, q, rf, rd, fx, fxv, rho, k, t = 100.0, 0.25, 0.01, 0.02, 0.04, 1.3, 0.12, -0.35, 95.0, 1.5
    b = oracle.make_basket([s], [v], [q], [[1.0]], [1.0], AVG, True, k, t, 0.0, 400_000)
    p, se = oracle.quanto_price(b, fx, fxv, [rho], rd, rf)
    f = s * math.exp((rf - q - rho * v * fxv) * t)
    want = fx * math.exp(-rd * t) * _black(f, k, v, t)
    assert abs(p - want) < 4 * se, (p, want, se)


def test_variants_zero_maturity_and_reference_case(oracle):
    b = oracle.make_basket([100.0, 95.0, 105.0], [0.2, 0.22, 0.21], [0.0] * 3,
                           [[1.0, 0.4, 0.35], [0.4, 1.0, 0.3], [0.35, 0.3, 1.0]], [0.0, 0.5, 0.5],
                           AVG, True, 1.0, 0.0, 0.02, 100)
    assert oracle.outperformance_price(b, 0) == (max(100.0 / 100.0 - 1.0, 0.0), 0.0)
    b = oracle.make_basket([100.0, 98.0], [0.2, 0.22], [0.01, 0.01], [[1.0, 0.4], [0.4, 1.0]],
                           [0.6, 0.4], AVG, True, 90.0, 0.0, 0.0, 100)
    assert oracle.quanto_price(b, 1.2, 0.12, [0.3, 0.25], 0.03, 0.02) == \
        (1.2 * (0.6 * 100 + 0.4 * 98 - 90.0), 0.0)
    # quanto_and_outperformance_are_priced (basket.rs:1298-1349)
    b = oracle.make_basket([100.0, 95.0, 105.0], [0.2, 0.22, 0.21], [0.0] * 3,
                           [[1.0, 0.4, 0.35], [0.4, 1.0, 0.3], [0.35, 0.3, 1.0]], [0.0, 0.5, 0.5],
                           AVG, True, 1.0, 1.0, 0.02, 120_000)
    p, _ = oracle.outperformance_price(b, 0)
    assert math.isfinite(p) and p >= 0.0
    b = oracle.make_basket([100.0, 98.0], [0.2, 0.22], [0.01, 0.01], [[1.0, 0.4], [0.4, 1.0]],
                           [0.6, 0.4], AVG, True, 100.0, 1.0, 0.0, 120_000)
    p, _ = oracle.quanto_price(b, 1.2, 0.12, [0.3, 0.25], 0.03, 0.02)
    assert math.isfinite(p) and p >= 0.0


# ---- factor-model baskets (basket.rs:162-205, 686-757) ----------------------

@pytest.mark.parametrize("betas", [(0.8, 0.5), (-0.6, 0.7)])
def test_one_factor_best_of_matches_margrabe(oracle, betas):
    s1, s2, q1, q2, r, t = 0.3, 0.2, 0.01, 0.03, 0.02, 1.0
    rho = betas[0] * betas[1]
    b = oracle.make_basket([100.0, 50.0], [s1, s2], [q1, q2], np.eye(2), None, BEST, True, 0.0,
                           t, r, 400_000)
    p, se = oracle.basket_factor_price(b, list(betas), True)
    sig = math.sqrt(s1 * s1 + s2 * s2 - 2 * rho * s1 * s2)
    a, bb = math.exp(-q1 * t), math.exp(-q2 * t)
    assert abs(p - (a + margrabe(bb, a, sig, t))) < 4 * se


def test_multi_factor_matches_cholesky_basket_on_implied_matrix(oracle):
    load = np.array([[0.6, 0.1], [0.5, -0.1], [0.4, 0.2], [0.35, -0.2], [0.25, 0.15]])
    implied = np.clip(load @ load.T, -1, 1)
    np.fill_diagonal(implied, 1.0)
    args = ([100.0, 101.0, 99.0, 97.0, 103.0], [0.2, 0.22, 0.18, 0.21, 0.19], [0.0] * 5)
    w = [0.3, 0.25, 0.2, 0.15, 0.1]
    bf = oracle.make_basket(*args, np.eye(5), w, AVG, True, 100.0, 1.0, 0.02, 200_000)
    bc = oracle.make_basket(*args, implied, w, AVG, True, 100.0, 1.0, 0.02, 200_000, seed=99)
    pf, sf = oracle.basket_factor_price(bf, load, False)
    pc, sc = oracle.basket_price(bc)
    assert abs(pf - pc) < 4 * math.hypot(sf, sc), (pf, pc)
    # factor_model_mc_prices_finite (basket.rs:1228-1260)
    assert math.isfinite(pf) and pf >= 0.0
    z = np.random.default_rng(3).standard_normal((1000, 7))
    bz = oracle.make_basket(*args, np.eye(5), w, AVG, True, 100.0, 1.0, 0.02, 1000)
    assert oracle.basket_factor_price(bz, load, False, z) != oracle.basket_factor_price(
        bz, load, False, z[:, ::-1].copy())   # the draw order matters: factors first


# ---- Student-t copula (basket.rs:120-160; math/correlation.rs:545-633) ------

def test_student_t_cdf_matches_scipy_and_closed_forms(oracle):
    """The oracle's restatement of statrs 0.18's StudentsT::cdf, beta_reg and
    ln_gamma against scipy's independent implementations, mpmath's
    regularized incomplete beta and the closed forms at 2 and 4 degrees of
    freedom."""
    from scipy import special, stats
    import mpmath as mp
    for nu in (2, 3, 4, 5, 7, 10, 30, 100, 1000):
        xs = np.concatenate([np.linspace(-30, 30, 241), np.geomspace(1e-2, 1e4, 40),
                             -np.geomspace(1e-2, 1e4, 40)])
        for x in xs:
            got, want = oracle.student_t_cdf(x, nu), stats.t.cdf(x, nu)
            if want > 1e-250:
                assert abs(got - want) <= 1e-11 * want, (nu, x, got, want)
            assert abs(got + oracle.student_t_cdf(-x, nu) - 1.0) <= 2e-16
    for x in (-7.0, -1.3, -0.2, 0.05, 0.9, 3.0, 25.0):
        f2 = 0.5 + x / (2.0 * math.sqrt(x * x + 2.0))
        t = x * x / 4.0
        f4 = 0.5 + 0.375 * (x / math.sqrt(1.0 + t)) * (1.0 - t / (3.0 * (1.0 + t)))
        assert abs(oracle.student_t_cdf(x, 2) - f2) <= 1e-14
        assert abs(oracle.student_t_cdf(x, 4) - f4) <= 1e-14
    # statrs evaluates I_h at h = nu / (nu + x^2), which rounds to 1 near 0
    assert oracle.student_t_cdf(1e-9, 100) == 0.5 and oracle.student_t_cdf(0.0, 5) == 0.5
    mp.mp.dps = 40
    rng = np.random.default_rng(11)
    for _ in range(200):
        a, b, x = rng.uniform(0.5, 60.0), rng.choice([0.5, rng.uniform(0.2, 8.0)]), rng.uniform()
        want = float(mp.betainc(a, b, 0, x, regularized=True))
        if want > 1e-250:
            assert abs(oracle.beta_reg(a, b, x) - want) <= 5e-12 * want, (a, b, x)
    for x in (0.5, 1.0, 1.5, 2.5, 7.0, 33.5, 150.0, 0.2):
        assert abs(oracle.ln_gamma(x) - special.gammaln(x)) <= 1e-12 * max(1.0, abs(special.gammaln(x)))


def _numpy_t_copula_price(spots, vols, divs, corr_or_none, load, one_factor, dof, weights, kind,
                          is_call, strike, t, r, z):
    """An independent restatement with scipy's t CDF and exact normal
    quantile, on the same draws: rows [factors..., assets..., chi-square...]."""
    from scipy import stats
    n = len(spots)
    if corr_or_none is not None:
        c = np.linalg.cholesky(np.asarray(corr_or_none)) @ z[:, :n].T
        k = 0
    else:
        L = np.asarray(load, dtype=np.float64).reshape(n, -1)
        k = L.shape[1]
        idio = np.sqrt(np.maximum(1.0 - (L * L).sum(axis=1), 0.0))
        c = L @ z[:, :k].T + idio[:, None] * z[:, k:k + n].T
    chi2 = (z[:, k + n:k + n + dof] ** 2).sum(axis=1)
    scale = np.sqrt(np.maximum(chi2 / dof, 1e-16))
    u = np.clip(stats.t.cdf(c / scale, dof), 1e-12, 1 - 1e-12)
    zz = stats.norm.ppf(u)
    st = np.asarray(spots)[:, None] * np.exp(
        ((r - np.asarray(divs) - 0.5 * np.asarray(vols) ** 2) * t)[:, None]
        + (np.asarray(vols) * math.sqrt(t))[:, None] * zz)
    if kind == AVG:
        u = (np.asarray(weights)[:, None] * st).sum(axis=0)
    else:
        rel = st / np.asarray(spots)[:, None]
        u = rel.max(axis=0) if kind == BEST else rel.min(axis=0)
    x = np.maximum(u - strike, 0.0) if is_call else np.maximum(strike - u, 0.0)
    return math.exp(-r * t) * x.mean()


@pytest.mark.parametrize("dof", [3, 8])
def test_t_copula_matches_independent_numpy_on_same_draws(oracle, dof):
    """Same draws, independent arithmetic (numpy Cholesky, scipy t CDF and
    exact normal quantile): agreement to the Acklam inverse's 1.2e-9."""
    spots, vols, divs = [100.0, 90.0, 110.0], [0.25, 0.3, 0.2], [0.01, 0.0, 0.02]
    corr = [[1, 0.4, 0.2], [0.4, 1, -0.1], [0.2, -0.1, 1]]
    w = [0.5, 0.3, 0.2]
    n = 4000
    z = np.random.default_rng(dof).standard_normal((n, 3 + dof))
    for kind, strike, is_call in ((AVG, 100.0, True), (WORST, 0.9, False), (BEST, 1.05, True)):
        b = oracle.make_basket(spots, vols, divs, corr, w if kind == AVG else None, kind, is_call,
                               strike, 1.25, 0.03, n, t_dof=dof)
        p, _ = oracle.basket_price(b, z)
        want = _numpy_t_copula_price(spots, vols, divs, corr, None, False, dof, w, kind, is_call,
                                     strike, 1.25, 0.03, z)
        assert abs(p - want) <= 1e-7 * max(want, 1e-3), (kind, p, want)
    load = [[0.6, 0.2], [0.5, -0.3], [0.4, 0.4]]
    zf = np.random.default_rng(40 + dof).standard_normal((n, 2 + 3 + dof))
    b = oracle.make_basket(spots, vols, divs, np.eye(3), w, AVG, True, 100.0, 1.0, 0.02, n,
                           t_dof=dof)
    p, _ = oracle.basket_factor_price(b, load, False, zf)
    want = _numpy_t_copula_price(spots, vols, divs, None, load, False, dof, w, AVG, True, 100.0,
                                 1.0, 0.02, zf)
    assert abs(p - want) <= 1e-7 * want, (p, want)


def test_t_copula_one_asset_marginal_is_black_scholes(oracle):
    """T = z / sqrt(chi2_nu / nu) is exactly t_nu, so Phi^-1(F_nu(T)) is
    exactly normal: one asset prices at Black-Scholes whatever the dof."""
    for dof, is_call, k in ((2, True, 100.0), (5, False, 105.0)):
        b = oracle.make_basket([100.0], [0.3], [0.01], [[1.0]], [1.0], AVG, is_call, k, 1.0, 0.03,
                               300_000, t_dof=dof)
        p, se = oracle.basket_price(b)
        want = bs_price(is_call, 100.0, k, 0.03, 0.01, 0.3, 1.0)
        assert abs(p - want) < 4 * se, (dof, p, want, se)


def test_t_copula_stream_order_and_tail_dependence(oracle):
    """The single stream draws each path's asset normals, then its dof
    chi-square normals; with zero correlation the t copula still couples the
    tails, so a best-of put (paid when every asset falls) is dearer than under
    the Gaussian copula and a worst-of put (paid when any falls) cheaper."""
    n, m, dof = 2_000, 3, 4
    b = oracle.make_basket([100, 90, 110], [0.2, 0.3, 0.25], [0.0] * 3, np.eye(3), None, WORST,
                           False, 1.0, 1.0, 0.01, n, seed=77, t_dof=dof)
    z = oracle.lsm_stream_normals(77, n * (m + dof)).reshape(n, m + dof)
    assert oracle.basket_price(b) == oracle.basket_price(b, z)
    for kind, sign in ((BEST, 1.0), (WORST, -1.0)):
        args = ([100.0, 100.0], [0.3, 0.3], [0.0, 0.0], np.eye(2), None, kind, False, 0.8, 1.0,
                0.02, 400_000)
        pt, st = oracle.basket_price(oracle.make_basket(*args, t_dof=3))
        pg, sg = oracle.basket_price(oracle.make_basket(*args))
        assert sign * (pt - pg) > 6 * math.hypot(st, sg), (kind, pt, pg)


# ---- correlation stress (basket.rs:447-454; correlation.rs:442-543) --------

def _numpy_stress(c, scen):
    """The scenario arithmetic restated in numpy: off-diagonal entries
    transformed, clamped to +-0.999999, the diagonal reset after each."""
    c = np.array(c, dtype=np.float64)
    n = c.shape[0]
    iu = np.triu_indices(n, 1)
    for kind, value, i, j in scen:
        if kind == "override":
            v = np.clip(value, -0.999999, 0.999999)
            c[i, j] = c[j, i] = v
        else:
            x = c[iu]
            x = {"scale": x * value, "shift": x + value, "floor": np.maximum(x, value),
                 "cap": np.minimum(x, value)}[kind]
            x = np.clip(x, -0.999999, 0.999999)
            c[iu] = x
            c[(iu[1], iu[0])] = x
        np.fill_diagonal(c, 1.0)
    return c


def test_stressed_correlation_matches_numpy_and_reference_test():
    from openferric_amd import CorrelationStressScenario as S, stressed_correlation_matrix
    from openferric_amd.core import InvalidInput
    base = [[1.0, 0.7, 0.6], [0.7, 1.0, 0.65], [0.6, 0.65, 1.0]]
    # stressed_correlation_is_repaired_to_psd (basket.rs:1212-1226)
    m = stressed_correlation_matrix(base, [S.AdditiveShift(0.35)])
    assert np.linalg.eigvalsh(m).min() >= -1e-8
    assert np.abs(m - higham_nearest(_numpy_stress(base, [("shift", 0.35, 0, 0)]))).max() < 1e-8
    rng = np.random.default_rng(4)
    a = rng.normal(size=(5, 8))
    c = a @ a.T
    d = np.sqrt(np.diag(c))
    c = c / np.outer(d, d)
    np.fill_diagonal(c, 1.0)
    c = (c + c.T) / 2
    cases = [[("scale", 0.5, 0, 0)], [("shift", -0.1, 0, 0)], [("floor", 0.2, 0, 0)],
             [("cap", 0.1, 0, 0)], [("override", 0.3, 1, 3)],
             [("scale", 1.3, 0, 0), ("cap", 0.6, 0, 0), ("override", -0.05, 4, 0)]]
    ctor = {"scale": S.ScaleOffDiagonal, "shift": S.AdditiveShift, "floor": S.FloorOffDiagonal,
            "cap": S.CapOffDiagonal}
    for sc in cases:
        want = _numpy_stress(c, sc)
        if np.linalg.eigvalsh(want).min() < -1e-10:
            want = higham_nearest(want)
            tol = 1e-8
        else:
            tol = 0.0
        objs = [S.OverridePair(i, j, v) if k == "override" else ctor[k](v) for k, v, i, j in sc]
        got = stressed_correlation_matrix(c, objs)
        assert np.abs(got - want).max() <= tol, (sc, np.abs(got - want).max())
    for bad, msg in (([S.ScaleOffDiagonal(float("nan"))], "scale factor must be finite"),
                     ([S.FloorOffDiagonal(1.5)], "floor must be finite"),
                     ([S.CapOffDiagonal(-2.0)], "cap must be finite"),
                     ([S.OverridePair(0, 0, 0.1)], "distinct and in-range"),
                     ([S.OverridePair(0, 3, 0.1)], "distinct and in-range"),
                     ([S.OverridePair(0, 1, 1.2)], "override pair value")):
        with pytest.raises(InvalidInput, match=msg):
            stressed_correlation_matrix(base, bad)
    with pytest.raises(InvalidInput, match="symmetric"):
        stressed_correlation_matrix([[1.0, 0.2], [0.3, 1.0]], [S.AdditiveShift(0.1)])
