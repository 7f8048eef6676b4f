"""Pins of the oracle's quasi-Monte Carlo restatement (CPU only).

* the reference's own assertions (src/math/sobol.rs tests, mc_qmc.rs tests):
  points in [0, 1), reproducibility, dimension 0's mean closer to 1/2 than
  xoshiro256++'s, QMC within 0.1 % of Black-Scholes at 10k points (seed 11),
  QMC standard error below 1.2x plain MC's at 50k points (seed 7);
* the Gray-code identity: the oracle's sequential next_into recurrence gives
  point n = XOR of the direction numbers at the set bits of n ^ (n >> 1)
  (then XOR the scrambler) -- the formula the GPU evaluates per path;
* the (0, m, 1)-net property of every coordinate: among Gray-code points
  0 .. 2^m - 1 each dyadic interval of length 2^-m holds exactly one point,
  which needs each direction number j to have its leading bit at 63 - j (a
  wrong mask or shift breaks it); checked for the scrambled uniforms too;
* the u64 -> (0, 1) map is exact (fma of the rounded word) and stays inside
  (0, 1).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_points_in_unit_interval_and_reproducible():
    _, u = O.sobol_points(8, 42, 1000)
    assert np.all(u >= 0.0) and np.all(u < 1.0)
    _, u2 = O.sobol_points(8, 42, 1000)
    assert np.array_equal(u, u2)
    _, u3 = O.sobol_points(8, 43, 1000)
    assert not np.array_equal(u[:, 1:], u3[:, 1:])


def test_first_dimension_mean_closer_to_half_than_prng():
    _, u = O.sobol_points(1, 7, 1000)
    s = (O.C.c_uint64 * 4)()
    O.lib().or_xoshiro_seed(7, s)
    prng = np.array([O.lib().or_xoshiro_next_f64(s) for _ in range(1000)])
    assert abs(u[:, 0].mean() - 0.5) <= abs(prng.mean() - 0.5)


def _u01(word):
    """fma(word as f64, 2^-64, 2^-65): the product is exact (a power of two),
    so one rounded addition is the fused result."""
    return float(word) * 2.0 ** -64 + 2.0 ** -65


def _gray_words(dims, seed, n):
    """Points 0..n-1 by the Gray-code formula on the oracle's directions."""
    out = np.zeros((n, dims), dtype=object)
    for d in range(dims):
        v = O.sobol_directions(d, seed)
        sc = O.sobol_scrambler(d, seed)
        for k in range(n):
            g = k ^ (k >> 1)
            x = 0
            j = 0
            while g:
                if g & 1:
                    x ^= v[j]
                g >>= 1
                j += 1
            out[k, d] = x ^ sc
    return out


@pytest.mark.parametrize("seed", [0, 42, 2**63 + 12345])
def test_gray_code_identity(seed):
    dims, n = 6, 300
    w, _ = O.sobol_points(dims, seed, n)
    g = _gray_words(dims, seed, n + 1)
    for k in range(1, n + 1):
        assert [int(x) for x in w[k - 1]] == list(g[k]), k


@pytest.mark.parametrize("dim", [0, 1, 2, 7, 31, 300])
def test_each_coordinate_is_a_0_m_1_net(dim):
    seed, m = 99, 10
    n = 1 << m
    v = O.sobol_directions(dim, seed)
    for j in range(64):
        assert (v[j] >> (63 - j)) & 1 == 1 and v[j] & ((1 << (63 - j)) - 1) == 0
    sc = O.sobol_scrambler(dim, seed)
    xs = []
    for k in range(n):
        g, x, j = k ^ (k >> 1), 0, 0
        while g:
            if g & 1:
                x ^= v[j]
            g >>= 1
            j += 1
        xs.append(x ^ sc)
    cells = sorted(x >> (64 - m) for x in xs)
    assert cells == list(range(n))
    # the same for the uniforms the oracle emits (points 1..n-1 plus point 0)
    _, u = O.sobol_points(dim + 1, seed, n - 1)
    u0 = _u01(sc)
    cells = sorted([int(math.floor(u0 * n))] + [int(math.floor(x * n)) for x in u[:, dim]])
    assert cells == list(range(n))


def test_uniform_map():
    w, u = O.sobol_points(4, 5, 200)
    for ww, uu in zip(w.ravel(), u.ravel()):
        assert uu == _u01(int(ww))
        assert 0.0 < uu < 1.0


def test_qmc_matches_black_scholes_at_10k_points():
    o = O.make_option(option_type=O.CALL, spot=100.0, strike=100.0, expiry=1.0, rate=0.05,
                      vol=0.2)
    p, se = O.mc_european_qmc(o, 10_000, 1, 11)
    bs = O.bs_price(O.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0, exact=True)
    assert abs(p - bs) / bs <= 0.001
    assert se is not None and se > 0.0


def test_qmc_stderr_competitive_with_prng_mc():
    o = O.make_option(option_type=O.CALL, spot=100.0, strike=100.0, expiry=1.0, rate=0.05,
                      vol=0.2)
    _, qse = O.mc_european_qmc(o, 50_000, 1, 7)
    _, mse = O.mc_run(o, 50_000, 1, 7)
    assert qse < mse * 1.2


def test_qmc_multi_step_converges_faster_than_prng():
    o = O.make_option(option_type=O.PUT, spot=100.0, strike=105.0, expiry=2.0, rate=0.03,
                      div_yield=0.01, vol=0.25)
    bs = O.bs_price(O.PUT, 100.0, 105.0, 0.03, 0.01, 0.25, 2.0, exact=True)
    p, _ = O.mc_european_qmc(o, 1 << 14, 8, 3)
    q, qse = O.mc_run(o, 1 << 14, 8, 3)
    assert abs(p - bs) < 4 * qse          # within the PRNG's own error bar
    assert abs(p - bs) < 0.02


def test_degenerate_inputs_follow_reference():
    o = O.make_option(option_type=O.CALL, spot=100.0, strike=90.0, expiry=1.0, rate=0.05,
                      vol=0.2)
    p, se = O.mc_european_qmc(o, 0, 4, 1)
    assert math.isnan(p) and se is None
    o0 = O.make_option(option_type=O.CALL, spot=100.0, strike=90.0, expiry=0.0, rate=0.05,
                       vol=0.2)
    assert O.mc_european_qmc(o0, 100, 4, 1) == (10.0, 0.0)
    ov = O.make_option(option_type=O.CALL, spot=100.0, strike=90.0, expiry=1.0, rate=0.05,
                       vol=0.0)
    p, se = O.mc_european_qmc(ov, 100, 4, 1)
    assert math.isnan(p) and se is None
