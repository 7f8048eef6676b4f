"""Pins the oracle's exact-terminal European pricers (oracle/oracle_exact.c):
mc_european_sequential / _parallel (src/engines/monte_carlo/mc_parallel.rs)
and mc_european_with_arena (mc_engine.rs:186-343) of the default build.

* Every payoff is the payoff of the i-th draw of the single stream, computed
  here independently from the oracle's pinned xoshiro256++ and Acklam
  functions (1e-14: Python has no fused multiply-add).
* The parallel pricer with one chunk IS the sequential pricer, bit for bit;
  with k chunks its paths are split_paths' chunks on the chunk seeds.
* The arena pricer draws the sequential pricer's stream with seed 42.
* Prices within 4 standard errors of Black-Scholes; the reference's tests
  (parallel vs sequential within 2 %, the arena within 1 % of Black-Scholes
  at 200,000 paths, the arena's repeatability).
* The reference's early returns: zero paths, zero expiry, vol <= 0,
  an invalid arena instrument.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

SEQ, PAR, ARENA = O.EXACT_SEQUENTIAL, O.EXACT_PARALLEL, O.EXACT_ARENA


def _stream_payoffs(seed, n, call, spot, strike, rate, q, vol, t):
    L = O.lib()
    s = (C.c_uint64 * 4)()
    L.or_xoshiro_seed(seed, s)
    drift = (rate - q - 0.5 * vol * vol) * t
    diff = vol * math.sqrt(t)
    out = np.empty(n)
    for i in range(n):
        z = L.or_acklam_inv_cdf(L.or_uniform_open01(L.or_xoshiro_next_f64(s)))
        st = spot * math.exp(diff * z + drift)
        out[i] = max(st - strike, 0.0) if call else max(strike - st, 0.0)
    return out


ARGS = dict(spot=100.0, strike=105.0, rate=0.04, q=0.015, vol=0.25, expiry=1.3)


def _run(kind, n, chunks=1, call=True, **kw):
    a = dict(ARGS, **kw)
    return O.mc_european_exact(kind, O.CALL if call else O.PUT, a["spot"], a["strike"], a["rate"],
                               a["q"], a["vol"], a["expiry"], n, chunks, want_payoffs=True)


@pytest.mark.parametrize("kind,seed", [(SEQ, 0xDEADBEEFCAFEBABE), (ARENA, 42)])
@pytest.mark.parametrize("call", [True, False])
def test_payoffs_are_the_single_streams(kind, seed, call):
    n = 1003
    _, _, pay = _run(kind, n, call=call)
    want = _stream_payoffs(seed, n, call, ARGS["spot"], ARGS["strike"], ARGS["rate"], ARGS["q"],
                           ARGS["vol"], ARGS["expiry"])
    assert np.allclose(pay, want, rtol=1e-14, atol=1e-12)


def test_parallel_one_chunk_is_sequential():
    p1, s1, a = _run(PAR, 4099, chunks=1)
    p2, s2, b = _run(SEQ, 4099)
    assert p1 == p2 and s1 == s2 and np.array_equal(a, b)


@pytest.mark.parametrize("n,chunks", [(1000, 3), (1003, 16), (10, 16)])
def test_parallel_chunks_follow_split_paths(n, chunks):
    _, _, pay = _run(PAR, n, chunks=chunks)
    base, rem = divmod(n, chunks)
    lens = [base + 1 if c < rem else base for c in range(chunks)]
    lens = [x for x in lens if x > 0]
    first = 0
    for i, ln in enumerate(lens):
        seed = (0xDEADBEEFCAFEBABE + i * 6364136223846793005) % (1 << 64)
        want = _stream_payoffs(seed, ln, True, ARGS["spot"], ARGS["strike"], ARGS["rate"],
                               ARGS["q"], ARGS["vol"], ARGS["expiry"])
        assert np.allclose(pay[first:first + ln], want, rtol=1e-14, atol=1e-12)
        first += ln
    assert first == n


@pytest.mark.parametrize("kind", [SEQ, PAR, ARENA])
def test_prices_against_black_scholes(kind):
    a = ARGS
    for call in (True, False):
        p, se, _ = _run(kind, 200_000, chunks=7, call=call)
        bs = O.bs_price(O.CALL if call else O.PUT, a["spot"], a["strike"], a["rate"], a["q"],
                        a["vol"], a["expiry"], exact=True)
        assert abs(p - bs) <= 4 * se, (kind, call, p, bs, se)


def test_reference_tests():
    # parallel_test.rs:49-72: parallel vs sequential within 2 % (100k paths)
    seq, _ = O.mc_european_exact(SEQ, O.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0, 100_000)
    par, _ = O.mc_european_exact(PAR, O.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0, 100_000, 16)
    assert abs(par - seq) / seq <= 0.02
    # mc_engine.rs:1170-1194: the arena within 1 % of Black-Scholes (200k paths)
    ar, _ = O.mc_european_exact(ARENA, O.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0, 200_000)
    assert abs(ar - 10.450583572185565) / 10.450583572185565 <= 0.01
    # mc_engine.rs:1197-1220: same seed and parameters give identical results
    a1 = O.mc_european_exact(ARENA, O.PUT, 95.0, 100.0, 0.02, 0.01, 0.3, 1.5, 4000)
    a2 = O.mc_european_exact(ARENA, O.PUT, 95.0, 100.0, 0.02, 0.01, 0.3, 1.5, 4000)
    assert a1 == a2 and math.isfinite(a1[0])


def test_early_returns():
    for kind in (SEQ, PAR, ARENA):
        p, se = O.mc_european_exact(kind, O.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0, 0, 4)
        assert math.isnan(p) and se is None
        p, se = O.mc_european_exact(kind, O.PUT, 100.0, 110.0, 0.05, 0.0, 0.2, 0.0, 10, 4)
        assert p == 10.0 and se == 0.0
        p, se = O.mc_european_exact(kind, O.CALL, 100.0, 100.0, 0.05, 0.0, 0.0, 1.0, 10, 4)
        assert math.isnan(p) and se is None
    p, se = O.mc_european_exact(ARENA, O.CALL, 100.0, -1.0, 0.05, 0.0, 0.2, 1.0, 10)
    assert math.isnan(p) and se is None
    # the sequential and parallel pricers do not validate the strike
    p, se = O.mc_european_exact(SEQ, O.PUT, 100.0, -1.0, 0.05, 0.0, 0.2, 1.0, 10)
    assert p == 0.0 and se == 0.0
