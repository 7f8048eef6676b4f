"""The exact-terminal European pricers on the GPU (ofg_european_exact)
against the oracle's restatement of the reference's default build
(mc_european_sequential / _parallel, mc_european_with_arena), on the
reference's own single streams: every path's payoff to 1e-12 (relative to
max(|payoff|, S0)), prices to 1e-11, standard errors to 1e-9; several
options and the degenerate cases in one call; 2^22 + 5 paths (many jump
bits); the Python mirrors and the reference's own tests through them."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEQ, PAR, ARENA = 0, 1, 2


def _opt(call=True, spot=100.0, strike=105.0, rate=0.04, q=0.015, vol=0.25, expiry=1.3,
         paths=100_003):
    from openferric_amd import _lib
    o = _lib.OfgOption()
    o.model, o.payoff = _lib.MODEL_GBM, _lib.PAYOFF_VANILLA
    o.option_type = _lib.CALL if call else _lib.PUT
    o.spot, o.strike, o.rate, o.div_yield, o.vol, o.expiry = spot, strike, rate, q, vol, expiry
    o.num_paths, o.num_steps = paths, 1
    return o


def _oracle(oracle, kind, o, chunks):
    return oracle.mc_european_exact(kind, oracle.CALL if o.option_type == 0 else oracle.PUT,
                                    o.spot, o.strike, o.rate, o.div_yield, o.vol, o.expiry,
                                    o.num_paths, chunks, want_payoffs=True)


def _check(got_r, got_pay, want, spot):
    p, se, pay = want
    if math.isnan(p):
        assert math.isnan(got_r.price) and not got_r.has_stderr
        return
    assert abs(got_r.price - p) <= 1e-11 * max(abs(p), 1e-12), (got_r.price, p)
    assert got_r.has_stderr and abs(got_r.stderr_ - se) <= 1e-9 * max(se, 1e-12)
    if got_pay is not None and pay is not None and len(pay):
        err = np.max(np.abs(got_pay - pay) / np.maximum(np.abs(pay), spot))
        assert err <= 1e-12, err


@pytest.mark.parametrize("kind,chunks", [(SEQ, 1), (PAR, 1), (PAR, 5), (PAR, 64), (ARENA, 1)])
def test_per_path_parity(oracle, gpu, kind, chunks):
    from openferric_amd import european_exact_raw
    opts = [_opt(True), _opt(False, strike=95.0, vol=0.4, expiry=0.7, paths=77_777)]
    res, pay = european_exact_raw(opts, kind, chunks, want_payoffs=True)
    first = 0
    for o, r in zip(opts, res):
        want = _oracle(oracle, kind, o, chunks)
        _check(r, pay[first:first + o.num_paths], want, o.spot)
        assert r.num_paths == o.num_paths and r.num_steps == 1.0 and r.vol == o.vol
        first += o.num_paths


def test_batch_with_degenerate_cases(oracle, gpu):
    from openferric_amd import european_exact_raw
    opts = [_opt(paths=5000), _opt(paths=0), _opt(expiry=0.0, strike=90.0, paths=10),
            _opt(vol=0.0, paths=10), _opt(strike=-1.0, paths=10), _opt(False, paths=131)]
    for kind in (SEQ, PAR, ARENA):
        res, pay = european_exact_raw(opts, kind, 3, want_payoffs=True)
        first = 0
        for o, r in zip(opts, res):
            want = _oracle(oracle, kind, o, 3)
            sim = not math.isnan(want[0]) and o.expiry > 0.0 and o.num_paths > 0
            _check(r, pay[first:first + o.num_paths] if sim else None, want, o.spot)
            first += o.num_paths
    # zero expiry: the payoff at the spot, stderr 0
    r = european_exact_raw([_opt(expiry=0.0, strike=90.0)], SEQ)[0]
    assert r.price == 10.0 and r.stderr_ == 0.0 and r.has_stderr


def test_many_paths(oracle, gpu):
    from openferric_amd import european_exact_raw
    o = _opt(paths=(1 << 22) + 5)
    for kind, chunks in ((SEQ, 1), (PAR, 12), (ARENA, 1)):
        r = european_exact_raw([o], kind, chunks)[0]
        p, se = oracle.mc_european_exact(kind, oracle.CALL, o.spot, o.strike, o.rate,
                                         o.div_yield, o.vol, o.expiry, o.num_paths, chunks)
        assert abs(r.price - p) <= 1e-11 * p and abs(r.stderr_ - se) <= 1e-9 * se


def test_mirrors_and_reference_tests(oracle, gpu):
    import openferric_amd as ofa
    m = ofa.Market.builder().spot(100.0).rate(0.05).dividend_yield(0.0).flat_vol(0.2).build()
    call = ofa.VanillaOption.european_call(100.0, 1.0)
    seq = ofa.mc_european_sequential(call, m, 100_000, 252)
    par = ofa.mc_european_parallel(call, m, 100_000, 252, n_chunks=16)
    p_seq, _ = oracle.mc_european_exact(SEQ, oracle.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0,
                                        100_000)
    p_par, _ = oracle.mc_european_exact(PAR, oracle.CALL, 100.0, 100.0, 0.05, 0.0, 0.2, 1.0,
                                        100_000, 16)
    assert abs(seq.price - p_seq) <= 1e-11 * p_seq and abs(par.price - p_par) <= 1e-11 * p_par
    assert par.diagnostics.get("num_threads") == 16.0
    # parallel_test.rs:49-72
    assert abs(par.price - seq.price) / seq.price <= 0.02
    # mc_engine.rs:1170-1194 and 1197-1220
    ar = ofa.mc_european_with_arena(call, m, 200_000, 1)
    assert abs(ar.price - 10.450583572185565) / 10.450583572185565 <= 0.01
    m2 = ofa.Market.builder().spot(95.0).rate(0.02).dividend_yield(0.01).flat_vol(0.3).build()
    put = ofa.VanillaOption.european_put(100.0, 1.5)
    a1 = ofa.mc_european_with_arena(put, m2, 4000, 32)
    a2 = ofa.mc_european_with_arena(put, m2, 4000, 96)
    a3 = ofa.mc_european_with_arena(put, m2, 4000, 32)
    assert math.isfinite(a2.price) and a1.price == a3.price and a1.stderr == a3.stderr
    # the throughput streams stay available
    fast = ofa.mc_european_sequential(call, m, 400_000, rng_mode=ofa.RngMode.Fast)
    assert abs(fast.price - 10.450583572185565) <= 4 * fast.stderr
