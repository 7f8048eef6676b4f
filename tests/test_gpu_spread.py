"""ofg_spread_price (SpreadMonteCarloEngine, spread_mc.rs:83-200) on the GPU
against the oracle's restatement on the same streams: every sample's payoff
to 1e-12 (relative to max(|payoff|, s1)), prices to 1e-11 relative, standard
errors to 1e-9, for the xoshiro256++ and Pcg64 kinds with antithetic on and
off; a batch of ragged path counts (1, 2, 3, a wave's 256 samples +- 1,
odd counts under antithetic) with a t = 0 option in it; 2^22 + 3 paths; the
validation messages; the Python mirror and the reference's own test; the
fast generator within 4 SE of Margrabe's closed form."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("s1", "s2", "k", "vol1", "vol2", "rho", "q1", "q2", "r", "t")


def _opt(**kw):
    o = dict(s1=100.0, s2=96.0, k=3.0, vol1=0.2, vol2=0.15, rho=0.5, q1=0.0, q2=0.0, r=0.05,
             t=0.5)
    o.update(kw)
    return o


def _struct(o, paths, seed):
    from openferric_amd import _lib
    s = _lib.OfgSpreadOption()
    for f in FIELDS:
        setattr(s, f, o[f])
    s.num_paths, s.seed = paths, seed
    return s


def _modes():
    from openferric_amd import _lib
    from oracle import oracle as O
    return [(_lib.RNG_REFERENCE, O.RNG_XOSHIRO), (_lib.RNG_PCG64, O.RNG_PCG64)]


def _check(r, pay, want, s1):
    p, se, wpay = want
    assert abs(r.price - p) <= 1e-11 * max(abs(p), 1e-12), (r.price, p)
    assert r.has_stderr and abs(r.stderr_ - se) <= 1e-9 * max(se, 1e-300), (r.stderr_, se)
    if pay is not None:
        err = np.max(np.abs(pay - wpay) / np.maximum(np.abs(wpay), s1))
        assert err <= 1e-12, err


CASES = [(_opt(), 100_003, 13), (_opt(k=0.0, rho=-0.7, q1=0.02, q2=0.04, t=2.0), 77_777, 5),
         (_opt(vol1=0.0, vol2=0.35, k=1.5, rho=1.0), 4099, 9)]


@pytest.mark.parametrize("anti", [True, False])
@pytest.mark.parametrize("mode", [0, 1])
def test_per_sample_parity(oracle, gpu, anti, mode):
    from openferric_amd import spread_price_raw
    lib_mode, or_kind = _modes()[mode]
    structs = [_struct(o, n, s) for o, n, s in CASES]
    res, pay = spread_price_raw(structs, anti, lib_mode, want_payoffs=True)
    for (o, n, seed), r, p in zip(CASES, res, pay):
        want = oracle.spread_mc(o, n, seed, antithetic=anti, rng_kind=or_kind, want_payoffs=True)
        _check(r, p, want, o["s1"])
        assert r.num_paths == n


def test_ragged_batch_with_intrinsic(oracle, gpu):
    from openferric_amd import _lib, spread_price_raw
    counts = [1, 2, 3, 255, 256, 257, 511, 513, 1001]
    opts = [_opt(rho=-0.2 + 0.05 * i, k=0.5 * i) for i in range(len(counts))]
    structs = [_struct(o, n, 100 + i) for i, (o, n) in enumerate(zip(opts, counts))]
    structs.insert(4, _struct(_opt(t=0.0, s1=104.0), 10, 3))
    for anti in (True, False):
        res, pay = spread_price_raw(structs, anti, _lib.RNG_REFERENCE, want_payoffs=True)
        r0 = res[4]
        assert r0.price == 5.0 and r0.stderr_ == 0.0 and r0.num_paths == 0 and pay[4] is None
        rest = [x for i, x in enumerate(zip(res, pay)) if i != 4]
        for i, ((r, p), o, n) in enumerate(zip(rest, opts, counts)):
            want = oracle.spread_mc(o, n, 100 + i, antithetic=anti, want_payoffs=True)
            _check(r, p, want, o["s1"])


def test_many_paths(oracle, gpu):
    from openferric_amd import spread_price_raw
    o = _opt(k=2.0)
    n = (1 << 22) + 3
    for lib_mode, or_kind in _modes():
        r = spread_price_raw([_struct(o, n, 31)], True, lib_mode)[0]
        p, se = oracle.spread_mc(o, n, 31, rng_kind=or_kind)
        assert abs(r.price - p) <= 1e-11 * p and abs(r.stderr_ - se) <= 1e-9 * se


def test_validation_messages(gpu):
    from openferric_amd import InvalidInput, _lib, spread_price_raw
    bad = [(dict(s1=-1.0, k=-1.0), "spread spots s1 and s2 must be > 0"),
           (dict(k=-0.5, vol1=-1.0), "spread strike k must be >= 0"),
           (dict(vol2=-0.1), "spread volatilities vol1 and vol2 must be >= 0"),
           (dict(rho=float("nan")), r"spread correlation rho must be in \[-1, 1\]"),
           (dict(t=-1.0), "spread maturity t must be >= 0")]
    for kw, msg in bad:
        with pytest.raises(InvalidInput, match=msg):
            spread_price_raw([_struct(_opt(), 10, 1), _struct(_opt(**kw), 10, 1)])
    with pytest.raises(InvalidInput, match="spread MC num_paths must be > 0"):
        spread_price_raw([_struct(_opt(), 0, 1)])
    with pytest.raises(InvalidInput, match="OFG_RNG_PCG64"):
        spread_price_raw([_struct(_opt(), 10, 1)], rng_mode=_lib.RNG_EXTERNAL)


def test_mirror_and_reference_test(oracle, gpu):
    """spread_mc.rs:206-241 through the Python mirror, and the binding's
    spread_price with rng kind strings."""
    import openferric_amd as ofa
    ins = ofa.SpreadOption(100.0, 96.0, 3.0, 0.2, 0.15, 0.5, 0.0, 0.0, 0.05, 0.5)
    res = ofa.SpreadMonteCarloEngine(300_000, 13).price(ins)
    kirk = oracle.kirk_spread_price(_opt())
    assert abs((res.price - kirk) / kirk) <= 0.01
    p, se = oracle.spread_mc(_opt(), 300_000, 13)
    assert abs(res.price - p) <= 1e-11 * p and abs(res.stderr - se) <= 1e-9 * se
    assert res.diagnostics.get("num_paths") == 300_000.0 and res.diagnostics.get("rho") == 0.5
    pc = ofa.spread_price(100.0, 96.0, 3.0, 0.2, 0.15, 0.5, 0.0, 0.0, 0.05, 0.5, 50_001, 7,
                          antithetic=False, rng_kind="pcg")
    p, se = oracle.spread_mc(_opt(), 50_001, 7, antithetic=False, rng_kind=oracle.RNG_PCG64)
    assert abs(pc.price - p) <= 1e-11 * p
    z = ofa.SpreadMonteCarloEngine(100, 1).price(
        ofa.SpreadOption(100.0, 90.0, 3.0, 0.2, 0.15, 0.5, 0.0, 0.0, 0.05, 0.0))
    assert z.price == 7.0 and z.stderr == 0.0 and z.diagnostics.get("num_paths") is None
    # randomized streams: independent draws, still within 4 SE of the seeded price
    rnd = ofa.SpreadMonteCarloEngine(300_000, 13).with_randomized_streams().price(ins)
    assert abs(rnd.price - res.price) <= 4 * math.hypot(rnd.stderr, res.stderr)
    with pytest.raises(ofa.InvalidInput, match="StdRng"):
        ofa.SpreadMonteCarloEngine(10, 1).with_rng_kind(ofa.FastRngKind.StdRng).price(ins)


def test_fast_generator_against_margrabe(oracle, gpu):
    from openferric_amd import RngMode, SpreadMonteCarloEngine, SpreadOption
    for o in (_opt(k=0.0, s2=105.0, q1=0.04, q2=0.06, t=1.0),
              _opt(k=0.0, s1=110.0, s2=100.0, vol1=0.22, vol2=0.18, rho=-0.6, q1=0.01, q2=0.02,
                   r=0.03, t=1.2)):
        ins = SpreadOption(*(o[f] for f in FIELDS))
        r = SpreadMonteCarloEngine(2_000_000, 3).with_rng_mode(RngMode.Fast).price(ins)
        m = oracle.margrabe_exchange_price(o)
        assert abs(r.price - m) <= 4 * r.stderr, (r.price, m, r.stderr)
