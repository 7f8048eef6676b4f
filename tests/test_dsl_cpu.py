"""Structured-product language on the CPU: the oracle (oracle/dsl_oracle.py)
pinned by closed forms, by the independently pinned autocallable oracle on
the same normals and by the reference's DSL unit tests; the library's front
end (ofg_dsl_compile, host only) against the reference's lexer / parser /
compiler behaviour and against the oracle's own front end."""
import glob
import math
import os

import numpy as np
import pytest
from scipy.stats import multivariate_normal, norm

from oracle import dsl_oracle as D
from tests import dsl_products as P
from tests.basket_numpy import bs_price

EQ = lambda s, v, q=0.0: dict(kind="equity", spot=s, vol=v, dividend_yield=q)


def _z(paths, steps, n, seed):
    return np.random.default_rng(seed).standard_normal((paths, steps, n))


# ---- oracle pins: closed forms ---------------------------------------------

def test_zero_coupon_is_discounted_notional():
    p, se = D.price(P.ZERO_COUPON, [EQ(100, 0.3)], [[1.0]], 0.04, 500, 24, _z(500, 24, 1, 1))
    # stderr: the reference's sum-of-squares close leaves cancellation noise only
    assert abs(p - 1e6 * math.exp(-0.08)) < 1e-6 and se < 1e-7 * p


def test_forward_and_call_match_closed_forms():
    z = _z(40_000, 3, 1, 2)
    p, se = D.price(P.FORWARD, [EQ(100, 0.25, 0.02)], [[1.0]], 0.03, 40_000, 3, z)
    assert abs(p - 100 * math.exp(-0.02 * 1.5)) < 4 * se
    z = _z(40_000, 4, 1, 3)
    p, se = D.price(P.CALL, [EQ(100, 0.2, 0.01)], [[1.0]], 0.03, 40_000, 4, z)
    assert abs(p - bs_price(True, 100, 100, 0.03, 0.01, 0.2, 1.0)) < 4 * se


@pytest.mark.parametrize("rho", [-0.4, 0.7])
def test_worst_of_digital_matches_bivariate_normal(rho):
    s1, s2, q1, q2, r = 0.25, 0.35, 0.01, 0.03, 0.02
    z = _z(40_000, 2, 2, 4)
    p, se = D.price(P.WORST_DIGITAL, [EQ(100, s1, q1), EQ(80, s2, q2)],
                    [[1.0, rho], [rho, 1.0]], r, 40_000, 2, z)
    d = [(r - q - 0.5 * s * s) / s for s, q in ((s1, q1), (s2, q2))]
    prob = multivariate_normal(mean=[0, 0], cov=[[1, rho], [rho, 1]]).cdf(d)
    assert abs(p - 100 * math.exp(-r) * prob) < 4 * se, (p, 100 * math.exp(-r) * prob, se)


def test_vasicek_and_schwartz_means():
    a, th, r0, vol = 0.8, 0.05, 0.02, 0.01
    rate = dict(kind="rate", spot=r0, vol=vol, mean_reversion=a, long_run_mean=th)
    p, se = D.price(P.TERMINAL_VALUE, [rate], [[1.0]], 0.0, 20_000, 8, _z(20_000, 8, 1, 5))
    assert abs(p - (th + (r0 - th) * math.exp(-a * 2.0))) < 4 * se
    k, mu, s0 = 1.2, math.log(70.0), 90.0
    com = dict(kind="commodity", spot=s0, vol=0.3, convenience_yield=0.0, kappa=k, mu=mu)
    p, se = D.price(P.TERMINAL_LOG, [com], [[1.0]], 0.0, 20_000, 8, _z(20_000, 8, 1, 6))
    assert abs(p - (mu + (math.log(s0) - mu) * math.exp(-k * 2.0))) < 4 * se


# ---- oracle pins: the autocallable oracle on the same normals ---------------

@pytest.mark.parametrize("n_und,barrier,coupon,maturity,steps", [
    (1, 1.0, 0.08, 1.0, 8), (3, 1.02, 0.1, 1.5, 12), (2, 0.95, 0.06, 2.0, 24)])
def test_standard_autocall_equals_autocallable_oracle(oracle, n_und, barrier, coupon, maturity,
                                                      steps):
    rng = np.random.default_rng(n_und)
    spots, vols = list(rng.uniform(80, 120, n_und)), list(rng.uniform(0.15, 0.35, n_und))
    c = np.full((n_und, n_und), 0.5)
    np.fill_diagonal(c, 1.0)
    paths, r, q = 3_000, 0.03, 0.0
    z = _z(paths, steps, n_und, 10 + n_und)
    src = P.standard_autocall(n_und, barrier, coupon, maturity)
    p, se = D.price(src, [EQ(s, v, q) for s, v in zip(spots, vols)], c.tolist(), r, paths, steps,
                    z)
    dates = D.compile_source(src).schedules[0][0]
    a = oracle.make_autocall(spots=spots, vols=vols, corr=c, dates=dates, notional=100.0,
                             maturity=maturity, autocall_barrier=barrier, coupon_rate=coupon,
                             ki_barrier=1e-9, ki_strike=1.0, rate=r, div_yield=q,
                             num_paths=paths, num_steps=steps)
    pa, sa = oracle.autocall_price(a, z)
    assert abs(p - pa) < 1e-10 * pa and abs(se - sa) < 1e-6 * sa


def test_phoenix_memory_equals_autocallable_oracle(oracle):
    n_und, paths, steps = 2, 3_000, 16
    spots, vols = [100.0, 90.0], [0.3, 0.25]
    c = [[1.0, 0.3], [0.3, 1.0]]
    z = _z(paths, steps, n_und, 21)
    src = P.phoenix_memory(n_und, 1.05, 0.8, 0.1, 1.0)
    p, se = D.price(src, [EQ(s, v) for s, v in zip(spots, vols)], c, 0.02, paths, steps, z)
    a = oracle.make_autocall(spots=spots, vols=vols, corr=c, dates=[0.25, 0.5, 0.75, 1.0],
                             notional=100.0, maturity=1.0, autocall_barrier=1.05,
                             coupon_rate=0.1, coupon_barrier=0.8, ki_barrier=1e-9,
                             ki_strike=1.0, rate=0.02, phoenix=True, memory=True,
                             num_paths=paths, num_steps=steps)
    pa, sa = oracle.autocall_price(a, z)
    assert abs(p - pa) < 1e-10 * pa and abs(se - sa) < 1e-6 * sa


def test_single_stream_is_path_step_asset_major(oracle):
    src = P.standard_autocall(2, 1.0, 0.05, 1.0)
    mk = [EQ(100, 0.2), EQ(100, 0.3)]
    c = [[1.0, 0.2], [0.2, 1.0]]
    z = oracle.lsm_stream_normals(42, 300 * 8 * 2).reshape(300, 8, 2)
    assert D.price(src, mk, c, 0.01, 300, 8, None, seed=42) == D.price(src, mk, c, 0.01, 300, 8, z)


# ---- the reference's own DSL tests (restated assertions) ---------------------

def test_reference_evaluator_unit_cases():
    """eval.rs:2171-2320: early redemption above the barrier, knock-in with
    final worst below one, no knock-in returns notional, worst-of uses the
    minimum performance -- on hand-built snapshots."""
    src = """\
product "T"
    notional: 100
    maturity: 1.0
    underlyings
        A = asset(0)
        B = asset(1)
    state
        ki: bool = false
    schedule semi_annual from 0.5 to 1.0
        let wof = worst_of(performances())
        if wof <= 0.6 then
            set ki = true
        if wof >= 1.0 and not is_final then
            pay notional * 0.05
            redeem notional
        if is_final then
            if ki and wof < 1.0 then
                redeem notional * wof
            else
                redeem notional
"""
    prod = D.compile_source(src)
    s0 = [100.0, 100.0]
    ev = lambda snaps: D.evaluate(prod, 2, 0.0, snaps, s0, 1)
    assert ev([[110.0, 105.0], [90.0, 90.0]]) == pytest.approx(105.0)     # called at 0.5
    assert ev([[55.0, 120.0], [80.0, 120.0]]) == pytest.approx(80.0)      # ki, final wof 0.8
    assert ev([[95.0, 99.0], [90.0, 97.0]]) == pytest.approx(100.0)       # no ki
    assert ev([[120.0, 95.0], [70.0, 150.0]]) == pytest.approx(100.0)     # min perf 0.95 < 1


def test_reference_engine_cases():
    """mod.rs / engine.rs tests: the forward product prices near the
    discounted notional, a two-asset product sees its correlation."""
    p, _ = D.price(P.ZERO_COUPON.replace("maturity: 2.0", "maturity: 1.0")
                   .replace("annual from 2.0 to 2.0", "annual from 1.0 to 1.0"),
                   [EQ(100, 0.2)], [[1.0]], 0.05, 2000, 100, _z(2000, 100, 1, 7))
    assert abs(p - 1e6 * math.exp(-0.05)) < 1.0
    lo, _ = D.price(P.WORST_DIGITAL, [EQ(100, 0.3), EQ(100, 0.3)], [[1, -0.5], [-0.5, 1]], 0.0,
                    8000, 4, _z(8000, 4, 2, 8))
    hi, _ = D.price(P.WORST_DIGITAL, [EQ(100, 0.3), EQ(100, 0.3)], [[1, 0.9], [0.9, 1]], 0.0,
                    8000, 4, _z(8000, 4, 2, 8))
    assert hi > lo + 10


def test_oracle_errors_match_reference_messages():
    with pytest.raises(D.PricingError, match="asset spot must be > 0"):
        D.price(P.CALL, [EQ(-1, 0.2)], [[1.0]], 0.0, 10, 2)
    with pytest.raises(D.PricingError, match="out of range"):
        D.price(P.OUT_OF_RANGE_PRICE, [EQ(100, 0.2)], [[1.0]], 0.0, 10, 2, _z(10, 2, 1, 0))
    assert D.price(P.MASKED_BAD_INDEX, [EQ(100, 1e-9)], [[1.0]], 0.0, 10, 2,
                   _z(10, 2, 1, 0))[0] == pytest.approx(100.0)
    with pytest.raises(D.PricingError, match="standalone"):
        D.price(P.CALL.replace("max(price(0) - 100, 0)", "performances()"), [EQ(100, 0.2)],
                [[1.0]], 0.0, 10, 2)


# ---- the library's front end (host only) -------------------------------------

def _lib_compile(src):
    from openferric_amd import parse_and_compile
    return parse_and_compile(src)


@pytest.mark.parametrize("src", [P.ZERO_COUPON, P.FORWARD, P.CALL, P.WORST_DIGITAL, P.KITCHEN_SINK,
                                 P.RANGE_ACCRUAL, P.standard_autocall(3, 1.0, 0.08, 1.5),
                                 P.phoenix_memory(2, 1.0, 0.7, 0.08, 2.0)])
def test_front_end_matches_oracle_front_end(src):
    p = _lib_compile(src)
    q = D.compile_source(src)
    assert (p.name, p.notional, p.maturity) == (q.name, q.notional, q.maturity)
    assert p.num_schedules == len(q.schedules) and p.num_state_vars == len(q.state)
    for k in range(p.num_schedules):
        assert p.schedule_dates(k) == q.schedules[k][0]
    assert p.max_local_slots() == max([s[2] for s in q.schedules] + [0])


HDR = 'product "x"\n    notional: 1\n    maturity: 1\n    schedule annual from 1 to 1\n'


def _at(src, needle, length=None):
    """The reference's "at start-end" span of the first `needle` in src."""
    i = src.index(needle)
    return f"{i}-{i + (len(needle) if length is None else length)}"


_DUP = 'product "x"\n    maturity: 1\n    notional: 1\n    notional: 2\n'
_UNDEF = HDR + "        pay foo\n"
_UNKNOWN = HDR + "        pay bar(1)\n"
_DOLLAR = 'product "x"\n    notional: 1 $\n'
_NOTNUM = 'product "x"\n    notional: x\n'
_NOCOLON = 'product "x"\n    notional 1\n'


@pytest.mark.parametrize("src,msg", [
    ('product "x"\n    notional: 1\n', "compile error: missing maturity declaration"),
    (_DUP, f"compile error at {_DUP.rindex('notional')}-{_DUP.rindex('2') + 1}: duplicate "
           "notional declaration"),
    (_UNDEF, f"compile error at {_at(_UNDEF, 'foo')}: undefined variable 'foo'"),
    (_UNKNOWN, f"compile error at {_at(_UNKNOWN, 'bar(1)')}: unknown function 'bar'"),
    (HDR + "        set s = 1\n", "undefined state variable 's'"),
    (_DOLLAR, f"lex error at {_at(_DOLLAR, '$')}: unexpected character '$'"),
    ('product "x"\n    notional: 1\n  maturity: 1\n', "inconsistent indentation"),
    ('product "x\n', "unterminated string literal"),
    ('product "x"\n    notional: 1e\n', "invalid number literal '1e'"),
    (_NOCOLON, f"parse error at {_at(_NOCOLON, '1')}: expected Colon, got Number(1.0)"),
    ('product "x"\n    schedule foo from 1 to 2\n', "expected schedule frequency"),
    (HDR + "        pay (1\n", "expected RParen, got Dedent"),
    (HDR + "        foo\n",
     'expected statement (let, if, pay, redeem, set, skip), got Some(Ident("foo"))'),
    ('product "x"\n    notional: 1\n    maturity: 1\n    state\n        a: float = x\n',
     "cannot use Ident(\"x\") as initial value for type 'float'"),
])
def test_front_end_errors(src, msg):
    from openferric_amd import InvalidInput
    with pytest.raises(InvalidInput) as e:
        _lib_compile(src)
    assert msg in str(e.value), (str(e.value), msg)


def test_front_end_error_spans_are_byte_offsets():
    from openferric_amd import InvalidInput
    with pytest.raises(InvalidInput) as e:
        _lib_compile(_NOTNUM)
    i = _NOTNUM.index(": x") + 2
    assert f'parse error at {i}-{i + 1}: expected number, got Ident("x")' in str(e.value)


def test_reference_example_files_compile():
    """tests/dsl_examples.rs compiles every examples/dsl/*.of file; both front
    ends accept all of them with identical schedules (read in place, when the
    reference tree is present)."""
    files = sorted(glob.glob("/root/reference/examples/dsl/*.of"))
    if not files:
        pytest.skip("reference examples not present")
    for f in files:
        src = open(f).read()
        p, q = _lib_compile(src), D.compile_source(src)
        assert [p.schedule_dates(k) for k in range(p.num_schedules)] == [s[0] for s in q.schedules]
        assert p.num_state_vars == len(q.state), f


# ---- native evaluation kernels: the generated source compiles (host only) ----

def _hiprtc():
    import ctypes as C
    # no on-disk code cache: the compiles of this test leave nothing behind
    os.environ["AMD_COMGR_CACHE"] = "0"
    path = "/opt/rocm/lib/libhiprtc.so"
    if not os.path.exists(path):
        pytest.skip("hipRTC is not installed")
    L = C.CDLL(path)
    L.hiprtcCreateProgram.argtypes = [C.POINTER(C.c_void_p), C.c_char_p, C.c_char_p, C.c_int,
                                      C.c_void_p, C.c_void_p]
    L.hiprtcCompileProgram.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p)]
    L.hiprtcGetProgramLogSize.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    L.hiprtcGetProgramLog.argtypes = [C.c_void_p, C.c_char_p]
    L.hiprtcGetCodeSize.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    L.hiprtcDestroyProgram.argtypes = [C.POINTER(C.c_void_p)]
    return L


def test_native_kernel_sources_compile_for_gfx950():
    """The source ofg_dsl_jit_compile would compile, for corpus products of
    1 to 4 underlyings: it compiles with hipRTC for gfx950 under the library's
    flags (no device needed), names the kernel, writes every constant as an
    exact hexadecimal literal and resolves the registers to named locals."""
    import ctypes as C
    from openferric_amd import parse_and_compile
    L = _hiprtc()
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-ffp-contract=off",
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    products = [(P.CALL, 1), (P.WORST_DIGITAL, 2), (P.KITCHEN_SINK, 4), (P.RANGE_ACCRUAL, 3),
                (P.standard_autocall(3, 1.0, 0.08, 1.5), 3),
                (P.phoenix_memory(4, 1.05, 0.75, 0.1, 2.0), 4)]
    for src, n in products:
        prod = parse_and_compile(src)
        text = prod.jit_source(n)
        assert 'extern "C" __global__' in text and "ofg_dsl_eval_jit" in text
        assert "double r0 = 0.0;" in text and "a.code" not in text
        h = C.c_void_p()
        assert L.hiprtcCreateProgram(C.byref(h), text.encode(), b"k.hip", 0, None, None) == 0
        rc = L.hiprtcCompileProgram(h, len(opts), (C.c_char_p * len(opts))(*opts))
        size = C.c_size_t(0)
        L.hiprtcGetProgramLogSize(h, C.byref(size))
        log = C.create_string_buffer(size.value + 1)
        L.hiprtcGetProgramLog(h, log)
        code = C.c_size_t(0)
        L.hiprtcGetCodeSize(h, C.byref(code))
        L.hiprtcDestroyProgram(C.byref(h))
        assert rc == 0, log.value.decode()[:2000]
        assert code.value > 1000
    # constants are exact: every hexadecimal literal of the source reads back
    # as one of the product's constants
    import re
    text = parse_and_compile(P.standard_autocall(3, 1.0, 0.08, 1.5)).jit_source(3)
    assert float.hex(0.08) in text
    src = P.CALL.replace("max(price(0) - 100, 0)", "0.1 + 0.3 * 123.456 - 1_000_000 / 7")
    text = parse_and_compile(src).jit_source(1)
    lits = {float.fromhex(m) for m in re.findall(r"-?0x[01]\.[0-9a-f]{13}p[+-]\d+", text)}
    assert {0.1, 0.3, 123.456, 1_000_000.0, 7.0} <= lits, lits
    with pytest.raises(Exception, match="n_assets must be in 1..16"):
        parse_and_compile(P.CALL).jit_source(17)
