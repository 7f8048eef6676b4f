"""The device math header built for the host (CPU only).

Checks the exact functions the kernels run: exp_fp64 against high-precision
exp, the Acklam restatement and the reference-draw stream against the oracle
(bitwise: same operations, and the host build uses the same libm).
"""
import ctypes as C
import os
import subprocess
import tempfile

import mpmath as mp
import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def shim():
    out = os.path.join(tempfile.gettempdir(), f"ofg_math_shim_{os.getpid()}.so")
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off",
           "-fPIC", "-shared", "-o", out, os.path.join(HERE, "cpu_shim", "math_shim.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("host build of ofg_math.hpp failed:\n" + r.stderr)
    L = C.CDLL(out)
    L.shim_exp.argtypes = [C.c_double]; L.shim_exp.restype = C.c_double
    L.shim_acklam.argtypes = [C.c_double]; L.shim_acklam.restype = C.c_double
    L.shim_reference_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_int64,
                                         C.POINTER(C.c_double)]
    L.shim_reference_normals.restype = None
    L.shim_log.argtypes = [C.c_double]; L.shim_log.restype = C.c_double
    L.shim_sqrt.argtypes = [C.c_double]; L.shim_sqrt.restype = C.c_double
    L.shim_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.shim_sincos.restype = None
    L.shim_fast_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.POINTER(C.c_double)]
    L.shim_fast_normals.restype = None
    L.shim_box_muller.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    L.shim_box_muller.restype = None
    yield L
    os.unlink(out)


def test_exp_fp64_accuracy(shim):
    mp.mp.dps = 30
    xs = np.concatenate([np.linspace(-30, 30, 4001), np.linspace(-0.5, 0.5, 2001),
                         np.random.default_rng(3).uniform(-700, 700, 2000)])
    worst = 0.0
    for x in xs:
        got = shim.shim_exp(float(x))
        ref = mp.exp(mp.mpf(float(x)))
        worst = max(worst, float(abs(mp.mpf(got) - ref) / ref))
    assert worst <= 3.5e-16, worst  # within ~1.6 ulp
    assert shim.shim_exp(0.0) == 1.0
    assert shim.shim_exp(-800.0) == 0.0
    assert shim.shim_exp(800.0) == float("inf")


def test_acklam_matches_oracle_bitwise(shim, oracle):
    L = oracle.lib()
    ps = np.concatenate([np.linspace(2.3e-16, 1 - 2.3e-16, 20001),
                         np.geomspace(2.3e-16, 0.03, 2000), 1 - np.geomspace(2.3e-16, 0.03, 2000)])
    for p in ps:
        p = float(L.or_uniform_open01(float(p)))
        assert shim.shim_acklam(p) == L.or_acklam_inv_cdf(p), p


def test_reference_draws_match_oracle_bitwise(shim, oracle):
    for seed, sample in [(42, 0), (42, 49_999), (7, 123), (2**64 - 5, 3)]:
        a = np.zeros(300)
        shim.shim_reference_normals(seed, sample, 300, a.ctypes.data_as(C.POINTER(C.c_double)))
        z1, _ = oracle.sample_normals(seed, sample, 300, 1)
        assert np.array_equal(a, z1)


def test_log_sqrt_sincos_accuracy(shim):
    mp.mp.dps = 30
    rng = np.random.default_rng(11)
    # domain (0, 1): the Box-Muller argument
    xs = np.concatenate([rng.uniform(0, 1, 3000), np.geomspace(2.2e-308, 0.999, 3000),
                         1 - np.geomspace(1.1e-16, 0.5, 1000)])
    worst = 0.0
    for x in xs:
        x = float(x)
        if x >= 1.0:
            continue
        ref = mp.log(mp.mpf(x))
        worst = max(worst, float(abs((mp.mpf(shim.shim_log(x)) - ref) / ref)))
    assert worst <= 4e-16, worst
    worst = 0.0
    for x in np.geomspace(1e-17, 80.0, 3000):
        ref = mp.sqrt(mp.mpf(float(x)))
        worst = max(worst, float(abs((mp.mpf(shim.shim_sqrt(float(x))) - ref) / ref)))
    assert worst <= 2.3e-16, worst
    ws = ws2 = 0.0
    s, c = C.c_double(), C.c_double()
    for t in np.linspace(-np.pi / 4, np.pi / 4, 4001):
        shim.shim_sincos(float(t), C.byref(s), C.byref(c))
        if t != 0:
            ws = max(ws, float(abs((mp.mpf(s.value) - mp.sin(float(t))) / mp.sin(float(t)))))
        ws2 = max(ws2, float(abs((mp.mpf(c.value) - mp.cos(float(t))) / mp.cos(float(t)))))
    assert ws <= 3e-16 and ws2 <= 3e-16, (ws, ws2)


def test_box_muller_extremes_and_symmetries(shim):
    z0, z1 = C.c_double(), C.c_double()
    # largest radius: k = 0 -> u = 2^-53, R = sqrt(106 ln 2)
    shim.shim_box_muller(0, 2**31, C.byref(z0), C.byref(z1))
    r = np.hypot(z0.value, z1.value)
    assert abs(r - np.sqrt(-2 * np.log(2.0**-53))) < 1e-14
    # angle 0 (wt = 2^31) and no symmetry bits -> (R, 0)
    assert abs(z1.value) < 1e-15 and z0.value > 0
    # smallest radius: k = 2^52 - 1 -> u = 1 - 2^-53
    shim.shim_box_muller(((1 << 52) - 1) << 12, 2**31, C.byref(z0), C.byref(z1))
    assert 0 < np.hypot(z0.value, z1.value) < 2e-8
    # the 8 symmetry codes map one angle to 8 distinct points of the circle
    pts = set()
    for code in range(8):
        shim.shim_box_muller((1 << 63) | (code << 8), 2**31 + 2**29, C.byref(z0), C.byref(z1))
        pts.add((round(z0.value, 12), round(z1.value, 12)))
    assert len(pts) == 8


def test_fast_normals_distribution(shim):
    """2^22 fast-mode normals: moments, tails and a KS test against N(0,1)."""
    from scipy import stats
    n4 = 1 << 18
    out = np.zeros(4 * n4 * 4)
    for sample in range(4):
        shim.shim_fast_normals(42, sample, n4,
                               out[sample * 4 * n4:].ctypes.data_as(C.POINTER(C.c_double)))
    z = out
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1) < 5 * np.sqrt(2.0 / n)
    assert abs(stats.skew(z)) < 5 * np.sqrt(6.0 / n)
    assert abs(stats.kurtosis(z)) < 5 * np.sqrt(24.0 / n)
    assert stats.kstest(z, "norm").pvalue > 1e-4
    # tail frequencies beyond 2 and 3 sigma
    for k in (2.0, 3.0):
        p = 2 * stats.norm.sf(k)
        f = np.mean(np.abs(z) > k)
        assert abs(f - p) < 5 * np.sqrt(p * (1 - p) / n)
    # consecutive normals of one stream are uncorrelated
    zz = out[: 4 * n4]
    assert abs(np.corrcoef(zz[:-1], zz[1:])[0, 1]) < 5 / np.sqrt(zz.size)
    assert abs(np.corrcoef(zz[0::2], zz[1::2])[0, 1]) < 5 / np.sqrt(zz.size / 2)
    assert abs(np.corrcoef(zz[0::2] ** 2, zz[1::2] ** 2)[0, 1]) < 5 / np.sqrt(zz.size / 2)


def test_xoshiro_inverse_walks_the_stream_back(shim):
    shim.shim_xoshiro_roundtrip.argtypes = [C.c_uint64, C.c_int, C.c_int64,
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    shim.shim_xoshiro_roundtrip.restype = C.c_int
    for plus in (0, 1):
        for seed in (0, 42, 2**64 - 1, 123456789):
            n = 1000
            f = (C.c_uint64 * n)()
            b = (C.c_uint64 * n)()
            assert shim.shim_xoshiro_roundtrip(seed, plus, n, f, b) == 1
            assert list(b) == list(f)[::-1]
