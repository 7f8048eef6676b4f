"""The device math header built for the host (CPU only).

Checks the exact functions the kernels run: exp_fp64 and the fast mode's
-2 log, sqrt, exp2 and sin/cos against mpmath, the Acklam restatement and the
reference-draw stream against the oracle (bitwise: same operations, and the
host build uses the same libm), the fast normals' distribution and the
xoshiro inverse.
"""
import ctypes as C

import mpmath as mp
import numpy as np
import pytest


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


def test_neg2log_sqrt_accuracy(shim):
    mp.mp.dps = 30
    rng = np.random.default_rng(11)
    # domain (0, 1): the Box-Muller argument, down to the smallest normal
    xs = np.concatenate([rng.uniform(0, 1, 3000), np.geomspace(2.2e-308, 0.999, 3000),
                         1 - np.geomspace(1.1e-16, 0.5, 1000), [2.0**-53, 0.5, 0.25]])
    worst = 0.0
    for x in xs:
        x = float(x)
        if x >= 1.0:
            continue
        ref = -2 * mp.log(mp.mpf(x))
        worst = max(worst, float(abs((mp.mpf(shim.shim_neg2log(x)) - ref) / ref)))
    assert worst <= 3.5e-16, worst
    # every table interval, both ends, at three binary exponents
    for e in (0, -1, -40):
        for i in range(256):
            for m in (0.5 + i / 512, 0.5 + (i + 1) / 512 - 2.0**-53):
                x = m * 2.0**e
                ref = -2 * mp.log(mp.mpf(x))
                assert abs((mp.mpf(shim.shim_neg2log(x)) - ref) / ref) <= 3.5e-16, x
    worst = 0.0
    for x in np.geomspace(1e-17, 80.0, 3000):
        ref = mp.sqrt(mp.mpf(float(x)))
        worst = max(worst, float(abs((mp.mpf(shim.shim_sqrt(float(x))) - ref) / ref)))
    assert worst <= 2.3e-16, worst


def test_exp2_256_accuracy_and_range(shim):
    """2^(y/256) against mpmath; y covers the exponents of path steps, every
    table index with both signs of k, and the saturated ends."""
    mp.mp.dps = 30
    ys = np.concatenate([np.linspace(-30, 30, 4001) * 369.3299304675746,
                         np.arange(-1024, 1024) + 0.5, np.arange(-1024, 1024, dtype=float),
                         np.random.default_rng(4).uniform(-700, 700, 2000) * 369.3299304675746])
    worst = 0.0
    for y in ys:
        got = shim.shim_exp2_256(float(y))
        ref = mp.power(2, mp.mpf(float(y)) / 256)
        worst = max(worst, float(abs(mp.mpf(got) - ref) / ref))
    assert worst <= 2.5e-16, worst
    assert shim.shim_exp2_256(0.0) == 1.0 and shim.shim_exp2_256(256.0) == 2.0
    assert shim.shim_exp2_256(-256.0) == 0.5
    for big in (1e6, 1e9, 2.0**30):
        assert shim.shim_exp2_256(big) == float("inf")
        assert shim.shim_exp2_256(-big) == 0.0


def test_sincos_of_a_turn(shim):
    """(sin, cos) of 2 pi (wt + 1/2) / 2^32 over random words and every
    table-interval boundary."""
    mp.mp.dps = 30
    rng = np.random.default_rng(12)
    wts = list(rng.integers(0, 2**32, 4000))
    for i in range(512):
        for off in (0, 1, (1 << 22) - 1, 1 << 22, (1 << 23) - 1):
            wts.append((i << 23) + off)
    wts += [0, 2**32 - 1, 2**30, 2**31, 3 * 2**30]
    s, c = C.c_double(), C.c_double()
    worst = 0.0
    for wt in wts:
        wt = int(wt)
        shim.shim_sincos_turn(wt, C.byref(s), C.byref(c))
        a = 2 * mp.pi * (mp.mpf(wt) + mp.mpf(1) / 2) / 2**32
        worst = max(worst, float(abs(s.value - mp.sin(a))), float(abs(c.value - mp.cos(a))))
    assert worst <= 1.7e-16, worst  # absolute, 1.5 ulp of 1


def test_box_muller_extremes_and_angles(shim):
    z0, z1 = C.c_double(), C.c_double()
    # largest radius: k = 0 -> u = 2^-53, R = sqrt(106 ln 2)
    rmax = np.sqrt(-2 * np.log(2.0**-53))
    shim.shim_box_muller(0, 0, C.byref(z0), C.byref(z1))
    assert abs(np.hypot(z0.value, z1.value) - rmax) < 1e-14
    # angle of wt = 0 is pi / 2^32: (R, R pi 2^-32), to sincos_turn's absolute bound
    assert abs(z0.value - rmax) < 1e-14 and abs(z1.value - rmax * np.pi / 2**32) < 1.7e-16 * rmax
    # quarter, half and three-quarter turns
    for wt, (cx, cy) in [(2**30, (0, 1)), (2**31, (-1, 0)), (3 * 2**30, (0, -1))]:
        shim.shim_box_muller(0, wt, C.byref(z0), C.byref(z1))
        assert abs(z0.value - cx * rmax) < 1e-8 and abs(z1.value - cy * rmax) < 1e-8
    # smallest radius: k = 2^52 - 1 -> u = 1 - 2^-53
    shim.shim_box_muller(((1 << 52) - 1) << 12, 2**31, C.byref(z0), C.byref(z1))
    assert 0 < np.hypot(z0.value, z1.value) < 2e-8
    # radius of a mid-range u, against mpmath
    k = 123456789012345
    shim.shim_box_muller(k << 12, 12345, C.byref(z0), C.byref(z1))
    u = (mp.mpf(k) + mp.mpf(1) / 2) * mp.mpf(2) ** -52
    assert abs(np.hypot(z0.value, z1.value) / float(mp.sqrt(-2 * mp.log(u))) - 1) < 5e-16


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
    for plus in (0, 1):
        for seed in (0, 42, 2**64 - 1, 123456789):
            n = 1000
            f = (C.c_uint64 * n)()
            b = (C.c_uint64 * n)()
            assert shim.shim_xoshiro_roundtrip(seed, plus, n, f, b) == 1
            assert list(b) == list(f)[::-1]


def test_xoroshiro128plus_unit_state_by_hand(shim):
    """xoroshiro128+ from state (1, 2), derived by hand from the published
    transition (a = 24, b = 16, c = 37): output s0 + s1 = 3; t = 1 ^ 2 = 3,
    s0' = rotl(1, 24) ^ 3 ^ (3 << 16) = 0x1030003, s1' = rotl(3, 37) =
    0x6000000000, so the second output is 0x6001030003."""
    out = (C.c_uint64 * 2)()
    shim.shim_xoroshiro_from_state(1, 2, 2, out)
    assert list(out) == [3, 0x6001030003]
    # from (2^63, 0): t = 2^63; s0' = rotl(2^63, 24) ^ 2^63 ^ 0 = 2^23 ^ 2^63;
    # s1' = rotl(2^63, 37) = 2^36
    shim.shim_xoroshiro_from_state(1 << 63, 0, 2, out)
    assert list(out) == [1 << 63, ((1 << 23) ^ (1 << 63)) + (1 << 36)]


@pytest.mark.parametrize("seed", [0, 1, 13, 7919, 0xDEADBEEFCAFEBABE, 2**64 - 1])
def test_device_pcg64_equals_oracle(shim, oracle, seed):
    """The device Pcg64 (two-word state, mul_hi) against the oracle's 128-bit
    restatement, itself pinned by numpy's PCG64."""
    import ctypes as C
    n = 3000
    out = (C.c_uint64 * n)()
    u = np.zeros(n)
    shim.shim_pcg64(seed, n, out, u.ctypes.data_as(C.POINTER(C.c_double)))
    want = oracle.pcg64_stream(seed, n)
    assert list(out) == want
    assert np.array_equal(u, np.array([(x >> 11) * 2.0 ** -53 for x in want]))
