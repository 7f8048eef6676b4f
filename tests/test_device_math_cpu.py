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
