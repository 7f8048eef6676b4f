import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: long-running statistical check")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


def _load_math_shim():
    import ctypes as C
    path = os.path.join(ROOT, "tests", "cpu_shim", "libofg_math_shim.so")
    srcs = [os.path.join(ROOT, "tests", "cpu_shim", "math_shim.hip")] + [
        os.path.join(ROOT, "openferric_amd", "csrc", f)
        for f in ("ofg_math.hpp", "ofg_fast_tables.inc")]
    if not os.path.exists(path) or any(os.path.getmtime(s) > os.path.getmtime(path)
                                       for s in srcs if os.path.exists(s)):
        from openferric_amd import build as b
        b.build_math_shim()
    L = C.CDLL(path)
    d, u64, i64, pd = C.c_double, C.c_uint64, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "shim_exp": ([d], d), "shim_acklam": ([d], d),
        "shim_reference_normals": ([u64, u64, i64, pd], None),
        "shim_neg2log": ([d], d), "shim_sqrt": ([d], d), "shim_exp2_256": ([d], d),
        "shim_sincos_turn": ([C.c_uint32, pd, pd], None),
        "shim_fast_normals": ([u64, u64, i64, pd], None),
        "shim_box_muller": ([u64, C.c_uint32, pd, pd], None),
        "shim_fast_gbm_terminal": ([u64, u64, d, d, d, d, C.c_int], d),
        "shim_fast_heston_terminal": ([u64, u64, d, d, d, d, d, d, d, d, C.c_int], d),
        "shim_xoshiro_roundtrip": ([u64, C.c_int, i64, C.POINTER(u64), C.POINTER(u64)],
                                   C.c_int),
        "shim_xoroshiro_from_state": ([u64, u64, i64, C.POINTER(u64)], None),
        "shim_pcg64": ([u64, i64, C.POINTER(u64), pd], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes, f.restype = args, res
    return L


@pytest.fixture(scope="session")
def shim():
    """Host-only build of the device math header (tests/cpu_shim)."""
    return _load_math_shim()


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Loads the HIP library; GPU tests fail loudly when it is missing."""
    import openferric_amd as ofa
    from openferric_amd import _lib
    _lib.load()
    _lib.context(0)
    return ofa
