"""Failure paths of the C ABI that need no GPU: every entry point refuses
null arguments with OFG_INVALID_INPUT and a message (no crash, nothing
launched), a context cannot be created without a HIP device and says why,
and the allocation-failure hook is callable. The device side of the
failure paths (an allocation that fails mid-call) is tests/test_gpu_failure.py."""
import ctypes as C

import pytest

OFG_OK, OFG_INVALID_INPUT, OFG_DEVICE_ERROR = 0, 1, 5

# entry points returning a status whose first argument is a context, a
# batch, a product or a host array; called with every argument NULL
STATUS_ENTRY_POINTS = [
    "ofg_price", "ofg_sample_payoffs", "ofg_price_greeks", "ofg_price_noalloc",
    "ofg_finite_difference_greeks", "ofg_bump_and_reprice", "ofg_mc_greeks_estimate",
    "ofg_european_exact", "ofg_spread_price", "ofg_autotune", "ofg_batch_run",
    "ofg_batch_results", "ofg_lsm_price", "ofg_lsm_last_timing", "ofg_qmc_price",
    "ofg_autocall_price", "ofg_basket_price", "ofg_repair_correlation",
    "ofg_stress_correlation", "ofg_dsl_compile", "ofg_dsl_get_info", "ofg_dsl_schedule_dates",
    "ofg_dsl_price", "ofg_stream_jump", "ofg_fill_normals", "ofg_price_device",
    "ofg_dsl_jit_compile", "ofg_dsl_jit_stats", "ofg_dsl_jit_source",
]


@pytest.fixture(scope="module")
def raw():
    """A fresh ctypes handle (no argtypes): the raw C calling convention."""
    from openferric_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    L.ofg_last_error.restype = C.c_char_p
    return L


@pytest.mark.parametrize("name", STATUS_ENTRY_POINTS)
def test_null_arguments_are_invalid_input(raw, name):
    f = getattr(raw, name)
    f.restype = C.c_int32
    rc = f(*([C.c_void_p(None)] * 9))
    assert rc == OFG_INVALID_INPUT, (name, rc)
    assert raw.ofg_last_error() in (b"null argument", b"null batch")


def test_handle_returning_entry_points_give_null(raw):
    raw.ofg_batch_create.restype = C.c_void_p
    assert raw.ofg_batch_create(*([C.c_void_p(None)] * 5)) is None
    assert raw.ofg_last_error() == b"null argument"
    # destructors and accessors accept NULL
    raw.ofg_batch_pricing_kernel_ms.restype = C.c_double
    assert raw.ofg_batch_pricing_kernel_ms(C.c_void_p(None)) == 0.0
    raw.ofg_batch_destroy(C.c_void_p(None))
    raw.ofg_dsl_free(C.c_void_p(None))
    raw.ofg_destroy(C.c_void_p(None))


def test_create_without_a_device_reports_the_hip_error(raw):
    raw.ofg_create.restype = C.c_void_p
    ctx = raw.ofg_create(0)
    if ctx is not None:  # a HIP device is visible here: nothing to check
        raw.ofg_destroy(C.c_void_p(ctx))
        pytest.skip("a HIP device is present")
    msg = raw.ofg_last_error()
    assert msg.startswith(b"hipGetDeviceCount") or msg == b"no such HIP device", msg
    # a status call after the failure starts with a clean message
    f = raw.ofg_price
    f.restype = C.c_int32
    assert f(*([C.c_void_p(None)] * 5)) == OFG_INVALID_INPUT
    assert raw.ofg_last_error() == b"null argument"


def test_out_of_range_device_is_refused(raw):
    raw.ofg_create.restype = C.c_void_p
    assert raw.ofg_create(-1) is None
    assert raw.ofg_last_error() != b""


def test_allocation_failure_hook_needs_no_gpu(raw):
    raw.ofg_inject_alloc_failure.restype = None
    raw.ofg_inject_alloc_failure(C.c_int32(3))
    raw.ofg_inject_alloc_failure(C.c_int32(0))   # disarmed again
    raw.ofg_inject_alloc_failure(C.c_int32(-5))  # negative = disarmed
