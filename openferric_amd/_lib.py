"""ctypes binding of libopenferric_amd.so (the C ABI in include/openferric_amd.h).

The product path has no CPU fallback: if the HIP library is missing or no GPU
is present, calls raise DeviceError.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopenferric_amd.so")

OFG_OK = 0
OFG_INVALID_INPUT = 1
OFG_CONVERGENCE_FAILURE = 2
OFG_MARKET_DATA_MISSING = 3
OFG_NUMERICAL_ERROR = 4
OFG_DEVICE_ERROR = 5

MODEL_GBM, MODEL_HESTON = 0, 1
PAYOFF_VANILLA, PAYOFF_ASIAN, PAYOFF_BARRIER = 0, 1, 2
CALL, PUT = 0, 1
ARITHMETIC, GEOMETRIC = 0, 1
FIXED, FLOATING = 0, 1
UP, DOWN = 0, 1
IN, OUT = 0, 1
RNG_REFERENCE, RNG_EXTERNAL, RNG_FAST, RNG_PCG64 = 0, 1, 2, 3
GREEKS_NONE, GREEKS_ADJOINT, GREEKS_FD_CRN, GREEKS_FORWARD = 0, 1, 2, 3
EXACT_SEQUENTIAL, EXACT_PARALLEL, EXACT_ARENA = 0, 1, 2
LSM_VANILLA, LSM_BERMUDAN, LSM_BARRIER = 0, 1, 2
EUROPEAN, AMERICAN, BERMUDAN = 0, 1, 2
LSM_GBM, LSM_LOCAL_VOL, LSM_HESTON = 0, 1, 2
BASKET_AVERAGE, BASKET_BEST_OF, BASKET_WORST_OF = 0, 1, 2

EXPORTED_SYMBOLS = [
    "ofg_create", "ofg_destroy", "ofg_last_error", "ofg_version", "ofg_price",
    "ofg_sample_payoffs", "ofg_price_greeks", "ofg_batch_create", "ofg_batch_run",
    "ofg_batch_results", "ofg_batch_pricing_kernel_ms", "ofg_batch_destroy", "ofg_autotune",
    "ofg_mc_greeks_estimate", "ofg_lsm_price", "ofg_stream_jump", "ofg_lsm_last_timing",
    "ofg_qmc_price", "ofg_autocall_price", "ofg_basket_price", "ofg_repair_correlation", "ofg_stress_correlation",
    "ofg_dsl_compile", "ofg_dsl_free", "ofg_dsl_get_info", "ofg_dsl_schedule_dates",
    "ofg_dsl_price", "ofg_price_noalloc", "ofg_finite_difference_greeks",
    "ofg_bump_and_reprice", "ofg_european_exact", "ofg_spread_price",
    "ofg_inject_alloc_failure", "ofg_fill_normals", "ofg_price_device",
    "ofg_dsl_jit_compile", "ofg_dsl_jit_stats", "ofg_dsl_jit_source",
]


class OfgOption(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("payoff", C.c_int32), ("option_type", C.c_int32),
        ("averaging", C.c_int32), ("strike_type", C.c_int32), ("barrier_dir", C.c_int32),
        ("barrier_style", C.c_int32), ("n_obs", C.c_int32),
        ("spot", C.c_double), ("strike", C.c_double), ("expiry", C.c_double),
        ("rate", C.c_double), ("div_yield", C.c_double), ("vol", C.c_double),
        ("kappa", C.c_double), ("theta", C.c_double), ("xi", C.c_double),
        ("rho", C.c_double), ("v0", C.c_double),
        ("barrier_level", C.c_double), ("rebate", C.c_double),
        ("obs_times", C.POINTER(C.c_double)),
        ("num_paths", C.c_int64), ("num_steps", C.c_int64), ("seed", C.c_uint64),
    ]


class OfgEngineCfg(C.Structure):
    _fields_ = [
        ("antithetic", C.c_int32), ("control_variate", C.c_int32),
        ("rng_mode", C.c_int32), ("greek_mode", C.c_int32),
        ("bump_rel", C.c_double),
        ("normals1", C.POINTER(C.c_double)), ("normals2", C.POINTER(C.c_double)),
        ("tune_block", C.c_int32), ("tune_item_steps", C.c_int32),
    ]


class OfgResult(C.Structure):
    _fields_ = [
        ("price", C.c_double), ("stderr_", C.c_double),
        ("has_stderr", C.c_int32), ("has_greeks", C.c_int32),
        ("delta", C.c_double), ("gamma", C.c_double), ("vega", C.c_double),
        ("theta", C.c_double), ("rho", C.c_double),
        ("heston_grad", C.c_double * 6),
        ("num_paths", C.c_double), ("num_steps", C.c_double), ("vol", C.c_double),
        ("kernel_ms", C.c_double),
    ]


class OfgDeviceIo(C.Structure):
    """ofg_device_io: device addresses (int or None) of the inputs and outputs."""
    _fields_ = [(k, C.c_void_p) for k in ("spot", "strike", "expiry", "rate", "vol", "v0", "kappa",
                                          "theta", "xi", "rho", "price", "stderr_", "grads")]


class OfgFdBumps(C.Structure):
    _fields_ = [("ds", C.c_double), ("dr", C.c_double), ("dv", C.c_double), ("dt", C.c_double)]


class OfgFdGreeks(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("delta", "gamma", "vega", "theta", "rho", "vanna",
                                          "volga", "kernel_ms")]


class OfgSpreadOption(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("s1", "s2", "k", "vol1", "vol2", "rho", "q1", "q2",
                                          "r", "t")] + [("num_paths", C.c_int64),
                                                        ("seed", C.c_uint64)]


class OfgMcGreeks(C.Structure):
    _fields_ = [
        ("pathwise_delta", C.c_double), ("pathwise_gamma", C.c_double),
        ("lr_delta", C.c_double), ("lr_gamma", C.c_double), ("lr_vega", C.c_double),
        ("num_samples", C.c_double), ("kernel_ms", C.c_double),
    ]


class OfgLsmOption(C.Structure):
    _fields_ = [
        ("instrument", C.c_int32), ("option_type", C.c_int32), ("exercise", C.c_int32),
        ("dynamics", C.c_int32), ("n_dates", C.c_int32), ("barrier_dir", C.c_int32),
        ("barrier_style", C.c_int32), ("pad", C.c_int32),
        ("spot", C.c_double), ("strike", C.c_double), ("expiry", C.c_double),
        ("rate", C.c_double), ("div_yield", C.c_double), ("vol", C.c_double),
        ("kappa", C.c_double), ("theta", C.c_double), ("xi", C.c_double),
        ("rho", C.c_double), ("v0", C.c_double),
        ("barrier_level", C.c_double), ("rebate", C.c_double),
        ("dates", C.POINTER(C.c_double)), ("strikes", C.POINTER(C.c_double)),
        ("num_paths", C.c_int64), ("num_steps", C.c_int64), ("seed", C.c_uint64),
    ]


class OfgLsmBoundary(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("strike", C.c_double), ("boundary_spot", C.c_double),
        ("has_boundary", C.c_int32), ("pad", C.c_int32),
        ("itm_paths", C.c_int64), ("exercised_paths", C.c_int64),
    ]


class OfgAutocall(C.Structure):
    _fields_ = [
        ("phoenix", C.c_int32), ("memory", C.c_int32), ("n_underlyings", C.c_int32),
        ("n_dates", C.c_int32), ("underlyings", C.POINTER(C.c_int32)),
        ("autocall_dates", C.POINTER(C.c_double)),
        ("notional", C.c_double), ("maturity", C.c_double), ("autocall_barrier", C.c_double),
        ("coupon_rate", C.c_double), ("coupon_barrier", C.c_double),
        ("ki_barrier", C.c_double), ("ki_strike", C.c_double),
        ("n_assets", C.c_int32), ("pad", C.c_int32),
        ("spots", C.POINTER(C.c_double)), ("vols", C.POINTER(C.c_double)),
        ("corr", C.POINTER(C.c_double)), ("rate", C.c_double), ("div_yield", C.c_double),
        ("num_paths", C.c_int64), ("num_steps", C.c_int64),
    ]


class OfgAutocallSens(C.Structure):
    _fields_ = [
        ("delta", C.c_double * 8), ("vega", C.c_double), ("cega", C.c_double),
        ("n_delta", C.c_int32), ("ok", C.c_int32),
    ]


class OfgBasket(C.Structure):
    _fields_ = [
        ("basket_type", C.c_int32), ("is_call", C.c_int32), ("n_weights", C.c_int32),
        ("n_assets", C.c_int32), ("weights", C.POINTER(C.c_double)),
        ("strike", C.c_double), ("maturity", C.c_double),
        ("spots", C.POINTER(C.c_double)), ("vols", C.POINTER(C.c_double)),
        ("dividends", C.POINTER(C.c_double)), ("corr", C.POINTER(C.c_double)),
        ("rate", C.c_double), ("num_paths", C.c_int64), ("seed", C.c_uint64),
        ("product", C.c_int32), ("leader_index", C.c_int32), ("fx_rate", C.c_double),
        ("fx_vol", C.c_double), ("domestic_rate", C.c_double), ("foreign_rate", C.c_double),
        ("asset_fx_corr", C.POINTER(C.c_double)), ("n_asset_fx_corr", C.c_int32),
        ("factor_kind", C.c_int32), ("n_loading_rows", C.c_int32), ("n_factors", C.c_int32),
        ("loadings", C.POINTER(C.c_double)), ("copula", C.c_int32), ("copula_dof", C.c_int32),
    ]


COPULA_GAUSSIAN, COPULA_STUDENT_T = 0, 1


class OfgCorrStress(C.Structure):
    _fields_ = [("kind", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("pad", C.c_int32),
                ("value", C.c_double)]


(STRESS_SCALE_OFF_DIAGONAL, STRESS_ADDITIVE_SHIFT, STRESS_FLOOR_OFF_DIAGONAL,
 STRESS_CAP_OFF_DIAGONAL, STRESS_OVERRIDE_PAIR) = range(5)


class OfgBasketSens(C.Structure):
    _fields_ = [
        ("delta", C.c_double * 16), ("vega", C.c_double), ("cega", C.c_double),
        ("rho", C.c_double), ("n_delta", C.c_int32), ("ok", C.c_int32),
        ("repaired", C.c_int32), ("pad", C.c_int32),
    ]


DSL_EQUITY, DSL_FX, DSL_COMMODITY, DSL_RATE = 0, 1, 2, 3


class OfgDslAsset(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("pad", C.c_int32), ("spot", C.c_double), ("vol", C.c_double),
        ("dividend_yield", C.c_double), ("domestic_rate", C.c_double),
        ("foreign_rate", C.c_double), ("convenience_yield", C.c_double), ("kappa", C.c_double),
        ("mu", C.c_double), ("mean_reversion", C.c_double), ("long_run_mean", C.c_double),
    ]


class OfgDslMarket(C.Structure):
    _fields_ = [
        ("n_assets", C.c_int32), ("pad", C.c_int32), ("assets", C.POINTER(OfgDslAsset)),
        ("corr", C.POINTER(C.c_double)), ("rate", C.c_double),
    ]


class OfgDslEngine(C.Structure):
    _fields_ = [("num_paths", C.c_int64), ("num_steps", C.c_int64), ("seed", C.c_uint64)]


class OfgDslInfo(C.Structure):
    _fields_ = [
        ("notional", C.c_double), ("maturity", C.c_double), ("num_underlyings", C.c_int32),
        ("n_state", C.c_int32), ("n_schedules", C.c_int32), ("n_locals", C.c_int32),
        ("name", C.c_char * 128),
    ]


class DeviceError(RuntimeError):
    """HIP library missing or device failure (OFG_DEVICE_ERROR)."""


_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Loads the HIP library; raises DeviceError when it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise DeviceError(
                f"openferric_amd: {LIB_PATH} is not built; run `python -m openferric_amd.build`")
        L = C.CDLL(LIB_PATH)
        pv = C.c_void_p
        L.ofg_create.argtypes = [C.c_int32]; L.ofg_create.restype = pv
        L.ofg_destroy.argtypes = [pv]; L.ofg_destroy.restype = None
        L.ofg_last_error.argtypes = []; L.ofg_last_error.restype = C.c_char_p
        L.ofg_version.argtypes = []; L.ofg_version.restype = C.c_char_p
        L.ofg_price.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t, C.POINTER(OfgEngineCfg),
                                C.POINTER(OfgResult)]
        L.ofg_price.restype = C.c_int32
        L.ofg_price_greeks.argtypes = L.ofg_price.argtypes
        L.ofg_price_greeks.restype = C.c_int32
        L.ofg_price_noalloc.argtypes = L.ofg_price.argtypes
        L.ofg_price_noalloc.restype = C.c_int32
        L.ofg_finite_difference_greeks.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t,
                                                   C.POINTER(OfgEngineCfg), C.POINTER(OfgFdBumps),
                                                   C.POINTER(OfgFdGreeks)]
        L.ofg_finite_difference_greeks.restype = C.c_int32
        L.ofg_bump_and_reprice.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t,
                                           C.POINTER(OfgEngineCfg), C.POINTER(OfgFdBumps),
                                           C.POINTER(C.c_double)]
        L.ofg_bump_and_reprice.restype = C.c_int32
        L.ofg_european_exact.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t, C.c_int32,
                                         C.c_int32, C.POINTER(OfgResult), C.POINTER(C.c_double)]
        L.ofg_european_exact.restype = C.c_int32
        L.ofg_spread_price.argtypes = [pv, C.POINTER(OfgSpreadOption), C.c_size_t, C.c_int32,
                                       C.c_int32, C.POINTER(OfgResult), C.POINTER(C.c_double)]
        L.ofg_spread_price.restype = C.c_int32
        L.ofg_sample_payoffs.argtypes = [pv, C.POINTER(OfgOption), C.POINTER(OfgEngineCfg),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ofg_sample_payoffs.restype = C.c_int32
        L.ofg_batch_create.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t,
                                       C.POINTER(OfgEngineCfg), C.c_int32]
        L.ofg_batch_create.restype = pv
        L.ofg_batch_run.argtypes = [pv, C.POINTER(C.c_double)]; L.ofg_batch_run.restype = C.c_int32
        L.ofg_batch_results.argtypes = [pv, C.POINTER(OfgResult)]
        L.ofg_batch_results.restype = C.c_int32
        L.ofg_batch_pricing_kernel_ms.argtypes = [pv]
        L.ofg_batch_pricing_kernel_ms.restype = C.c_double
        L.ofg_batch_destroy.argtypes = [pv]; L.ofg_batch_destroy.restype = None
        L.ofg_autotune.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t, C.POINTER(OfgEngineCfg),
                                   C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.ofg_autotune.restype = C.c_int32
        L.ofg_mc_greeks_estimate.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t,
                                             C.POINTER(OfgEngineCfg), C.c_double,
                                             C.POINTER(OfgMcGreeks)]
        L.ofg_mc_greeks_estimate.restype = C.c_int32
        L.ofg_lsm_price.argtypes = [pv, C.POINTER(OfgLsmOption), C.c_size_t,
                                    C.POINTER(OfgEngineCfg), C.POINTER(OfgResult),
                                    C.POINTER(OfgLsmBoundary), C.POINTER(C.c_int64)]
        L.ofg_lsm_price.restype = C.c_int32
        L.ofg_stream_jump.argtypes = [C.POINTER(C.c_uint64), C.c_uint64, C.POINTER(C.c_uint64)]
        L.ofg_stream_jump.restype = C.c_int32
        L.ofg_fill_normals.argtypes = [pv, C.c_int32, C.c_uint64, C.c_uint64, C.c_int64,
                                       C.c_int64, C.c_int32, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.ofg_fill_normals.restype = C.c_int32
        L.ofg_price_device.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t,
                                       C.POINTER(OfgEngineCfg), C.POINTER(OfgDeviceIo)]
        L.ofg_price_device.restype = C.c_int32
        L.ofg_inject_alloc_failure.argtypes = [C.c_int32]
        L.ofg_inject_alloc_failure.restype = None
        L.ofg_lsm_last_timing.argtypes = [pv, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int32)]
        L.ofg_lsm_last_timing.restype = C.c_int32
        L.ofg_qmc_price.argtypes = [pv, C.POINTER(OfgOption), C.c_size_t, C.POINTER(OfgResult)]
        L.ofg_qmc_price.restype = C.c_int32
        L.ofg_autocall_price.argtypes = [pv, C.POINTER(OfgAutocall), C.c_size_t,
                                         C.POINTER(OfgEngineCfg), C.c_int32,
                                         C.POINTER(OfgResult), C.POINTER(OfgAutocallSens)]
        L.ofg_autocall_price.restype = C.c_int32
        L.ofg_basket_price.argtypes = [pv, C.POINTER(OfgBasket), C.c_size_t,
                                       C.POINTER(OfgEngineCfg), C.c_int32,
                                       C.POINTER(OfgResult), C.POINTER(OfgBasketSens)]
        L.ofg_basket_price.restype = C.c_int32
        L.ofg_repair_correlation.argtypes = [C.POINTER(C.c_double), C.c_int32,
                                             C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.ofg_repair_correlation.restype = C.c_int32
        L.ofg_stress_correlation.argtypes = [C.POINTER(C.c_double), C.c_int32,
                                             C.POINTER(OfgCorrStress), C.c_int32,
                                             C.POINTER(C.c_double)]
        L.ofg_stress_correlation.restype = C.c_int32
        L.ofg_dsl_compile.argtypes = [C.c_char_p, C.POINTER(pv)]
        L.ofg_dsl_compile.restype = C.c_int32
        L.ofg_dsl_free.argtypes = [pv]
        L.ofg_dsl_free.restype = None
        L.ofg_dsl_get_info.argtypes = [pv, C.POINTER(OfgDslInfo)]
        L.ofg_dsl_get_info.restype = C.c_int32
        L.ofg_dsl_schedule_dates.argtypes = [pv, C.c_int32, C.POINTER(C.c_double), C.c_int32,
                                             C.POINTER(C.c_int32)]
        L.ofg_dsl_schedule_dates.restype = C.c_int32
        L.ofg_dsl_price.argtypes = [pv, pv, C.POINTER(OfgDslMarket), C.c_size_t,
                                    C.POINTER(OfgDslEngine), C.POINTER(OfgEngineCfg),
                                    C.POINTER(OfgResult)]
        L.ofg_dsl_price.restype = C.c_int32
        L.ofg_dsl_jit_compile.argtypes = [pv, pv, C.c_int32, C.POINTER(C.c_double)]
        L.ofg_dsl_jit_compile.restype = C.c_int32
        L.ofg_dsl_jit_stats.argtypes = [pv, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ofg_dsl_jit_stats.restype = C.c_int32
        L.ofg_dsl_jit_source.argtypes = [pv, C.c_int32, C.c_char_p, C.c_size_t,
                                         C.POINTER(C.c_size_t)]
        L.ofg_dsl_jit_source.restype = C.c_int32
        _lib = L
        return L


def last_error() -> str:
    return load().ofg_last_error().decode()


_ctx = {}


def context(device: int = 0):
    """Process-wide context per device (the engine's OnceLock analogue)."""
    with _lock:
        c = _ctx.get(device)
    if c:
        return c
    L = load()
    c = L.ofg_create(device)
    if not c:
        raise DeviceError(f"ofg_create({device}) failed: {last_error()}")
    with _lock:
        _ctx[device] = c
    return c
