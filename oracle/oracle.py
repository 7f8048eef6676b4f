"""ctypes binding of the CPU oracle (liboracle.so), and the oracle's Python
restatements of the finite-difference Greeks harness (src/greeks/bsm.rs).

TEST INFRASTRUCTURE ONLY: the tests' reference for the HIP product path and the timed
CPU baseline. Only tests/, the smoke check and the benchmark's cpu_baseline
leg import this module. The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

GBM, HESTON = 0, 1
VANILLA, ASIAN, BARRIER = 0, 1, 2
CALL, PUT = 0, 1
ARITH, GEO = 0, 1
FIXED, FLOATING = 0, 1
UP, DOWN = 0, 1
IN, OUT = 0, 1


class OrOption(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("payoff", C.c_int32), ("option_type", C.c_int32),
        ("averaging", C.c_int32), ("strike_type", C.c_int32),
        ("barrier_dir", C.c_int32), ("barrier_style", C.c_int32), ("n_obs", C.c_int32),
        ("spot", C.c_double), ("strike", C.c_double), ("expiry", C.c_double),
        ("rate", C.c_double), ("div_yield", C.c_double), ("vol", C.c_double),
        ("kappa", C.c_double), ("theta", C.c_double), ("xi", C.c_double),
        ("rho", C.c_double), ("v0", C.c_double),
        ("barrier_level", C.c_double), ("rebate", C.c_double),
        ("obs_times", C.POINTER(C.c_double)),
    ]


class OrCfg(C.Structure):
    _fields_ = [
        ("num_paths", C.c_int64), ("num_steps", C.c_int64), ("seed", C.c_uint64),
        ("antithetic", C.c_int32), ("control_variate", C.c_int32),
        ("cv_expected", C.c_double), ("cv_kind", C.c_int32), ("pad", C.c_int32),
    ]


LSM_VANILLA, LSM_BERMUDAN, LSM_BARRIER = 0, 1, 2
EUROPEAN, AMERICAN, BERMUDAN = 0, 1, 2
LSM_GBM, LSM_LOCAL_VOL, LSM_HESTON = 0, 1, 2


class OrLsmOption(C.Structure):
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


class OrLsmBoundary(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("strike", C.c_double), ("boundary_spot", C.c_double),
        ("has_boundary", C.c_int32), ("pad", C.c_int32),
        ("itm_paths", C.c_int64), ("exercised_paths", C.c_int64),
    ]


class OrAutocall(C.Structure):
    _fields_ = [
        ("n_assets", C.c_int32), ("phoenix", C.c_int32), ("memory", C.c_int32),
        ("n_dates", C.c_int32),
        ("notional", C.c_double), ("maturity", C.c_double), ("autocall_barrier", C.c_double),
        ("coupon_rate", C.c_double), ("coupon_barrier", C.c_double), ("ki_barrier", C.c_double),
        ("ki_strike", C.c_double), ("rate", C.c_double), ("div_yield", C.c_double),
        ("spots", C.POINTER(C.c_double)), ("vols", C.POINTER(C.c_double)),
        ("corr", C.POINTER(C.c_double)), ("dates", C.POINTER(C.c_double)),
        ("num_paths", C.c_int64), ("num_steps", C.c_int64), ("seed", C.c_uint64),
    ]


class OrBasket(C.Structure):
    _fields_ = [
        ("basket_type", C.c_int32), ("is_call", C.c_int32), ("n_assets", C.c_int32),
        ("t_dof", C.c_int32), ("weights", C.POINTER(C.c_double)), ("strike", C.c_double),
        ("maturity", C.c_double), ("rate", C.c_double), ("spots", C.POINTER(C.c_double)),
        ("vols", C.POINTER(C.c_double)), ("dividends", C.POINTER(C.c_double)),
        ("corr", C.POINTER(C.c_double)), ("num_paths", C.c_int64), ("seed", C.c_uint64),
    ]


BASKET_AVERAGE, BASKET_BEST_OF, BASKET_WORST_OF = 0, 1, 2


def build(force: bool = False) -> str:
    """Compile liboracle.so with make (gcc); returns the library path. Without
    `force`, make rebuilds only when a source is newer than the library (the
    sources are absent where only the built library was shipped)."""
    if force or not os.path.exists(_LIB_PATH) or os.path.exists(os.path.join(_HERE, "oracle.c")):
        import fcntl
        # one build at a time: parallel test workers may call this together
        with open(os.path.join(_HERE, ".build.lock"), "w") as lock:
            fcntl.flock(lock, fcntl.LOCK_EX)
            subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                           check=True, capture_output=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        d, i32, i64, u64 = C.c_double, C.c_int32, C.c_int64, C.c_uint64
        pd = C.POINTER(C.c_double)
        pu64 = C.POINTER(C.c_uint64)
        L.or_splitmix64_next.argtypes = [pu64]; L.or_splitmix64_next.restype = u64
        L.or_xoshiro_seed.argtypes = [u64, pu64]; L.or_xoshiro_seed.restype = None
        L.or_xoshiro_next_u64.argtypes = [pu64]; L.or_xoshiro_next_u64.restype = u64
        L.or_xoshiro_next_f64.argtypes = [pu64]; L.or_xoshiro_next_f64.restype = d
        L.or_xoshiro_advance.argtypes = [pu64, u64]; L.or_xoshiro_advance.restype = None
        L.or_stream_seed.argtypes = [u64, u64]; L.or_stream_seed.restype = u64
        L.or_uniform_open01.argtypes = [d]; L.or_uniform_open01.restype = d
        L.or_acklam_inv_cdf.argtypes = [d]; L.or_acklam_inv_cdf.restype = d
        L.or_hart_norm_cdf.argtypes = [d]; L.or_hart_norm_cdf.restype = d
        L.or_fast_norm_pdf.argtypes = [d]; L.or_fast_norm_pdf.restype = d
        L.or_gbm_generate.argtypes = [d, d, d, d, i64, pd, pd]; L.or_gbm_generate.restype = None
        L.or_heston_generate.argtypes = [d, d, d, d, d, d, d, d, i64, pd, pd, pd]
        L.or_heston_generate.restype = None
        L.or_heston_step_euler.argtypes = [d, d, d, d, d, d, d, d, d, d, pd, pd]
        L.or_heston_step_euler.restype = None
        L.or_payoff.argtypes = [C.POINTER(OrOption), pd, i64]; L.or_payoff.restype = d
        L.or_cv_eval.argtypes = [C.POINTER(OrOption), i32, pd, i64]; L.or_cv_eval.restype = d
        L.or_sample_normals.argtypes = [u64, i64, i64, i32, pd, pd]
        L.or_sample_normals.restype = None
        L.or_mc_run.argtypes = [C.POINTER(OrOption), C.POINTER(OrCfg), i32, pd, pd, pd, pd, pd, pd]
        L.or_mc_run.restype = C.c_int
        L.or_bs_price.argtypes = [i32, d, d, d, d, d, d]; L.or_bs_price.restype = d
        L.or_bs_price_exact.argtypes = [i32, d, d, d, d, d, d]; L.or_bs_price_exact.restype = d
        L.or_bs_greeks.argtypes = [i32, d, d, d, d, d, d, pd]; L.or_bs_greeks.restype = None
        L.or_geometric_asian_discrete.argtypes = [i32, d, d, d, d, d, pd, i64]
        L.or_geometric_asian_discrete.restype = d
        L.or_builtin_cv_expected.argtypes = [C.POINTER(OrOption), C.POINTER(i32)]
        L.or_builtin_cv_expected.restype = d
        L.or_mc_european_pathwise_aad.argtypes = [C.POINTER(OrOption), i64, i64, u64, i32, pd]
        L.or_mc_european_pathwise_aad.restype = C.c_int
        L.or_heston_price_delta_aad.argtypes = [C.POINTER(OrOption), i64, i64, u64, pd]
        L.or_heston_price_delta_aad.restype = C.c_int
        L.or_mc_european_exact.argtypes = [i32, i32, d, d, d, d, d, d, i64, i32, pd, pd]
        L.or_mc_european_exact.restype = C.c_int
        L.or_pcg64_seed.argtypes = [u64, pu64]; L.or_pcg64_seed.restype = None
        L.or_pcg64_next_u64.argtypes = [pu64]; L.or_pcg64_next_u64.restype = u64
        L.or_pcg64_next_f64.argtypes = [pu64]; L.or_pcg64_next_f64.restype = d
        L.or_set_rng_kind.argtypes = [i32]; L.or_set_rng_kind.restype = None
        L.or_rng_kind.argtypes = []; L.or_rng_kind.restype = i32
        L.or_spread_validate.argtypes = [pd]; L.or_spread_validate.restype = i32
        L.or_spread_mc.argtypes = [pd, i64, u64, i32, i32, pd, pd]; L.or_spread_mc.restype = i32
        L.or_pathwise_greeks.argtypes = [C.POINTER(OrOption), i64, i64, u64, i32, pd, pd, i32,
                                         i32, pd, pd]
        L.or_pathwise_greeks.restype = C.c_int
        L.or_mc_greeks_estimate.argtypes = [C.POINTER(OrOption), i64, u64, i32, d, pd]
        L.or_mc_greeks_estimate.restype = C.c_int
        L.or_lu3_solve.argtypes = [pd, pd, pd]; L.or_lu3_solve.restype = C.c_int
        L.or_autocall_price.argtypes = [C.POINTER(OrAutocall), pd, pd, pd]
        L.or_autocall_price.restype = C.c_int
        L.or_cholesky_lower.argtypes = [pd, i32, pd]; L.or_cholesky_lower.restype = C.c_int
        L.or_bump_corr_matrix.argtypes = [pd, i32, d, pd]; L.or_bump_corr_matrix.restype = None
        L.or_basket_price.argtypes = [C.POINTER(OrBasket), pd, pd, pd]
        L.or_basket_price.restype = C.c_int
        L.or_basket_factor_price.argtypes = [C.POINTER(OrBasket), C.c_int, C.c_int32, pd, pd, pd,
                                             pd]
        L.or_basket_factor_price.restype = C.c_int
        L.or_ln_gamma.argtypes = [C.c_double]; L.or_ln_gamma.restype = C.c_double
        L.or_beta_reg.argtypes = [C.c_double] * 3; L.or_beta_reg.restype = C.c_double
        L.or_student_t_cdf.argtypes = [C.c_double] * 2; L.or_student_t_cdf.restype = C.c_double
        L.or_outperformance_price.argtypes = [C.POINTER(OrBasket), C.c_int32, pd, pd, pd]
        L.or_outperformance_price.restype = C.c_int
        L.or_quanto_price.argtypes = [C.POINTER(OrBasket), C.c_double, C.c_double, pd, C.c_double,
                                      C.c_double, pd, pd, pd]
        L.or_quanto_price.restype = C.c_int
        L.or_sobol_directions.argtypes = [u64, u64, pu64]; L.or_sobol_directions.restype = None
        L.or_sobol_scrambler.argtypes = [u64, u64]; L.or_sobol_scrambler.restype = u64
        L.or_sobol_points.argtypes = [i64, u64, i64, pu64, pd]; L.or_sobol_points.restype = None
        L.or_mc_european_qmc.argtypes = [C.POINTER(OrOption), i64, i64, u64, pd, pd,
                                         C.POINTER(i32)]
        L.or_mc_european_qmc.restype = None
        L.or_lsm_price.argtypes = [C.POINTER(OrLsmOption), pd, pd, pd,
                                   C.POINTER(OrLsmBoundary), C.POINTER(i64)]
        L.or_lsm_price.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_option(model=GBM, payoff=VANILLA, option_type=CALL, spot=100.0, strike=100.0,
                expiry=1.0, rate=0.05, div_yield=0.0, vol=0.2, kappa=0.0, theta=0.0, xi=0.0,
                rho=0.0, v0=0.0, averaging=ARITH, strike_type=FIXED, barrier_dir=UP,
                barrier_style=OUT, barrier_level=0.0, rebate=0.0, obs_times=None):
    """Builds an OrOption; the returned object keeps obs_times alive."""
    o = OrOption()
    o.model, o.payoff, o.option_type = model, payoff, option_type
    o.averaging, o.strike_type = averaging, strike_type
    o.barrier_dir, o.barrier_style = barrier_dir, barrier_style
    o.spot, o.strike, o.expiry, o.rate, o.div_yield, o.vol = spot, strike, expiry, rate, div_yield, vol
    o.kappa, o.theta, o.xi, o.rho, o.v0 = kappa, theta, xi, rho, v0
    o.barrier_level, o.rebate = barrier_level, rebate
    if obs_times is not None:
        arr = np.ascontiguousarray(obs_times, dtype=np.float64)
        o._obs = arr
        o.obs_times = _ptr(arr)
        o.n_obs = len(arr)
    else:
        o.obs_times = None
        o.n_obs = 0
    return o


def builtin_cv_expected(o):
    kind = C.c_int32(0)
    e = lib().or_builtin_cv_expected(C.byref(o), C.byref(kind))
    return e, kind.value


def mc_run(o, num_paths, num_steps, seed, antithetic=False, control_variate=False,
           nthreads=1, normals1=None, normals2=None, want_samples=False):
    """MonteCarloEngine::run with the instrument's payoff (and built-in CV).

    Returns (price, stderr) or (price, stderr, xs, ys) with want_samples."""
    cfg = OrCfg()
    cfg.num_paths, cfg.num_steps, cfg.seed = num_paths, num_steps, seed
    cfg.antithetic = 1 if antithetic else 0
    cfg.control_variate = 0
    if control_variate:
        e, kind = builtin_cv_expected(o)
        if kind >= 0:
            cfg.control_variate, cfg.cv_expected, cfg.cv_kind = 1, e, kind
    samples = (num_paths + 1) // 2 if antithetic else num_paths
    xs = np.zeros(samples) if want_samples else None
    ys = np.zeros(samples) if want_samples else None
    n1 = np.ascontiguousarray(normals1, dtype=np.float64) if normals1 is not None else None
    n2 = np.ascontiguousarray(normals2, dtype=np.float64) if normals2 is not None else None
    price, se = C.c_double(), C.c_double()
    rc = lib().or_mc_run(C.byref(o), C.byref(cfg), nthreads, _ptr(n1), _ptr(n2),
                         C.byref(price), C.byref(se), _ptr(xs), _ptr(ys))
    if rc != 0:
        raise ValueError(f"or_mc_run failed rc={rc}")
    if want_samples:
        return price.value, se.value, xs, ys
    return price.value, se.value


def sample_normals(seed, sample, steps, streams):
    z1 = np.zeros(steps)
    z2 = np.zeros(steps)
    lib().or_sample_normals(seed, sample, steps, streams, _ptr(z1), _ptr(z2))
    return z1, z2


def normals_matrix(seed, samples, steps, streams):
    """[sample][step] normals exactly as the reference draws them."""
    z1 = np.zeros((samples, steps))
    z2 = np.zeros((samples, steps))
    for i in range(samples):
        a, b = sample_normals(seed, i, steps, streams)
        z1[i] = a
        z2[i] = b
    return z1, z2


def bs_price(option_type, spot, strike, rate, q, vol, expiry, exact=False):
    f = lib().or_bs_price_exact if exact else lib().or_bs_price
    return f(option_type, spot, strike, rate, q, vol, expiry)


def bs_greeks(option_type, spot, strike, rate, q, vol, expiry):
    out = np.zeros(5)
    lib().or_bs_greeks(option_type, spot, strike, rate, q, vol, expiry, _ptr(out))
    return dict(zip(["delta", "gamma", "vega", "theta", "rho"], out.tolist()))


def geometric_asian_discrete(option_type, s0, k, r, q, sigma, obs):
    arr = np.ascontiguousarray(obs, dtype=np.float64)
    return lib().or_geometric_asian_discrete(option_type, s0, k, r, q, sigma, _ptr(arr), len(arr))


def mc_european_pathwise_aad(o, num_paths, num_steps, seed, antithetic):
    out = np.zeros(6)
    rc = lib().or_mc_european_pathwise_aad(C.byref(o), num_paths, num_steps, seed,
                                           1 if antithetic else 0, _ptr(out))
    if rc:
        raise ValueError("aad failed")
    return dict(zip(["price", "stderr", "delta", "vega", "theta", "rho"], out.tolist()))


def heston_price_delta_aad(o, num_paths, num_steps, seed):
    out = np.zeros(2)
    rc = lib().or_heston_price_delta_aad(C.byref(o), num_paths, num_steps, seed, _ptr(out))
    if rc:
        raise ValueError("heston aad failed")
    return float(out[0]), float(out[1])


EXACT_SEQUENTIAL, EXACT_PARALLEL, EXACT_ARENA = 0, 1, 2


def mc_european_exact(kind, option_type, spot, strike, rate, q, vol, expiry, n_paths,
                      n_chunks=1, want_payoffs=False):
    """mc_european_sequential / _parallel / _with_arena of the reference's
    default build (oracle_exact.c). Returns (price, stderr or None[, payoffs])."""
    out = np.zeros(3)
    pay = np.full(n_paths, np.nan) if want_payoffs else None
    rc = lib().or_mc_european_exact(kind, option_type, spot, strike, rate, q, vol, expiry,
                                    n_paths, n_chunks, _ptr(out), _ptr(pay))
    if rc:
        raise ValueError(f"or_mc_european_exact failed rc={rc}")
    se = float(out[1]) if out[2] else None
    return (float(out[0]), se, pay) if want_payoffs else (float(out[0]), se)


RNG_XOSHIRO, RNG_PCG64 = 0, 1
SPREAD_MESSAGES = {
    -1: "spread spots s1 and s2 must be > 0",
    -2: "spread strike k must be >= 0",
    -3: "spread volatilities vol1 and vol2 must be >= 0",
    -4: "spread correlation rho must be in [-1, 1]",
    -5: "spread maturity t must be >= 0",
    -6: "spread MC num_paths must be > 0",
}
SPREAD_FIELDS = ("s1", "s2", "k", "vol1", "vol2", "rho", "q1", "q2", "r", "t")


class rng_kind:
    """Context manager: the oracle's MonteCarloEngine streams (or_mc_run,
    or_sample_normals, or_mc_greeks_estimate) draw Pcg64 or xoshiro256++."""

    def __init__(self, kind):
        self.kind = kind

    def __enter__(self):
        self.prev = lib().or_rng_kind()
        lib().or_set_rng_kind(self.kind)
        return self

    def __exit__(self, *exc):
        lib().or_set_rng_kind(self.prev)
        return False


def pcg64_stream(seed, n):
    """The first n u64 draws of Pcg64::seed_from_u64(seed) (fast_rng.rs:93-137)."""
    st = (C.c_uint64 * 4)()
    lib().or_pcg64_seed(seed, st)
    return [lib().or_pcg64_next_u64(st) for _ in range(n)]


def spread_mc(option, num_paths, seed, antithetic=True, rng_kind=RNG_XOSHIRO,
              want_payoffs=False):
    """SpreadMonteCarloEngine::price (spread_mc.rs:83-175) with reproducible
    streams; option is a dict of SPREAD_FIELDS. Returns (price, stderr[,
    payoffs]); raises ValueError with the reference's message on its errors."""
    p = np.array([float(option[f]) for f in SPREAD_FIELDS])
    out = np.zeros(2)
    samples = (num_paths + 1) // 2 if antithetic else num_paths
    pay = np.full(max(samples, 0), np.nan) if want_payoffs else None
    rc = lib().or_spread_mc(_ptr(p), num_paths, seed, 1 if antithetic else 0, rng_kind,
                            _ptr(out), _ptr(pay))
    if rc:
        raise ValueError(SPREAD_MESSAGES.get(rc, f"or_spread_mc failed rc={rc}"))
    return (float(out[0]), float(out[1]), pay) if want_payoffs else (float(out[0]), float(out[1]))


def _ncdf(x):
    return 0.5 * math.erfc(-x / math.sqrt(2.0))


def margrabe_exchange_price(o):
    """margrabe_exchange_price (src/engines/analytic/spread.rs:69-98), k = 0,
    with the exact normal CDF."""
    t = o["t"]
    if t <= 0.0:
        return max(o["s1"] - o["s2"], 0.0)
    var = o["vol1"] ** 2 - 2.0 * o["rho"] * o["vol1"] * o["vol2"] + o["vol2"] ** 2
    sigma = math.sqrt(max(var, 0.0))
    s1_df = o["s1"] * math.exp(-o["q1"] * t)
    s2_df = o["s2"] * math.exp(-o["q2"] * t)
    if sigma <= 0.0:
        return max(s1_df - s2_df, 0.0)
    sst = sigma * math.sqrt(t)
    d1 = (math.log(o["s1"] / o["s2"]) + (o["q2"] - o["q1"] + 0.5 * sigma * sigma) * t) / sst
    return s1_df * _ncdf(d1) - s2_df * _ncdf(d1 - sst)


def kirk_spread_price(o):
    """kirk_spread_price (src/engines/analytic/spread.rs:100-147) with the
    exact normal CDF."""
    t = o["t"]
    if t <= 0.0:
        return max(o["s1"] - o["s2"] - o["k"], 0.0)
    f1 = o["s1"] * math.exp((o["r"] - o["q1"]) * t)
    f2 = o["s2"] * math.exp((o["r"] - o["q2"]) * t)
    den = f2 + o["k"]
    beta = f2 / den
    v1, v2, rho = o["vol1"], o["vol2"], o["rho"]
    sk2 = v1 * v1 - 2.0 * beta * rho * v1 * v2 + beta * beta * v2 * v2
    sk = math.sqrt(max(sk2, 0.0))
    disc = math.exp(-o["r"] * t)
    if sk <= 0.0:
        return disc * max(f1 - den, 0.0)
    sst = sk * math.sqrt(t)
    d1 = (math.log(f1 / den) + 0.5 * sk2 * t) / sst
    return disc * (f1 * _ncdf(d1) - den * _ncdf(d1 - sst))


PATHWISE_SLOTS = ["s0", "sigma_or_v0", "r", "T", "kappa", "theta", "xi", "rho"]


def pathwise_greeks(o, num_paths, num_steps, seed, antithetic=False, normals1=None,
                    normals2=None, nthreads=1, mutate=0, want_samples=False):
    """Pathwise estimator of MonteCarloEngine::run's price and its derivatives
    (forward tangents, oracle_pathwise.c). Returns dict(price, stderr,
    g=array of 8: d/d{S0, sigma|v0, r, T, kappa, theta, xi, rho}) and, with
    want_samples, `samples`: [sample][pv, g0..g7]."""
    out = np.zeros(10)
    samples = (num_paths + 1) // 2 if antithetic else num_paths
    per = np.zeros((samples, 9)) if want_samples else None
    n1 = np.ascontiguousarray(normals1, dtype=np.float64) if normals1 is not None else None
    n2 = np.ascontiguousarray(normals2, dtype=np.float64) if normals2 is not None else None
    rc = lib().or_pathwise_greeks(C.byref(o), num_paths, num_steps, seed, 1 if antithetic else 0,
                                  _ptr(n1), _ptr(n2), nthreads, mutate, _ptr(out), _ptr(per))
    if rc:
        raise ValueError(f"or_pathwise_greeks failed rc={rc}")
    r = dict(price=float(out[0]), stderr=float(out[1]), g=out[2:].copy())
    if want_samples:
        r["samples"] = per
    return r


def mc_greeks_estimate(o, num_paths, seed, antithetic=True, spot_bump_rel=1e-2):
    """MonteCarloGreeksEngine estimators: dict of pathwise delta/gamma and
    likelihood-ratio delta/gamma/vega (mc_greeks.rs:131-243)."""
    out = np.zeros(5)
    rc = lib().or_mc_greeks_estimate(C.byref(o), num_paths, seed, 1 if antithetic else 0,
                                     spot_bump_rel, _ptr(out))
    if rc:
        raise ValueError("invalid input")
    return dict(zip(["pw_delta", "pw_gamma", "lr_delta", "lr_gamma", "lr_vega"], out.tolist()))


def make_lsm_option(instrument=LSM_VANILLA, option_type=PUT, exercise=AMERICAN,
                    dynamics=LSM_GBM, spot=36.0, strike=40.0, expiry=1.0, rate=0.06,
                    div_yield=0.0, vol=0.2, kappa=0.0, theta=0.0, xi=0.0, rho=0.0, v0=0.0,
                    barrier_dir=UP, barrier_style=OUT, barrier_level=0.0, rebate=0.0,
                    dates=None, strikes=None, num_paths=1000, num_steps=50, seed=42):
    """Builds an OrLsmOption; the returned object keeps dates/strikes alive."""
    o = OrLsmOption()
    o.instrument, o.option_type, o.exercise, o.dynamics = instrument, option_type, exercise, dynamics
    o.barrier_dir, o.barrier_style = barrier_dir, barrier_style
    o.spot, o.strike, o.expiry, o.rate, o.div_yield, o.vol = spot, strike, expiry, rate, div_yield, vol
    o.kappa, o.theta, o.xi, o.rho, o.v0 = kappa, theta, xi, rho, v0
    o.barrier_level, o.rebate = barrier_level, rebate
    o.num_paths, o.num_steps, o.seed = num_paths, num_steps, seed
    d = np.ascontiguousarray(dates if dates is not None else [], dtype=np.float64)
    k = np.ascontiguousarray(strikes if strikes is not None else [], dtype=np.float64)
    o._keep = (d, k)
    o.dates, o.strikes = _ptr(d), _ptr(k)
    o.n_dates = len(d)
    return o


def lu3_solve(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(3)
    ok = lib().or_lu3_solve(_ptr(a), _ptr(b), _ptr(x))
    return x if ok else None


def lsm_price(o, normals=None, want_boundary=False):
    """LongstaffSchwartzEngine::price. normals, when given, replace the
    stream's draws in its order ([path][step], Heston: [path][step][z1, z2]).
    Returns (price, stderr) or (price, stderr, boundary list of dicts)."""
    z = np.ascontiguousarray(normals, dtype=np.float64).ravel() if normals is not None else None
    price, se = C.c_double(), C.c_double()
    cap = max(int(o.num_steps), 1)
    bnd = (OrLsmBoundary * cap)() if want_boundary else None
    nb = C.c_int64(0)
    rc = lib().or_lsm_price(C.byref(o), _ptr(z), C.byref(price), C.byref(se), bnd, C.byref(nb))
    if rc != 0:
        raise ValueError(f"or_lsm_price failed rc={rc}")
    if not want_boundary:
        return price.value, se.value
    pts = [dict(time=b.time, strike=b.strike,
                boundary_spot=b.boundary_spot if b.has_boundary else None,
                itm_paths=b.itm_paths, exercised_paths=b.exercised_paths)
           for b in bnd[:nb.value]]
    return price.value, se.value, pts


def lsm_stream_normals(seed, count):
    """The first `count` draws of the engine's single stream (Acklam of the
    clamped uniform of xoshiro256++ seeded by seed_from_u64(seed))."""
    s = (C.c_uint64 * 4)()
    L = lib()
    L.or_xoshiro_seed(seed, s)
    out = np.empty(count)
    for i in range(count):
        out[i] = L.or_acklam_inv_cdf(L.or_uniform_open01(L.or_xoshiro_next_f64(s)))
    return out


def xoshiro_advance(state, k):
    """The xoshiro256++ state after k draws, by stepping."""
    s = (C.c_uint64 * 4)(*[int(x) & (2**64 - 1) for x in state])
    lib().or_xoshiro_advance(s, int(k))
    return list(s)


def sobol_points(dims, seed, n):
    """(words uint64 [n][dims], uniforms [n][dims]) of SobolSequence::new(dims,
    seed) stepped n times."""
    w = np.zeros((n, dims), dtype=np.uint64)
    u = np.zeros((n, dims))
    lib().or_sobol_points(dims, seed, n, w.ctypes.data_as(C.POINTER(C.c_uint64)), _ptr(u))
    return w, u


def sobol_directions(dim, seed):
    v = (C.c_uint64 * 64)()
    lib().or_sobol_directions(dim, seed, v)
    return [int(x) for x in v]


def sobol_scrambler(dim, seed):
    return int(lib().or_sobol_scrambler(dim, seed))


def mc_european_qmc(o, n_paths, n_steps, seed):
    """mc_european_qmc_with_seed: (price, stderr or None)."""
    p, se, has = C.c_double(), C.c_double(), C.c_int32()
    lib().or_mc_european_qmc(C.byref(o), n_paths, n_steps, seed, C.byref(p), C.byref(se),
                             C.byref(has))
    return p.value, (se.value if has.value else None)


def make_autocall(spots, vols, corr, dates, notional=100.0, maturity=1.0, autocall_barrier=1.0,
                  coupon_rate=0.08, coupon_barrier=0.0, ki_barrier=0.2, ki_strike=1.0, rate=0.01,
                  div_yield=0.0, phoenix=False, memory=False, num_paths=1000, num_steps=64,
                  seed=94_210):
    """OrAutocall over the contract's (already selected) underlyings."""
    a = OrAutocall()
    sp = np.ascontiguousarray(spots, dtype=np.float64)
    vo = np.ascontiguousarray(vols, dtype=np.float64)
    co = np.ascontiguousarray(corr, dtype=np.float64).ravel()
    dt = np.ascontiguousarray(dates, dtype=np.float64)
    a._keep = (sp, vo, co, dt)
    a.n_assets, a.phoenix, a.memory, a.n_dates = len(sp), int(phoenix), int(memory), len(dt)
    a.notional, a.maturity, a.autocall_barrier = notional, maturity, autocall_barrier
    a.coupon_rate, a.coupon_barrier, a.ki_barrier, a.ki_strike = coupon_rate, coupon_barrier, \
        ki_barrier, ki_strike
    a.rate, a.div_yield = rate, div_yield
    a.spots, a.vols, a.corr, a.dates = _ptr(sp), _ptr(vo), _ptr(co), _ptr(dt)
    a.num_paths, a.num_steps, a.seed = num_paths, num_steps, seed
    return a


def autocall_price(a, normals=None):
    """simulate_autocallable_paths: (mean, stderr). normals [path][step][asset]."""
    z = np.ascontiguousarray(normals, dtype=np.float64).ravel() if normals is not None else None
    p, se = C.c_double(), C.c_double()
    rc = lib().or_autocall_price(C.byref(a), _ptr(z), C.byref(p), C.byref(se))
    if rc != 0:
        raise ValueError(f"or_autocall_price failed rc={rc}")
    return p.value, se.value


def cholesky_lower(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    l = np.zeros((n, n))
    ok = lib().or_cholesky_lower(_ptr(a.ravel()), n, _ptr(l))
    return l if ok else None


def make_basket(spots, vols, dividends, corr, weights=None, basket_type=BASKET_AVERAGE,
                is_call=True, strike=100.0, maturity=1.0, rate=0.02, num_paths=10_000,
                seed=3_148_159, t_dof=0):
    b = OrBasket()
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in
            (spots, vols, dividends, np.asarray(corr, dtype=np.float64).ravel(),
             weights if weights is not None else np.zeros(len(spots)))]
    b._keep = arrs
    b.spots, b.vols, b.dividends, b.corr, b.weights = [_ptr(a) for a in arrs]
    b.n_assets, b.basket_type, b.is_call = len(arrs[0]), basket_type, int(is_call)
    b.strike, b.maturity, b.rate = strike, maturity, rate
    b.num_paths, b.seed, b.t_dof = num_paths, seed, t_dof
    return b


def ln_gamma(x):
    return lib().or_ln_gamma(float(x))


def beta_reg(a, b, x):
    return lib().or_beta_reg(float(a), float(b), float(x))


def student_t_cdf(x, nu):
    """statrs 0.18 StudentsT::new(0, 1, nu).cdf(x)."""
    return lib().or_student_t_cdf(float(x), float(nu))


def basket_price(b, normals=None):
    """simulate_basket_once_matrix: (price, stderr). normals [path][asset]
    (Student-t copula: [path][asset + dof])."""
    z = np.ascontiguousarray(normals, dtype=np.float64).ravel() if normals is not None else None
    p, se = C.c_double(), C.c_double()
    rc = lib().or_basket_price(C.byref(b), _ptr(z), C.byref(p), C.byref(se))
    if rc != 0:
        raise ValueError(f"or_basket_price failed rc={rc}")
    return p.value, se.value


def outperformance_price(b, leader, normals=None):
    """price_outperformance_basket_mc: (price, stderr); b.weights = lagger weights."""
    z = np.ascontiguousarray(normals, dtype=np.float64).ravel() if normals is not None else None
    p, se = C.c_double(), C.c_double()
    rc = lib().or_outperformance_price(C.byref(b), int(leader), _ptr(z), C.byref(p), C.byref(se))
    if rc != 0:
        raise ValueError(f"or_outperformance_price failed rc={rc}")
    return p.value, se.value


def quanto_price(b, fx_rate, fx_vol, asset_fx_corr, domestic_rate, foreign_rate, normals=None):
    """price_quanto_basket_mc: (price, stderr)."""
    z = np.ascontiguousarray(normals, dtype=np.float64).ravel() if normals is not None else None
    rho = np.ascontiguousarray(asset_fx_corr, dtype=np.float64)
    p, se = C.c_double(), C.c_double()
    rc = lib().or_quanto_price(C.byref(b), fx_rate, fx_vol, _ptr(rho), domestic_rate,
                               foreign_rate, _ptr(z), C.byref(p), C.byref(se))
    if rc != 0:
        raise ValueError(f"or_quanto_price failed rc={rc}")
    return p.value, se.value


def basket_factor_price(b, loadings, one_factor, normals=None):
    """simulate_basket_once_factor: (price, stderr). normals [path][k + n]
    (Student-t copula: [path][k + n + dof])."""
    L = np.ascontiguousarray(loadings, dtype=np.float64)
    k = 1 if one_factor else L.shape[1]
    z = np.ascontiguousarray(normals, dtype=np.float64).ravel() if normals is not None else None
    p, se = C.c_double(), C.c_double()
    rc = lib().or_basket_factor_price(C.byref(b), int(one_factor), k, _ptr(L.ravel()), _ptr(z),
                                      C.byref(p), C.byref(se))
    if rc != 0:
        raise ValueError(f"or_basket_factor_price failed rc={rc}")
    return p.value, se.value


# ---- finite_difference_greeks / bump_and_reprice (src/greeks/bsm.rs) ------
# The reference's harness restated over a GenericPricer callable
# price(s, k, r, sigma, t) (bsm.rs:44-55); Python evaluates each expression
# left to right, as Rust does.
FD_FIELDS = ("delta", "gamma", "vega", "theta", "rho", "vanna", "volga")


def bump_and_reprice(pricer, s, k, r, sigma, t, bump_s, bump_r, bump_sigma, bump_t):
    """bump_and_reprice, bsm.rs:130-163: (base, s_up, r_up, v_up, t_up,
    s_dn); t_up at max(t + bump_t, 1e-8), s_dn not floored."""
    base = pricer(s, k, r, sigma, t)
    s_up = pricer(s + bump_s, k, r, sigma, t)
    r_up = pricer(s, k, r + bump_r, sigma, t)
    v_up = pricer(s, k, r, sigma + bump_sigma, t)
    t_up = pricer(s, k, r, sigma, max(t + bump_t, 1e-8))
    return (base, s_up, r_up, v_up, t_up, pricer(s - bump_s, k, r, sigma, t))


def finite_difference_greeks(pricer, s, k, r, sigma, t, ds, dr, dv, dt):
    """finite_difference_greeks, bsm.rs:165-212: dict of FD_FIELDS."""
    s_dn = max(s - ds, 1e-8)
    v_dn = max(sigma - dv, 1e-8)
    p0 = pricer(s, k, r, sigma, t)
    p_s_up = pricer(s + ds, k, r, sigma, t)
    p_s_dn = pricer(s_dn, k, r, sigma, t)
    p_r_up = pricer(s, k, r + dr, sigma, t)
    p_r_dn = pricer(s, k, r - dr, sigma, t)
    p_v_up = pricer(s, k, r, sigma + dv, t)
    p_v_dn = pricer(s, k, r, v_dn, t)
    p_t_up = pricer(s, k, r, sigma, t + dt)
    p_t_dn = pricer(s, k, r, sigma, max(t - dt, 1e-8))
    p_s_up_v_up = pricer(s + ds, k, r, sigma + dv, t)
    p_s_up_v_dn = pricer(s + ds, k, r, v_dn, t)
    p_s_dn_v_up = pricer(s_dn, k, r, sigma + dv, t)
    p_s_dn_v_dn = pricer(s_dn, k, r, v_dn, t)
    return dict(
        delta=(p_s_up - p_s_dn) / (2.0 * ds),
        gamma=(p_s_up - 2.0 * p0 + p_s_dn) / (ds * ds),
        rho=(p_r_up - p_r_dn) / (2.0 * dr),
        vega=(p_v_up - p_v_dn) / (2.0 * dv),
        theta=-(p_t_up - p_t_dn) / (2.0 * dt),
        vanna=(p_s_up_v_up - p_s_up_v_dn - p_s_dn_v_up + p_s_dn_v_dn) / (4.0 * ds * dv),
        volga=(p_v_up - 2.0 * p0 + p_v_dn) / (dv * dv),
    )


def bsm_greeks_closed_form(call, s, k, r, q, sigma, t):
    """black_scholes_merton_greeks (bsm.rs:63-128) with the exact normal
    pdf/cdf: delta, gamma, vega, theta (-dV/dT), rho, vanna (d2V/dS dsigma)
    and volga (d2V/dsigma2)."""
    import math
    if s <= 0.0 or k <= 0.0 or t <= 0.0 or sigma <= 0.0:
        return dict.fromkeys(FD_FIELDS, 0.0)
    pdf = lambda x: math.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)
    cdf = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))
    vt = sigma * math.sqrt(t)
    d1 = (math.log(s / k) + (r - q + 0.5 * sigma * sigma) * t) / vt
    d2 = d1 - vt
    sqrt_t = math.sqrt(t)
    df_q, df_r = math.exp(-q * t), math.exp(-r * t)
    delta = df_q * cdf(d1) if call else df_q * (cdf(d1) - 1.0)
    gamma = df_q * pdf(d1) / (s * sigma * sqrt_t)
    vega = s * df_q * pdf(d1) * sqrt_t
    if call:
        theta = (-s * df_q * pdf(d1) * sigma / (2.0 * sqrt_t) - r * k * df_r * cdf(d2)
                 + q * s * df_q * cdf(d1))
        rho = k * t * df_r * cdf(d2)
    else:
        theta = (-s * df_q * pdf(d1) * sigma / (2.0 * sqrt_t) + r * k * df_r * cdf(-d2)
                 - q * s * df_q * cdf(-d1))
        rho = -k * t * df_r * cdf(-d2)
    vanna = -df_q * pdf(d1) * d2 / sigma
    volga = vega * d1 * d2 / sigma
    return dict(delta=delta, gamma=gamma, vega=vega, theta=theta, rho=rho, vanna=vanna,
                volga=volga)
