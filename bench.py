#!/usr/bin/env python3
"""Benchmark: batched FP64 Monte Carlo pricing on one MI355X.

Workload (one "step" = one replay of the captured batch): 4096 GBM options in
one persistent-kernel launch, half European and half arithmetic-average Asian
(observed at every step) calls/puts alternating; per option paths
floor(10^U[4,6]), steps U{12..504}, K U[70,130], sigma U[0.1,0.6],
r U[0,0.08], T U[0.1,3], S0 = 100, q = 0; FP64; plain Monte Carlo (no
antithetic, no control variate), fast RNG mode. Inputs are synthetic, drawn
from a seeded host RNG.

Prints one JSON line: FP64 path-steps/s, the roofline of the pricing kernel
against MI355X FP64 vector peak (canonical 65 FLOP per GBM path-step, 66 with
the Asian running sum), and the CPU baseline (the C restatement of the
reference algorithm, all host threads, on a bounded sample of the workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 vector peak
FLOP_GBM_STEP = 65            # canonical FLOP per GBM path-step
FLOP_ASIAN_STEP = 66          # + arithmetic running sum
WORKLOAD_SEED = 0x0FE22C1C
# HBM bytes per launch of the pricing kernel on the default workload, for
# each launch shape the autotuner chooses from, from rocprofv3 FETCH_SIZE
# (x2.000) and WRITE_SIZE (x1.000) scaled by a 1 GiB copy kernel
# (profiles/price_kernel_fast_gbm_batch4096_counters.md).
PRICE_KERNEL_HBM_BYTES = {   # (threads per block, lane-steps per item) -> bytes
    (512, 8192): 156.7e6, (256, 8192): 155.1e6, (128, 8192): 158.4e6, (64, 8192): 141.5e6,
    (512, 2048): 415.3e6, (256, 2048): 417.7e6, (128, 2048): 417.5e6, (64, 2048): 352.9e6,
    (512, 32768): 73.2e6, (256, 32768): 74.0e6, (128, 32768): 75.2e6, (64, 32768): 71.6e6}


def make_workload(n_opts=4096, seed=WORKLOAD_SEED, path_scale=1.0):
    """Config: batched GBM, half European / half arithmetic Asian."""
    rng = np.random.default_rng(seed)
    u_paths = rng.uniform(4.0, 6.0, n_opts)
    paths = np.floor(10.0 ** u_paths).astype(np.int64)
    steps = rng.integers(12, 505, n_opts)
    strike = rng.uniform(70.0, 130.0, n_opts)
    sigma = rng.uniform(0.1, 0.6, n_opts)
    rate = rng.uniform(0.0, 0.08, n_opts)
    expiry = rng.uniform(0.1, 3.0, n_opts)
    opts = []
    for i in range(n_opts):
        asian = i >= n_opts // 2
        n = int(steps[i])
        T = float(expiry[i])
        d = dict(payoff="asian" if asian else "vanilla",
                 option_type="call" if i % 2 == 0 else "put",
                 spot=100.0, strike=float(strike[i]), expiry=T, rate=float(rate[i]), q=0.0,
                 vol=float(sigma[i]), num_steps=n,
                 num_paths=max(64, int(paths[i] * path_scale)),
                 obs_times=[T * (j / n) for j in range(1, n + 1)] if asian else None)
        opts.append(d)
    return opts


def to_ofg(d, seed=42):
    from openferric_amd import _lib
    from openferric_amd.engines import _Keep
    keep = _Keep()
    o = _lib.OfgOption()
    o.model = _lib.MODEL_HESTON if d.get("model") == "heston" else _lib.MODEL_GBM
    o.payoff = {"asian": _lib.PAYOFF_ASIAN, "barrier": _lib.PAYOFF_BARRIER}.get(
        d["payoff"], _lib.PAYOFF_VANILLA)
    if o.model == _lib.MODEL_HESTON:
        o.kappa, o.theta, o.xi, o.rho, o.v0 = d["kappa"], d["theta"], d["xi"], d["rho"], d["v0"]
    if d["payoff"] == "barrier":
        o.barrier_dir, o.barrier_style = _lib.UP, _lib.OUT
        o.barrier_level, o.rebate = d["barrier_level"], 0.0
    o.option_type = _lib.CALL if d["option_type"] == "call" else _lib.PUT
    o.spot, o.strike, o.expiry = d["spot"], d["strike"], d["expiry"]
    o.rate, o.div_yield, o.vol = d["rate"], d["q"], d["vol"]
    if d["obs_times"] is not None:
        o.obs_times = keep.arr(d["obs_times"])
        o.n_obs = len(d["obs_times"])
        o.averaging, o.strike_type = _lib.ARITHMETIC, _lib.FIXED
    o.num_paths, o.num_steps, o.seed = d["num_paths"], d["num_steps"], seed
    return o, keep


def greeks_workload(n_opts=64, seed=WORKLOAD_SEED + 4, paths=1 << 20, steps=252):
    """Adjoint Greeks workload: GBM arithmetic-Asian calls, daily observation,
    S0 = 100, K U[80,120], sigma U[0.15,0.45], r = 0.05, q = 0, T = 1."""
    rng = np.random.default_rng(seed)
    strike = rng.uniform(80.0, 120.0, n_opts)
    sigma = rng.uniform(0.15, 0.45, n_opts)
    obs = [1.0 * (j / steps) for j in range(1, steps + 1)]
    return [dict(payoff="asian", option_type="call", spot=100.0, strike=float(strike[i]),
                 expiry=1.0, rate=0.05, q=0.0, vol=float(sigma[i]), num_steps=steps,
                 num_paths=paths, obs_times=obs) for i in range(n_opts)]


def heston_workload(n_opts=256, paths=1 << 20, steps=365, seed=WORKLOAD_SEED + 3):
    """Heston up-and-out calls, daily monitoring: kappa U[0.5,5], theta
    U[0.01,0.1], xi U[0.1,1], rho U[-0.9,0], v0 U[0.01,0.1], barrier = spot x
    U[1.1,1.5], S0 = 100, K U[80,120], r U[0,0.08], q = 0, T = 1."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_opts):
        out.append(dict(model="heston", payoff="barrier", option_type="call", spot=100.0,
                        strike=float(rng.uniform(80, 120)), expiry=1.0,
                        rate=float(rng.uniform(0, 0.08)), q=0.0, vol=0.2,
                        kappa=float(rng.uniform(0.5, 5)), theta=float(rng.uniform(0.01, 0.1)),
                        xi=float(rng.uniform(0.1, 1.0)), rho=float(rng.uniform(-0.9, 0.0)),
                        v0=float(rng.uniform(0.01, 0.1)),
                        barrier_level=100.0 * float(rng.uniform(1.1, 1.5)),
                        num_steps=steps, num_paths=paths, obs_times=None))
    return out


def heston_greeks_workload(n_opts=1024, paths=1 << 18, seed=WORKLOAD_SEED + 5):
    """Batched Heston calibration gradients: alternating European / up-and-out
    calls, T U[0.05,5], steps ceil(252 T), Heston ranges as above."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_opts):
        T = float(rng.uniform(0.05, 5.0))
        barrier = i % 2 == 1
        out.append(dict(model="heston", payoff="barrier" if barrier else "vanilla",
                        option_type="call", spot=100.0, strike=float(rng.uniform(80, 120)),
                        expiry=T, rate=float(rng.uniform(0, 0.08)), q=0.0, vol=0.2,
                        kappa=float(rng.uniform(0.5, 5)), theta=float(rng.uniform(0.01, 0.1)),
                        xi=float(rng.uniform(0.1, 1.0)), rho=float(rng.uniform(-0.9, 0.0)),
                        v0=float(rng.uniform(0.01, 0.1)),
                        barrier_level=100.0 * float(rng.uniform(1.1, 1.5)),
                        num_steps=int(np.ceil(252 * T)), num_paths=paths, obs_times=None))
    return out


def path_steps_and_flops(opts):
    ps = sum(d["num_paths"] * d["num_steps"] for d in opts)
    fl = sum(d["num_paths"] * d["num_steps"] * (FLOP_ASIAN_STEP if d["payoff"] == "asian"
                                                else FLOP_GBM_STEP) for d in opts)
    return ps, fl


FLOP_HESTON_STEP = 124       # canonical FLOP per Heston path-step


def greeks_leg(rng_mode):
    """Adjoint vs bump-and-revalue on the Greeks workload (same GPU, same
    draws). Greeks: delta, vega, rho, theta (4 per option); bump-and-revalue
    = central differences with common random numbers, 2k+1 = 9 pricing passes
    in one batched launch."""
    from openferric_amd import engine_cfg, price_raw, _lib
    opts = greeks_workload()
    structs, keeps = zip(*[to_ofg(d) for d in opts])
    structs = list(structs)
    base = engine_cfg(rng_mode=rng_mode)
    price_raw(structs, base)  # warm-up
    p = price_raw(structs, base)
    adj = price_raw(structs, base, greeks=True)
    adj = price_raw(structs, base, greeks=True)
    fd = price_raw(structs, engine_cfg(rng_mode=rng_mode, greek_mode=_lib.GREEKS_FD_CRN,
                                       bump_rel=1e-4))
    fwd_cfg = engine_cfg(rng_mode=rng_mode, greek_mode=_lib.GREEKS_FORWARD)
    price_raw(structs, fwd_cfg)
    fwd = price_raw(structs, fwd_cfg)
    t_price, t_adj, t_fd = p[0].kernel_ms, adj[0].kernel_ms, fd[0].kernel_ms
    t_fwd = fwd[0].kernel_ms
    n_greeks = 4 * len(opts)
    rel = []
    for a, f in zip(adj, fd):
        for k in ("delta", "vega", "rho", "theta"):
            fa, ff = getattr(a, k), getattr(f, k)
            rel.append(abs(fa - ff) / max(abs(ff), 1e-12))
    rel_fwd = [abs(getattr(a, k) - getattr(f, k)) / max(abs(getattr(a, k)), 1e-12)
               for a, f in zip(adj, fwd) for k in ("delta", "vega", "rho", "theta")]
    ps = sum(d["num_paths"] * d["num_steps"] for d in opts)
    return {"workload": "64 GBM arithmetic-Asian calls (daily obs), 2^20 paths x 252 steps, "
                        "K U[80,120], sigma U[0.15,0.45]",
            "greeks": ["delta", "vega", "rho", "theta"],
            "pricing_ms": t_price, "adjoint_ms": t_adj, "bump_ms": t_fd,
            "adjoint_cost_in_pricing_passes": t_adj / t_price,
            "adjoint_greeks_per_s": n_greeks / (t_adj * 1e-3),
            "bump_greeks_per_s": n_greeks / (t_fd * 1e-3),
            "adjoint_speedup_vs_bump": t_fd / t_adj,
            "pricing_path_steps_per_s": ps / (t_price * 1e-3),
            "adjoint_vs_bump_max_rel_diff": max(rel), "adjoint_vs_bump_median_rel_diff":
                float(np.median(rel)),
            # the same pathwise estimator in one forward sweep (GBM running sums)
            "forward_ms": t_fwd, "forward_cost_in_pricing_passes": t_fwd / t_price,
            "forward_greeks_per_s": n_greeks / (t_fwd * 1e-3),
            "forward_speedup_vs_bump": t_fd / t_fwd,
            "forward_vs_adjoint_max_rel_diff": max(rel_fwd)}


def heston_leg(rng_mode, cpu=True):
    """Heston up-and-out pricing: 256 options x 2^20 paths x 365 steps, one
    persistent launch replayed from a hipGraph; fast draws, and the
    reference's own draws."""
    from openferric_amd import BatchPricer, RngMode, engine_cfg
    opts = heston_workload()
    structs, keeps = zip(*[to_ofg(d) for d in opts])
    ps = sum(d["num_paths"] * d["num_steps"] for d in opts)
    out = {"workload": "256 Heston up-and-out calls, 2^20 paths x 365 steps", "path_steps": ps}
    for name, mode in (("fast", rng_mode), ("reference_draws", RngMode.Reference)):
        b = BatchPricer(list(structs), engine_cfg(rng_mode=mode), use_graph=True)
        b.run()
        ms = [b.run() for _ in range(2)]
        b.close()
        k_ms = float(np.mean(ms))
        v = ps / (k_ms * 1e-3)
        leg = {"kernel_ms": k_ms, "path_steps_per_s": v,
               "frac_fp64_peak": v * FLOP_HESTON_STEP / 1e12 / FP64_PEAK_TFLOPS}
        if name == "fast":
            out.update(leg)
        else:
            out[name] = leg
    if cpu:
        out["cpu_baseline"] = heston_cpu_baseline()
        out["speedup_vs_cpu_reference_draws"] = (out["reference_draws"]["path_steps_per_s"]
                                                 / out["cpu_baseline"]["value"])
    return out


def heston_greeks_leg(rng_mode):
    """Batched Heston sensitivities: 1024 options, 2^18 paths, steps
    ceil(252 T); d price / d {S0, v0, kappa, theta, xi, rho} by the adjoint
    kernel, by the forward tangent and by bump-and-revalue (central
    differences with common random numbers, 2k+1 = 13 pricing passes of the
    whole batch in one launch, relative bump 1e-4)."""
    from openferric_amd import _lib, engine_cfg, price_raw
    opts = heston_greeks_workload()
    structs, keeps = zip(*[to_ofg(d) for d in opts])
    structs = list(structs)
    cfg = engine_cfg(rng_mode=rng_mode)
    price_raw(structs, cfg)  # warm-up
    p = price_raw(structs, cfg)
    a = price_raw(structs, cfg, greeks=True)
    fwd_cfg = engine_cfg(rng_mode=rng_mode, greek_mode=_lib.GREEKS_FORWARD)
    f = price_raw(structs, fwd_cfg)
    fd = price_raw(structs, engine_cfg(rng_mode=rng_mode, greek_mode=_lib.GREEKS_FD_CRN,
                                       bump_rel=1e-4))
    rel = [abs(y - x) / max(abs(x), 1e-12) for ra, rf in zip(a, f)
           for x, y in zip(ra.heston_grad, rf.heston_grad)]
    # adjoint vs bump on the Europeans (a barrier's pathwise estimate omits
    # the crossing term by construction), per sensitivity, relative to the
    # option's price scale
    rel_fd = {}
    names = ["s0", "v0", "kappa", "theta", "xi", "rho"]
    for k, name in enumerate(names):
        rel_fd[name] = float(np.median([
            abs(ra.heston_grad[k] - rb.heston_grad[k]) / max(abs(rb.heston_grad[k]), 1e-3)
            for d, ra, rb in zip(opts, a, fd) if d["payoff"] == "vanilla"]))
    t_p, t_a, t_f, t_fd = p[0].kernel_ms, a[0].kernel_ms, f[0].kernel_ms, fd[0].kernel_ms
    ps = sum(d["num_paths"] * d["num_steps"] for d in opts)
    n_greeks = 6 * len(opts)
    return {"workload": "1024 Heston European/up-and-out calls, 2^18 paths, steps ceil(252T)",
            "path_steps": ps, "pricing_ms": t_p, "adjoint_ms": t_a,
            "pricing_path_steps_per_s": ps / (t_p * 1e-3),
            "pricing_frac_fp64_peak": ps / (t_p * 1e-3) * FLOP_HESTON_STEP / 1e12
            / FP64_PEAK_TFLOPS,
            "adjoint_cost_in_pricing_passes": t_a / t_p,
            "adjoint_greeks_per_s": n_greeks / (t_a * 1e-3),
            # the same sensitivities by a forward tangent of the Euler step
            "forward_ms": t_f,
            "forward_cost_in_pricing_passes": t_f / t_p,
            "forward_greeks_per_s": n_greeks / (t_f * 1e-3),
            "forward_vs_adjoint_max_rel_diff": max(rel),
            "bump_ms": t_fd, "bump_passes": 13, "bump_greeks_per_s": n_greeks / (t_fd * 1e-3),
            "adjoint_speedup_vs_bump": t_fd / t_a, "forward_speedup_vs_bump": t_fd / t_f,
            "adjoint_vs_bump_median_rel_diff_europeans": rel_fd}


# Longstaff & Schwartz (2001) Table 1: American puts, K = 40, r = 6 %, 50
# exercise steps (tests/lsm_reference.rs of the reference)
LSM_TABLE1 = [(s, v, t) for s in (36.0, 38.0, 40.0, 42.0, 44.0) for v in (0.2, 0.4)
              for t in (1.0, 2.0)]
LSM_STEPS = 50
LSM_ROUND_BYTES = 32      # per path and backward round: spots at two dates + value in/out
LSM_FINAL_BYTES = 16      # last round: spot at the first date + value in


def lsm_structs(paths):
    from openferric_amd import LongstaffSchwartzEngine, Market, VanillaOption
    from openferric_amd.engines import _Keep, lsm_option_struct
    keep = _Keep()
    out = []
    for s, v, t in LSM_TABLE1:
        m = Market.builder().spot(s).rate(0.06).flat_vol(v).build()
        out.append(lsm_option_struct(VanillaOption.american_put(40.0, t), m, paths, LSM_STEPS,
                                     42, keep))
    return out, keep


def lsm_leg(paths=1 << 20, reps=3, cpu=True):
    """Batched Longstaff-Schwartz: the 20 Table 1 American puts, 50 steps."""
    from openferric_amd import RngMode, engine_cfg, lsm_last_timing, lsm_price_raw
    structs, keep = lsm_structs(paths)
    n_ex = LSM_STEPS - 1
    round_bytes = len(structs) * paths * (LSM_ROUND_BYTES * (n_ex - 1) + LSM_FINAL_BYTES)
    ps = len(structs) * paths * LSM_STEPS
    out = {"workload": f"20 American puts (Longstaff-Schwartz Table 1), {paths} paths x "
                       f"{LSM_STEPS} steps each, one batched call", "path_steps": ps,
           "round_launches": n_ex, "round_bytes": round_bytes}
    for name, mode in (("fast", RngMode.Fast), ("reference", RngMode.Reference)):
        cfg = engine_cfg(rng_mode=mode)
        lsm_price_raw(structs, cfg)   # warm-up (allocations)
        tot, pth, rnd = [], [], []
        for _ in range(reps):
            res = lsm_price_raw(structs, cfg)
            t = lsm_last_timing()
            tot.append(res[0].kernel_ms)
            pth.append(t["paths_ms"])
            rnd.append(t["rounds_ms"])
        k, pm, rm = float(np.mean(tot)), float(np.mean(pth)), float(np.mean(rnd))
        gbs = round_bytes / (rm * 1e-3) / 1e9
        out[name] = {"kernel_ms": k, "paths_ms": pm, "rounds_ms": rm,
                     "options_per_s": len(structs) / (k * 1e-3), "path_steps_per_s": ps / (k * 1e-3),
                     "path_kernel_path_steps_per_s": ps / (pm * 1e-3),
                     "rounds_achieved_gb_s": gbs, "rounds_frac_hbm_peak": gbs / 8000.0,
                     "prices_first3": [r.price for r in res[:3]]}
    # the reference test's own size (tests/lsm_reference.rs: 100k paths, seed
    # 42), reference draws: the same paths the reference engine simulates
    small, keep_small = lsm_structs(100_000)
    cfg = engine_cfg(rng_mode=RngMode.Reference)
    lsm_price_raw(small, cfg)
    ms = [lsm_price_raw(small, cfg)[0].kernel_ms for _ in range(reps)]
    out["table1_reference_size"] = {
        "workload": "the 20 Table 1 puts at 100000 paths x 50 steps, reference draws",
        "kernel_ms": float(np.mean(ms)),
        "path_steps_per_s": len(small) * 100_000 * LSM_STEPS / (float(np.mean(ms)) * 1e-3)}
    if cpu:
        out["cpu_baseline"] = lsm_cpu_baseline()
        out["speedup_vs_cpu_reference_draws"] = (
            out["reference"]["path_steps_per_s"] / out["cpu_baseline"]["value"])
    return out


def lsm_cpu_baseline(paths=100_000):
    """The oracle's LongstaffSchwartzEngine restatement (sequential, as the
    reference) on Table 1 options at the reference test's size, one option
    per host thread."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    O.build()
    threads = cpu_threads()
    cases = [LSM_TABLE1[i % len(LSM_TABLE1)] for i in range(4 * threads)]
    structs = [O.make_lsm_option(spot=s, strike=40.0, vol=v, expiry=t, rate=0.06,
                                 num_paths=paths, num_steps=LSM_STEPS, seed=42)
               for s, v, t in cases]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(O.lsm_price, structs))
    dt = time.perf_counter() - t0
    ps = len(structs) * paths * LSM_STEPS
    return {"value": ps / dt, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(structs)} Table 1 American puts at {paths} paths x {LSM_STEPS} "
                      f"steps, {threads} threads (one option each at a time), {dt:.2f} s, "
                      f"oracle C restatement"}


QMC_OPTS, QMC_POINTS, QMC_DIMS = 64, 1 << 20, 64


def qmc_workload(seed=WORKLOAD_SEED + 6):
    """European calls/puts, K U[80,120], sigma U[0.1,0.5], r U[0,0.06],
    T U[0.25,2], S0 = 100, each on its own Sobol seed."""
    rng = np.random.default_rng(seed)
    return [dict(option_type="call" if i % 2 == 0 else "put", spot=100.0,
                 strike=float(rng.uniform(80, 120)), vol=float(rng.uniform(0.1, 0.5)),
                 rate=float(rng.uniform(0.0, 0.06)), expiry=float(rng.uniform(0.25, 2.0)),
                 seed=1000 + i) for i in range(QMC_OPTS)]


def qmc_leg(reps=3, cpu=True):
    """Sobol quasi-Monte Carlo: 64 European options, 2^20 points x 64
    dimensions each, one launch (mc_european_qmc_with_seed per option)."""
    from openferric_amd import _lib, qmc_price_raw
    opts = qmc_workload()
    structs = []
    for d in opts:
        o = _lib.OfgOption()
        o.model, o.payoff = _lib.MODEL_GBM, _lib.PAYOFF_VANILLA
        o.option_type = _lib.CALL if d["option_type"] == "call" else _lib.PUT
        o.spot, o.strike, o.expiry, o.rate, o.vol = d["spot"], d["strike"], d["expiry"], \
            d["rate"], d["vol"]
        o.num_paths, o.num_steps, o.seed = QMC_POINTS, QMC_DIMS, d["seed"]
        structs.append(o)
    qmc_price_raw(structs)
    ms = [qmc_price_raw(structs)[0].kernel_ms for _ in range(reps)]
    k = float(np.mean(ms))
    ps = QMC_OPTS * QMC_POINTS * QMC_DIMS
    v = ps / (k * 1e-3)
    out = {"workload": f"{QMC_OPTS} European options, 2^20 Sobol points x {QMC_DIMS} "
                       "dimensions each (reference Sobol points), one launch",
           "path_steps": ps, "kernel_ms": k, "path_steps_per_s": v,
           "frac_fp64_peak": v * FLOP_GBM_STEP / 1e12 / FP64_PEAK_TFLOPS}
    if cpu:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle as O
        O.build()
        threads = cpu_threads()
        pts = 1 << 17
        sample = opts[:2 * threads]

        def one(d):
            oo = O.make_option(option_type=O.CALL if d["option_type"] == "call" else O.PUT,
                               spot=d["spot"], strike=d["strike"], expiry=d["expiry"],
                               rate=d["rate"], vol=d["vol"])
            return O.mc_european_qmc(oo, pts, QMC_DIMS, d["seed"])

        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(one, sample))
        dt = time.perf_counter() - t0
        cps = len(sample) * pts * QMC_DIMS / dt
        out["cpu_baseline"] = {"value": cps, "unit": "path-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{len(sample)} options at 2^17 points x {QMC_DIMS} "
                                         f"dimensions, one per thread, {dt:.2f} s, oracle C "
                                         "restatement (sequential next_into as the reference)"}
        out["speedup_vs_cpu"] = v / cps
    return out


AC_CONTRACTS, AC_UNDERLYINGS, AC_PATHS, AC_STEPS = 16, 4, 1 << 18, 252
BK_BASKETS, BK_ASSETS, BK_PATHS = 32, 8, 1 << 20


def flop_multi_asset_step(n):
    """Canonical FLOP per asset-step of a correlated path: the GBM step's 65
    (normal, exp, move) + the average Cholesky row, (n + 1) / 2 FMAs."""
    return FLOP_GBM_STEP + (n + 1)


def _corr_matrix(n, rho, seed):
    rng = np.random.default_rng(seed)
    c = np.full((n, n), rho) + rng.uniform(-0.1, 0.1, (n, n))
    c = (c + c.T) / 2
    np.fill_diagonal(c, 1.0)
    assert np.linalg.eigvalsh(c).min() > 0.05
    return c


def autocall_workload(seed=WORKLOAD_SEED + 7):
    """Worst-of Phoenix notes with memory on 4 of 12 underlyings, quarterly
    observations over 3 years, daily steps."""
    from openferric_amd import PhoenixAutocallable
    rng = np.random.default_rng(seed)
    market = dict(spots=list(rng.uniform(50, 150, 12)), vols=list(rng.uniform(0.15, 0.4, 12)),
                  corr=_corr_matrix(12, 0.45, seed), r=0.03, q=0.01)
    contracts = []
    for i in range(AC_CONTRACTS):
        und = [int(x) for x in rng.choice(12, AC_UNDERLYINGS, replace=False)]
        contracts.append(PhoenixAutocallable(
            und, 100.0, [0.25 * k for k in range(1, 13)], float(rng.uniform(1.0, 1.1)),
            float(rng.uniform(0.6, 0.8)), float(rng.uniform(0.06, 0.12)), True,
            float(rng.uniform(0.5, 0.65)), 1.0, 3.0))
    return market, contracts


def basket_workload(seed=WORKLOAD_SEED + 8):
    """Average / best-of / worst-of calls and puts on 8 assets each."""
    from openferric_amd import BasketOption
    rng = np.random.default_rng(seed)
    out = []
    for i in range(BK_BASKETS):
        kind = i % 3
        w = list(rng.dirichlet(np.ones(BK_ASSETS))) if kind == 0 else []
        strike = float(rng.uniform(90, 110)) if kind == 0 else float(rng.uniform(0.9, 1.1))
        out.append(dict(basket=BasketOption(w, strike, float(rng.uniform(0.5, 3.0)),
                                            bool(i % 2 == 0), kind),
                        spots=list(rng.uniform(50, 150, BK_ASSETS)),
                        vols=list(rng.uniform(0.15, 0.45, BK_ASSETS)),
                        divs=list(rng.uniform(0.0, 0.03, BK_ASSETS)),
                        corr=_corr_matrix(BK_ASSETS, 0.35, seed + i), r=0.02))
    return out


def multi_leg(reps=3, cpu=True):
    """Multi-asset contracts, each priced with its bump-and-reprice
    sensitivities (every bump a scenario on the base price's draws) in one
    launch: worst-of Phoenix autocallables (fast draws) and baskets (the
    reference's single stream, split by jump-ahead)."""
    from openferric_amd import RngMode, autocall_price_raw, autocall_struct, basket_price_raw
    from openferric_amd import basket_struct
    from openferric_amd.engines import _Keep
    out = {}
    # ---- autocallables
    m, contracts = autocall_workload()
    keep = _Keep()
    structs = [autocall_struct(c, m["spots"], m["vols"], m["corr"], m["r"], m["q"], AC_PATHS,
                               AC_STEPS, keep) for c in contracts]
    autocall_price_raw(structs, RngMode.Fast, None, True)
    ms = [autocall_price_raw(structs, RngMode.Fast, None, True)[0][0].kernel_ms
          for _ in range(reps)]
    k = float(np.mean(ms))
    scen = AC_CONTRACTS * (1 + 2 * AC_UNDERLYINGS + 4)
    paths = scen * AC_PATHS
    nominal = paths * AC_STEPS * AC_UNDERLYINGS
    ac = {"workload": f"{AC_CONTRACTS} worst-of Phoenix notes with memory, {AC_UNDERLYINGS} "
                      f"underlyings, 12 quarterly observations, {AC_STEPS} steps, {AC_PATHS} "
                      f"paths, price + delta/vega/cega ({scen} scenarios), fast draws, one launch",
          "kernel_ms": k, "scenario_paths_per_s": paths / (k * 1e-3),
          "nominal_asset_steps_per_s": nominal / (k * 1e-3),
          "note": "called paths stop early, so nominal asset-steps overstate the work"}
    if cpu:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle as O
        O.build()
        threads = cpu_threads()
        cpaths = 1 << 16
        sample = (contracts * 4)[:2 * threads]

        def one(c):
            u = list(c.underlyings)
            a = O.make_autocall(spots=[m["spots"][i] for i in u], vols=[m["vols"][i] for i in u],
                                corr=np.asarray(m["corr"])[np.ix_(u, u)],
                                dates=list(c.autocall_dates), notional=c.notional,
                                maturity=c.maturity, autocall_barrier=c.autocall_barrier,
                                coupon_rate=c.coupon_rate, coupon_barrier=c.coupon_barrier,
                                ki_barrier=c.ki_barrier, ki_strike=c.ki_strike, rate=m["r"],
                                div_yield=m["q"], phoenix=True, memory=True,
                                num_paths=cpaths, num_steps=AC_STEPS)
            return O.autocall_price(a, None)

        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(one, sample))
        dt = time.perf_counter() - t0
        cps = len(sample) * cpaths / dt
        ac["cpu_baseline"] = {"value": cps, "unit": "scenario-paths/s", "cores": threads,
                              "kind": "port",
                              "sample": f"{len(sample)} base prices at {cpaths} paths, one per "
                                        f"thread, {dt:.2f} s, oracle C restatement (the "
                                        "reference's single stream)"}
        ac["speedup_vs_cpu"] = ac["scenario_paths_per_s"] / cps
    out["autocallable"] = ac
    # ---- baskets
    bw = basket_workload()
    keep = _Keep()
    bstructs = [basket_struct(d["basket"], d["spots"], d["vols"], d["corr"], d["r"], d["divs"],
                              BK_PATHS, keep) for d in bw]
    bscen = BK_BASKETS * (1 + 2 * BK_ASSETS + 4)
    bk = {"workload": f"{BK_BASKETS} baskets (average / best-of / worst-of) on {BK_ASSETS} "
                      f"assets, {BK_PATHS} paths, price + basket_sensitivities ({bscen} "
                      "scenarios), one launch"}
    for mode, name in ((RngMode.Reference, "reference_draws"), (RngMode.Fast, "fast_draws")):
        basket_price_raw(bstructs, mode, None, True)
        ms = [basket_price_raw(bstructs, mode, None, True)[0][0].kernel_ms for _ in range(reps)]
        k = float(np.mean(ms))
        sp = bscen * BK_PATHS / (k * 1e-3)
        bk[name] = {"kernel_ms": k, "scenario_paths_per_s": sp}
    if cpu:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle as O
        threads = cpu_threads()
        cpaths = 1 << 22
        sample = (bw * 4)[:2 * threads]

        def oneb(d):
            b = d["basket"]
            ob = O.make_basket(d["spots"], d["vols"], d["divs"], d["corr"],
                               b.weights if b.weights else None, b.basket_type, b.is_call,
                               b.strike, b.maturity, d["r"], cpaths)
            return O.basket_price(ob, None)

        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(oneb, sample))
        dt = time.perf_counter() - t0
        cps = len(sample) * cpaths / dt
        bk["cpu_baseline"] = {"value": cps, "unit": "scenario-paths/s", "cores": threads,
                              "kind": "port",
                              "sample": f"{len(sample)} base prices at {cpaths} paths, one per "
                                        f"thread, {dt:.2f} s, oracle C restatement (the "
                                        "reference's single stream)"}
        bk["speedup_vs_cpu"] = bk["reference_draws"]["scenario_paths_per_s"] / cps
    bk["note"] = ("the sensitivities share each path's draws, correlated normals and terminal "
                  "factors across the 21 scenarios of a basket, so scenario-paths/s counts "
                  "scenarios priced, not independent simulations")
    out["basket"] = bk
    return out


DSL_PATHS, DSL_STEPS = 1 << 20, 378
DSL_PRODUCT = """\
product "WoF Autocall 18m"
    notional: 1_000_000
    maturity: 1.5
    underlyings
        SPX = asset(0)
        SX5E = asset(1)
        NKY = asset(2)
    state
        ki_hit: bool = false
    schedule quarterly from 0.25 to 1.5
        let wof = worst_of(performances())
        if wof <= 0.60 then
            set ki_hit = true
        if wof >= 1.0 and not is_final then
            pay notional * 0.08 * observation_date
            redeem notional
        if is_final then
            pay notional * 0.08 * 1.5
            if ki_hit and wof < 1.0 then
                redeem notional * wof
            else
                redeem notional
"""


def dsl_leg(reps=3, cpu=True):
    """Structured-product language: a worst-of autocallable on three
    correlated equities, 2^20 paths x 378 daily steps, the reference's
    single stream (jump-ahead) and fast draws; price alone and with
    greeks_multi_asset (5 markets in one launch)."""
    from openferric_amd import (AssetMarketData as A, DslMonteCarloEngine, MultiAssetMarket,
                                RngMode, parse_and_compile)
    prod = parse_and_compile(DSL_PRODUCT)
    corr = [[1.0, 0.7, 0.5], [0.7, 1.0, 0.6], [0.5, 0.6, 1.0]]
    mkt = MultiAssetMarket([A.Equity(100.0, 0.20, 0.02), A.Equity(100.0, 0.22, 0.03),
                            A.Equity(100.0, 0.25, 0.01)], corr, 0.03)
    asset_steps = DSL_PATHS * DSL_STEPS * 3
    out = {"workload": f"worst-of autocallable (3 correlated equities, quarterly over 1.5 y, "
                       f"knock-in), {DSL_PATHS} paths x {DSL_STEPS} steps",
           "asset_steps": asset_steps}
    def timed(mode):
        eng = DslMonteCarloEngine(DSL_PATHS, DSL_STEPS, 42, rng_mode=mode)
        eng.price_multi_asset(prod, mkt)
        ms = []
        for _ in range(reps):
            ms.append(eng.price_multi_asset(prod, mkt).kernel_ms)
        k = float(np.mean(ms))
        eng.greeks_multi_asset(prod, mkt, 0)
        gms = []
        for _ in range(reps):
            eng.greeks_multi_asset(prod, mkt, 0)
            gms.append(eng.last_kernel_ms)
        g = float(np.mean(gms))
        return {"price_ms": k, "asset_steps_per_s": asset_steps / (k * 1e-3),
                "greeks_5_markets_ms": g,
                "greeks_asset_steps_per_s": 5 * asset_steps / (g * 1e-3),
                "frac_fp64_peak": asset_steps / (k * 1e-3) * flop_multi_asset_step(3) / 1e12
                / FP64_PEAK_TFLOPS}

    out["frac_note"] = ("frac_fp64_peak counts the reference's step (an exponential per "
                        "asset-step); the fast mode's GBM-only kernels step log growths and "
                        "take the exponential per stored snapshot")
    modes = ((RngMode.Reference, "reference_draws"), (RngMode.Fast, "fast_draws"))
    # the program interpreter (the default), then the product's native
    # evaluation kernel (compiled at run time, the same prices bit for bit)
    for mode, name in modes:
        out[name] = timed(mode)
    try:
        native = {"compile_ms": prod.jit_compile(3)}
        for mode, name in modes:
            native[name] = timed(mode)
    except Exception as e:   # a missing run-time compiler must not cost the bench line
        native = {"error": f"{type(e).__name__}: {e}"[:300]}
    out["native_kernel"] = native
    if cpu:
        # the oracle's compiled correlated-path pricer (or_autocall_price: the
        # same three correlated GBM assets over 378 steps drawn path after
        # path from the reference's single xoshiro256++ stream, Acklam
        # normals, Cholesky factor, exact exponentials), one contract per
        # host thread; its autocall barrier is out of reach so that every
        # path is simulated to maturity, as the product-language engine
        # simulates each path whole before it runs the program on it
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle as O
        O.build()
        threads = cpu_threads()
        cpaths = 1 << 15
        dates = [0.25 * k for k in range(1, 7)]

        def one(i):
            a = O.make_autocall(spots=[100.0] * 3, vols=[0.20, 0.22, 0.25], corr=corr,
                                dates=dates, notional=1.0, maturity=1.5, autocall_barrier=1e300,
                                coupon_rate=0.08, ki_barrier=0.60, ki_strike=1.0, rate=0.03,
                                div_yield=0.02, num_paths=cpaths, num_steps=DSL_STEPS,
                                seed=1000 + i)
            return O.autocall_price(a, None)

        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(one, range(threads)))
        dt = time.perf_counter() - t0
        cps = threads * cpaths * DSL_STEPS * 3 / dt
        out["cpu_baseline"] = {"value": cps, "unit": "asset-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{threads} x {cpaths} paths x {DSL_STEPS} steps x 3 "
                                         f"assets, one per thread, {dt:.2f} s: the oracle's "
                                         "compiled correlated-path pricer (or_autocall_price) on "
                                         "the reference's single stream, every path simulated to "
                                         "maturity; the product's program (a few operations "
                                         "per observation date) is not run"}
        out["speedup_vs_cpu"] = out["reference_draws"]["asset_steps_per_s"] / cps
    return out


def reference_draws_leg(opts, reps=2):
    """The batched GBM workload on the reference's own draws
    (OFG_RNG_REFERENCE: per-sample xoshiro256++ streams and the Acklam
    inverse CDF, per-path parity with MonteCarloEngine::run): one persistent
    launch replayed from a hipGraph, as the headline."""
    from openferric_amd import BatchPricer, RngMode, engine_cfg
    structs, keeps = zip(*[to_ofg(d) for d in opts])
    b = BatchPricer(list(structs), engine_cfg(rng_mode=RngMode.Reference), use_graph=True)
    b.run()
    ms = [b.run() for _ in range(reps)]
    b.close()
    ps, flops = path_steps_and_flops(opts)
    k = float(np.mean(ms))
    return {"workload": "the batched GBM workload (4096 options) on the reference's draws",
            "kernel_ms": k, "path_steps_per_s": ps / (k * 1e-3),
            "frac_fp64_peak": flops / (k * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}


def host_call_leg(structs, cfg, ps, batch_prices, reps=2):
    """The same batch through one ofg_price call on host arrays, as a caller
    holding the descriptors in host memory makes it: the call uploads the
    descriptors and its work list, allocates its device buffers, launches the
    persistent kernel and the reduction and copies the results back. Wall
    time per call after one untimed call (PCIe and allocation included)."""
    from openferric_amd.engines import price_raw
    price_raw(list(structs), cfg)
    walls, res = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        res = price_raw(list(structs), cfg)
        walls.append(time.perf_counter() - t0)
    w = float(np.mean(walls))
    prices = np.array([r.price for r in res])
    return {"workload": "the headline batch (4096 options) through one ofg_price call on host "
                        "buffers per repetition",
            "wall_ms": 1000.0 * w, "path_steps_per_s": ps / w,
            "max_abs_diff_vs_batch": float(np.max(np.abs(prices - batch_prices)))}


CONFIG1 = dict(payoff="vanilla", option_type="call", spot=100.0, strike=100.0, expiry=1.0,
               rate=0.05, q=0.0, vol=0.2, num_steps=252, num_paths=100_000, obs_times=None)


def config1_leg(cpu=True, reps=5):
    """The reference's own CPU-runnable case: a European call under GBM,
    S0 = K = 100, r = 5 %, sigma = 20 %, T = 1, 100,000 antithetic paths x 252
    steps, seed 42, on the reference's draws (OFG_RNG_REFERENCE), checked
    against the oracle on the same draws and against Black-Scholes; the CPU
    baseline is the oracle on the whole case."""
    from openferric_amd import RngMode, engine_cfg, price_raw
    o, keep = to_ofg(CONFIG1, seed=42)
    out = {"workload": "European call under GBM, 100000 antithetic paths x 252 steps, "
                       "seed 42, reference draws"}
    ps = CONFIG1["num_paths"] * CONFIG1["num_steps"]
    for name, mode in (("reference_draws", RngMode.Reference), ("fast_draws", RngMode.Fast)):
        cfg = engine_cfg(antithetic=True, rng_mode=mode)
        price_raw([o], cfg)
        rs = [price_raw([o], cfg)[0] for _ in range(reps)]
        k = float(np.mean([r.kernel_ms for r in rs]))
        out[name] = {"price": rs[0].price, "stderr": rs[0].stderr_, "kernel_ms": k,
                     "path_steps_per_s": ps / (k * 1e-3)}
    if cpu:
        from oracle import oracle as O
        O.build()
        threads = cpu_threads()
        oo = O.make_option(spot=100.0, strike=100.0, expiry=1.0, rate=0.05, vol=0.2)
        O.mc_run(oo, 1000, 252, 42, antithetic=True, nthreads=threads)
        t0 = time.perf_counter()
        n_rep = 0
        while True:
            p_or, se_or = O.mc_run(oo, 100_000, 252, 42, antithetic=True, nthreads=threads)
            n_rep += 1
            dt = time.perf_counter() - t0
            if dt >= 2.0:
                break
        v = n_rep * ps / dt
        out["cpu_baseline"] = {
            "value": v, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"the whole case, {n_rep} repetitions in {dt:.2f} s, oracle C "
                      f"restatement of MonteCarloEngine::run, OpenMP {threads} threads",
            "gflops_canonical": v * FLOP_GBM_STEP / 1e9}
        out["oracle_price"] = p_or
        out["gpu_vs_oracle_rel_diff"] = abs(out["reference_draws"]["price"] - p_or) / p_or
        out["speedup_vs_cpu_reference_draws"] = out["reference_draws"]["path_steps_per_s"] / v
    out["black_scholes"] = 10.450583572185565
    return out


def heston_cpu_baseline(target_s=10.0):
    """The oracle (C restatement of MonteCarloEngine::run with
    HestonPathGenerator, all host threads) on a fixed subset of the Heston
    barrier workload: every 32nd option, paths scaled to ~target_s."""
    from oracle import oracle as O
    O.build()
    sub = heston_workload()[::32]
    threads = cpu_threads()

    def run(paths):
        t0 = time.perf_counter()
        for d in sub:
            oo = O.make_option(model=O.HESTON, payoff=O.BARRIER, spot=d["spot"],
                               strike=d["strike"], expiry=d["expiry"], rate=d["rate"],
                               kappa=d["kappa"], theta=d["theta"], xi=d["xi"], rho=d["rho"],
                               v0=d["v0"], barrier_dir=O.UP, barrier_style=O.OUT,
                               barrier_level=d["barrier_level"])
            O.mc_run(oo, paths, d["num_steps"], 42, nthreads=threads)
        return len(sub) * paths * 365, time.perf_counter() - t0

    paths = 1 << 12
    for _ in range(6):
        ps, dt = run(paths)
        if dt >= 0.5 * target_s or paths >= (1 << 20):
            break
        paths = min(1 << 20, int(paths * target_s / max(dt, 1e-3)))
    v = ps / dt
    return {"value": v, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(sub)} Heston up-and-out calls (every 32nd of the workload) at "
                      f"{paths} paths x 365 steps, {ps:.3e} path-steps in {dt:.1f} s, oracle C "
                      f"restatement, OpenMP {threads} threads",
            "gflops_canonical": v * FLOP_HESTON_STEP / 1e9}


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB (a plain copy kernel of known
    size for scaling the FETCH_SIZE / WRITE_SIZE counters)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n}), flush=True)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(opts, target_s=15.0):
    """Times the oracle (C restatement of MonteCarloEngine::run, all host
    threads) on a fixed subset of the workload: every 8th option, paths
    scaled so the measured sample is ~target_s of CPU work (the scale is
    found by short calibration runs, the last run is the one reported)."""
    from oracle import oracle as O
    O.build()
    sub = opts[::8]
    threads = cpu_threads()

    def run(scale):
        ds = [dict(d, num_paths=max(64, int(d["num_paths"] * scale))) for d in sub]
        ps = 0
        t0 = time.perf_counter()
        for d in ds:
            oo = O.make_option(payoff=O.ASIAN if d["payoff"] == "asian" else O.VANILLA,
                               option_type=O.CALL if d["option_type"] == "call" else O.PUT,
                               spot=d["spot"], strike=d["strike"], expiry=d["expiry"],
                               rate=d["rate"], div_yield=d["q"], vol=d["vol"],
                               obs_times=d["obs_times"])
            O.mc_run(oo, d["num_paths"], d["num_steps"], 42, nthreads=threads)
            ps += d["num_paths"] * d["num_steps"]
        return ps, time.perf_counter() - t0

    scale = 1.0 / 256
    for _ in range(6):
        ps, dt = run(scale)
        if dt >= 0.5 * target_s or scale >= 1.0:
            break
        scale = min(1.0, scale * target_s / max(dt, 1e-3))
    return {"value": ps / dt, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(sub)} options (every 8th of the batch), paths x{scale:.4f}, "
                      f"{ps:.3e} path-steps in {dt:.1f} s, oracle C restatement, "
                      f"OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default launch shape (block 256, 8192 lane-steps per item)")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--path-scale", type=float, default=1.0,
                    help="scale per-option path counts (profiling runs only)")
    ap.add_argument("--no-greeks", action="store_true", help="skip the adjoint Greeks leg")
    ap.add_argument("--no-heston", action="store_true", help="skip the Heston barrier pricing leg")
    ap.add_argument("--no-heston-greeks", action="store_true",
                    help="skip the batched Heston sensitivities leg")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the single European (100000 antithetic paths x 252 steps) leg")
    ap.add_argument("--no-reference-draws", action="store_true",
                    help="skip the batch on the reference's draws")
    ap.add_argument("--no-host-call", action="store_true",
                    help="skip the whole-batch ofg_price call from host buffers")
    ap.add_argument("--no-lsm", action="store_true", help="skip the Longstaff-Schwartz leg")
    ap.add_argument("--lsm-only", action="store_true",
                    help="run only the Longstaff-Schwartz leg (profiling)")
    ap.add_argument("--no-qmc", action="store_true", help="skip the quasi-Monte Carlo leg")
    ap.add_argument("--qmc-only", action="store_true",
                    help="run only the quasi-Monte Carlo leg (profiling)")
    ap.add_argument("--no-multi", action="store_true",
                    help="skip the autocallable / basket leg")
    ap.add_argument("--multi-only", action="store_true",
                    help="run only the autocallable / basket leg (profiling)")
    ap.add_argument("--no-dsl", action="store_true", help="skip the structured-product leg")
    ap.add_argument("--dsl-only", action="store_true",
                    help="run only the structured-product leg (profiling)")
    ap.add_argument("--heston-only", action="store_true",
                    help="run only the two Heston legs (profiling)")
    ap.add_argument("--shape", default="",
                    help="BLOCK,ITEM: launch shape of the batch instead of the autotuner's "
                         "(counter runs of one instantiation)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run one D2D copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 "
                                   "(a batch would shard by work item across GPUs, each "
                                   "GPU running the persistent kernel on its share; not "
                                   "measured here)"}))
        sys.exit(2)

    from openferric_amd import BatchPricer, autotune, engine_cfg, RngMode
    from openferric_amd import _lib

    if args.lsm_only:
        print(json.dumps({"lsm": lsm_leg(cpu=not args.no_cpu)}), flush=True)
        return
    if args.qmc_only:
        print(json.dumps({"qmc": qmc_leg(cpu=not args.no_cpu)}), flush=True)
        return
    if args.multi_only:
        print(json.dumps({"multi": multi_leg(cpu=not args.no_cpu)}), flush=True)
        return
    if args.dsl_only:
        print(json.dumps({"dsl": dsl_leg(cpu=not args.no_cpu)}), flush=True)
        return
    if args.heston_only:
        print("[bench] heston barrier leg", file=sys.stderr, flush=True)
        h = heston_leg(RngMode.Fast, cpu=not args.no_cpu)
        print("[bench] heston sensitivities leg", file=sys.stderr, flush=True)
        print(json.dumps({"heston": h, "heston_greeks": heston_greeks_leg(RngMode.Fast)}),
              flush=True)
        return

    opts = make_workload(path_scale=args.path_scale)
    structs, keeps = zip(*[to_ofg(d) for d in opts])
    cfg = engine_cfg(antithetic=False, control_variate=False, rng_mode=RngMode.Fast)
    tuned = None
    if args.shape:
        tuned = tuple(int(x) for x in args.shape.split(","))
        cfg = engine_cfg(antithetic=False, control_variate=False, rng_mode=RngMode.Fast,
                         tune_block=tuned[0], tune_item_steps=tuned[1])
    elif not args.no_autotune:
        # block size x work-item size, timed on a scaled copy of the batch; the
        # choice is remembered by the context and used by the batch below
        print("[bench] autotuning launch shape", file=sys.stderr, flush=True)
        tuned = autotune(list(structs), cfg)
        print(f"[bench] autotuned: block {tuned[0]}, item {tuned[1]} lane-steps",
              file=sys.stderr, flush=True)
    print(f"[bench] building batch of {len(opts)} options", file=sys.stderr, flush=True)
    batch = BatchPricer(list(structs), cfg, use_graph=not args.no_graph)
    for w in range(args.warmup):
        batch.run()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    kernel_ms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        kernel_ms.append(batch.run())
        print(f"[bench] step {s + 1}/{args.steps}: {kernel_ms[-1]:.1f} ms", file=sys.stderr,
              flush=True)
    wall = time.perf_counter() - t0
    res = batch.results()
    prices = np.array([r.price for r in res])
    if not np.all(np.isfinite(prices)):
        print("[bench] non-finite prices", file=sys.stderr)
        sys.exit(3)

    ps, flops = path_steps_and_flops(opts)
    ms_per_step = 1000.0 * wall / args.steps
    value = ps / (wall / args.steps)
    k_ms = float(np.mean(kernel_ms))
    achieved = flops / (k_ms * 1e-3) / 1e12
    out = {
        "metric": "FP64 path-steps/s (frac of VALU peak)",
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched GBM: 4096 options (2048 European + 2048 arithmetic "
                               "Asian, daily observation), paths floor(10^U[4,6]), steps "
                               "U{12..504}, plain MC, fast RNG (xoroshiro128+ + Box-Muller), "
                               "one persistent launch replayed from a hipGraph",
                   "options": len(opts), "path_steps": ps, "graph": not args.no_graph,
                   "path_scale": args.path_scale,
                   "autotuned": None if tuned is None else
                   {"block": tuned[0], "item_lane_steps": tuned[1]}},
        "roofline": {"bound": "valu_fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": PRICE_KERNEL_HBM_BYTES.get((256, 8192) if tuned is None
                                                           else tuple(tuned))
                     if args.path_scale == 1.0 else None,
                     # the instantiation launched: threads per block from the autotuner
                     # (the library's default shape, 256, without it)
                     "kernel": f"price_kernel<RNG_FAST, GBM, {256 if tuned is None else tuned[0]}>",
                     "kernel_ms_avg": k_ms,
                     "flop_model": "65 FLOP per GBM path-step, 66 with Asian running sum"},
    }
    if not args.no_host_call:
        print("[bench] host-buffer call", file=sys.stderr, flush=True)
        out["host_call"] = host_call_leg(structs, cfg, ps, prices)
    if not args.no_reference_draws:
        print("[bench] batch on the reference's draws", file=sys.stderr, flush=True)
        out["reference_draws"] = reference_draws_leg(opts)
    if not args.no_config1:
        print("[bench] single European leg", file=sys.stderr, flush=True)
        out["config1"] = config1_leg(cpu=not args.no_cpu)
    if not args.no_greeks:
        print("[bench] adjoint greeks leg", file=sys.stderr, flush=True)
        out["greeks"] = greeks_leg(RngMode.Fast)
    if not args.no_heston:
        print("[bench] heston leg", file=sys.stderr, flush=True)
        out["heston"] = heston_leg(RngMode.Fast, cpu=not args.no_cpu)
    if not args.no_heston_greeks:
        print("[bench] heston sensitivities leg", file=sys.stderr, flush=True)
        out["heston_greeks"] = heston_greeks_leg(RngMode.Fast)
    if not args.no_lsm:
        print("[bench] longstaff-schwartz leg", file=sys.stderr, flush=True)
        out["lsm"] = lsm_leg(cpu=not args.no_cpu)
    if not args.no_qmc:
        print("[bench] quasi-monte-carlo leg", file=sys.stderr, flush=True)
        out["qmc"] = qmc_leg(cpu=not args.no_cpu)
    if not args.no_multi:
        print("[bench] multi-asset leg", file=sys.stderr, flush=True)
        out["multi"] = multi_leg(cpu=not args.no_cpu)
    if not args.no_dsl:
        print("[bench] structured-product leg", file=sys.stderr, flush=True)
        out["dsl"] = dsl_leg(cpu=not args.no_cpu)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(opts)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    batch.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
