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
    o.model = _lib.MODEL_GBM
    o.payoff = _lib.PAYOFF_ASIAN if d["payoff"] == "asian" else _lib.PAYOFF_VANILLA
    o.option_type = _lib.CALL if d["option_type"] == "call" else _lib.PUT
    o.spot, o.strike, o.expiry = d["spot"], d["strike"], d["expiry"]
    o.rate, o.div_yield, o.vol = d["rate"], d["q"], d["vol"]
    if d["obs_times"] is not None:
        o.obs_times = keep.arr(d["obs_times"])
        o.n_obs = len(d["obs_times"])
        o.averaging, o.strike_type = _lib.ARITHMETIC, _lib.FIXED
    o.num_paths, o.num_steps, o.seed = d["num_paths"], d["num_steps"], seed
    return o, keep


def path_steps_and_flops(opts):
    ps = sum(d["num_paths"] * d["num_steps"] for d in opts)
    fl = sum(d["num_paths"] * d["num_steps"] * (FLOP_ASIAN_STEP if d["payoff"] == "asian"
                                                else FLOP_GBM_STEP) for d in opts)
    return ps, fl


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(opts, target_s=12.0):
    """Times the oracle (C restatement of MonteCarloEngine::run, all host
    threads) on a fixed subset of the workload: every 32nd option, paths
    scaled so the sample is ~target_s of CPU work."""
    from oracle import oracle as O
    O.build()
    sub = opts[::32]
    threads = cpu_threads()
    # calibrate on a small slice
    probe = [dict(d, num_paths=max(64, d["num_paths"] // 200)) for d in sub[:8]]

    def run(ds):
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

    ps, dt = run(probe)
    rate = ps / max(dt, 1e-9)
    full_ps, _ = path_steps_and_flops(sub)
    scale = min(1.0, target_s * rate / max(full_ps, 1))
    sample = [dict(d, num_paths=max(64, int(d["num_paths"] * scale))) for d in sub]
    ps, dt = run(sample)
    return {"value": ps / dt, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} options (every 32nd of the batch), paths x{scale:.4f}, "
                      f"{ps:.3e} path-steps in {dt:.1f} s, oracle C restatement, "
                      f"OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--path-scale", type=float, default=1.0,
                    help="scale per-option path counts (profiling runs only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md)"}))
        sys.exit(2)

    from openferric_amd import BatchPricer, engine_cfg, RngMode
    from openferric_amd import _lib

    opts = make_workload(path_scale=args.path_scale)
    structs, keeps = zip(*[to_ofg(d) for d in opts])
    cfg = engine_cfg(antithetic=False, control_variate=False, rng_mode=RngMode.Fast)
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
                               "U{12..504}, plain MC, fast RNG (xoshiro256++ + Box-Muller), "
                               "one persistent launch replayed from a hipGraph",
                   "options": len(opts), "path_steps": ps, "graph": not args.no_graph,
                   "path_scale": args.path_scale},
        "roofline": {"bound": "valu_fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "price_kernel<RNG_FAST,256>",
                     "kernel_ms_avg": k_ms,
                     "flop_model": "65 FLOP per GBM path-step, 66 with Asian running sum"},
    }
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(opts)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    batch.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
