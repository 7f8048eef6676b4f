"""Parity on subsets of the benchmark configurations, priced as batches in one
call through the C ABI (the configurations themselves are too large for the
CPU oracle):

* config 2 (batched GBM): every 64th option of the 4096-option workload
  (32 European, 32 arithmetic Asian with daily observation, steps 12..504,
  mixed calls and puts) at 4096 paths, reference draws, one launch: each
  option's price within 1e-11 of the oracle; 8 of them with external normals
  (the reference streams' draws laid out [step][sample]) likewise;
* config 3 (Heston up-and-out, daily monitoring): 8 options of the workload
  at 2^14 paths x 365 steps, reference draws, one launch, against the oracle.
"""
import numpy as np
import pytest

from tests.cases import make_pair, step_major

pytestmark = pytest.mark.gpu

PRICE_TOL = 1e-11


def _pair(d, paths, seed):
    kw = {k: v for k, v in d.items() if v is not None}
    kw.update(num_paths=paths, seed=seed)
    return make_pair(**kw)


def _check(oracle, pairs, got, paths, steps_of, antithetic=False, normals=None):
    for i, ((oo, g, keep), r) in enumerate(zip(pairs, got)):
        n1 = n2 = None
        if normals is not None:
            n1, n2 = normals[i]
        p_ref, se_ref = oracle.mc_run(oo, paths, steps_of(i), g.seed, antithetic=antithetic,
                                      normals1=n1, normals2=n2)
        scale = max(abs(p_ref), 1e-3)
        assert abs(r.price - p_ref) <= PRICE_TOL * scale, (i, r.price, p_ref)
        assert abs(r.stderr_ - se_ref) <= 1e-8 * max(se_ref, 1e-6), (i, r.stderr_, se_ref)


def test_config2_subset_reference_draws(oracle, gpu):
    import bench
    from openferric_amd import RngMode, engine_cfg, price_raw
    opts = bench.make_workload()[::64]
    assert len(opts) == 64 and sum(d["payoff"] == "asian" for d in opts) == 32
    paths = 4096
    pairs = [_pair(d, paths, 40 + i) for i, d in enumerate(opts)]
    got = price_raw([p[1] for p in pairs], engine_cfg(rng_mode=RngMode.Reference))
    _check(oracle, pairs, got, paths, lambda i: opts[i]["num_steps"])


def test_config2_subset_external_normals(oracle, gpu):
    import bench
    from openferric_amd import RngMode, engine_cfg, price_raw
    opts = bench.make_workload()[::512]
    paths = 512
    pairs, z_all, normals = [], [], []
    for i, d in enumerate(opts):
        pairs.append(_pair(d, paths, 900 + i))
        z1, _ = oracle.normals_matrix(900 + i, paths, d["num_steps"], 1)
        z_all.append(step_major(z1).ravel())
        normals.append((z1, None))
    cfg = engine_cfg(rng_mode=RngMode.External, normals1=np.concatenate(z_all))
    got = price_raw([p[1] for p in pairs], cfg)
    _check(oracle, pairs, got, paths, lambda i: opts[i]["num_steps"], normals=normals)


def test_config3_subset_heston_barriers(oracle, gpu):
    import bench
    from openferric_amd import RngMode, engine_cfg, price_raw
    opts = bench.heston_workload(n_opts=8, paths=1 << 14)
    paths = 1 << 14
    pairs = [_pair(dict(d, barrier_dir="up", barrier_style="out", rebate=0.0), paths, 70 + i)
             for i, d in enumerate(opts)]
    got = price_raw([p[1] for p in pairs], engine_cfg(rng_mode=RngMode.Reference))
    _check(oracle, pairs, got, paths, lambda i: 365)
