"""ofg_price_noalloc (pricing in the library's fixed per-GPU region, no device
allocation in the call) against ofg_price through the C ABI: two calls in a
row on the region, the second with a smaller input, each equal to ofg_price
bit for bit (the same kernels, work queue and summation order); the limits
are refused before anything runs and the region is usable afterwards."""
import numpy as np
import pytest

from tests.cases import daily_obs, make_pair

pytestmark = pytest.mark.gpu


def _batch(n, seed, heston_every=4):
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        steps = int(rng.integers(8, 160))
        T = float(rng.uniform(0.2, 2.0))
        kind = ["vanilla", "asian", "barrier"][i % 3]
        kw = dict(payoff=kind, option_type="call" if i % 2 else "put",
                  strike=float(rng.uniform(80, 120)), expiry=T, rate=float(rng.uniform(0, 0.06)),
                  vol=float(rng.uniform(0.1, 0.5)), num_steps=steps,
                  num_paths=int(rng.integers(1000, 60000)), seed=int(rng.integers(1, 1000)))
        if kind == "asian":
            kw["obs_times"] = daily_obs(T, steps)
        if kind == "barrier":
            kw.update(barrier_dir="up", barrier_style="out", barrier_level=140.0, rebate=0.5)
        if i % heston_every == 0:
            kw.update(model="heston", kappa=1.5, theta=0.04, xi=0.5, rho=-0.6, v0=0.05)
        if i == 5:
            kw["expiry"] = 0.0  # maturity-zero shortcut, no simulation
            kw.pop("obs_times", None)
            kw["payoff"] = "vanilla"
        pairs.append(make_pair(**kw))
    return pairs


@pytest.mark.parametrize("rng,anti", [("reference", False), ("fast", True)])
def test_noalloc_equals_price_on_two_calls(gpu, rng, anti):
    from openferric_amd import RngMode, engine_cfg, price_noalloc_raw, price_raw
    mode = {"reference": RngMode.Reference, "fast": RngMode.Fast}[rng]
    cfg = engine_cfg(antithetic=anti, rng_mode=mode)
    for n, seed in ((40, 1), (9, 2)):   # the second call is smaller
        pairs = _batch(n, seed)
        opts = [p[1] for p in pairs]
        want = price_raw(opts, cfg)
        got = price_noalloc_raw(opts, cfg)
        for w, g in zip(want, got):
            assert g.price == w.price and g.stderr_ == w.stderr_
            assert (g.num_paths, g.num_steps, g.vol) == (w.num_paths, w.num_steps, w.vol)
        assert got[5].price == want[5].price and got[5].stderr_ == 0.0


def test_noalloc_limits_refused_before_any_launch(gpu):
    from openferric_amd import (InvalidInput, RngMode, engine_cfg, price_noalloc_raw,
                                price_raw, _lib)
    _, g, keep = make_pair(num_paths=5000, num_steps=20)
    cfg = engine_cfg(rng_mode=RngMode.Fast)
    with pytest.raises(InvalidInput, match="at most 4096 options"):
        price_noalloc_raw([g] * 4097, cfg)
    with pytest.raises(InvalidInput, match="OFG_GREEKS_NONE"):
        price_noalloc_raw([g], engine_cfg(rng_mode=RngMode.Fast, greek_mode=_lib.GREEKS_ADJOINT))
    with pytest.raises(InvalidInput, match="generated draws"):
        price_noalloc_raw([g], engine_cfg(rng_mode=RngMode.External, normals1=np.zeros(10)))
    # observation weights: 2049 daily-fixed Asians x 512 steps > 2^20 weights
    _, ga, keep_a = make_pair(payoff="asian", obs_times=daily_obs(1.0, 511), num_steps=511,
                              num_paths=64)
    with pytest.raises(InvalidInput, match="observation weights"):
        price_noalloc_raw([ga] * 2049, cfg)
    # the region still prices afterwards, equal to ofg_price
    assert price_noalloc_raw([g], cfg)[0].price == price_raw([g], cfg)[0].price


def test_noalloc_large_batch_uses_bigger_work_items(gpu):
    """More than 2^18 work items at the default item size: the item size is
    raised until the queue fits; prices equal ofg_price at that shape."""
    from openferric_amd import RngMode, engine_cfg, price_noalloc_raw, price_raw
    _, g, keep = make_pair(num_paths=1 << 22, num_steps=2048, seed=3)
    # 80 x 2^22 samples at 256 samples per item (8192 lane-steps / 2048 steps)
    # is 1.3M items; doubling the item size three times gives 164k
    opts = [g] * 80
    cfg = engine_cfg(rng_mode=RngMode.Fast, tune_item_steps=8192)
    got = price_noalloc_raw(opts, cfg)
    want = price_raw(opts, engine_cfg(rng_mode=RngMode.Fast, tune_item_steps=65536))
    assert [r.price for r in got] == [r.price for r in want]
