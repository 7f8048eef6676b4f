"""torch autograd path and the device-resident entry point ofg_price_device.

* GPU tensors are read and written in place (ofg_price_device); prices and
  the full gradient equal the C ABI's ofg_price (OFG_GREEKS_FORWARD, the
  same pathwise estimator as ofg_price_greeks) on the same draws, for GBM
  vanilla / Asian / barrier batches and for Heston (checked against the
  C ABI's adjoint gradient, ofg_price_greeks, which the forward tangent
  equals to a few 1e-12);
* no argument requiring grad (or torch.no_grad): the price launch alone,
  equal to ofg_price's price;
* broadcasting: a scalar argument's gradient is the sum over the options;
* CPU tensors (ofg_price with host copies) agree with GPU tensors;
* ofg_price_device on its own: the device-side validation (reference
  messages naming the option, nothing written), its refusals, and
  equality with ofg_price on host values.
"""
import ctypes as C

import numpy as np
import pytest

from tests.cases import daily_obs, make_pair

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"


def _t(v, grad=False):
    return torch.tensor(v, dtype=torch.float64, device=DEV, requires_grad=grad)


SPOT, VOL, RATE, T, K = [100.0, 95.0, 110.0], [0.2, 0.3, 0.25], [0.03, 0.05, 0.01], \
    [1.0, 0.5, 2.0], [100.0, 90.0, 120.0]


def _c_abi(payoff, i, greeks, **extra):
    from openferric_amd import RngMode, _lib, engine_cfg, price_raw
    kw = dict(spot=SPOT[i], vol=VOL[i], rate=RATE[i], expiry=T[i], strike=K[i], num_paths=30_000,
              num_steps=48, seed=3, payoff=payoff)
    if payoff == "asian":
        kw["obs_times"] = daily_obs(T[i], 48)
    if payoff == "barrier":
        kw.update(barrier_level=140.0, barrier_dir="up", barrier_style="out")
    kw.update(extra)
    _, g, keep = make_pair(**kw)
    cfg = engine_cfg(rng_mode=RngMode.Reference,
                     greek_mode=_lib.GREEKS_FORWARD if greeks else _lib.GREEKS_NONE)
    return price_raw([g], cfg)[0]


@pytest.mark.parametrize("payoff", ["vanilla", "asian", "barrier"])
def test_gbm_gradients_equal_c_abi(gpu, payoff):
    from openferric_amd import RngMode
    from openferric_amd.torch_ops import gbm_mc_price
    spot, vol, rate, T_ = _t(SPOT, True), _t(VOL, True), _t(RATE, True), _t(T, True)
    p = gbm_mc_price(spot, vol, rate, T_, _t(K), payoff=payoff, num_paths=30_000, num_steps=48,
                     seed=3, barrier_level=140.0, rng_mode=RngMode.Reference)
    w = _t([1.0, 2.0, -0.5])
    (p * w).sum().backward()
    for i in range(3):
        r = _c_abi(payoff, i, True)
        assert abs(p[i].item() - r.price) <= 1e-13 * max(abs(r.price), 1.0)
        for got, want in ((spot.grad[i], r.delta), (vol.grad[i], r.vega),
                          (rate.grad[i], r.rho), (T_.grad[i], -r.theta)):
            assert abs(got.item() - w[i].item() * want) <= 1e-12 * max(abs(want), 1.0), \
                (payoff, i, got.item(), w[i].item() * want)


def test_no_grad_is_the_price_launch(gpu):
    from openferric_amd import RngMode
    from openferric_amd.torch_ops import gbm_mc_price
    with torch.no_grad():
        p = gbm_mc_price(_t(SPOT, True), _t(VOL), _t(RATE), _t(T), _t(K), num_paths=30_000,
                         num_steps=48, seed=3, rng_mode=RngMode.Reference)
    assert not p.requires_grad
    for i in range(3):
        r = _c_abi("vanilla", i, False)
        assert abs(p[i].item() - r.price) <= 1e-13 * r.price


def test_heston_gradients_equal_c_abi_adjoint(gpu):
    from openferric_amd import RngMode, engine_cfg, price_raw
    from openferric_amd.torch_ops import heston_mc_price
    prm = dict(spot=[100.0, 105.0], v0=[0.04, 0.06], kappa=[1.5, 2.5], theta=[0.05, 0.03],
               xi=[0.4, 0.7], rho=[-0.6, -0.3], rate=[0.02, 0.04], expiry=[1.0, 0.75])
    ts = {k: _t(v, True) for k, v in prm.items()}
    for payoff in ("vanilla", "barrier"):
        for t in ts.values():
            t.grad = None
        p = heston_mc_price(ts["spot"], ts["v0"], ts["kappa"], ts["theta"], ts["xi"], ts["rho"],
                            ts["rate"], ts["expiry"], _t([100.0, 100.0]), payoff=payoff,
                            num_paths=20_000, num_steps=40, seed=5, barrier_level=150.0,
                            rng_mode=RngMode.Reference)
        p.sum().backward()
        for i in range(2):
            kw = {k: v[i] for k, v in prm.items()}
            kw.update(model="heston", payoff=payoff, strike=100.0, num_paths=20_000, num_steps=40,
                      seed=5, q=0.0)
            if payoff == "barrier":
                kw.update(barrier_level=150.0, barrier_dir="up", barrier_style="out")
            _, g, keep = make_pair(**kw)
            r = price_raw([g], engine_cfg(rng_mode=RngMode.Reference), greeks=True)[0]
            assert abs(p[i].item() - r.price) <= 1e-12 * r.price
            want = dict(spot=r.heston_grad[0], v0=r.heston_grad[1], kappa=r.heston_grad[2],
                        theta=r.heston_grad[3], xi=r.heston_grad[4], rho=r.heston_grad[5],
                        rate=r.rho, expiry=-r.theta)
            for k, wv in want.items():
                got = ts[k].grad[i].item()
                assert abs(got - wv) <= 1e-9 * max(abs(wv), 1.0), (payoff, i, k, got, wv)


def test_broadcast_scalar_gradient_is_the_sum(gpu):
    from openferric_amd import RngMode
    from openferric_amd.torch_ops import gbm_mc_price
    vol = _t(0.25, True)
    spot = _t(SPOT, True)
    p = gbm_mc_price(spot, vol, 0.03, 1.0, _t(K), num_paths=20_000, num_steps=16, seed=9,
                     rng_mode=RngMode.Reference)
    assert p.shape == (3,)
    p.sum().backward()
    assert vol.grad.shape == ()
    total = 0.0
    for i in range(3):
        total += _c_abi("vanilla", i, True, vol=0.25, rate=0.03, expiry=1.0, num_paths=20_000,
                        num_steps=16, seed=9).vega
    assert abs(vol.grad.item() - total) <= 1e-11 * abs(total)


def test_cpu_tensors_agree_with_gpu_tensors(gpu):
    from openferric_amd import RngMode
    from openferric_amd.torch_ops import gbm_mc_price
    out = []
    for dev in ("cpu", DEV):
        s = torch.tensor(SPOT, dtype=torch.float64, device=dev, requires_grad=True)
        p = gbm_mc_price(s, torch.tensor(VOL, dtype=torch.float64, device=dev), 0.02, 1.0,
                         torch.tensor(K, dtype=torch.float64, device=dev), num_paths=10_000,
                         num_steps=24, seed=1, rng_mode=RngMode.Reference)
        p.sum().backward()
        out.append((p.detach().cpu().numpy(), s.grad.cpu().numpy()))
    assert np.allclose(out[0][0], out[1][0], rtol=1e-13, atol=0)
    assert np.allclose(out[0][1], out[1][1], rtol=1e-12, atol=1e-14)


def test_strike_gradient_is_refused(gpu):
    from openferric_amd.torch_ops import gbm_mc_price
    with pytest.raises(ValueError, match="strike"):
        gbm_mc_price(_t(SPOT), _t(VOL), _t(RATE), _t(T), _t(K, True))


def _device_call(opts, cfg, **arrays):
    from openferric_amd import _lib
    L = _lib.load()
    n = len(opts)
    io = _lib.OfgDeviceIo()
    keep = {}
    for k, v in arrays.items():
        keep[k] = torch.tensor(v, dtype=torch.float64, device=DEV)
        setattr(io, k, keep[k].data_ptr())
    price = torch.full((n,), -7.0, dtype=torch.float64, device=DEV)
    se = torch.full((n,), -7.0, dtype=torch.float64, device=DEV)
    io.price, io.stderr_ = price.data_ptr(), se.data_ptr()
    torch.cuda.synchronize()
    arr = (_lib.OfgOption * n)(*opts)
    rc = L.ofg_price_device(_lib.context(0), arr, n, C.byref(cfg), C.byref(io))
    return rc, price.cpu().numpy(), se.cpu().numpy()


def test_price_device_equals_price_on_host_values(gpu):
    from openferric_amd import RngMode, engine_cfg, price_raw
    pairs = [make_pair(spot=s, vol=v, rate=r, expiry=t, strike=k, num_paths=8000, num_steps=30,
                       seed=11) for s, v, r, t, k in zip(SPOT, VOL, RATE, T, K)]
    pairs.append(make_pair(model="heston", kappa=2.0, theta=0.05, xi=0.5, rho=-0.5, v0=0.04,
                           payoff="barrier", barrier_level=130.0, num_paths=6000, num_steps=20))
    opts = [g for _, g, _ in pairs]
    for mode in (RngMode.Reference, RngMode.Fast, RngMode.Pcg64):
        cfg = engine_cfg(antithetic=True, rng_mode=mode)
        ref = price_raw(opts, cfg)
        # values from the structs themselves (no device override), then all
        # of spot / strike / vol / rate / expiry from device arrays
        rc, p, se = _device_call(opts, cfg)
        assert rc == 0
        rc2, p2, se2 = _device_call(opts, cfg, spot=[o.spot for o in opts],
                                    strike=[o.strike for o in opts], vol=[o.vol for o in opts],
                                    rate=[o.rate for o in opts], expiry=[o.expiry for o in opts])
        assert rc2 == 0
        for i, r in enumerate(ref):
            for got in (p[i], p2[i]):
                assert abs(got - r.price) <= 1e-13 * max(abs(r.price), 1.0), (mode, i)
            assert abs(se2[i] - r.stderr_) <= 1e-10 * r.stderr_


def test_price_device_validation_and_refusals(gpu):
    from openferric_amd import RngMode, _lib, engine_cfg
    opts = [make_pair(num_paths=1000, num_steps=8)[1] for _ in range(3)]
    cfg = engine_cfg(rng_mode=RngMode.Reference)
    rc, p, se = _device_call(opts, cfg, strike=[100.0, -5.0, 100.0])
    assert rc == 1 and "vanilla strike must be > 0 (option 1)" in _lib.last_error()
    assert (p == -7.0).all() and (se == -7.0).all()   # nothing written
    rc, p, _ = _device_call(opts, cfg, vol=[0.2, 0.2, 0.0])
    assert rc == 1 and "market volatility must be > 0 (option 2)" in _lib.last_error()
    rc, p, _ = _device_call(opts, cfg, expiry=[1.0, float("nan"), 1.0])
    assert rc == 1 and "finite" in _lib.last_error()
    h = [make_pair(model="heston", kappa=1.0, theta=0.04, xi=0.5, rho=-0.5, v0=0.04,
                   num_paths=500, num_steps=4)[1]]
    rc, p, _ = _device_call(h, cfg, rho=[1.0])
    assert rc == 1 and "invalid Heston parameters" in _lib.last_error()
    for bad in (engine_cfg(rng_mode=RngMode.Reference, control_variate=True),
                engine_cfg(rng_mode=RngMode.Reference, greek_mode=_lib.GREEKS_FD_CRN)):
        rc, _, _ = _device_call(opts, bad)
        assert rc == 1
    a, g, keep = make_pair(payoff="asian", obs_times=daily_obs(1.0, 8), num_paths=500, num_steps=8)
    rc, _, _ = _device_call([g], cfg, expiry=[1.0])
    assert rc == 1 and "every step" in _lib.last_error()
