"""Device failure paths through the C ABI: an allocation that fails in the
middle of a call (ofg_inject_alloc_failure arms the k-th allocation to fail
as an out-of-memory hipMalloc) makes the entry point return
OFG_DEVICE_ERROR with a message naming the allocation; nothing is written
past what was allocated, and the same context then prices the same batch
bit for bit as a context that never failed. Every allocation point of each
call is failed in turn, on a fresh context each time, until the call makes
fewer allocations than the armed count."""
import ctypes as C

import pytest

from tests.cases import daily_obs, make_pair

pytestmark = pytest.mark.gpu

OFG_OK, OFG_DEVICE_ERROR = 0, 5


def _batch():
    keep, opts = [], []
    for kw in (dict(payoff="vanilla", num_paths=3000, num_steps=24),
               dict(payoff="asian", obs_times=daily_obs(1.0, 16), num_paths=2000, num_steps=16),
               dict(payoff="barrier", barrier_level=130.0, num_paths=2500, num_steps=20),
               dict(model="heston", kappa=1.5, theta=0.04, xi=0.4, rho=-0.6, v0=0.05,
                    num_paths=2200, num_steps=18)):
        _, g, k = make_pair(**kw)
        opts.append(g)
        keep.append(k)
    return opts, keep


def _calls():
    """(name, fn(L, ctx) -> (status, comparable result))."""
    from openferric_amd import RngMode, _lib, engine_cfg
    opts, keep = _batch()
    n = len(opts)
    arr = (_lib.OfgOption * n)(*opts)

    def price(mode):
        cfg = engine_cfg(antithetic=True, rng_mode=RngMode.Reference, greek_mode=mode)

        def run(L, ctx):
            out = (_lib.OfgResult * n)()
            rc = L.ofg_price(ctx, arr, n, C.byref(cfg), out)
            return rc, [(r.price, r.stderr_, r.delta, r.vega, r.rho, r.theta,
                         tuple(r.heston_grad)) for r in out]
        return run

    gbm = (_lib.OfgOption * 3)(*opts[:3])  # finite_difference_greeks takes GBM options
    bumps = (_lib.OfgFdBumps * 3)(*[_lib.OfgFdBumps(0.5, 1e-4, 1e-3, 1.0 / 365.0)] * 3)

    def fd(L, ctx):
        cfg = engine_cfg(rng_mode=RngMode.Reference)
        out = (_lib.OfgFdGreeks * 3)()
        rc = L.ofg_finite_difference_greeks(ctx, gbm, 3, C.byref(cfg), bumps, out)
        return rc, [tuple(getattr(r, f) for f, _ in _lib.OfgFdGreeks._fields_ if f != "kernel_ms")
                    for r in out]

    def sample(L, ctx):
        cfg = engine_cfg(antithetic=False, rng_mode=RngMode.Fast)
        xs = (C.c_double * int(opts[0].num_paths))()
        ys = (C.c_double * int(opts[0].num_paths))()
        rc = L.ofg_sample_payoffs(ctx, C.byref(opts[0]), C.byref(cfg), xs, ys)
        return rc, (list(xs), list(ys))

    return [("price", price(_lib.GREEKS_NONE)), ("adjoint", price(_lib.GREEKS_ADJOINT)),
            ("forward", price(_lib.GREEKS_FORWARD)), ("fd_crn", price(_lib.GREEKS_FD_CRN)),
            ("finite_difference_greeks", fd), ("sample_payoffs", sample)], keep


def test_every_allocation_failure_is_a_device_error_and_recoverable(gpu):
    from openferric_amd import _lib
    L = _lib.load()
    calls, keep = _calls()
    for name, run in calls:
        c0 = L.ofg_create(0)
        assert c0
        rc, clean = run(L, c0)
        L.ofg_destroy(c0)
        assert rc == OFG_OK, (name, _lib.last_error())
        failed = 0
        for k in range(1, 64):
            ctx = L.ofg_create(0)
            assert ctx
            try:
                print(f"{name}: allocation {k} fails", flush=True)
                L.ofg_inject_alloc_failure(k)
                rc, res = run(L, ctx)
                L.ofg_inject_alloc_failure(0)
                if rc == OFG_OK:  # the call makes fewer than k allocations
                    assert res == clean, name
                    break
                assert rc == OFG_DEVICE_ERROR, (name, k, rc, _lib.last_error())
                assert "injected" in _lib.last_error(), _lib.last_error()
                failed += 1
                rc, res = run(L, ctx)
                assert rc == OFG_OK, (name, k, _lib.last_error())
                assert res == clean, (name, k)
            finally:
                L.ofg_inject_alloc_failure(0)
                L.ofg_destroy(ctx)
        assert failed >= 1, name
