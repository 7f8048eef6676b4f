"""torch.autograd bindings of the Monte Carlo pricer and its pathwise Greeks.

`gbm_mc_price(spot, vol, rate, expiry, strike, ...)` and
`heston_mc_price(spot, v0, kappa, theta, xi, rho, rate, expiry, strike, ...)`
return a float64 tensor of Monte Carlo prices, one per option, and are
differentiable in every tensor argument except the strike (the kernels carry
no strike tangent: a strike that requires grad is an error, not a silent
zero). Tensor arguments broadcast against each other (a scalar vol with a
vector of spots is one vol for every option); gradients are summed back to
each argument's own shape.

Forward pass: when no argument requires grad (or grad mode is off) it is the
plain pricing launch; otherwise the forward pathwise evaluation
(`OFG_GREEKS_FORWARD`: one sweep per path, the adjoint's estimator, about
1.05 pricing passes under GBM and 1.26 under Heston) fills the price and the
full gradient in the same launch, and backward scales that gradient by
grad_output. Arguments on a HIP device are read in place and the results
written in place through `ofg_price_device` (no host copy of the data; the
caller's stream is synchronized first, the results are complete on return);
CPU tensors go through `ofg_price`. With observation times given, the
maturities are read on the host (the schedule is mapped to steps there).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import torch

from . import _lib
from .engines import RngMode, _Keep, _raise, engine_cfg, price_raw

# gradient slots of the kernels: 0 S0, 1 sigma | v0, 2 r, 3 T, 4 kappa,
# 5 theta, 6 xi, 7 rho
_SLOTS = 8


def _template(n, payoff, option_type, num_paths, num_steps, seed, obs_times, barrier_level,
              heston, host):
    """Contract descriptors; market fields from `host` (lists) where given."""
    keep = _Keep()
    opts = []
    for i in range(n):
        o = _lib.OfgOption()
        o.model = _lib.MODEL_HESTON if heston else _lib.MODEL_GBM
        o.payoff = {"vanilla": _lib.PAYOFF_VANILLA, "asian": _lib.PAYOFF_ASIAN,
                    "barrier": _lib.PAYOFF_BARRIER}[payoff]
        o.option_type = _lib.CALL if option_type == "call" else _lib.PUT
        o.spot, o.strike, o.expiry, o.rate, o.vol = 100.0, 100.0, 1.0, 0.0, 0.2
        o.div_yield = 0.0
        o.v0, o.kappa, o.theta, o.xi, o.rho = 0.04, 1.0, 0.04, 0.5, 0.0
        for name, vals in host.items():
            setattr(o, name, vals[i])
        if payoff == "asian":
            if obs_times is not None:
                o.obs_times = keep.arr(obs_times)
                o.n_obs = len(obs_times)
            elif "expiry" in host:   # every step, on the host maturity
                obs = [host["expiry"][i] * (j / num_steps) for j in range(1, num_steps + 1)]
                o.obs_times = keep.arr(obs)
                o.n_obs = len(obs)
        if payoff == "barrier":
            o.barrier_dir, o.barrier_style = _lib.UP, _lib.OUT
            o.barrier_level = barrier_level
        o.num_paths, o.num_steps, o.seed = num_paths, num_steps, seed
        opts.append(o)
    return opts, keep


def _flat(t):
    return t.detach().to(torch.float64).reshape(-1).contiguous()


def _price(names, tensors, heston, payoff, option_type, num_paths, num_steps, seed, obs_times,
           barrier_level, rng_mode, want_grad):
    """Prices (n,) and raw gradients (n, 8) or None, on the inputs' device."""
    dev = tensors[0].device
    n = tensors[0].numel()
    greek = _lib.GREEKS_FORWARD if want_grad else _lib.GREEKS_NONE
    cfg = engine_cfg(rng_mode=rng_mode, greek_mode=greek)
    flat = [_flat(t) for t in tensors]
    if dev.type == "cuda":
        host_names = {"expiry"} if obs_times is not None else set()
        host = {k: v.cpu().tolist() for k, v in zip(names, flat) if k in host_names}
        opts, keep = _template(n, payoff, option_type, num_paths, num_steps, seed, obs_times,
                               barrier_level, heston, host)
        price = torch.empty(n, dtype=torch.float64, device=dev)
        grads = torch.zeros((n, _SLOTS), dtype=torch.float64, device=dev) if want_grad else None
        io = _lib.OfgDeviceIo()
        for k, v in zip(names, flat):
            if k not in host_names:
                setattr(io, k, v.data_ptr())
        io.price = price.data_ptr()
        io.grads = grads.data_ptr() if want_grad else None
        torch.cuda.current_stream(dev).synchronize()   # the inputs are ready
        L = _lib.load()
        arr = (_lib.OfgOption * n)(*opts)
        _raise(L.ofg_price_device(_lib.context(dev.index or 0), arr, n, C.byref(cfg),
                                  C.byref(io)))
        return price, grads
    host = {k: v.tolist() for k, v in zip(names, flat)}
    opts, keep = _template(n, payoff, option_type, num_paths, num_steps, seed, obs_times,
                           barrier_level, heston, host)
    res = price_raw(opts, cfg)
    price = torch.tensor([r.price for r in res], dtype=torch.float64)
    if not want_grad:
        return price, None
    rows = []
    for r in res:
        if heston:
            h = list(r.heston_grad)
            rows.append([h[0], h[1], r.rho, -r.theta, h[2], h[3], h[4], h[5]])
        else:
            rows.append([r.delta, r.vega, r.rho, -r.theta, 0.0, 0.0, 0.0, 0.0])
    return price, torch.tensor(rows, dtype=torch.float64)


class _McPrice(torch.autograd.Function):
    """Inputs arrive expanded to the broadcast shape (torch's expand records
    the broadcast, and its backward sums each gradient back to the argument's
    own shape)."""

    @staticmethod
    def forward(ctx, spec, *tensors):
        names, slots, heston, kw, want_grad, bshape = spec
        price, grads = _price(names, tensors, heston, want_grad=want_grad, **kw)
        ctx.slots, ctx.bshape = slots, bshape
        ctx.save_for_backward(grads if grads is not None else price.new_zeros(0))
        return price.reshape(bshape)

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        out = [None]
        for slot, need in zip(ctx.slots, ctx.needs_input_grad[1:]):
            if slot is None or not need:
                out.append(None)
                continue
            g = grad_out.reshape(-1).to(grads.dtype) * grads[:, slot].to(grad_out.device)
            out.append(g.reshape(ctx.bshape))
        return tuple(out)


def _apply(names, slots, args, heston, kw):
    like = next((a for a in args if isinstance(a, torch.Tensor)), None)
    dev = like.device if like is not None else torch.device("cpu")
    tensors = [a if isinstance(a, torch.Tensor) else torch.tensor(a, dtype=torch.float64, device=dev)
               for a in args]
    for name, t in zip(names, tensors):
        if t.device != dev:
            raise ValueError(f"{name} is on {t.device}, the other arguments on {dev}")
    k = names.index("strike")
    if tensors[k].requires_grad:
        raise ValueError("the Monte Carlo kernels carry no strike sensitivity: "
                         "strike must not require grad")
    bshape = torch.broadcast_shapes(*[t.shape for t in tensors])
    expanded = [t.expand(bshape) for t in tensors]
    want = torch.is_grad_enabled() and any(t.requires_grad for t in tensors)
    spec = (names, slots, heston, kw, want, bshape)
    return _McPrice.apply(spec, *expanded)


def gbm_mc_price(spot, vol, rate, expiry, strike, payoff="vanilla", option_type="call",
                 num_paths=1 << 16, num_steps=64, seed=42,
                 obs_times: Optional[Sequence[float]] = None, barrier_level=0.0,
                 rng_mode=RngMode.Fast):
    """Monte Carlo prices under GBM (MonteCarloPricingEngine, mu = rate),
    differentiable in spot, vol, rate and expiry."""
    names = ("spot", "vol", "rate", "expiry", "strike")
    kw = dict(payoff=payoff, option_type=option_type, num_paths=num_paths, num_steps=num_steps,
              seed=seed, obs_times=obs_times, barrier_level=barrier_level, rng_mode=rng_mode)
    return _apply(names, (0, 1, 2, 3, None), (spot, vol, rate, expiry, strike), False, kw)


def heston_mc_price(spot, v0, kappa, theta, xi, rho, rate, expiry, strike, payoff="vanilla",
                    option_type="call", num_paths=1 << 16, num_steps=64, seed=42,
                    barrier_level=0.0, rng_mode=RngMode.Fast):
    """Monte Carlo prices under Heston (drift = rate), differentiable in spot,
    v0, kappa, theta, xi, rho, rate and expiry."""
    names = ("spot", "v0", "kappa", "theta", "xi", "rho", "rate", "expiry", "strike")
    kw = dict(payoff=payoff, option_type=option_type, num_paths=num_paths, num_steps=num_steps,
              seed=seed, obs_times=None, barrier_level=barrier_level, rng_mode=rng_mode)
    return _apply(names, (0, 1, 4, 5, 6, 7, 2, 3, None),
                  (spot, v0, kappa, theta, xi, rho, rate, expiry, strike), True, kw)
