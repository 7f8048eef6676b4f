"""torch.autograd bindings of the adjoint Greeks kernel.

`gbm_mc_price(spot, vol, rate, expiry, ...)` and
`heston_mc_price(spot, v0, kappa, theta, xi, rho, ...)` return a float64
tensor of Monte Carlo prices, one per option, differentiable with respect to
every tensor argument. The forward pass runs the adjoint kernel once (price
and the full pathwise gradient in the same launch); backward scales the saved
gradient by grad_output. Contract data stay on the host (a few doubles per
option); the result lives on the inputs' device.
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from . import _lib
from .engines import RngMode, _Keep, engine_cfg, price_raw


def _options(n, spot, vol, rate, expiry, strike, payoff, option_type, num_paths, num_steps, seed,
             obs_times, barrier_level, heston=None):
    keep = _Keep()
    opts = []
    for i in range(n):
        o = _lib.OfgOption()
        o.model = _lib.MODEL_HESTON if heston is not None else _lib.MODEL_GBM
        o.payoff = {"vanilla": _lib.PAYOFF_VANILLA, "asian": _lib.PAYOFF_ASIAN,
                    "barrier": _lib.PAYOFF_BARRIER}[payoff]
        o.option_type = _lib.CALL if option_type == "call" else _lib.PUT
        o.spot, o.strike, o.expiry = spot[i], strike[i], expiry[i]
        o.rate, o.div_yield, o.vol = rate[i], 0.0, vol[i] if vol is not None else 0.2
        if heston is not None:
            o.v0, o.kappa, o.theta, o.xi, o.rho = (h[i] for h in heston)
        if payoff == "asian":
            obs = obs_times if obs_times is not None else \
                [expiry[i] * (j / num_steps) for j in range(1, num_steps + 1)]
            o.obs_times = keep.arr(obs)
            o.n_obs = len(obs)
        if payoff == "barrier":
            o.barrier_dir, o.barrier_style = _lib.UP, _lib.OUT
            o.barrier_level = barrier_level
        o.num_paths, o.num_steps, o.seed = num_paths, num_steps, seed
        opts.append(o)
    return opts, keep


def _host(t):
    return t.detach().to("cpu", torch.float64).reshape(-1).tolist()


class _GbmMcPrice(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spot, vol, rate, expiry, strike, payoff, option_type, num_paths, num_steps,
                seed, obs_times, barrier_level, rng_mode):
        n = spot.numel()
        opts, keep = _options(n, _host(spot), _host(vol), _host(rate), _host(expiry),
                              _host(strike), payoff, option_type, num_paths, num_steps, seed,
                              obs_times, barrier_level)
        res = price_raw(opts, engine_cfg(rng_mode=rng_mode), greeks=True)
        dev = spot.device
        price = torch.tensor([r.price for r in res], dtype=torch.float64, device=dev)
        # d price / d (spot, vol, rate, expiry): theta is -d/dT
        grads = torch.tensor([[r.delta, r.vega, r.rho, -r.theta] for r in res],
                             dtype=torch.float64, device=dev)
        ctx.save_for_backward(grads)
        return price

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        g = grad_out.reshape(-1, 1).to(grads.dtype) * grads
        return (g[:, 0], g[:, 1], g[:, 2], g[:, 3], None, None, None, None, None, None, None,
                None, None)


class _HestonMcPrice(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spot, v0, kappa, theta, xi, rho, rate, expiry, strike, payoff, option_type,
                num_paths, num_steps, seed, barrier_level, rng_mode):
        n = spot.numel()
        opts, keep = _options(n, _host(spot), None, _host(rate), _host(expiry), _host(strike),
                              payoff, option_type, num_paths, num_steps, seed, None, barrier_level,
                              heston=[_host(v0), _host(kappa), _host(theta), _host(xi), _host(rho)])
        res = price_raw(opts, engine_cfg(rng_mode=rng_mode), greeks=True)
        dev = spot.device
        price = torch.tensor([r.price for r in res], dtype=torch.float64, device=dev)
        grads = torch.tensor([list(r.heston_grad) + [r.rho] for r in res], dtype=torch.float64,
                             device=dev)
        ctx.save_for_backward(grads)
        return price

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        g = grad_out.reshape(-1, 1).to(grads.dtype) * grads
        # order: S0, v0, kappa, theta, xi, rho, rate
        return (g[:, 0], g[:, 1], g[:, 2], g[:, 3], g[:, 4], g[:, 5], g[:, 6], None, None, None,
                None, None, None, None, None, None)


def gbm_mc_price(spot, vol, rate, expiry, strike, payoff="vanilla", option_type="call",
                 num_paths=1 << 16, num_steps=64, seed=42,
                 obs_times: Optional[Sequence[float]] = None, barrier_level=0.0,
                 rng_mode=RngMode.Fast):
    """Monte Carlo prices under GBM, differentiable in spot, vol, rate, expiry
    (each a tensor of shape [n]; strike a tensor of shape [n])."""
    return _GbmMcPrice.apply(spot, vol, rate, expiry, strike, payoff, option_type, num_paths,
                             num_steps, seed, obs_times, barrier_level, rng_mode)


def heston_mc_price(spot, v0, kappa, theta, xi, rho, rate, expiry, strike, payoff="vanilla",
                    option_type="call", num_paths=1 << 16, num_steps=64, seed=42,
                    barrier_level=0.0, rng_mode=RngMode.Fast):
    """Monte Carlo prices under Heston (drift = rate), differentiable in spot,
    v0, kappa, theta, xi, rho and rate."""
    return _HestonMcPrice.apply(spot, v0, kappa, theta, xi, rho, rate, expiry, strike, payoff,
                                option_type, num_paths, num_steps, seed, barrier_level, rng_mode)
