"""Heston calibration to option prices with Monte Carlo gradients on the GPU.

The reference calibrates Heston to implied vols with Carr-Madan FFT prices and
a Levenberg-Marquardt optimiser whose Jacobian is a finite difference of the
residuals (src/calibration/heston.rs:96-255, src/calibration/optimizers.rs:
101-282). Here the model prices are Monte Carlo prices of the Heston Euler
scheme and the Jacobian is analytic: one batched launch prices every quote
AND returns dPrice/d{v0, kappa, theta, xi, rho} (forward tangent or adjoint
kernel), so an LM iteration costs one launch instead of 1 + 5 residual
evaluations. Seeds stay fixed across iterations (common random numbers), so
the objective is a smooth function of the parameters.

The optimiser restates the reference's LM: damping a_ii += lambda (|a_ii| + 1),
step accepted when it lowers 0.5 |r|^2, lambda scaled by lambda_down /
lambda_up, candidates clamped to the box, and the same tolerances, defaults
and termination reasons.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .core import InvalidInput, OptionType
from .engines import GreekMethod, RngMode, _Keep, engine_cfg, price_raw


@dataclass
class HestonCalibrationParams:
    """calibration/heston.rs:23-49 (vector order v0, kappa, theta, sigma_v, rho)."""
    v0: float
    kappa: float
    theta: float
    sigma_v: float
    rho: float

    def to_vec(self) -> List[float]:
        return [self.v0, self.kappa, self.theta, self.sigma_v, self.rho]

    @staticmethod
    def from_slice(x: Sequence[float]) -> "HestonCalibrationParams":
        if len(x) != 5:
            raise InvalidInput("heston parameter vector must have length 5")
        return HestonCalibrationParams(*[float(v) for v in x])


@dataclass
class BoxConstraints:
    lower: List[float]
    upper: List[float]

    def __post_init__(self):
        if len(self.lower) != len(self.upper) or any(
                not (lo <= hi) for lo, hi in zip(self.lower, self.upper)):
            raise InvalidInput("invalid box constraints")

    def dimension(self) -> int:
        return len(self.lower)

    def clamp(self, x: Sequence[float]) -> List[float]:
        return [min(max(float(v), lo), hi) for v, lo, hi in zip(x, self.lower, self.upper)]


@dataclass
class LmOptions:
    """optimizers.rs:24-50 (finite_diff_epsilon unused: the Jacobian is analytic)."""
    max_iterations: int = 80
    initial_lambda: float = 1e-2
    lambda_up: float = 3.0
    lambda_down: float = 0.35
    gradient_tolerance: float = 1e-6
    step_tolerance: float = 1e-7
    objective_tolerance: float = 1e-10
    finite_diff_epsilon: float = 1e-4
    max_stagnation: int = 20


class TerminationReason(enum.Enum):
    GradientTolerance = "gradient_tolerance"
    StepTolerance = "step_tolerance"
    ObjectiveTolerance = "objective_tolerance"
    Stagnation = "stagnation"
    MaxIterations = "max_iterations"
    NumericalFailure = "numerical_failure"


@dataclass
class ConvergenceInfo:
    iterations: int
    objective_evaluations: int
    gradient_norm: float
    step_norm: float
    converged: bool
    reason: TerminationReason


@dataclass
class OptimisationResult:
    x: List[float]
    objective: float
    residuals: np.ndarray
    jacobian: np.ndarray
    convergence: ConvergenceInfo


def _objective(r: np.ndarray) -> float:
    return 0.5 * float(np.dot(r, r))


def levenberg_marquardt(initial: Sequence[float], bounds: BoxConstraints, options: LmOptions,
                        residual_and_jacobian: Callable[[List[float]], Tuple[np.ndarray, np.ndarray]]
                        ) -> OptimisationResult:
    """levenberg_marquardt (optimizers.rs:144-282) with the Jacobian supplied
    together with the residuals by `residual_and_jacobian(x) -> (r, J)`."""
    if len(initial) != bounds.dimension():
        raise InvalidInput("LM initial vector dimension does not match bounds")
    x = bounds.clamp(initial)
    evals = 1
    r, jac = residual_and_jacobian(x)
    if r.size == 0:
        raise InvalidInput("LM residual function returned empty residual vector")
    objective = _objective(r)
    if not math.isfinite(objective):
        raise InvalidInput("LM objective is not finite at initial point")
    lam = max(options.initial_lambda, 1e-12)
    iterations = 0
    grad_norm = step_norm = math.inf
    reason = TerminationReason.MaxIterations
    converged = False
    stagnation = 0
    for it in range(options.max_iterations):
        iterations = it + 1
        a = jac.T @ jac
        g = jac.T @ r
        grad_norm = float(np.linalg.norm(g))
        if not math.isfinite(grad_norm):
            reason = TerminationReason.NumericalFailure
            break
        if grad_norm <= options.gradient_tolerance:
            converged, reason = True, TerminationReason.GradientTolerance
            break
        a = a + np.diag(lam * (np.abs(np.diag(a)) + 1.0))
        try:
            delta = np.linalg.solve(a, -g)
        except np.linalg.LinAlgError:
            lam = min(lam * options.lambda_up, 1e12)
            stagnation += 1
            if stagnation >= options.max_stagnation:
                reason = TerminationReason.Stagnation
                break
            continue
        step_norm = float(np.linalg.norm(delta))
        if step_norm <= options.step_tolerance:
            converged, reason = True, TerminationReason.StepTolerance
            break
        cand = bounds.clamp([xi + di for xi, di in zip(x, delta)])
        rc, jc = residual_and_jacobian(cand)
        evals += 1
        cobj = _objective(rc)
        if math.isfinite(cobj) and cobj + 1e-16 < objective:
            improvement = abs(objective - cobj)
            x, r, jac, objective = cand, rc, jc, cobj
            lam = max(lam * options.lambda_down, 1e-12)
            stagnation = 0
            if improvement <= options.objective_tolerance:
                converged, reason = True, TerminationReason.ObjectiveTolerance
                break
        else:
            lam = min(lam * options.lambda_up, 1e12)
            stagnation += 1
            if stagnation >= options.max_stagnation:
                reason = TerminationReason.Stagnation
                break
    return OptimisationResult(list(x), objective, r, jac,
                              ConvergenceInfo(iterations, evals, grad_norm, step_norm, converged,
                                              reason))


@dataclass
class OptionPriceQuote:
    """A European quote: strike, maturity, market price, LM weight."""
    strike: float
    maturity: float
    market_price: float
    weight: float = 1.0
    option_type: OptionType = OptionType.Call


def bs_price(option_type: OptionType, spot: float, strike: float, rate: float, q: float,
             vol: float, t: float) -> float:
    """Black-Scholes price with the exact normal CDF (for converting vol quotes)."""
    n = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))
    sq = vol * math.sqrt(t)
    d1 = (math.log(spot / strike) + (rate - q + 0.5 * vol * vol) * t) / sq
    d2 = d1 - sq
    if option_type is OptionType.Call:
        return spot * math.exp(-q * t) * n(d1) - strike * math.exp(-rate * t) * n(d2)
    return strike * math.exp(-rate * t) * n(-d2) - spot * math.exp(-q * t) * n(-d1)


@dataclass
class HestonMcCalibrationResult:
    params: HestonCalibrationParams
    objective: float
    model_prices: List[float]
    residuals: List[float]
    jacobian: List[List[float]]
    condition_number: float
    convergence: ConvergenceInfo
    launches: int
    kernel_ms: float


@dataclass
class HestonMcCalibrator:
    """Heston calibration to prices with GPU Monte Carlo prices and gradients.

    Defaults follow HestonCalibrator (calibration/heston.rs:52-85): bounds on
    (v0, kappa, theta, sigma_v, rho) and 28 LM iterations. Each quote is a
    European option priced with `num_paths` paths, ceil(steps_per_year T)
    Euler steps, seed `seed + index`."""
    spot: float = 100.0
    rate: float = 0.02
    dividend_yield: float = 0.0
    bounds: BoxConstraints = field(default_factory=lambda: BoxConstraints(
        [1e-4, 1e-3, 1e-4, 1e-3, -0.999], [2.0, 20.0, 2.0, 5.0, 0.999]))
    lm_options: LmOptions = field(default_factory=lambda: LmOptions(max_iterations=28))
    num_paths: int = 1 << 16
    steps_per_year: int = 252
    seed: int = 2024
    rng_mode: RngMode = RngMode.Fast
    greek_method: GreekMethod = GreekMethod.Forward
    device: int = 0

    def initial_guess(self, quotes: Sequence[OptionPriceQuote]) -> List[float]:
        """heston.rs:88-100 with the ATM implied vol of the nearest-strike quote."""
        q = min(quotes, key=lambda x: abs(x.strike - self.spot))
        atm = max(self._implied_vol(q), 1e-4)
        return self.bounds.clamp([atm * atm, 1.5, atm * atm, 0.6, -0.5])

    def _implied_vol(self, q: OptionPriceQuote) -> float:
        lo, hi = 1e-4, 5.0
        for _ in range(100):
            mid = 0.5 * (lo + hi)
            if bs_price(q.option_type, self.spot, q.strike, self.rate, self.dividend_yield, mid,
                        q.maturity) < q.market_price:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    def _options(self, x: Sequence[float], quotes: Sequence[OptionPriceQuote]):
        v0, kappa, theta, xi, rho = x
        out = []
        for i, q in enumerate(quotes):
            o = _lib.OfgOption()
            o.model = _lib.MODEL_HESTON
            o.payoff = _lib.PAYOFF_VANILLA
            o.option_type = _lib.CALL if q.option_type is OptionType.Call else _lib.PUT
            o.spot, o.strike, o.expiry = self.spot, q.strike, q.maturity
            o.rate, o.div_yield, o.vol = self.rate, self.dividend_yield, 0.2
            o.kappa, o.theta, o.xi, o.rho, o.v0 = kappa, theta, xi, rho, v0
            o.num_paths = self.num_paths
            o.num_steps = max(1, int(math.ceil(self.steps_per_year * q.maturity)))
            o.seed = (self.seed + i) & (2**64 - 1)
            out.append(o)
        return out

    def model_prices_and_gradients(self, x: Sequence[float],
                                   quotes: Sequence[OptionPriceQuote]):
        """One batched launch: prices and dPrice/d(v0, kappa, theta, xi, rho)."""
        cfg = engine_cfg(rng_mode=self.rng_mode, greek_mode=self.greek_method.value)
        res = price_raw(self._options(x, quotes), cfg, self.device)
        prices = np.array([r.price for r in res])
        # heston_grad order: S0, v0, kappa, theta, xi, rho
        grads = np.array([[r.heston_grad[k] for k in (1, 2, 3, 4, 5)] for r in res])
        return prices, grads, res[0].kernel_ms

    def calibrate(self, quotes: Sequence[OptionPriceQuote],
                  initial: Optional[Sequence[float]] = None) -> HestonMcCalibrationResult:
        if not quotes:
            raise InvalidInput("heston calibration requires non-empty instrument set")
        if any(q.strike <= 0.0 or q.maturity <= 0.0 or not math.isfinite(q.market_price)
               or q.market_price < 0.0 for q in quotes):
            raise InvalidInput("invalid Heston quote set")
        market = np.array([q.market_price for q in quotes])
        sw = np.sqrt(np.maximum([q.weight for q in quotes], 1e-12))
        stats = {"launches": 0, "ms": 0.0}

        def fn(x):
            p, g, ms = self.model_prices_and_gradients(x, quotes)
            stats["launches"] += 1
            stats["ms"] += ms
            return (p - market) * sw, g * sw[:, None]

        x0 = list(initial) if initial is not None else self.initial_guess(quotes)
        opt = levenberg_marquardt(x0, self.bounds, self.lm_options, fn)
        prices = market + opt.residuals / sw
        s = np.linalg.svd(opt.jacobian, compute_uv=False)
        cond = float(s[0] / s[-1]) if s.size and s[-1] > 0 else math.inf
        return HestonMcCalibrationResult(
            params=HestonCalibrationParams.from_slice(opt.x), objective=opt.objective,
            model_prices=prices.tolist(), residuals=opt.residuals.tolist(),
            jacobian=opt.jacobian.tolist(), condition_number=cond,
            convergence=opt.convergence, launches=stats["launches"], kernel_ms=stats["ms"])
