"""Heston calibration with GPU Monte Carlo prices and gradients (C ABI calls).

* Self-consistency: quotes priced by the engine itself at known parameters,
  with the calibrator's own seeds, are matched exactly at those parameters
  (common random numbers), so LM from the reference's initial guess must
  recover them; a wrong Jacobian sign or column mapping cannot.
* Semi-analytic quotes (Lewis formula, oracle/heston_analytic.py): the
  calibrated parameters reproduce the quotes to within the Monte Carlo and
  Euler error of the model prices.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TRUE = [0.045, 2.0, 0.055, 0.55, -0.65]   # v0, kappa, theta, sigma_v, rho
GRID_T = [0.25, 0.5, 1.0, 2.0]
GRID_K = [80.0, 90.0, 95.0, 100.0, 105.0, 110.0, 125.0]


def _quotes(prices, option_type=None):
    from openferric_amd import OptionType
    from openferric_amd.calibration import OptionPriceQuote
    out, i = [], 0
    for t in GRID_T:
        for k in GRID_K:
            out.append(OptionPriceQuote(k, t, float(prices[i]), 1.0,
                                        option_type or OptionType.Call))
            i += 1
    return out


@pytest.mark.parametrize("method", ["forward", "adjoint"])
def test_calibration_recovers_parameters_under_common_random_numbers(gpu, method):
    from openferric_amd import GreekMethod
    from openferric_amd.calibration import HestonMcCalibrator
    cal = HestonMcCalibrator(num_paths=1 << 14, greek_method={
        "forward": GreekMethod.Forward, "adjoint": GreekMethod.Adjoint}[method])
    placeholder = _quotes(np.ones(len(GRID_T) * len(GRID_K)))
    market, _, _ = cal.model_prices_and_gradients(TRUE, placeholder)
    quotes = _quotes(market)
    res = cal.calibrate(quotes)
    got = res.params.to_vec()
    assert res.objective < 1e-12, (res.objective, got, res.convergence)
    for g, t in zip(got, TRUE):
        assert abs(g - t) <= 2e-3 * max(abs(t), 0.05), (got, TRUE)
    assert res.launches <= cal.lm_options.max_iterations + 1


def test_calibration_to_semi_analytic_quotes(gpu):
    from oracle.heston_analytic import heston_call
    from openferric_amd.calibration import HestonMcCalibrator
    cal = HestonMcCalibrator(num_paths=1 << 17)
    v0, kappa, theta, xi, rho = TRUE
    market = [heston_call(100.0, k, t, cal.rate, 0.0, kappa, theta, xi, rho, v0)
              for t in GRID_T for k in GRID_K]
    res = cal.calibrate(_quotes(market))
    # the fit: price errors within a few cents (MC noise with 2^17 paths and
    # the Euler bias at 252 steps per year are both of that order)
    err = np.abs(np.array(res.model_prices) - np.array(market))
    assert err.max() < 0.08, err.max()
    assert np.sqrt(np.mean(err ** 2)) < 0.03
    p = res.params
    assert abs(p.v0 - v0) < 0.006 and abs(p.theta - theta) < 0.012
    assert abs(p.rho - rho) < 0.15 and abs(p.sigma_v - xi) < 0.2
