"""openferric_amd: MI355X (gfx950) Monte Carlo pricing engine for OpenFerric's
MC hot path (GBM / Heston paths, payoffs, reduction, pathwise Greeks).

The compute path is libopenferric_amd.so (hand-written HIP kernels behind a C
ABI, include/openferric_amd.h). This package mirrors the reference's Rust
PricingEngine contract on top of it.
"""
from .core import (AsianOption, AsianSpec, Averaging, BarrierDirection, BarrierOption,
                   BarrierSpec, BarrierStyle, BermudanLsmOutput, BermudanOption,
                   ExerciseBoundaryPoint, ConvergenceFailure, Diagnostics, ExerciseStyle, Gbm,
                   Greeks, Heston, InvalidInput, Market, MarketDataMissing, NumericalError,
                   OptionType, PricingError, PricingResult, StrikeType, VanillaOption)
from .engines import (BatchPricer, GpuArithmeticAsianMC, GpuMonteCarloPricingEngine, GreekMethod,
                      HestonAadConfig, HestonMonteCarloPricingEngine, RngMode, VarianceReduction,
                      MonteCarloGreeksEngine, autotune, engine_cfg, heston_price_delta_aad,
                      mc_european_parallel, mc_european_pathwise_aad, mc_european_sequential,
                      mc_european_with_arena, LongstaffSchwartzEngine, LsmDynamics,
                      lsm_last_timing, lsm_option_struct, lsm_price_raw, stream_jump,
                      mc_european_qmc, mc_european_qmc_with_seed, qmc_price_raw,
                      option_struct, price_noalloc_raw, price_raw, sample_payoffs, fill_normals,
                      european_exact_raw)
from .multi import (Autocallable, AutocallableSensitivities, BasketCopula, BasketOption,
                    FactorCorrelationModel, price_basket_mc_with_factor_model,
                    price_basket_mc_with_copula, CorrelationStressScenario,
                    stressed_correlation_matrix,
                    BasketSensitivities, BasketType, OutperformanceBasketOption,
                    PhoenixAutocallable, QuantoBasketOption, autocall_price_raw,
                    autocall_struct, autocallable_sensitivities, basket_price_raw,
                    basket_sensitivities, basket_struct, phoenix_autocallable_sensitivities,
                    price_autocallable, price_basket_mc, price_outperformance_basket_mc,
                    price_phoenix_autocallable, price_quanto_basket_mc,
                    validate_or_repair_correlation_matrix)
from .dsl import (AssetMarketData, CompiledProduct, CrossGreeks, DslMonteCarloEngine, DslProduct,
                  ExtendedGreeks, MultiAssetMarket, parse_and_compile)
from ._lib import DeviceError

from .greeks import (FiniteDifferenceGreeks, McPricer, bump_and_reprice, bump_and_reprice_raw,
                     fd_greeks_raw, finite_difference_greeks)
from .engines import FastRngKind, kind_rng_mode, rng_kind_from_str
from .spread import (SpreadMonteCarloEngine, SpreadOption, spread_price, spread_price_raw,
                     spread_struct)
__all__ = [
    "FastRngKind", "kind_rng_mode", "SpreadMonteCarloEngine", "SpreadOption", "rng_kind_from_str", "spread_price",
    "spread_price_raw", "spread_struct",
    "FiniteDifferenceGreeks", "McPricer", "bump_and_reprice", "bump_and_reprice_raw",
    "fd_greeks_raw", "finite_difference_greeks",
    "AsianOption", "AsianSpec", "Averaging", "BarrierDirection", "BarrierOption", "BarrierSpec",
    "BarrierStyle", "ConvergenceFailure", "Diagnostics", "ExerciseStyle", "Gbm", "Greeks",
    "Heston", "InvalidInput", "Market", "MarketDataMissing", "NumericalError", "OptionType",
    "PricingError", "PricingResult", "StrikeType", "VanillaOption", "BatchPricer",
    "GpuArithmeticAsianMC", "GpuMonteCarloPricingEngine", "HestonAadConfig",
    "HestonMonteCarloPricingEngine", "RngMode", "VarianceReduction", "autotune", "engine_cfg",
    "heston_price_delta_aad", "mc_european_pathwise_aad", "option_struct", "price_raw", "price_noalloc_raw", "european_exact_raw",
    "sample_payoffs", "fill_normals", "DeviceError", "MonteCarloGreeksEngine", "mc_european_parallel",
    "mc_european_sequential", "mc_european_with_arena", "GreekMethod", "BermudanOption",
    "BermudanLsmOutput", "ExerciseBoundaryPoint", "LongstaffSchwartzEngine", "LsmDynamics",
    "lsm_last_timing", "lsm_option_struct", "lsm_price_raw", "stream_jump",
    "mc_european_qmc", "mc_european_qmc_with_seed", "qmc_price_raw",
    "Autocallable", "AutocallableSensitivities", "PhoenixAutocallable", "autocall_price_raw",
    "BasketOption", "BasketSensitivities", "BasketType", "basket_price_raw",
    "basket_sensitivities", "basket_struct", "price_basket_mc", "OutperformanceBasketOption",
    "QuantoBasketOption", "price_outperformance_basket_mc", "price_quanto_basket_mc",
    "BasketCopula", "FactorCorrelationModel", "price_basket_mc_with_factor_model",
    "price_basket_mc_with_copula", "CorrelationStressScenario", "stressed_correlation_matrix",
    "validate_or_repair_correlation_matrix",
    "AssetMarketData", "CompiledProduct", "CrossGreeks", "DslMonteCarloEngine", "DslProduct",
    "ExtendedGreeks", "MultiAssetMarket", "parse_and_compile",
    "autocall_struct", "autocallable_sensitivities", "phoenix_autocallable_sensitivities",
    "price_autocallable", "price_phoenix_autocallable",
]
