"""The Python mirror of the reference's contract types (CPU only): validation
messages, builders, diagnostics, and the C ABI structs' layout."""
import ctypes as C

import pytest

import openferric_amd as ofa
from openferric_amd import _lib


def test_market_builder_validation():
    # market/market.rs:367-414
    with pytest.raises(ofa.InvalidInput, match="market spot is required"):
        ofa.Market.builder().flat_vol(0.2).build()
    with pytest.raises(ofa.InvalidInput, match="market spot must be > 0"):
        ofa.Market.builder().spot(0.0).flat_vol(0.2).build()
    with pytest.raises(ofa.InvalidInput, match="either market flat_vol or vol_surface is required"):
        ofa.Market.builder().spot(100.0).build()
    with pytest.raises(ofa.InvalidInput, match="market flat_vol must be > 0"):
        ofa.Market.builder().spot(100.0).flat_vol(-0.1).build()
    m = ofa.Market.builder().spot(100.0).rate(0.05).flat_vol(0.2).build()
    assert m.dividend_yield == 0.0 and m.vol_for(90.0, 1.0) == 0.2


def test_instrument_validation_messages():
    with pytest.raises(ofa.InvalidInput, match="vanilla strike must be > 0"):
        ofa.VanillaOption.european_call(0.0, 1.0).validate()
    with pytest.raises(ofa.InvalidInput, match="vanilla expiry must be >= 0"):
        ofa.VanillaOption.european_put(100.0, -1.0).validate()
    spec = ofa.AsianSpec(ofa.Averaging.Arithmetic, ofa.StrikeType.Fixed, [])
    with pytest.raises(ofa.InvalidInput, match="asian observation_times cannot be empty"):
        ofa.AsianOption(ofa.OptionType.Call, 100.0, 1.0, spec).validate()
    spec = ofa.AsianSpec(ofa.Averaging.Arithmetic, ofa.StrikeType.Fixed, [0.5, 1.2])
    with pytest.raises(ofa.InvalidInput, match=r"must lie in \[0, expiry\]"):
        ofa.AsianOption(ofa.OptionType.Call, 100.0, 1.0, spec).validate()
    with pytest.raises(ofa.InvalidInput, match="barrier option type is required"):
        ofa.BarrierOption.builder().strike(100.0).expiry(1.0).up_and_out(120.0).build()
    with pytest.raises(ofa.InvalidInput, match="barrier level must be > 0"):
        ofa.BarrierOption.builder().call().strike(100.0).expiry(1.0).up_and_out(0.0).build()
    b = ofa.BarrierOption.builder().put().strike(100.0).expiry(1.0).down_and_in(80.0).rebate(2.0).build()
    assert b.barrier.style is ofa.BarrierStyle.In and b.barrier.rebate == 2.0


def test_error_display_matches_reference():
    # core/engine.rs:304-313
    assert str(ofa.InvalidInput("x")) == "invalid input: x"
    assert str(ofa.NumericalError("y")) == "numerical error: y"


def test_diagnostics_capacity_and_keys():
    d = ofa.Diagnostics()
    for k in ["num_paths", "num_steps", "vol", "delta", "d1", "d2", "npv", "rho"]:
        d.insert(k, 1.0)
    assert d.insert("vol", 2.0) == 1.0
    with pytest.raises(OverflowError):
        d.insert("u", 1.0)
    with pytest.raises(KeyError):
        ofa.Diagnostics().insert("not_a_key", 1.0)


def test_heston_validate():
    assert ofa.Heston(0.03, 1.8, 0.04, 0.6, -0.5, 0.04).validate()
    assert not ofa.Heston(0.03, 0.0, 0.04, 0.6, -0.5, 0.04).validate()
    assert not ofa.Heston(0.03, 1.0, 0.04, 0.6, 1.0, 0.04).validate()


def test_ctypes_structs_match_header_layout():
    # sizes and offsets of the C ABI structs as declared in include/openferric_amd.h
    assert C.sizeof(_lib.OfgOption) == 8 * 4 + 13 * 8 + 8 + 3 * 8
    assert _lib.OfgOption.obs_times.offset == 32 + 13 * 8
    assert C.sizeof(_lib.OfgEngineCfg) == 4 * 4 + 8 + 2 * 8 + 2 * 4
    assert C.sizeof(_lib.OfgResult) == 2 * 8 + 2 * 4 + 5 * 8 + 6 * 8 + 4 * 8


def test_gpu_calls_fail_loudly_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    m = ofa.Market(100.0, 0.05, 0.0, 0.2)
    with pytest.raises(ofa.DeviceError):
        ofa.GpuMonteCarloPricingEngine(1000, 10, 1).price(
            ofa.VanillaOption.european_call(100.0, 1.0), m)
