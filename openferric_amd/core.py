"""Contract types mirroring OpenFerric's pricing interface.

Mirrors src/core/{engine,types}.rs, src/instruments/{vanilla,asian,barrier}.rs,
src/market/market.rs (flat-vol Market) and src/models/stochastic.rs (Gbm,
Heston): same names, argument meaning and validation messages.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence


# ---------------------------------------------------------------- errors ---
class PricingError(Exception):
    """PricingError (src/core/engine.rs:292-302)."""

    prefix = "pricing error"

    def __str__(self) -> str:
        return f"{self.prefix}: {self.args[0] if self.args else ''}"


class InvalidInput(PricingError):
    prefix = "invalid input"


class ConvergenceFailure(PricingError):
    prefix = "convergence failure"


class MarketDataMissing(PricingError):
    prefix = "market data missing"


class NumericalError(PricingError):
    prefix = "numerical error"


# ----------------------------------------------------------------- enums ---
class OptionType(enum.Enum):
    """core/types.rs:13-38"""
    Call = 0
    Put = 1

    def sign(self) -> float:
        return 1.0 if self is OptionType.Call else -1.0


class ExerciseStyle(enum.Enum):
    European = "european"
    American = "american"
    Bermudan = "bermudan"


class BarrierDirection(enum.Enum):
    Up = 0
    Down = 1


class BarrierStyle(enum.Enum):
    In = 0
    Out = 1


class Averaging(enum.Enum):
    Arithmetic = 0
    Geometric = 1


class StrikeType(enum.Enum):
    Fixed = 0
    Floating = 1


@dataclass
class BarrierSpec:
    """core/types.rs:66-96"""
    direction: BarrierDirection
    style: BarrierStyle
    level: float
    rebate: float = 0.0


@dataclass
class AsianSpec:
    """core/types.rs:114-122"""
    averaging: Averaging
    strike_type: StrikeType
    observation_times: List[float]


# --------------------------------------------------------------- results ---
@dataclass
class Greeks:
    """core/engine.rs:15-28"""
    delta: float
    gamma: float
    vega: float
    theta: float
    rho: float


DIAG_KEYS = (
    "barrier_level", "conversion_value", "credit_spread", "d", "d1", "d2", "delta",
    "discount_factor", "double_knockout_base", "effective_vol", "exercise_dates",
    "fair_variance", "fair_volatility", "inside_barriers", "integral", "max_exercises",
    "min_exercises", "npv", "num_paths", "num_threads", "observation_count", "num_space_steps",
    "num_steps", "num_time_steps", "pd", "pm", "pu", "pv_forward", "rho", "rho_domestic",
    "rho_foreign", "s_max", "series_terms", "survival_digital", "u", "var_of_var", "vol",
    "vol_adj",
)


class Diagnostics:
    """Fixed 8-slot diagnostics map (core/engine.rs:191-274)."""

    CAPACITY = 8

    def __init__(self):
        self._entries: Dict[str, float] = {}

    def insert(self, key: str, value: float) -> Optional[float]:
        if key not in DIAG_KEYS:
            raise KeyError(f"unsupported diagnostics key `{key}`; add it to core::DiagKey")
        prev = self._entries.get(key)
        if prev is None and len(self._entries) >= self.CAPACITY:
            raise OverflowError(f"diagnostics capacity exceeded ({self.CAPACITY})")
        self._entries[key] = value
        return prev

    def get(self, key: str) -> Optional[float]:
        return self._entries.get(key)

    def contains_key(self, key: str) -> bool:
        return key in self._entries

    def __len__(self) -> int:
        return len(self._entries)

    def is_empty(self) -> bool:
        return not self._entries

    def items(self):
        return self._entries.items()

    def __repr__(self) -> str:
        return f"Diagnostics({self._entries})"


@dataclass
class PricingResult:
    """core/engine.rs:277-287"""
    price: float
    stderr: Optional[float] = None
    greeks: Optional[Greeks] = None
    diagnostics: Diagnostics = field(default_factory=Diagnostics)
    # Heston parameter gradient dPrice/d{S0, v0, kappa, theta, xi, rho}
    # (extension: the reference Greeks type has no slot for these).
    heston_gradient: Optional[Dict[str, float]] = None
    kernel_ms: Optional[float] = None


# ----------------------------------------------------------- instruments ---
@dataclass
class VanillaOption:
    """instruments/vanilla.rs:16-96"""
    option_type: OptionType
    strike: float
    expiry: float
    exercise: ExerciseStyle = ExerciseStyle.European
    bermudan_dates: Sequence[float] = ()

    @staticmethod
    def european_call(strike: float, expiry: float) -> "VanillaOption":
        return VanillaOption(OptionType.Call, strike, expiry)

    @staticmethod
    def european_put(strike: float, expiry: float) -> "VanillaOption":
        return VanillaOption(OptionType.Put, strike, expiry)

    @staticmethod
    def american_call(strike: float, expiry: float) -> "VanillaOption":
        return VanillaOption(OptionType.Call, strike, expiry, ExerciseStyle.American)

    @staticmethod
    def american_put(strike: float, expiry: float) -> "VanillaOption":
        return VanillaOption(OptionType.Put, strike, expiry, ExerciseStyle.American)

    def validate(self) -> None:
        if not self.strike > 0.0:
            raise InvalidInput("vanilla strike must be > 0")
        if self.expiry < 0.0:
            raise InvalidInput("vanilla expiry must be >= 0")
        if self.exercise is ExerciseStyle.Bermudan:
            if not self.bermudan_dates:
                raise InvalidInput("bermudan exercise dates cannot be empty")
            if any(d <= 0.0 or d > self.expiry for d in self.bermudan_dates):
                raise InvalidInput("bermudan exercise dates must lie in (0, expiry]")

    def instrument_type(self) -> str:
        return "VanillaOption"


@dataclass
class AsianOption:
    """instruments/asian.rs:16-73"""
    option_type: OptionType
    strike: float
    expiry: float
    asian: AsianSpec

    def validate(self) -> None:
        if self.expiry < 0.0:
            raise InvalidInput("asian expiry must be >= 0")
        if not self.strike > 0.0:
            raise InvalidInput("asian strike must be > 0")
        if not self.asian.observation_times:
            raise InvalidInput("asian observation_times cannot be empty")
        if any(t < 0.0 or t > self.expiry for t in self.asian.observation_times):
            raise InvalidInput("asian observation_times must lie in [0, expiry]")

    def instrument_type(self) -> str:
        return "AsianOption"


@dataclass
class BarrierOption:
    """instruments/barrier.rs:18-59"""
    option_type: OptionType
    strike: float
    expiry: float
    barrier: BarrierSpec

    @staticmethod
    def builder() -> "BarrierOptionBuilder":
        return BarrierOptionBuilder()

    def validate(self) -> None:
        if not self.strike > 0.0:
            raise InvalidInput("barrier strike must be > 0")
        if self.expiry < 0.0:
            raise InvalidInput("barrier expiry must be >= 0")
        if not self.barrier.level > 0.0:
            raise InvalidInput("barrier level must be > 0")
        if self.barrier.rebate < 0.0:
            raise InvalidInput("barrier rebate must be >= 0")

    def instrument_type(self) -> str:
        return "BarrierOption"


@dataclass
class BermudanOption:
    """instruments/bermudan.rs:18-144: exercise dates with a per-date strike."""
    option_type: OptionType
    expiry: float
    exercise_dates: Sequence[float]
    strike_schedule: Sequence[float]

    @staticmethod
    def new(option_type: OptionType, expiry: float, exercise_dates: Sequence[float],
            strike_schedule: Sequence[float]) -> "BermudanOption":
        return BermudanOption(option_type, expiry, list(exercise_dates), list(strike_schedule))

    @staticmethod
    def with_constant_strike(option_type: OptionType, strike: float, expiry: float,
                             exercise_dates: Sequence[float]) -> "BermudanOption":
        return BermudanOption(option_type, expiry, list(exercise_dates),
                              [strike] * len(exercise_dates))

    def num_exercise_dates(self) -> int:
        return len(self.exercise_dates)

    def effective_schedule(self):
        """(time, strike) pairs; the expiry with the last strike is appended
        when the last date falls before it (bermudan.rs:85-103)."""
        self.validate()
        out = list(zip(self.exercise_dates, self.strike_schedule))
        if self.expiry - out[-1][0] > 1.0e-12:
            out.append((self.expiry, out[-1][1]))
        return out

    def validate(self) -> None:
        if not (self.expiry > 0.0 and math.isfinite(self.expiry)):
            raise InvalidInput("bermudan expiry must be finite and > 0")
        if not self.exercise_dates:
            raise InvalidInput("bermudan exercise_dates cannot be empty")
        if len(self.exercise_dates) != len(self.strike_schedule):
            raise InvalidInput("bermudan strike_schedule must match exercise_dates length")
        if any(not math.isfinite(t) or t <= 0.0 or t > self.expiry for t in self.exercise_dates):
            raise InvalidInput("bermudan exercise_dates must be finite and in (0, expiry]")
        if any(b <= a for a, b in zip(self.exercise_dates, self.exercise_dates[1:])):
            raise InvalidInput("bermudan exercise_dates must be strictly increasing")
        if any(not math.isfinite(k) or k <= 0.0 for k in self.strike_schedule):
            raise InvalidInput("bermudan strike_schedule entries must be finite and > 0")

    def instrument_type(self) -> str:
        return "BermudanOption"


@dataclass
class ExerciseBoundaryPoint:
    """engines/lsm/longstaff_schwartz.rs:42-54"""
    time: float
    strike: float
    boundary_spot: Optional[float]
    itm_paths: int
    exercised_paths: int


@dataclass
class BermudanLsmOutput:
    """engines/lsm/longstaff_schwartz.rs:57-63"""
    result: PricingResult
    exercise_boundary: List[ExerciseBoundaryPoint]


class BarrierOptionBuilder:
    """instruments/barrier.rs:66-178"""

    def __init__(self):
        self._type = self._strike = self._expiry = None
        self._dir = self._style = self._level = None
        self._rebate = None

    def call(self):
        self._type = OptionType.Call
        return self

    def put(self):
        self._type = OptionType.Put
        return self

    def strike(self, k):
        self._strike = k
        return self

    def expiry(self, t):
        self._expiry = t
        return self

    def _set(self, d, s, level):
        self._dir, self._style, self._level = d, s, level
        return self

    def up_and_out(self, level):
        return self._set(BarrierDirection.Up, BarrierStyle.Out, level)

    def up_and_in(self, level):
        return self._set(BarrierDirection.Up, BarrierStyle.In, level)

    def down_and_out(self, level):
        return self._set(BarrierDirection.Down, BarrierStyle.Out, level)

    def down_and_in(self, level):
        return self._set(BarrierDirection.Down, BarrierStyle.In, level)

    def rebate(self, r):
        self._rebate = r
        return self

    def build(self) -> BarrierOption:
        if self._type is None:
            raise InvalidInput("barrier option type is required")
        if self._strike is None:
            raise InvalidInput("barrier strike is required")
        if self._expiry is None:
            raise InvalidInput("barrier expiry is required")
        if self._dir is None:
            raise InvalidInput("barrier direction is required")
        if self._style is None:
            raise InvalidInput("barrier style is required")
        if self._level is None:
            raise InvalidInput("barrier level is required")
        opt = BarrierOption(self._type, self._strike, self._expiry,
                            BarrierSpec(self._dir, self._style, self._level,
                                        self._rebate if self._rebate is not None else 0.0))
        opt.validate()
        return opt


# ---------------------------------------------------------------- market ---
@dataclass
class Market:
    """Market snapshot with a flat vol source (market/market.rs:202-302)."""
    spot: float
    rate: float = 0.0
    dividend_yield: float = 0.0
    flat_vol: float = 0.0

    @staticmethod
    def builder() -> "MarketBuilder":
        return MarketBuilder()

    def vol_for(self, strike: float, expiry: float) -> float:
        return self.flat_vol

    def has_discrete_dividends(self) -> bool:
        return False

    def effective_dividend_yield(self, maturity: float) -> float:
        """market.rs:275-282 with an empty dividend schedule
        (dividends.rs:151-217): r - ln(F/S)/T with F = S exp((r - q) T), which
        equals q up to rounding."""
        if maturity <= 0.0 or self.spot <= 0.0:
            return self.dividend_yield
        forward = self.spot * math.exp((self.rate - self.dividend_yield) * maturity)
        ratio = max(forward / self.spot, 1.0e-16)
        return self.rate - math.log(ratio) / maturity


class MarketBuilder:
    """market/market.rs:367-414 (flat vol, no discrete dividends)."""

    def __init__(self):
        self._spot = None
        self._rate = None
        self._q = None
        self._vol = None

    def spot(self, s):
        self._spot = s
        return self

    def rate(self, r):
        self._rate = r
        return self

    def dividend_yield(self, q):
        self._q = q
        return self

    def flat_vol(self, v):
        self._vol = v
        return self

    def build(self) -> Market:
        if self._spot is None:
            raise InvalidInput("market spot is required")
        if self._spot <= 0.0:
            raise InvalidInput("market spot must be > 0")
        if self._vol is None:
            raise InvalidInput("either market flat_vol or vol_surface is required")
        if self._vol <= 0.0:
            raise InvalidInput("market flat_vol must be > 0")
        return Market(self._spot, self._rate or 0.0, self._q or 0.0, self._vol)


# ---------------------------------------------------------------- models ---
@dataclass
class Gbm:
    """models/stochastic.rs:12-34"""
    mu: float
    sigma: float


@dataclass
class Heston:
    """models/stochastic.rs:37-72"""
    mu: float
    kappa: float
    theta: float
    xi: float
    rho: float
    v0: float

    def validate(self) -> bool:
        return (self.kappa > 0.0 and self.theta >= 0.0 and self.xi >= 0.0 and self.v0 >= 0.0
                and -1.0 < self.rho < 1.0)
