"""Structured products written in the product language for the tests (test
data; the products follow the families the language documents: notes,
autocallables, Phoenix coupons, accumulators, range accruals, digitals)."""

ZERO_COUPON = """\
product "Zero"
    notional: 1_000_000
    maturity: 2.0
    underlyings
        A = asset(0)
    schedule annual from 2.0 to 2.0
        redeem notional
"""

FORWARD = """\
product "Fwd"
    notional: 100
    maturity: 1.5
    underlyings
        A = asset(0)
    schedule 1.5 from 1.5 to 1.5
        redeem notional * price(0) / 100
"""

CALL = """\
product "Call"
    notional: 1
    maturity: 1.0
    underlyings
        A = asset(0)
    schedule annual from 1.0 to 1.0
        pay max(price(0) - 100, 0)
"""

WORST_DIGITAL = """\
product "Worst digital"
    notional: 100
    maturity: 1.0
    underlyings
        A = asset(0)
        B = asset(1)
    schedule annual from 1.0 to 1.0
        if worst_of(performances()) >= 1.0 then
            pay notional
"""

TERMINAL_VALUE = """\
product "Terminal value"
    notional: 1
    maturity: 2.0
    underlyings
        R = rate(0)
    schedule 2.0 from 2.0 to 2.0
        redeem price(0)
"""

TERMINAL_LOG = """\
product "Terminal log"
    notional: 1
    maturity: 2.0
    underlyings
        C = commodity(0)
    schedule 2.0 from 2.0 to 2.0
        redeem log(price(0))
"""


def standard_autocall(n_und, barrier, coupon, maturity, freq="quarterly", first=None):
    """A standard worst-of autocallable without knock-in, written so that it
    matches the autocallable pricer's standard note: called at the first
    observation with worst >= barrier for notional * (1 + c t), otherwise
    notional * (1 + c T) at maturity."""
    und = "\n".join(f"        U{i} = asset({i})" for i in range(n_und))
    first = first if first is not None else {"quarterly": 0.25, "semi_annual": 0.5}[freq]
    return f"""\
product "Standard autocall"
    notional: 100
    maturity: {maturity}
    underlyings
{und}
    schedule {freq} from {first} to {maturity}
        let wof = worst_of(performances())
        if wof >= {barrier} then
            redeem notional * (1 + {coupon} * observation_date)
        if is_final then
            redeem notional * (1 + {coupon} * {maturity})
"""


def phoenix_memory(n_und, barrier, coupon_barrier, coupon, maturity):
    """A worst-of Phoenix note with memory coupons over quarterly dates,
    matching the autocallable pricer's Phoenix note without knock-in."""
    und = "\n".join(f"        U{i} = asset({i})" for i in range(n_und))
    return f"""\
product "Phoenix memory"
    notional: 100
    maturity: {maturity}
    underlyings
{und}
    state
        missed: float = 0.0
    schedule quarterly from 0.25 to {maturity}
        let wof = worst_of(performances())
        let accr = notional * {coupon} * 0.25
        if wof >= {coupon_barrier} then
            pay missed + accr
            set missed = 0.0
        else
            set missed = missed + accr
        if wof >= {barrier} then
            redeem notional
        if is_final then
            redeem notional
"""


KITCHEN_SINK = """\
product "Kitchen sink"
    notional: 1_000
    maturity: 2.0
    underlyings
        SPX = asset(0)
        EUR = fx(1)
        OIL = commodity(2)
        RT = rate(3)
    state
        ki: bool = false
        acc: float = 0.0
        hits: float = 0
        flag: bool = 1
    // monthly accrual with a target and a range count
    schedule monthly from 0.0833 to 1.0
        let perf = worst_of(performances())
        let b = best_of(price(0) / 100, price(1) / 1.1, 0.5)
        if perf <= 0.7 then
            set ki = true
        if acc < 0.5 then
            if perf >= 0.9 and perf <= 1.1 then
                set acc = acc + 0.0833
                pay notional * 0.0833 * (perf - 1.0)
            else if perf > 1.1 or not flag then
                set hits = hits + 1
            else
                pay -notional * 0.001 * abs(perf - 1)
        if b == b and b != 0.25 then
            set flag = min(b, 2) > 0.95
    schedule semi_annual from 1.5 to 2.0
        let r = price(3)
        let s = exp(log(price(2) / 80) * 0.5)
        if is_final then
            if ki and worst_of(performances()) < 1.0 then
                redeem notional * max(worst_of(performances()), 0.2)
            else
                redeem notional * (1 + hits * 0.01 + max(r, 0) + 0.1 * min(s, 1.5))
        else
            if r > 0.2 then
                skip
            pay notional * 0.01 * (r - 0.02) / (observation_date - 1.0)
"""

RANGE_ACCRUAL = """\
product "Range accrual"
    notional: 1_000_000
    maturity: 1.0
    underlyings
        A = asset(0)
        B = asset(1)
        C = asset(2)
    state
        days: float = 0.0
    schedule monthly from 0.0833 to 1.0
        let perf = worst_of(performances())
        if perf >= 0.8 and perf <= 1.2 then
            set days = days + 1.0
        if is_final then
            pay notional * 0.06 * days / 12
            redeem notional
"""

OUT_OF_RANGE_PRICE = """\
product "Bad index"
    notional: 1
    maturity: 1.0
    underlyings
        A = asset(0)
    schedule annual from 1.0 to 1.0
        if price(0) > 0 then
            pay price(5)
"""

MASKED_BAD_INDEX = """\
product "Masked index"
    notional: 1
    maturity: 1.0
    underlyings
        A = asset(0)
    schedule annual from 1.0 to 1.0
        if price(0) < 0 then
            pay price(5)
        redeem price(0)
"""
