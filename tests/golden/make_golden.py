#!/usr/bin/env python3
"""Writes the golden fixtures used to pin the oracle.

Sources (no reference code is imported or executed):
  * the reference's own test fixture tables (CSV data files under
    tests/fixtures of rosssaunders/openferric), converted to JSON:
    QuantLib Heston Lewis prices, QuantLib arithmetic Asian cases, QuantLib /
    Haug European values;
  * high-precision values computed here with mpmath: the inverse normal CDF
    on a grid spanning Acklam's three branches.

Run from the repository root: python tests/golden/make_golden.py
"""
import csv
import json
import os
import re

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = "/root/reference/tests/fixtures"


def rows(name):
    with open(os.path.join(FIX, name)) as f:
        lines = [l for l in f if not l.startswith("#")]
    return list(csv.DictReader(lines))


def frac(s):
    if "/" in s:
        a, b = s.split("/")
        return float(a) / float(b)
    return float(s)


def main():
    out = {}
    # Heston Lewis (QuantLib hestonmodel.cpp via the reference fixtures)
    params = rows("heston_params.csv")[0]
    out["heston_lewis"] = {
        "params": {k: float(v) for k, v in params.items()},
        "cases": [{"strike": float(r["strike"]), "put": float(r["expected_put"]),
                   "call": float(r["expected_call"])} for r in rows("heston_expected_put_call.csv")],
    }
    out["asian_arithmetic_cases4"] = [
        {"line": int(r["source_line"]), "option_type": r["option_type"].split("::")[-1],
         "underlying": float(r["underlying"]), "strike": float(r["strike"]),
         "dividend_yield": float(r["dividend_yield"]), "rate": float(r["risk_free_rate"]),
         "first": frac(r["first"]), "length": frac(r["length"]), "fixings": int(r["fixings"]),
         "vol": float(r["volatility"]), "expected": float(r["expected"])}
        for r in rows("asian_arithmetic_cases4.csv")]
    out["european_haug"] = [
        {"line": int(r["source_line"]), "option_type": r["option_type"],
         "strike": float(r["strike"]), "spot": float(r["spot"]), "q": float(r["dividend"]),
         "rate": float(r["rate"]), "expiry": float(r["expiry"]), "vol": float(r["vol"]),
         "expected": float(r["expected"]), "tol": float(r["tolerance"])}
        for r in rows("european_haug_values.csv")]
    # The inline Lewis table of tests/quantlib_reference.rs:1641-1707 (the CSV
    # fixture's rows for K >= 90 do not match the stated parameters; the
    # inline table does, to 1e-14, with the semi-analytic pricer here).
    src = open("/root/reference/tests/quantlib_reference.rs").read()
    block = src[src.index("fn heston_alan_lewis_cases"):]
    block = block[:block.index("(spot, v0, rho")]
    cases = re.findall(r"OptionType::(Call|Put),\s*strike:\s*([0-9.]+),\s*expected:\s*([0-9_.]+)",
                       block)
    out["heston_lewis_inline"] = [{"option_type": t, "strike": float(k),
                                   "expected": float(e.replace("_", ""))} for t, k, e in cases]
    # Fang-Oosterlee (tests/fang_oosterlee_heston.rs:34-57): T=1 and T=10, K=100
    out["fang_oosterlee"] = {"params": {"spot": 100.0, "rate": 0.0, "q": 0.0, "kappa": 1.5768,
                                        "theta": 0.0398, "xi": 0.5751, "rho": -0.5711,
                                        "v0": 0.0175},
                             "cases": [{"expiry": 1.0, "strike": 100.0, "call": 5.785155450},
                                       {"expiry": 10.0, "strike": 100.0, "call": 22.318945791}]}
    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)

    mp.mp.dps = 40
    grid = [1e-15, 1e-12, 1e-9, 1e-6, 1e-4, 1e-3, 5e-3, 0.01, 0.02, 0.0242, 0.02424999,
            0.02425, 0.02425001, 0.03, 0.05, 0.1, 0.2, 0.3, 0.4, 0.45, 0.49, 0.4999, 0.5,
            0.5001, 0.51, 0.55, 0.6, 0.7, 0.8, 0.9, 0.95, 0.97, 0.97574999, 0.97575,
            0.97575001, 0.98, 0.99, 0.999, 0.9999, 1 - 1e-6, 1 - 1e-9, 1 - 2.220446049250313e-16]
    grid += [i / 997.0 for i in range(1, 997)]
    inv = [[p, float(mp.sqrt(2) * mp.erfinv(2 * mp.mpf(p) - 1))] for p in grid]
    with open(os.path.join(HERE, "inverse_normal_mpmath.json"), "w") as f:
        json.dump(inv, f)


if __name__ == "__main__":
    main()
