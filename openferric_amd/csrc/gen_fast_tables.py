"""Writes ofg_fast_tables.inc: the constant tables of the fast RNG mode's math
(staged once per block into LDS by the kernels).

    python openferric_amd/csrc/gen_fast_tables.py

Every entry is computed with mpmath at 50 digits and rounded once to double.
Layout of OFG_FAST_TABLE (doubles):
  [0, 512)     log: 256 pairs (T1, 2 log T1). Interval i of m in [0.5, 1) has
               midpoint c_i = 0.5 + (i + 1/2)/512 and T1 = 1/c_i rounded; the
               last interval uses T1 = 1 so that r = m - 1 is exact near 1.
  [512, 768)   exp2: 2^(j/256), j = 0..255.
  [768, 1792)  sincos: 512 pairs (sin a_i, cos a_i), a_i = 2 pi (i + 1/2) / 512.
It also prints the fitted coefficient of the cos correction polynomial.
"""
import os

import mpmath as mp

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def _d(x) -> str:
    return repr(float(x))


def log_pairs():
    out = []
    for i in range(256):
        if i == 255:
            t1 = 1.0
        else:
            c = mp.mpf(1) / 2 + (mp.mpf(i) + mp.mpf(1) / 2) / 512
            t1 = float(1 / c)
        out.append((t1, 2 * mp.log(mp.mpf(t1))))
    return out


def exp2_entries():
    return [mp.power(2, mp.mpf(j) / 256) for j in range(256)]


def sincos_pairs():
    out = []
    for i in range(512):
        a = 2 * mp.pi * (mp.mpf(i) + mp.mpf(1) / 2) / 512
        out.append((mp.sin(a), mp.cos(a)))
    return out


def cos_fit():
    """c in cos d - 1 ~ w (-1/2 + c w), w = d^2, |d| <= pi/512: the value that
    minimises the maximum absolute error (grid search on the error's
    equioscillation; the Taylor value 1/24 gives 7.4e-17)."""
    W = (mp.pi / 512) ** 2
    ws = mp.linspace(0, W, 300)

    def err(c):
        return max(abs(w * (-mp.mpf(1) / 2 + c * w) - (mp.cos(mp.sqrt(w)) - 1)) for w in ws)

    lo, hi = mp.mpf(0.8), mp.mpf(1.0)
    for _ in range(40):  # ternary search on k, c = 1/24 - k W / 720
        k1, k2 = lo + (hi - lo) / 3, hi - (hi - lo) / 3
        if err(mp.mpf(1) / 24 - k1 * W / 720) < err(mp.mpf(1) / 24 - k2 * W / 720):
            hi = k2
        else:
            lo = k1
    c = mp.mpf(1) / 24 - (lo + hi) / 2 * W / 720
    return float(c), float(err(mp.mpf(float(c))))


def main():
    lines = ["// Generated by gen_fast_tables.py (mpmath, 50 digits, rounded once).",
             "// Layout: [0,512) log pairs (T1, 2 log T1); [512,768) 2^(j/256);",
             "// [768,1792) (sin, cos) of 2 pi (i + 1/2) / 512.",
             "#define OFG_FAST_TABLE_LEN 1792",
             "static constexpr double OFG_FAST_TABLE[OFG_FAST_TABLE_LEN] = {"]
    for t1, t2 in log_pairs():
        lines.append(f"    {_d(t1)}, {_d(t2)},")
    for v in exp2_entries():
        lines.append(f"    {_d(v)},")
    for s, c in sincos_pairs():
        lines.append(f"    {_d(s)}, {_d(c)},")
    lines.append("};")
    with open(os.path.join(HERE, "ofg_fast_tables.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")
    c, e = cos_fit()
    print(f"cos correction coefficient {c!r} (max abs error {e:.3g})")


if __name__ == "__main__":
    main()
