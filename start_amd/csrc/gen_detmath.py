"""Constants of detmath.h, computed from their definitions with Python's
decimal arithmetic at 80 digits and rounded to nearest double.

    python start_amd/csrc/gen_detmath.py          # print the constant block
    python start_amd/csrc/gen_detmath.py --check  # exit 1 if detmath.h differs

pi comes from Machin's formula, atan(k/8) from the Taylor series after two
argument halvings (atan x = 2 atan(x / (1 + sqrt(1 + x^2)))), ln 2 from the
decimal module's ln.  Splits: "hi" values rounded to a stated number of
significant bits (so that k * hi is exact for the small integers k the
reductions multiply it by), "lo" values the nearest double to the rest.
"""
from __future__ import annotations

import decimal
import math
import os
import re
import sys
from fractions import Fraction

D = decimal.Decimal
decimal.getcontext().prec = 80


def atan_series(x: D) -> D:
    """Taylor series of atan for |x| < 0.25."""
    s, term, n, x2 = D(0), x, 1, x * x
    while True:
        t = term / n
        if abs(t) < D(10) ** -78:
            return s
        s += t
        term *= -x2
        n += 2


def atan(x: D) -> D:
    for _ in range(2):
        x = x / (1 + (1 + x * x).sqrt())
    return 4 * atan_series(x)


PI = 16 * atan_series(D(1) / 5) - 4 * atan_series(D(1) / 239)
LN2 = D(2).ln()


def nearest(v: D) -> float:
    return float(v)  # decimal -> float conversion rounds to nearest


def bits(v: D, n: int) -> float:
    """v rounded to n significant bits."""
    e = math.frexp(float(v))[1]
    scale = D(2) ** (n - e)
    m = (v * scale).to_integral_value(rounding=decimal.ROUND_HALF_EVEN)
    return float(Fraction(int(m)) / Fraction(2) ** (n - e))


def split(v: D, hi_bits: int | None = None) -> tuple[float, float]:
    hi = nearest(v) if hi_bits is None else bits(v, hi_bits)
    return hi, nearest(v - D(hi))


def inv_factorial(n: int) -> float:
    return float(Fraction(1, math.factorial(n)))


def block() -> str:
    out = []

    def c(name, value, note):
        out.append(f"#define {name} ({value.hex()})  /* {note} */")

    p1 = bits(PI / 2, 33)
    p2 = bits(PI / 2 - D(p1), 33)
    p3 = nearest(PI / 2 - D(p1) - D(p2))
    c("DM_PIO2_1", p1, "pi/2, first 33 bits")
    c("DM_PIO2_2", p2, "pi/2, next 33 bits")
    c("DM_PIO2_3", p3, "pi/2, the rest")
    c("DM_TWO_OVER_PI", nearest(2 / PI), "2/pi")
    ln_hi, ln_lo = split(LN2, 42)
    c("DM_LN2_HI", ln_hi, "ln 2, first 42 bits")
    c("DM_LN2_LO", ln_lo, "ln 2, the rest")
    c("DM_INV_LN2", nearest(1 / LN2), "1/ln 2")
    c("DM_SQRT_HALF", nearest(D("0.5").sqrt()), "sqrt(1/2)")
    # atan(k/8) and the three angles built on it, as hi + lo
    rows = []
    for name, f in (("A", lambda a: a), ("B", lambda a: PI / 2 - a), ("C", lambda a: PI - a),
                    ("D", lambda a: PI / 2 + a)):
        hi, lo = zip(*(split(f(atan(D(k) / 8))) for k in range(9)))
        rows.append((name, hi, lo))
    out.append("/* atan(k/8) (row 0), pi/2 - atan(k/8), pi - atan(k/8), pi/2 + atan(k/8), k = 0..8 */")
    for part in ("HI", "LO"):
        out.append(f"DM_TABLE double dm_atan_{part.lower()}[4][9] = {{")
        for name, hi, lo in rows:
            vals = hi if part == "HI" else lo
            out.append("    {" + ", ".join(v.hex() for v in vals) + "},")
        out.append("};")
    # Taylor coefficients
    out.append("/* Taylor coefficients: sin (-1)^k/(2k+1)!, cos (-1)^k/(2k)!, exp 1/k!, atan (-1)^k/(2k+1), atanh 1/(2k+1) */")
    sin_c = [(-1) ** k * inv_factorial(2 * k + 1) for k in range(1, 9)]  # r^3 .. r^17
    cos_c = [(-1) ** k * inv_factorial(2 * k) for k in range(2, 10)]     # r^4 .. r^18
    exp_c = [inv_factorial(k) for k in range(2, 15)]                     # r^2 .. r^14
    atan_c = [float(Fraction((-1) ** k, 2 * k + 1)) for k in range(1, 9)]  # s^3 .. s^17
    atanh_c = [float(Fraction(1, 2 * k + 1)) for k in range(1, 12)]       # q^3 .. q^23
    for name, vals in (("SIN", sin_c), ("COS", cos_c), ("EXP", exp_c), ("ATAN", atan_c), ("ATANH", atanh_c)):
        for i, v in enumerate(vals):
            out.append(f"#define DM_{name}{i} ({v.hex()})")
    return "\n".join(out) + "\n"


BEGIN, END = "/* BEGIN generated constants */\n", "/* END generated constants */\n"

if __name__ == "__main__":
    text = block()
    if "--check" in sys.argv:
        h = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "detmath.h")).read()
        m = re.search(re.escape(BEGIN) + "(.*)" + re.escape(END), h, re.S)
        if not m or m.group(1) != text:
            print("detmath.h: the generated constants differ from gen_detmath.py", file=sys.stderr)
            sys.exit(1)
        print("detmath.h constants ok")
    else:
        sys.stdout.write(BEGIN + text + END)
