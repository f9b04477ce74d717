"""Derive the FP64 constants used by the sincos/exp routines of the oracle and the
device library from exact rational arithmetic (no library math involved).

pi comes from Machin's formula pi = 16 atan(1/5) - 4 atan(1/239); ln 2 from the
series ln 2 = sum_{k>=1} 1/(k 2^k).  Cody-Waite splits: the high parts keep 33
significant bits so that k*P1 and k*P2 are exact for |k| < 2^20.
Run: python3 oracle/gen_math_constants.py
"""
from fractions import Fraction
import math

PREC = 200  # bits of working precision (scaled integers)
ONE = 1 << PREC


def atan_inv(n):
    # atan(1/n) * 2^PREC by the alternating Taylor series
    total, term, k, sign = 0, ONE // n, 1, 1
    n2 = n * n
    while term:
        total += sign * (term // k)
        term //= n2
        k += 2
        sign = -sign
    return total


PI = Fraction(16 * atan_inv(5) - 4 * atan_inv(239), ONE)
LN2 = Fraction(sum(ONE // (k << k) for k in range(1, PREC + 10)), ONE)


def round_bits(x: Fraction, bits: int) -> float:
    """Round x to `bits` significant bits (nearest), returned as an exact double."""
    if x == 0:
        return 0.0
    e = math.floor(math.log2(abs(float(x))))
    scale = Fraction(2) ** (bits - 1 - e)
    m = round(x * scale)
    return float(Fraction(m) / scale)


def split3(x: Fraction):
    p1 = round_bits(x, 33)
    p2 = round_bits(x - Fraction(p1), 33)
    p3 = float(x - Fraction(p1) - Fraction(p2))
    return p1, p2, p3


pio2 = PI / 2
P1, P2, P3 = split3(pio2)
LN2_HI = round_bits(LN2, 32)
LN2_LO = float(LN2 - Fraction(LN2_HI))
consts = {
    "TWO_OVER_PI": float(2 / PI),
    "PIO2_1": P1, "PIO2_2": P2, "PIO2_3": P3,
    "LOG2E": float(1 / LN2),
    "LN2_HI": LN2_HI, "LN2_LO": LN2_LO,
}
inv_fact = {n: float(Fraction(1, math.factorial(n))) for n in range(1, 19)}

if __name__ == "__main__":
    for k, v in consts.items():
        print(f"#define {k} {v.hex()} /* {v!r} */")
    for n in range(1, 19):
        v = inv_fact[n]
        print(f"/* 1/{n}! */ {v.hex()}  /* {v!r} */")
    print("pi check", float(PI) == math.pi, "ln2 check", float(LN2) == math.log(2))
