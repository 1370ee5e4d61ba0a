"""Write fmath_consts.h: the coefficients of the portable math functions as exact hexadecimal doubles.

Each coefficient is a rational rounded once to double (1/n!, (-1)^n/(2n+1)!, 1/(2n+1), the asin series terms
(2n)! / (4^n n!^2 (2n+1)), the Stirling terms B_2n / (2n (2n-1))); the range-reduction splits keep 32 / 30
significant bits of ln 2 and pi/2. Run: python gen_consts.py > fmath_consts.h
"""
from fractions import Fraction
import math

SPLITS = {
    "FM_LN2_HI": "0x1.62e42fee00000p-1", "FM_LN2_LO": "0x1.a39ef35793c76p-33", "FM_INV_LN2": "0x1.71547652b82fep+0",
    "FM_PIO2_1": "0x1.921fb54000000p+0", "FM_PIO2_2": "0x1.10b4611800000p-30", "FM_PIO2_3": "0x1.313198a2e0370p-61",
}
SCALARS = {"FM_TWO_OVER_PI": 0.6366197723675814, "FM_PI_2": 1.5707963267948966, "FM_TWO_PI": 6.283185307179586,
           "FM_SQRT_HALF": 0.7071067811865476, "FM_EXP_HI": 709.78, "FM_EXP_LO": -745.2,
           "FM_LN_SQRT_2PI": 0.9189385332046728, "FM_E": math.e}
ARRAYS = {
    "FM_EXP_C": [Fraction(1, math.factorial(n)) for n in range(14)],
    "FM_LOG_C": [Fraction(1, 2 * n + 1) for n in range(12)],
    "FM_SIN_C": [Fraction((-1) ** n, math.factorial(2 * n + 1)) for n in range(9)],
    "FM_COS_C": [Fraction((-1) ** n, math.factorial(2 * n)) for n in range(10)],
    "FM_ASIN_C": [Fraction(math.factorial(2 * n), 4 ** n * math.factorial(n) ** 2 * (2 * n + 1)) for n in range(25)],
    "FM_STIRLING": [Fraction(1, 12), Fraction(-1, 360), Fraction(1, 1260), Fraction(-1, 1680), Fraction(1, 1188),
                    Fraction(-691, 360360)],
    "FM_FACT": [Fraction(math.factorial(k)) for k in range(10)],
}


def main():
    print("// Generated by gen_consts.py: exact doubles of the portable math coefficients. Do not edit.")
    print("#pragma once")
    for k, v in SPLITS.items():
        print(f"#define {k} {v}")
    for k, v in SCALARS.items():
        print(f"#define {k} {float(v).hex()}")
    for k, vals in ARRAYS.items():
        body = ", ".join(float(v).hex() for v in vals)
        print(f"#define {k}_N {len(vals)}")
        print(f"#define {k}_INIT {{{body}}}")


if __name__ == "__main__":
    main()
