"""Portable double-precision exp, log, sin, cos and asin for the oracle (test infrastructure only).

The decision paths of the CRV response (SiPM firing u < p_fire(v), Poisson counts, Box-Muller noise, the
Gumbel fit) call transcendentals. libm and the GPU's math library may differ by an ulp, which flips knife-edge
integer decisions. The oracle and the HIP kernels therefore evaluate the same published methods with the same
sequence of IEEE-754 double operations (no fused multiply-add), so both give identical bits:

* exp: Cody & Waite (1980) argument reduction x = k ln2 + r with a 32-bit-split ln2, then the Taylor series of
  e^r to degree 13 (|r| <= ln2/2, truncation < 5e-18), scaled by 2^k (ldexp).
* log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1), odd series to s^23.
* sin / cos: reduction by pi/2 in three 30-bit parts (Cody-Waite), Taylor series on |r| <= pi/4, quadrant swap.
* asin: the Maclaurin series of asin to z^24 (z = x^2 <= 1/4) on |x| <= 1/2, asin x = pi/2 - 2 asin(sqrt((1-x)/2))
  above.
* acos: pi/2 - asin x on |x| <= 1/2; 2 asin(sqrt((1-|x|)/2)) above (pi minus it for x < 0), the same series.

Each function works elementwise on numpy arrays with one numpy ufunc per IEEE operation (numpy never contracts
a*b+c), so the op order is the one written here. The coefficients are rationals rounded once to double
(1/n!, 1/(2n+1), the asin series terms); `constants()` exports them for the C header of the HIP build, and a test
checks that header against these values.
"""
from fractions import Fraction
import math

import numpy as np

# Cody-Waite splits (high parts keep 32 / 30 significant bits, so k * HI is exact for the |k| that occur)
LN2_HI = float.fromhex("0x1.62e42fee00000p-1")
LN2_LO = float.fromhex("0x1.a39ef35793c76p-33")
INV_LN2 = float.fromhex("0x1.71547652b82fep+0")
PIO2_1 = float.fromhex("0x1.921fb54000000p+0")
PIO2_2 = float.fromhex("0x1.10b4611800000p-30")
PIO2_3 = float.fromhex("0x1.313198a2e0370p-61")
TWO_OVER_PI = 0.6366197723675814
PI_2 = 1.5707963267948966
TWO_PI = 6.283185307179586
SQRT_HALF = 0.7071067811865476
EXP_HI = 709.78          # above: +inf
EXP_LO = -745.2          # below: 0

EXP_C = [float(Fraction(1, math.factorial(n))) for n in range(14)]                      # 1/n!, n = 0..13
LOG_C = [float(Fraction(1, 2 * n + 1)) for n in range(12)]                            # 1/(2n+1), n = 0..11
SIN_C = [float(Fraction((-1) ** n, math.factorial(2 * n + 1))) for n in range(9)]      # (-1)^n/(2n+1)!
COS_C = [float(Fraction((-1) ** n, math.factorial(2 * n))) for n in range(10)]         # (-1)^n/(2n)!
ASIN_C = [float(Fraction(math.factorial(2 * n), 4 ** n * math.factorial(n) ** 2 * (2 * n + 1)))
          for n in range(25)]                                                           # asin Maclaurin terms


def constants():
    """Name -> value (or list of values) of every constant above, for the C header generator check."""
    return {"LN2_HI": LN2_HI, "LN2_LO": LN2_LO, "INV_LN2": INV_LN2, "PIO2_1": PIO2_1, "PIO2_2": PIO2_2,
            "PIO2_3": PIO2_3, "TWO_OVER_PI": TWO_OVER_PI, "PI_2": PI_2, "TWO_PI": TWO_PI, "SQRT_HALF": SQRT_HALF,
            "EXP_HI": EXP_HI, "EXP_LO": EXP_LO, "EXP_C": EXP_C, "LOG_C": LOG_C, "SIN_C": SIN_C, "COS_C": COS_C,
            "ASIN_C": ASIN_C}


def _horner(coef, z):
    p = np.full(np.shape(z), coef[-1])
    for c in coef[-2::-1]:
        p = p * z + c
    return p


def exp(x):
    x = np.asarray(x, np.float64)
    xc = np.clip(np.where(np.isnan(x), 0.0, x), EXP_LO, EXP_HI)
    k = np.rint(xc * INV_LN2)
    r = (xc - k * LN2_HI) - k * LN2_LO
    y = np.ldexp(_horner(EXP_C, r), k.astype(np.int64))
    y = np.where(x > EXP_HI, np.inf, y)
    y = np.where(x < EXP_LO, 0.0, y)
    return np.where(np.isnan(x), np.nan, y)


def log(x):
    x = np.asarray(x, np.float64)
    ok = np.isfinite(x) & (x > 0)
    xs = np.where(ok, x, 1.0)
    m, e = np.frexp(xs)
    small = m < SQRT_HALF
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e).astype(np.float64)
    f = m - 1.0
    s = f / (2.0 + f)
    z = s * s
    q = z * _horner(LOG_C[1:], z)
    lm = (s + s * q) * 2.0
    y = e * LN2_HI + (lm + e * LN2_LO)
    y = np.where(x == 0.0, -np.inf, y)
    y = np.where(x == np.inf, np.inf, y)
    return np.where((x < 0) | np.isnan(x), np.nan, y)


def _reduce(x):
    k = np.rint(x * TWO_OVER_PI)
    r = ((x - k * PIO2_1) - k * PIO2_2) - k * PIO2_3
    return r, k.astype(np.int64) & 3


def _sinp(r):
    return r * _horner(SIN_C, r * r)


def _cosp(r):
    return _horner(COS_C, r * r)


def sin(x):
    x = np.asarray(x, np.float64)
    r, q = _reduce(x)
    s, c = _sinp(r), _cosp(r)
    return np.select([q == 0, q == 1, q == 2], [s, c, -s], -c)


def cos(x):
    x = np.asarray(x, np.float64)
    r, q = _reduce(x)
    s, c = _sinp(r), _cosp(r)
    return np.select([q == 0, q == 1, q == 2], [c, -s, -c], s)


def asin(x):
    x = np.asarray(x, np.float64)
    a = np.abs(x)
    lo = a * _horner(ASIN_C, a * a)
    w = (1.0 - a) * 0.5
    s = np.sqrt(np.maximum(w, 0.0))
    hi = PI_2 - 2.0 * (s * _horner(ASIN_C, w))
    r = np.where(a <= 0.5, lo, hi)
    r = np.where(x < 0, -r, r)
    return np.where(a > 1.0, np.nan, r)


def acos(x):
    x = np.asarray(x, np.float64)
    a = np.abs(x)
    mid = PI_2 - x * _horner(ASIN_C, x * x)
    w = (1.0 - a) * 0.5
    s = np.sqrt(np.maximum(w, 0.0))
    v = 2.0 * (s * _horner(ASIN_C, w))
    r = np.where(a <= 0.5, mid, np.where(x > 0, v, (2.0 * PI_2) - v))
    return np.where(a > 1.0, np.nan, r)
