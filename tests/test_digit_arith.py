"""The FP64 arithmetic of the GEMM epilogue's CRT digit (crt_digit, ozablas_amd/csrc/oza_device.hpp):

    x = fma(v, RN(1/m), 2^31 + 2^-10);  c = trunc(fma(fract(x), m, 0.5))

is claimed to equal v mod m for every int32 v and every modulus.  Here each IEEE operation is
evaluated exactly in rationals and rounded to nearest double (Python's int / int true division is
correctly rounded), and the result is compared with the integer definition of v mod m, on the
int32 extremes, the multiples of m (and their neighbours) at both ends of the range, a dense
window around 0 and seeded random values, for all 20 moduli.  Host only; the device code itself is
checked bit for bit through every Scheme II GPU parity test."""
from fractions import Fraction

import numpy as np

BIAS = Fraction(2 ** 31) + Fraction(1, 2 ** 10)


def rn(q: Fraction) -> float:
    return q.numerator / q.denominator  # correctly rounded


def digit(v: int, m: int, rcp=None) -> int:
    rcp = 1.0 / m if rcp is None else rcp  # RN(1/m), as the CRT tables hold it
    x = Fraction(rn(Fraction(v) * Fraction(rcp) + BIAS))
    f = x - (x.numerator // x.denominator)  # fract, exact
    y = rn(f * m + Fraction(1, 2))
    return int(y)  # truncation (y > 0)


def values(m, rng):
    lo, hi = -(2 ** 31), 2 ** 31 - 1
    vs = set(range(-600, 601)) | {lo, lo + 1, hi, hi - 1}
    for end in (lo, hi):
        q = end // m
        for k in range(q - 3, q + 4):
            for d in (-1, 0, 1):
                v = k * m + d
                if lo <= v <= hi:
                    vs.add(v)
    vs |= set(int(x) for x in rng.integers(lo, hi, 1500, endpoint=True))
    return sorted(vs)


def test_digit_is_v_mod_m(orc):
    rng = np.random.default_rng(2026)
    for m in [int(x) for x in orc.moduli()]:
        for v in values(m, rng):
            assert digit(v, m) == v % m, (v, m)


def test_digit_arith_catches_a_coarse_reciprocal():
    """Planted on the CPU: with 1/m held in single precision the quotient is off for large |v|,
    which the check above would report."""
    m = 251
    rcp32 = float(np.float32(1.0 / m))
    vs = values(m, np.random.default_rng(5))
    assert any(digit(v, m, rcp32) != v % m for v in vs)
