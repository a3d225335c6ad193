"""The quotient by m of the modularity gains (rule_formulas.h div_m).

Device code computes x / m for integer x as q0 = RN(x * RN(1/m)),
res = fma(-q0, m, x), q = fma(res, RN(1/m), q0) instead of dividing.  These
tests replay that sequence in exact rational arithmetic (an fma is the exact
a*b + c rounded once; CPython's int / int and float(Fraction) round
correctly) and require it to equal IEEE division, x / m, bit for bit: over
every x up to 2000 for every m up to 400, over the boundaries of the int32
range, over quotients next to rounding midpoints and over random pairs.  The
reference's gains (rules.cpp:34-43, 45-56, 24-32) divide by m and by 2m; the
kernels take x / (2m) as (x / m) * 0.5, which the last test checks too.
"""
import random
from fractions import Fraction

import pytest


def rn(q):
    """q (a Fraction) rounded to the nearest double, ties to even."""
    return q.numerator / q.denominator


def fma(a, b, c):
    return rn(Fraction(a) * Fraction(b) + Fraction(c))


def device_quotient(x, m):
    y = 1.0 / m
    xd, md = float(x), float(m)
    q0 = xd * y
    res = fma(-q0, md, xd)
    return fma(res, y, q0)


def check(x, m):
    got = device_quotient(x, m)
    want = x / m
    assert got == want and got.hex() == want.hex(), (x, m, got.hex(), want.hex())


def test_every_small_pair():
    for m in range(1, 401):
        for x in range(0, 2001):
            check(x, m)


def test_int32_boundaries():
    big = [2**31 - 1, 2**31 - 2, 2**30 + 1, 2**30, 2**30 - 1, 3 * 2**29, 2**29 + 3, 1_000_000_007, 999_999_937]
    for m in big + [1, 2, 3, 5, 7, 1023, 1025, 65535, 65537]:
        for x in big + [0, 1, 2, 3, m - 1, m, m + 1, 2 * m - 1 if 2 * m - 1 < 2**31 else 5]:
            if 0 <= x < 2**31:
                check(x, m)
                check(-x, m)


def test_quotients_next_to_midpoints():
    """x / m within a few 2^-60 ulp of a rounding midpoint: the hardest cases
    for a correction step.  For a midpoint t* = (2j+1) 2^(e-53), the nearest
    quotients x / m are found by taking x = round(t* m)."""
    rng = random.Random(7)
    hits = 0
    for _ in range(20000):
        m = rng.randrange(3, 2**31, 2)
        e = rng.randrange(-31, 31)
        # a random odd multiple of 2^(e-53) in [2^e, 2^(e+1))
        j = rng.randrange(2**52, 2**53)
        mid = Fraction(2 * j + 1, 2 ** (53 - e)) if e < 53 else Fraction((2 * j + 1) * 2 ** (e - 53))
        x0 = round(mid * m)
        for x in (x0 - 1, x0, x0 + 1):
            if 0 < x < 2**31:
                check(x, m)
                hits += 1
    assert hits > 1000


def test_random_pairs():
    rng = random.Random(11)
    for _ in range(50000):
        m = rng.randrange(1, 2**31)
        x = rng.randrange(0, 2**31)
        check(x, m)
        check(rng.randrange(0, 1 << rng.randrange(1, 31)), rng.randrange(1, 1 << rng.randrange(1, 31)))


@pytest.mark.parametrize("m", [1, 3, 7, 12345, 2**31 - 1])
def test_half_of_the_quotient_is_the_quotient_by_2m(m):
    rng = random.Random(m)
    for _ in range(2000):
        x = rng.randrange(0, 2**31)
        assert device_quotient(x, m) * 0.5 == x / (2.0 * m)
