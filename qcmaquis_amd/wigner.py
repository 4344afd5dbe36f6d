"""Wigner 6j / 9j symbols on doubled spins (the GSL argument convention).

The reference obtains these from GSL (gsl_sf_coupling_6j/9j, called from
dmrg/framework/dmrg/block_matrix/symmetry/gsl_coupling.h:68-71,156-158).
GSL is not part of this build; the symbols are computed here from the
Racah single-sum formula for the 6j symbol and the standard identity
    {a b c; d e f; g h i} = sum_x (-1)^{2x} (2x+1)
                            {a b c; f i x} {d e f; b x h} {g h i; x a d}
(Edmonds, Angular Momentum in QM, eq. 6.4.3), with exact integer/rational
arithmetic and a single float conversion at the end.
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from math import factorial, sqrt


def _tri(a, b, c):
    return ((a + b + c) % 2 == 0) and abs(a - b) <= c <= a + b


def triangle(a, b, c):
    return _tri(a, b, c)


def _delta2(a, b, c):
    # a, b, c doubled spins satisfying the triangle rule
    return Fraction(factorial((a + b - c) // 2) * factorial((a - b + c) // 2) * factorial((-a + b + c) // 2),
                    factorial((a + b + c) // 2 + 1))


@lru_cache(maxsize=None)
def _sixj_exact(a, b, c, d, e, f):
    """Return (sum, delta_sq) with 6j = sum * sqrt(delta_sq), or None if zero."""
    if not (_tri(a, b, c) and _tri(a, e, f) and _tri(d, b, f) and _tri(d, e, c)):
        return None
    dsq = _delta2(a, b, c) * _delta2(a, e, f) * _delta2(d, b, f) * _delta2(d, e, c)
    s1 = (a + b + c) // 2
    s2 = (a + e + f) // 2
    s3 = (d + b + f) // 2
    s4 = (d + e + c) // 2
    p1 = (a + b + d + e) // 2
    p2 = (b + c + e + f) // 2
    p3 = (c + a + f + d) // 2
    tmin = max(s1, s2, s3, s4)
    tmax = min(p1, p2, p3)
    tot = 0
    for t in range(tmin, tmax + 1):
        num = factorial(t + 1)
        den = (factorial(t - s1) * factorial(t - s2) * factorial(t - s3) * factorial(t - s4)
               * factorial(p1 - t) * factorial(p2 - t) * factorial(p3 - t))
        term = Fraction(num, den)
        tot += -term if (t % 2) else term
    return Fraction(tot), dsq


def sixj(a, b, c, d, e, f):
    r = _sixj_exact(a, b, c, d, e, f)
    if r is None:
        return 0.0
    s, dsq = r
    return float(s) * sqrt(float(dsq))


@lru_cache(maxsize=None)
def ninej(a, b, c, d, e, f, g, h, i):
    if not (_tri(a, b, c) and _tri(d, e, f) and _tri(g, h, i)
            and _tri(a, d, g) and _tri(b, e, h) and _tri(c, f, i)):
        return 0.0
    lo = max(abs(a - i), abs(b - f), abs(d - h))
    hi = min(a + i, b + f, d + h)
    tot = 0.0
    for x in range(lo, hi + 1):
        if (x + a + i) % 2:
            continue
        w = (x + 1) * (-1 if (x % 2) else 1)
        tot += w * sixj(a, b, c, f, i, x) * sixj(d, e, f, b, x, h) * sixj(g, h, i, x, a, d)
    return tot


def mod_coupling(a, b, c, d, e, f, g, h, i):
    """sqrt((g+1)(h+1)(c+1)(f+1)) * 9j  (gsl_coupling.h:166-175)."""
    return sqrt((g + 1.0) * (h + 1.0) * (c + 1.0) * (f + 1.0)) * ninej(a, b, c, d, e, f, g, h, i)
