"""Exactness of the two root shortcuts the kernels take, checked in exact rational arithmetic.

- Correctly rounded single-precision roots (pass 2's final distance, the closest-atom pass):
  the hardware root, within one ulp of the correctly rounded root r, is corrected by Tuckerman's
  rounding test on the residuals x - s^-.s and x - s^+.s of its two neighbours (exact FMAs, so
  only their signs matter).  From any start in {r - 1 ulp, r, r + 1 ulp} the test returns r.
- Integer z offsets in pass 2's trace-back: dz = c^2 exactly with c <= 9, and the hardware root of
  float(dz), within one ulp of c, rounded to the nearest integer is c.
"""
import math
import random
from fractions import Fraction

import numpy as np


def _bits(f):
    return int(np.array(f, np.float32).view(np.uint32))


def _from_bits(b):
    return float(np.array(b, np.uint32).view(np.float32))


def _rn_sqrt(x):
    """correctly rounded float32 root of a positive float32 x, by exact comparison with midpoints"""
    X = Fraction(x)
    r = float(np.float32(math.sqrt(x)))
    cands = [_from_bits(_bits(r) + d) for d in (-2, -1, 0, 1, 2)]
    cands = [c for c in cands if c > 0 and math.isfinite(c)]
    best = cands[0]
    for c in cands[1:]:
        lo, hi = min(best, c), max(best, c)
        m = (Fraction(lo) + Fraction(hi)) / 2
        if m * m < X:
            best = hi
        elif m * m > X:
            best = lo
        else:  # exact tie: the even significand
            best = lo if _bits(lo) % 2 == 0 else hi
    return best


def _tuckerman(x, s):
    """the kernels' correction, with exact residual signs"""
    X, S = Fraction(x), Fraction(s)
    sm, sp = _from_bits(_bits(s) - 1), _from_bits(_bits(s) + 1)
    r = s
    if X - Fraction(sm) * S <= 0:
        r = sm
    if X - Fraction(sp) * S > 0:
        r = sp
    return r


def _inputs(n, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        k = rng.random()
        if k < 0.4:
            x = rng.uniform(1e-4, 1e3)
        elif k < 0.7:  # squared grid distances: small integers times dx^2
            x = rng.randint(1, 400) * rng.choice([0.04, 0.0625, 0.25, 0.09, 1.0])
        else:  # squares of midpoints between neighbouring floats, and their neighbours
            r = float(np.float32(rng.uniform(0.1, 30)))
            h = (Fraction(r) + Fraction(_from_bits(_bits(r) + 1))) / 2
            x = _from_bits(_bits(float(np.float32(float(h * h)))) + rng.choice([-1, 0, 1]))
        out.append(float(np.float32(x)))
    return out


def test_rounding_test_recovers_the_correct_root():
    for x in _inputs(4000, 11) + [1.0, 2.0, 4.0, 0.04, 1.17549435e-38, 3.4e38]:
        r = _rn_sqrt(x)
        for d in (-1, 0, 1):
            s = _from_bits(_bits(r) + d)
            assert _bits(_tuckerman(x, s)) == _bits(r), (x, r, s)


def test_rounding_test_catches_a_wrong_start():
    """the check is not vacuous: a start two ulps off is not always corrected"""
    misses = 0
    for x in _inputs(500, 12):
        r = _rn_sqrt(x)
        if _bits(_tuckerman(x, _from_bits(_bits(r) + 2))) != _bits(r):
            misses += 1
    assert misses > 0


def test_integer_root_of_a_perfect_square():
    for c in range(0, 10):
        x = float(np.float32(c * c))
        for d in (-1, 0, 1):
            s = _from_bits(_bits(float(np.float32(math.sqrt(x)))) + d) if c else abs(d) * 1.4e-45
            assert int(np.float32(s) + np.float32(0.5)) == c, (c, s)
