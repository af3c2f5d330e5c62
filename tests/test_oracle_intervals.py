"""The oracle's interval operations: multiplication against the IEEE-1788 sign
case table and against exact rational products, squares and scalar products
against their closed forms."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import binding as orc


def rand_interval(rng):
    k = rng.integers(0, 6)
    if k == 0:
        return (0.0, 0.0)
    if k == 1:
        a = float(rng.uniform(0, 5))
        return (0.0 if rng.random() < 0.5 else -0.0, a)
    if k == 2:
        a = float(rng.uniform(-5, 0))
        return (a, 0.0 if rng.random() < 0.5 else -0.0)
    a, b = sorted(float(v) for v in rng.uniform(-5, 5, 2))
    if k == 3:
        return (a, a)
    return (a, b)


def test_mul_equals_case_table():
    rng = np.random.default_rng(11)
    lib = orc.lib()
    for _ in range(20000):
        x, y = rand_interval(rng), rand_interval(rng)
        got = orc.iv("orc_iv_mul", x, y)
        ref = orc.iv("orc_iv_mul_case", x, y)
        # equal as values (the case table may return -0 where the min/max form reports +0)
        assert got[0] == ref[0] and got[1] == ref[1], (x, y, got, ref)
        # zero bounds are reported as +0
        for v in got:
            if v == 0.0:
                assert math.copysign(1.0, v) > 0


def test_mul_contains_exact_products():
    rng = np.random.default_rng(12)
    for _ in range(3000):
        x, y = rand_interval(rng), rand_interval(rng)
        lo, hi = orc.iv("orc_iv_mul", x, y)
        for _ in range(5):
            px = float(rng.uniform(x[0], x[1])) if x[1] > x[0] else x[0]
            py = float(rng.uniform(y[0], y[1])) if y[1] > y[0] else y[0]
            q = Fraction(px) * Fraction(py)
            assert Fraction(lo) <= q <= Fraction(hi)
        # tightness: the bounds are the directed roundings of endpoint products
        ends = [Fraction(a) * Fraction(b) for a in x for b in y]
        assert Fraction(lo) <= min(ends) and Fraction(hi) >= max(ends)
        assert lo == orc.lib().orc_mul_rd(*min(((a, b) for a in x for b in y),
                                              key=lambda ab: Fraction(ab[0]) * Fraction(ab[1]))) or lo == 0.0


def test_sqr_closed_form():
    rng = np.random.default_rng(13)
    lib = orc.lib()
    for _ in range(5000):
        x = rand_interval(rng)
        lo, hi = orc.iv("orc_iv_sqr", x)
        if x[0] >= 0:
            want = (lib.orc_mul_rd(x[0], x[0]), lib.orc_mul_ru(x[1], x[1]))
        elif x[1] <= 0:
            want = (lib.orc_mul_rd(x[1], x[1]), lib.orc_mul_ru(x[0], x[0]))
        else:
            want = (0.0, max(lib.orc_mul_ru(x[0], x[0]), lib.orc_mul_ru(x[1], x[1])))
        assert lo == want[0] and hi == want[1]
        assert lo >= 0.0 and math.copysign(1.0, lo) > 0


@pytest.mark.parametrize("c", [2.5, -1.75, 0.0, -0.0, 1e-300, -3e300])
def test_scalar_mul_matches_point_interval_mul(c):
    rng = np.random.default_rng(14)
    for _ in range(2000):
        x = rand_interval(rng)
        a = orc.iv("orc_iv_mul_s", x, c)
        b = orc.iv("orc_iv_mul", x, (c, c))
        assert a[0].hex() == b[0].hex() and a[1].hex() == b[1].hex(), (x, c, a, b)
