"""Pins the oracle's extended op set (Div, PowN, Sqrt, Exp, Log, Sin, Cos,
Tanh, Abs, Min, Max) where the reference's interval library (cuinterval, an
empty submodule in the snapshot) cannot: against exact rational arithmetic and
mpmath at 200 bits.

* every constant of the enclosures is the stated directed rounding;
* the directed fma, division and square root are correctly rounded;
* every interval result contains the exact range of the function over the
  argument interval (soundness) and exceeds it by at most a few ulps
  (tightness), over random and edge-case intervals;
* the IEEE-1788 special cases (empty results, half-lines, [0, 0]);
* the round-to-nearest point versions are within a few ulps.
"""
from __future__ import annotations

import math
import pathlib
import random
import re
from fractions import Fraction

import mpmath
import numpy as np
import pytest

from cuminlpsearch_amd import capi as C
from oracle import binding as orc

mpmath.mp.prec = 200
ROOT = pathlib.Path(__file__).resolve().parent.parent
INF = math.inf


def directed(q: Fraction, up: bool) -> float:
    f = float(q)
    if Fraction(f) != q and (Fraction(f) < q) == up:
        f = math.nextafter(f, INF if up else -INF)
    return f


def mpf_fraction(x) -> Fraction:
    m, e = mpmath.mpf(x).man_exp
    return Fraction(int(m)) * (Fraction(2) ** e if e >= 0 else Fraction(1, 2 ** -e))


def mp(x: float):
    return mpmath.mpf(x)


# -- constants ---------------------------------------------------------------------

def generated_block(path: pathlib.Path) -> str:
    text = path.read_text()
    a = text.index("// BEGIN GENERATED CONSTANTS")
    b = text.index("// END GENERATED CONSTANTS")
    return text[a:b]


def parse_constants(block: str) -> dict:
    out = {}
    for m in re.finditer(r"CMI_CONST double (\w+)(?:\[\d+\])? = \{?([^;}]*)\}?;", block):
        out[m.group(1)] = [float.fromhex(v.strip()) for v in m.group(2).split(",")]
    return out


def test_device_and_oracle_constants_are_the_same_block():
    dev = generated_block(ROOT / "cuminlpsearch_amd" / "csrc" / "interval_ops.hpp")
    orc_ = generated_block(ROOT / "oracle" / "enclosure_constants.inc")
    assert dev == orc_


def test_constants_are_the_stated_roundings():
    k = parse_constants(generated_block(ROOT / "oracle" / "enclosure_constants.inc"))
    ln2 = mpf_fraction(mpmath.log(2))
    pio2 = mpf_fraction(mpmath.pi / 2)
    fact = math.factorial
    # reduction splits: the high parts have <= 32 significant bits, the error
    # bounds cover the rest
    for hi, parts, err, value, bits in (("kLn2Hi", ["kLn2Lo"], "kLn2Err", ln2, 80),
                                        ("kPio2A", ["kPio2B", "kPio2C"], "kPio2Err", pio2, 130)):
        h = Fraction(k[hi][0])
        assert (h * 2 ** 31).denominator == 1 or (h * 2 ** 32).denominator == 1
        rest = value - h - sum(Fraction(k[p][0]) for p in parts)
        assert abs(rest) <= Fraction(k[err][0])
        assert abs(rest) < Fraction(1, 2 ** bits)
    assert k["kLn2Up"][0] == directed(ln2, True) and k["kLn2Dn"][0] == directed(ln2, False)
    assert k["kExpUp"] == [directed(Fraction(1, fact(j)), True) for j in range(18)]
    assert k["kExpDn"] == [directed(Fraction(1, fact(j)), False) for j in range(18)]
    assert k["kAtanhUp"] == [directed(Fraction(1, 2 * j + 1), True) for j in range(12)]
    assert k["kAtanhDn"] == [directed(Fraction(1, 2 * j + 1), False) for j in range(12)]
    assert k["kSinUp"] == [directed(Fraction((-1) ** j, fact(2 * j + 1)), True) for j in range(10)]
    assert k["kSinDn"] == [directed(Fraction((-1) ** j, fact(2 * j + 1)), False) for j in range(10)]
    assert k["kCosUp"] == [directed(Fraction((-1) ** j, fact(2 * j)), True) for j in range(11)]
    assert k["kCosDn"] == [directed(Fraction((-1) ** j, fact(2 * j)), False) for j in range(11)]
    assert k["kExpm1Up"] == [directed(Fraction(1, fact(j + 1)), True) for j in range(19)]
    assert k["kExpm1Dn"] == [directed(Fraction(1, fact(j + 1)), False) for j in range(19)]
    # tail bounds
    assert Fraction(k["kExpTail"][0]) >= Fraction(2014, 1000) * Fraction(7, 10) ** 18 / fact(18)
    w = Fraction(295, 10000)
    assert Fraction(k["kAtanhTail"][0]) >= w ** 12 / 25 / (1 - w)
    assert Fraction(k["kSinTail"][0]) >= Fraction(62, 100) ** 10 / fact(21)
    assert Fraction(k["kCosTail"][0]) >= Fraction(62, 100) ** 11 / fact(22)
    assert Fraction(k["kExpm1Tail"][0]) >= Fraction(272, 100) / fact(20)
    assert (mp(math.pi / 4) + mp(1e-12)) ** 2 < mp(0.62)  # the sin/cos series range
    assert float(Fraction(k["kExpm1Tail"][0])) < 1e-17


# -- directed primitives --------------------------------------------------------------

def rand_double(rng, lo=-60, hi=60):
    return rng.choice([-1, 1]) * rng.uniform(1, 2) * 2.0 ** rng.randint(lo, hi)


def test_directed_fma_is_correctly_rounded():
    rng = random.Random(7)
    for _ in range(2000):
        a, b, c = rand_double(rng), rand_double(rng), rand_double(rng, -120, 120)
        exact = Fraction(a) * Fraction(b) + Fraction(c)
        assert orc.lib().orc_fma_ru(a, b, c) == directed(exact, True)
        assert orc.lib().orc_fma_rd(a, b, c) == directed(exact, False)


def test_directed_division_and_sqrt_are_correctly_rounded():
    rng = random.Random(8)
    for _ in range(2000):
        a, b = rand_double(rng, -300, 300), rand_double(rng, -300, 300)
        q = Fraction(a) / Fraction(b)
        assert orc.lib().orc_div_ru(a, b) == directed(q, True)
        assert orc.lib().orc_div_rd(a, b) == directed(q, False)
        v = abs(a)
        up, dn = orc.lib().orc_sqrt_ru(v), orc.lib().orc_sqrt_rd(v)
        assert Fraction(dn) ** 2 <= Fraction(v) <= Fraction(up) ** 2
        assert math.nextafter(dn, INF) >= up  # adjacent or equal


# -- interval enclosures: soundness and tightness -------------------------------------

def slack(x, rel=2.0 ** -46):
    """Allowed overestimate beyond the exact bound."""
    return rel * abs(x) + mp(2.0 ** -1060)


def check_bounds(lo, hi, true_lo, true_hi, rel=2.0 ** -46):
    assert mp(lo) <= true_lo, (lo, true_lo)
    assert mp(hi) >= true_hi, (hi, true_hi)
    if true_lo != -mpmath.inf and lo != -INF:
        assert true_lo - mp(lo) <= slack(true_lo, rel), (lo, true_lo)
    if true_hi != mpmath.inf and hi != INF:
        assert mp(hi) - true_hi <= slack(true_hi, rel), (hi, true_hi)


def intervals(rng, n, lo, hi, log_scale=False):
    out = []
    for _ in range(n):
        if log_scale:
            a = 10.0 ** rng.uniform(lo, hi)
            w = a * 10.0 ** rng.uniform(-17, 1)
        else:
            a = rng.uniform(lo, hi)
            w = rng.choice([0.0, 1e-15, 1e-9, 1e-4, 0.1, 1.0, 7.0]) * rng.random()
        out.append((a, a + w))
    return out


def run(op, ivs, n_exp=0):
    a = np.array(ivs)
    return orc.iv_ext(op, a[:, 0], a[:, 1], n_exp=n_exp)


def test_exp_encloses():
    rng = random.Random(1)
    ivs = intervals(rng, 600, -745.0, 709.0) + intervals(rng, 300, -3.0, 3.0)
    ivs += [(0.0, 0.0), (-1e-300, 1e-300), (708.0, 709.78), (-800.0, -700.0), (-INF, 0.0), (1.0, INF), (710.0, 720.0)]
    lo, hi, ok = run(C.OP_EXP, ivs)
    assert ok.all()
    for (a, b), l, h in zip(ivs, lo, hi):
        ta = mpmath.exp(mp(a)) if a != -INF else mp(0)
        tb = mpmath.exp(mp(b)) if b != INF else mpmath.inf
        if tb > mp(1.7976931348623157e308):
            assert h == INF
            tb = mpmath.inf
        if ta > mp(1.7976931348623157e308):
            assert l == 1.7976931348623157e308  # RD of an overflowing value
            continue
        check_bounds(l, h, ta, tb)


def test_log_encloses():
    rng = random.Random(2)
    ivs = intervals(rng, 600, -320.0, 300.0, log_scale=True) + intervals(rng, 300, 0.5, 2.0)
    ivs += [(1.0, 1.0), (0.999999, 1.000001), (5e-324, 1e-310), (1e-300, 1.7e308), (0.7071067811865476, 1.4142135623730951)]
    lo, hi, ok = run(C.OP_LOG, ivs)
    assert ok.all()
    for (a, b), l, h in zip(ivs, lo, hi):
        check_bounds(l, h, mpmath.log(mp(a)), mpmath.log(mp(b)))


def sin_range(a, b, shift):
    """Exact min/max of sin(x + shift*pi/2) over [a, b] (mpmath)."""
    f = (lambda x: mpmath.sin(x)) if shift == 0 else (lambda x: mpmath.cos(x))
    ma, mb = mp(a), mp(b)
    vals = [f(ma), f(mb)]
    half = mpmath.pi / 2
    k0 = int(mpmath.floor(ma / half)) - 1
    k = k0
    while k * half <= mb + 1:
        x = k * half
        if ma <= x <= mb:
            vals.append(f(x))
        k += 1
    return min(vals), max(vals)


@pytest.mark.parametrize("op,shift", [(C.OP_SIN, 0), (C.OP_COS, 1)])
def test_sin_cos_enclose(op, shift):
    rng = random.Random(3 + shift)
    ivs = intervals(rng, 400, -20.0, 20.0) + intervals(rng, 200, -1e6, 1e6)
    # endpoints at and within an ulp of multiples of pi/2
    for k in range(-9, 10):
        c = k * math.pi / 2
        ivs += [(c, c), (math.nextafter(c, -INF), math.nextafter(c, INF)), (c - 0.3, c), (c, c + 0.3)]
    ivs += [(0.0, 6.3), (-1e7, 1e7), (1e-300, 2e-300)]
    lo, hi, ok = run(op, ivs)
    assert ok.all()
    for (a, b), l, h in zip(ivs, lo, hi):
        if b - a >= 6.28 or abs(a) > 1048576 or abs(b) > 1048576:
            assert (l, h) == (-1.0, 1.0)
            continue
        tl, th = sin_range(a, b, shift)
        assert mp(l) <= tl and mp(h) >= th, (a, b, l, h, tl, th)
        assert tl - mp(l) <= 1e-12 and mp(h) - th <= 1e-12, (a, b, l, h, tl, th)


def test_tanh_encloses():
    rng = random.Random(5)
    ivs = intervals(rng, 500, -25.0, 25.0) + intervals(rng, 200, -0.6, 0.6)
    ivs += [(0.0, 0.0), (-1e-300, 1e-300), (39.0, 60.0), (-INF, INF), (0.49, 0.51)]
    lo, hi, ok = run(C.OP_TANH, ivs)
    for (a, b), l, h in zip(ivs, lo, hi):
        ta = mpmath.tanh(mp(a)) if a != -INF else mp(-1)
        tb = mpmath.tanh(mp(b)) if b != INF else mp(1)
        check_bounds(l, h, ta, tb, rel=2.0 ** -44)


def test_sqrt_is_correctly_rounded_and_restricted_to_its_domain():
    rng = random.Random(6)
    ivs = intervals(rng, 500, -300.0, 300.0, log_scale=True) + [(-4.0, 9.0), (0.0, 2.0), (5e-324, 1e-320)]
    lo, hi, ok = run(C.OP_SQRT, ivs)
    assert ok.all()
    for (a, b), l, h in zip(ivs, lo, hi):
        assert l == (orc.lib().orc_sqrt_rd(a) if a > 0 else 0.0)
        assert h == orc.lib().orc_sqrt_ru(b)
    lo, hi, ok = run(C.OP_SQRT, [(-3.0, -1.0)])
    assert not ok[0]


def test_division_cases():
    rng = random.Random(9)
    cases = []
    for _ in range(800):
        xl = rng.uniform(-50, 50)
        yl = rng.uniform(-50, 50)
        cases.append(((xl, xl + rng.uniform(0, 20)), (yl, yl + rng.uniform(0, 20))))
    a = np.array([c[0] for c in cases])
    b = np.array([c[1] for c in cases])
    lo, hi, ok = orc.iv_ext(C.OP_DIV, a[:, 0], a[:, 1], b[:, 0], b[:, 1])
    for (x, y), l, h, k in zip(cases, lo, hi, ok):
        assert k
        if y[0] > 0 or y[1] < 0:
            q = [Fraction(u) / Fraction(v) for u in x for v in y]
            assert l == directed(min(q), False) and h == directed(max(q), True), (x, y, l, h)
        else:
            # 0 in y: a half-line or the whole line, always containing the
            # quotients of sample points
            for u in np.linspace(x[0], x[1], 5):
                for v in np.linspace(y[0], y[1], 7):
                    if v != 0:
                        assert l <= u / v <= h
    # IEEE 1788 special cases
    def one(x, y):
        r = orc.iv_ext(C.OP_DIV, [x[0]], [x[1]], [y[0]], [y[1]])
        return r[0][0], r[1][0], bool(r[2][0])
    assert one((1.0, 2.0), (0.0, 0.0))[2] is False
    assert one((0.0, 0.0), (-1.0, 1.0)) == (0.0, 0.0, True)
    assert one((1.0, 2.0), (0.0, 4.0)) == (0.25, INF, True)
    assert one((1.0, 2.0), (-4.0, 0.0)) == (-INF, -0.25, True)
    assert one((-2.0, -1.0), (0.0, 4.0)) == (-INF, -0.25, True)
    assert one((-2.0, -1.0), (-4.0, 0.0)) == (0.25, INF, True)
    assert one((1.0, 2.0), (-1.0, 1.0)) == (-INF, INF, True)
    assert one((-1.0, 2.0), (0.0, 1.0)) == (-INF, INF, True)
    assert one((1.0, INF), (2.0, 4.0)) == (0.25, INF, True)
    assert one((-3.0, -2.0), (1.0, INF))[1] == 0.0


@pytest.mark.parametrize("n", [-3, -2, -1, 0, 1, 2, 3, 4, 7, 10, 17, 50])
def test_pown_encloses(n):
    rng = random.Random(100 + n)
    ivs = intervals(rng, 200, -3.0, 3.0) + [(0.0, 0.0), (-1.0, 1.0), (0.0, 2.0), (-2.0, 0.0), (1.5, 1.5)]
    lo, hi, ok = run(C.OP_POWN, ivs, n_exp=n)
    for (a, b), l, h, k in zip(ivs, lo, hi, ok):
        if n < 0 and a == 0.0 and b == 0.0:
            assert not k
            continue
        assert k
        pts = [mp(a), mp(b)] + ([mp(0)] if a < 0 < b else [])
        if n < 0 and a <= 0 <= b:
            # unbounded on the side(s) where 0 is approached
            vals = [p ** n for p in pts if p != 0]
            if n % 2 == 0 or (a < 0 < b):
                assert h == INF
            if vals:
                assert mp(l) <= min(vals)
            continue
        vals = [p ** n for p in pts]
        tl, th = min(vals), max(vals)
        assert mp(l) <= tl and mp(h) >= th
        tol = abs(n) * 2.0 ** -50
        assert tl - mp(l) <= tol * abs(tl) + mp(2.0 ** -1060)
        assert mp(h) - th <= tol * abs(th) + mp(2.0 ** -1060)


def test_abs_min_max_are_exact():
    r = orc.iv_ext(C.OP_ABS, [-3.0, 1.0, -5.0], [2.0, 4.0, -1.0])
    assert r[0].tolist() == [0.0, 1.0, 1.0] and r[1].tolist() == [3.0, 4.0, 5.0]
    r = orc.iv_ext(C.OP_MIN, [-3.0, 1.0], [2.0, 4.0], [0.0, -2.0], [1.0, 0.5])
    assert r[0].tolist() == [-3.0, -2.0] and r[1].tolist() == [1.0, 0.5]
    r = orc.iv_ext(C.OP_MAX, [-3.0, 1.0], [2.0, 4.0], [0.0, -2.0], [1.0, 0.5])
    assert r[0].tolist() == [0.0, 1.0] and r[1].tolist() == [2.0, 4.0]


def test_log_and_sqrt_domains():
    lo, hi, ok = run(C.OP_LOG, [(-2.0, -1.0), (-1.0, 0.0), (0.0, 1.0), (-1.0, math.e)])
    assert ok.tolist() == [False, False, True, True]
    assert lo[2] == -INF and hi[2] >= 0.0 and lo[3] == -INF and hi[3] >= 1.0


# -- point versions ---------------------------------------------------------------------

@pytest.mark.parametrize("op,fn,lo,hi", [
    (C.OP_EXP, mpmath.exp, -740.0, 709.0),
    (C.OP_LOG, mpmath.log, 1e-300, 1e300),
    (C.OP_SIN, mpmath.sin, -1e5, 1e5),
    (C.OP_COS, mpmath.cos, -1e5, 1e5),
    (C.OP_TANH, mpmath.tanh, -30.0, 30.0),
])
def test_point_versions_are_accurate(op, fn, lo, hi):
    rng = np.random.default_rng(11)
    if op == C.OP_LOG:
        x = 10.0 ** rng.uniform(-300, 300, 3000)
    else:
        x = rng.uniform(lo, hi, 3000)
    x = np.concatenate([x, [0.5, 1.0, 2.0, -0.25]]) if op != C.OP_LOG else x
    got = orc.pt_ext(op, x)
    for xi, g in zip(x, got):
        t = fn(mp(xi))
        err = abs(mp(g) - t)
        # absolute error for sin/cos near a zero (argument reduction to 2^-100)
        bound = 2.0 ** -50 * abs(t) + (mp(2.0 ** -60) * (1 + abs(mp(xi))) * mp(2.0 ** -40)
                                       if op in (C.OP_SIN, C.OP_COS) else mp(2.0 ** -1070))
        assert err <= bound, (xi, g, t)


def test_point_pown():
    x = np.linspace(-3.0, 3.0, 61)
    for n in (2, 3, 7, -2, -3):
        got = orc.pt_ext(C.OP_POWN, x, n_exp=n)
        with np.errstate(divide="ignore"):
            want = x.astype(object) ** n if n > 0 else np.array([1 / v ** -n if v != 0 else math.inf for v in x])
        for g, w, xi in zip(got, want, x):
            if xi == 0.0 and n < 0:
                assert math.isinf(g)
                continue
            assert abs(g - float(w)) <= abs(n) * 2.0 ** -51 * abs(float(w))
