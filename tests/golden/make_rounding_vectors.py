"""Generate tests/golden/directed_rounding.json: IEEE-754 binary64 add, sub and mul
rounded toward -inf and +inf, computed from exact rational arithmetic.

The exact result q = a op b is a Fraction; float(q) is the correctly rounded
nearest double (CPython rounds int/int true division correctly, subnormals
included), and one nextafter step toward the required side fixes the
direction whenever float(q) lands on the wrong side of q. Zero results take
the IEEE sign rules: a product's sign is the xor of the operand signs; an
exact zero sum of operands of opposite sign is -0 toward -inf and +0 toward
+inf; the sum of two equal-signed zeros keeps that sign. Values are written
as float.hex strings.

    python tests/golden/make_rounding_vectors.py
"""
from __future__ import annotations

import json
import math
import pathlib
import random
import struct
from fractions import Fraction

OUT = pathlib.Path(__file__).resolve().parent / "directed_rounding.json"
DBL_MAX = 1.7976931348623157e308


def directed(q: Fraction, down: bool) -> float:
    if q == 0:
        raise AssertionError
    if abs(q) > Fraction(DBL_MAX):
        if q > 0:
            return DBL_MAX if down else math.inf
        return -math.inf if down else -DBL_MAX
    d = float(q)
    if down and Fraction(d) > q:
        d = math.nextafter(d, -math.inf)
    if not down and Fraction(d) < q:
        d = math.nextafter(d, math.inf)
    return d


def signbit(x: float) -> bool:
    return math.copysign(1.0, x) < 0


def add(a: float, b: float, down: bool) -> float:
    q = Fraction(a) + Fraction(b)
    if q != 0:
        return directed(q, down)
    if a == 0 and b == 0 and signbit(a) == signbit(b):
        return a
    return -0.0 if down else 0.0


def mul(a: float, b: float, down: bool) -> float:
    q = Fraction(a) * Fraction(b)
    if q != 0:
        return directed(q, down)
    return -0.0 if signbit(a) != signbit(b) else 0.0


def rand_double(rng: random.Random) -> float:
    kind = rng.randrange(9)
    if kind == 0:
        return rng.choice([0.0, -0.0, 1.0, -1.0, 0.5, 3.0, 0.1, -0.1, 1e-300, 2.0 ** -1074])
    if kind == 1:  # subnormal
        return struct.unpack("<d", struct.pack("<Q", rng.getrandbits(52) | (rng.getrandbits(1) << 63)))[0]
    if kind == 2:  # tiny normal
        return math.ldexp(rng.uniform(1, 2), rng.randint(-1060, -1000)) * rng.choice([1, -1])
    if kind == 3:  # huge
        return math.ldexp(rng.uniform(1, 2), rng.randint(900, 1023)) * rng.choice([1, -1])
    if kind == 4:  # small integer
        return float(rng.randint(-1000, 1000))
    if kind == 5:  # short decimal
        return round(rng.uniform(-10, 10), rng.randint(1, 4))
    while True:  # any finite pattern
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if math.isfinite(x):
            return x


def main() -> None:
    rng = random.Random(20261203)
    rows = []
    for _ in range(1000):
        a = rand_double(rng)
        b = rand_double(rng)
        if rng.random() < 0.1:
            b = -a
        if rng.random() < 0.1:  # cancellation-prone: b close to -a
            b = -math.nextafter(a, math.inf)
        row = {"a": a.hex(), "b": b.hex()}
        row["add_rd"] = add(a, b, True).hex()
        row["add_ru"] = add(a, b, False).hex()
        row["sub_rd"] = add(a, -b, True).hex()
        row["sub_ru"] = add(a, -b, False).hex()
        row["mul_rd"] = mul(a, b, True).hex()
        row["mul_ru"] = mul(a, b, False).hex()
        rows.append(row)
    OUT.write_text(json.dumps({"generator": "exact rational arithmetic (fractions.Fraction)", "rows": rows}))
    print(f"wrote {len(rows)} rows to {OUT}")


if __name__ == "__main__":
    main()
