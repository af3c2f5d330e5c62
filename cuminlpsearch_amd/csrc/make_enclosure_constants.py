"""Write the constants of the transcendental interval enclosures.

    python cuminlpsearch_amd/csrc/make_enclosure_constants.py

Every constant is the directed (or nearest) IEEE binary64 rounding of an exact
real: rationals through Fractions, ln 2 and pi/2 through mpmath at 400 bits.
The block between the GENERATED markers of interval_ops.hpp (device) and of
oracle/enclosure_constants.inc (the CPU checker) is rewritten in place; the
CPU tests recompute every value independently.
"""
from __future__ import annotations

import math
import pathlib
import struct
from fractions import Fraction

import mpmath

mpmath.mp.prec = 400

HERE = pathlib.Path(__file__).resolve().parent
TARGETS = [HERE / "interval_ops.hpp", HERE.parent.parent / "oracle" / "enclosure_constants.inc"]
BEGIN = "// BEGIN GENERATED CONSTANTS (make_enclosure_constants.py)"
END = "// END GENERATED CONSTANTS"


def exact(x: float) -> Fraction:
    return Fraction(x)


def directed(q: Fraction, up: bool) -> float:
    """RU (up) or RD of a rational to binary64 (finite, normal range)."""
    f = float(q)  # round to nearest
    if Fraction(f) == q:
        return f
    if (Fraction(f) < q) == up:
        f = math.nextafter(f, math.inf if up else -math.inf)
    return f


def from_mpf(x) -> Fraction:
    m, e = mpmath.mpf(x).man_exp
    return Fraction(int(m)) * (Fraction(2) ** e if e >= 0 else Fraction(1, 2 ** -e))


def truncate_bits(q: Fraction, bits: int) -> Fraction:
    """q (positive) truncated to `bits` significant bits."""
    e = 0
    while q >= 2:
        q /= 2
        e += 1
    while q < 1:
        q *= 2
        e -= 1
    t = Fraction(math.floor(q * 2 ** (bits - 1)), 2 ** (bits - 1))
    return t * Fraction(2) ** e


def hexf(x: float) -> str:
    return x.hex() if x != 0 else "0x0p+0"


def constants() -> list[tuple[str, list[float], str]]:
    out = []
    ln2 = from_mpf(mpmath.log(2))
    pio2 = from_mpf(mpmath.pi / 2)

    out.append(("kInvLn2", [float(1 / ln2)], "1/ln 2 (nearest); only picks the reduction integer"))
    l1 = truncate_bits(ln2, 32)
    l2 = float(ln2 - l1)
    d = abs(ln2 - l1 - Fraction(l2))
    out.append(("kLn2Hi", [float(l1)], "ln 2 to 32 bits: k * kLn2Hi is exact for |k| < 2^21"))
    out.append(("kLn2Lo", [l2], "ln 2 - kLn2Hi (nearest)"))
    out.append(("kLn2Err", [directed(d, True)], "bound on |ln 2 - kLn2Hi - kLn2Lo|"))
    out.append(("kLn2Up", [directed(ln2, True)], "RU(ln 2)"))
    out.append(("kLn2Dn", [directed(ln2, False)], "RD(ln 2)"))

    out.append(("kTwoOverPi", [float(1 / pio2)], "2/pi (nearest); only picks the quadrant"))
    p1 = truncate_bits(pio2, 32)
    rem = pio2 - p1
    p2 = float(rem)
    rem -= Fraction(p2)
    p3 = float(rem)
    rem -= Fraction(p3)
    out.append(("kPio2A", [float(p1)], "pi/2 to 32 bits: k * kPio2A is exact for |k| < 2^21"))
    out.append(("kPio2B", [p2], "next 53 bits of pi/2"))
    out.append(("kPio2C", [p3], "next 53 bits of pi/2"))
    out.append(("kPio2Err", [directed(abs(rem), True)], "bound on |pi/2 - A - B - C|"))

    # exp: sum_{j<=17} r^j/j! on 0 <= r <= 0.7; tail <= e^0.7 0.7^18/18!
    fact = [math.factorial(j) for j in range(30)]
    n_exp = 17
    out.append(("kExpUp", [directed(Fraction(1, fact[j]), True) for j in range(n_exp + 1)], "RU(1/j!)"))
    out.append(("kExpDn", [directed(Fraction(1, fact[j]), False) for j in range(n_exp + 1)], "RD(1/j!)"))
    tail = Fraction(2014, 1000) * Fraction(7, 10) ** (n_exp + 1) / fact[n_exp + 1]  # e^0.7 < 2.014
    out.append(("kExpTail", [directed(tail, True)], "bound on the exp series tail for r <= 0.7"))
    # exp for points: |r| <= 0.35, degree 13
    out.append(("kExpRn", [float(Fraction(1, fact[j])) for j in range(14)], "RN(1/j!)"))

    # atanh: u * sum_{j<=11} w^j/(2j+1), w = u^2 <= 0.0295; tail <= w^12/25/(1-w)
    n_at = 11
    out.append(("kAtanhUp", [directed(Fraction(1, 2 * j + 1), True) for j in range(n_at + 1)], "RU(1/(2j+1))"))
    out.append(("kAtanhDn", [directed(Fraction(1, 2 * j + 1), False) for j in range(n_at + 1)], "RD(1/(2j+1))"))
    w = Fraction(295, 10000)
    tail = w ** (n_at + 1) / (2 * n_at + 3) / (1 - w)
    out.append(("kAtanhTail", [directed(tail, True)], "bound on the atanh series tail (sum/u) for u^2 <= 0.0295"))
    out.append(("kAtanhRn", [float(Fraction(1, 2 * j + 1)) for j in range(n_at + 1)], "RN(1/(2j+1))"))

    # sin(r) = r * sum_{j<=9} (-1)^j w^j/(2j+1)!, cos(r) = sum_{j<=10} (-1)^j w^j/(2j)!, w <= 0.62
    n_s, n_c = 9, 10
    sc = [Fraction((-1) ** j, fact[2 * j + 1]) for j in range(n_s + 1)]
    cc = [Fraction((-1) ** j, fact[2 * j]) for j in range(n_c + 1)]
    out.append(("kSinUp", [directed(c, True) for c in sc], "RU((-1)^j/(2j+1)!)"))
    out.append(("kSinDn", [directed(c, False) for c in sc], "RD((-1)^j/(2j+1)!)"))
    out.append(("kCosUp", [directed(c, True) for c in cc], "RU((-1)^j/(2j)!)"))
    out.append(("kCosDn", [directed(c, False) for c in cc], "RD((-1)^j/(2j)!)"))
    wmax = Fraction(62, 100)
    out.append(("kSinTail", [directed(wmax ** (n_s + 1) / fact[2 * n_s + 3], True)],
                "bound on the sin(r)/r series tail for r^2 <= 0.62 (alternating, decreasing)"))
    out.append(("kCosTail", [directed(wmax ** (n_c + 1) / fact[2 * n_c + 2], True)],
                "bound on the cos series tail for r^2 <= 0.62"))
    out.append(("kSinRn", [float(c) for c in sc], "RN((-1)^j/(2j+1)!)"))
    out.append(("kCosRn", [float(c) for c in cc], "RN((-1)^j/(2j)!)"))

    # expm1(u)/u = sum_{j<=18} u^j/(j+1)! on 0 <= u <= 1; tail <= e/20!
    n_m = 18
    out.append(("kExpm1Up", [directed(Fraction(1, fact[j + 1]), True) for j in range(n_m + 1)], "RU(1/(j+1)!)"))
    out.append(("kExpm1Dn", [directed(Fraction(1, fact[j + 1]), False) for j in range(n_m + 1)], "RD(1/(j+1)!)"))
    out.append(("kExpm1Tail", [directed(Fraction(272, 100) / fact[n_m + 2], True)],
                "bound on the expm1(u)/u series tail for 0 <= u <= 1"))
    out.append(("kExpm1Rn", [float(Fraction(1, fact[j + 1])) for j in range(n_m + 1)], "RN(1/(j+1)!)"))
    return out


def block() -> str:
    lines = [BEGIN]
    for name, vals, doc in constants():
        lines.append(f"// {doc}")
        if len(vals) == 1:
            lines.append(f"CMI_CONST double {name} = {hexf(vals[0])};")
        else:
            body = ", ".join(hexf(v) for v in vals)
            lines.append(f"CMI_CONST double {name}[{len(vals)}] = {{{body}}};")
    lines.append(END)
    return "\n".join(lines)


def main() -> int:
    text = block()
    for path in TARGETS:
        src = path.read_text() if path.exists() else f"{BEGIN}\n{END}\n"
        a = src.index(BEGIN)
        b = src.index(END) + len(END)
        path.write_text(src[:a] + text + src[b:])
        print(f"wrote {path}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
