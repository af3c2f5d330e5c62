"""The oracle's directed add/sub/mul against exact-rational golden vectors, and
the golden vectors against mpmath's correctly rounded binary64 arithmetic."""
import json
import math
import pathlib

import pytest

from oracle import binding as orc

ROWS = json.loads((pathlib.Path(__file__).parent / "golden" / "directed_rounding.json").read_text())["rows"]


def bits(x: float) -> str:
    return float.hex(x)


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
def test_oracle_matches_exact_rounding(op):
    lib = orc.lib()
    bad = []
    for r in ROWS:
        a, b = float.fromhex(r["a"]), float.fromhex(r["b"])
        for mode in ("rd", "ru"):
            got = getattr(lib, f"orc_{op}_{mode}")(a, b)
            want = float.fromhex(r[f"{op}_{mode}"])
            if bits(got) != bits(want):
                bad.append((r["a"], r["b"], mode, bits(got), bits(want)))
    assert not bad, bad[:5]


def test_golden_vectors_agree_with_mpmath():
    """Cross-check the generator in the normal range, where mpmath's 53-bit
    rounding equals binary64 rounding (no subnormal result, no overflow)."""
    mpmath = pytest.importorskip("mpmath")
    from mpmath.libmp import from_float, mpf_add, mpf_mul, to_float

    checked = 0
    for r in ROWS:
        a, b = float.fromhex(r["a"]), float.fromhex(r["b"])
        for op, fn in (("add", mpf_add), ("mul", mpf_mul)):
            for mode, rnd in (("rd", "f"), ("ru", "c")):
                want = float.fromhex(r[f"{op}_{mode}"])
                if want == 0.0 or not math.isfinite(want) or abs(want) < 2.2250738585072014e-308 * 4:
                    continue
                if abs(want) > 1.7976931348623157e308 / 2:
                    continue
                got = to_float(fn(from_float(a), from_float(b), 53, rnd))
                assert bits(got) == bits(want), (r["a"], r["b"], op, mode)
                checked += 1
    assert checked > 1000
