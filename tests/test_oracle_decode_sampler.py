"""Decode and sampler of the oracle against the reference's own known answers
(test/source/decode_test.cpp:48-188, sample_discrete_test.cu:173-202) and the
closed-form SplitMix64 hash."""
import numpy as np
import pytest

from cuminlpsearch_amd import capi as C
from oracle import binding as orc

M64 = (1 << 64) - 1


def sb(kind, parent, part, fan):
    return orc.iv("orc_slot_bounds", kind, parent, part, fan) if False else _sb(kind, parent, part, fan)


def _sb(kind, parent, part, fan):
    out = np.zeros(2)
    orc.lib().orc_slot_bounds(kind, orc._p(np.array(parent, np.float64)), part, fan, orc._p(out))
    return out


def test_continuous_children_cover_parent():
    parent = (0.0, 10.0)
    ch = [_sb(C.SLOT_CONTINUOUS, parent, p, 4) for p in range(4)]
    assert ch[0][0] == parent[0] and ch[-1][1] == parent[1]
    for a, b in zip(ch, ch[1:]):
        assert a[1] >= b[0]
    for c in ch:
        assert parent[0] <= c[0] and c[1] <= parent[1]


def test_degenerate_parent_and_fanout_one():
    assert list(_sb(C.SLOT_CONTINUOUS, (5.0, 5.0), 0, 4)) == [5.0, 5.0]
    assert list(_sb(C.SLOT_CONTINUOUS, (-3.0, 7.0), 0, 1)) == [-3.0, 7.0]


def test_integer_enumerate_known_answers():
    for part in range(4):
        assert list(_sb(C.SLOT_INTEGER_ENUMERATE, (0.0, 3.0), part, 50)) == [float(part)] * 2
    for part in range(3):
        assert list(_sb(C.SLOT_INTEGER_ENUMERATE, (2.4, 5.6), part, 50)) == [3.0 + part] * 2
    assert list(_sb(C.SLOT_INTEGER_ENUMERATE, (0.0, 3.0), 10, 50)) == [3.0, 3.0]
    assert list(_sb(C.SLOT_BINARY_ENUMERATE, (0.0, 1.0), 0, 2)) == [0.0, 0.0]
    assert list(_sb(C.SLOT_BINARY_ENUMERATE, (0.0, 1.0), 1, 2)) == [1.0, 1.0]


@pytest.mark.parametrize("lb,ub,fan", [(0.0, 10.0, 4), (0.1, 0.3, 3), (-3.7, 12.9, 7), (1.0, 1.0 + 1e-10, 5),
                                       (0.0, 1.0, 97)])
def test_top_child_ub_is_parent_ub_bitwise(lb, ub, fan):
    assert _sb(C.SLOT_CONTINUOUS, (lb, ub), fan - 1, fan)[1].hex() == ub.hex()


def test_interior_boundaries_are_unfused():
    # lb' = lb + RN(width * part): unfused round-to-nearest, as the host decode
    lb, ub, fan = -3.7, 12.9, 7
    width = (ub - lb) / fan
    for part in range(fan - 1):
        c = _sb(C.SLOT_CONTINUOUS, (lb, ub), part, fan)
        assert c[0] == lb + width * part
        assert c[1] == lb + width * (part + 1)


def hash_ref(vid, i, region, salt):
    x = (vid * 0x9E3779B97F4A7C15) & M64
    x ^= (i * 0xBF58476D1CE4E5B9) & M64
    x ^= (region * 0x94D049BB133111EB) & M64
    x ^= (salt * 0xD6E8FEB86659FD93) & M64
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def test_sample_hash_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(500):
        args = [int(v) for v in rng.integers(0, 2**40, 4)]
        assert orc.lib().orc_sample_hash(*args) == hash_ref(*args)


def test_draws_stay_in_bounds_and_on_lattice():
    lib = orc.lib()
    seen = set()
    for region in range(4):
        for i in range(50):
            v = lib.orc_draw(1, 3.0, 9.0, 1, i, region, 7)
            assert v == np.floor(v) and 3.0 <= v <= 9.0
            seen.add(v)
            b = lib.orc_draw(1, 0.0, 1.0, 2, i, region, 7)
            assert b in (0.0, 1.0)
            c = lib.orc_draw(0, 0.0, 10.0, 0, i, region, 7)
            assert 0.0 <= c <= 10.0
    assert len(seen) > 1
    # a sub-box narrower than one unit with no integer inside clamps to ceil(lb)
    assert lib.orc_draw(1, 2.2, 2.8, 0, 0, 0, 0) == 3.0


def test_continuous_draw_formula():
    lib = orc.lib()
    for i in range(100):
        x = hash_ref(4, i, 9, 11)
        u = (x >> 11) * (1.0 / 9007199254740992.0)
        want = float(np.fma(u, 2.5 - (-1.0), -1.0)) if hasattr(np, "fma") else None
        got = lib.orc_draw(0, -1.0, 2.5, 4, i, 9, 11)
        assert -1.0 <= got <= 2.5
        if want is not None:
            assert got == want


def test_oracle_argmin_takes_the_lowest_index_among_equal_minima():
    """The oracle's enumerator breaks ties as CUB ArgMin does (lowest index;
    reference replay.cuh:439-474), checked by brute force over a lattice where
    a whole line of points ties: (x - 150)^2 does not depend on y."""
    from cuminlpsearch_amd import capi as C
    from cuminlpsearch_amd.model import Problem
    from oracle import binding as orc
    p = Problem()
    x, y = p.int_var(0, 300), p.int_var(0, 199)
    p.set_objective((x - 150.0) * (x - 150.0))
    a = p.arrays()
    for vids, widths, lowest in (([0, 1], [301, 200], 150), ([1, 0], [200, 301], 150 * 200)):
        found, value, index, point = orc.candidate(a, a.box, vids, widths, [C.SLOT_INTEGER_ENUMERATE] * 2, 0, 0)
        # brute force: the lexicographic minimum of (value, index)
        def xval(r):
            return r % widths[0] if vids[0] == 0 else r // widths[0]
        best = min((xval(r) - 150.0) ** 2 for r in range(301 * 200))
        ties = [r for r in range(301 * 200) if (xval(r) - 150.0) ** 2 == best]
        assert found and value == 0.0 and index == min(ties) == lowest
        assert list(point) == [150.0, 0.0]
