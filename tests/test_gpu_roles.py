"""The three backend roles (bounder, sampler, enumerator) on the device against
the CPU oracle, plus the reference's role-contract checks: determinism per
salt, NaN witness when nothing is feasible, lattice draws, over-budget facts
before allocation, shape errors."""
import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd.model import Problem
from oracle import binding as orc

import problems as P

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def mixed_problem(seed):
    """Random DAG with two integer variables, so every slot kind has a target."""
    return cmi.EngineProblem.from_problem(P.random_dag(seed, n_vars=5, n_ops=50, n_con=3, int_vars=2))


CASES = [
    # (slot kinds, var ids, widths)
    ([C.SLOT_CONTINUOUS], [2], [4]),
    ([C.SLOT_CONTINUOUS, C.SLOT_CONTINUOUS], [3, 2], [5, 3]),
    ([C.SLOT_INTEGER_ENUMERATE, C.SLOT_CONTINUOUS], [0, 4], [7, 6]),
    ([C.SLOT_INTEGER_PARTITION, C.SLOT_INTEGER_ENUMERATE, C.SLOT_CONTINUOUS], [1, 0, 2], [3, 7, 20]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bounder_matches_oracle(ctx, case, seed):
    kinds, vids, widths = CASES[case]
    eng = mixed_problem(seed)
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    n = int(np.prod(widths))
    role = cmi.Role(tape, C.ROLE_BOUNDER, kinds, n)
    box = a.box
    f_d, lb_d = role.bound(box, vids, widths)
    f_o, lb_o = orc.bound(a, box, vids, widths, kinds)
    assert np.array_equal(f_d, f_o)
    assert same_bits(lb_d, lb_o)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("samples", [1, 7, 64])
def test_sampler_matches_oracle(ctx, case, samples):
    kinds, vids, widths = CASES[case]
    eng = mixed_problem(10 + case)
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    n = int(np.prod(widths))
    role = cmi.Role(tape, C.ROLE_SAMPLER, kinds, n, samples=samples)
    for salt in (1, 2, 99):
        d = role.sample(a.box, vids, widths, salt)
        found, value, index, point = orc.candidate(a, a.box, vids, widths, kinds, samples, salt)
        assert d.found == found
        assert same_bits([d.value], [value])
        assert d.index == index
        assert same_bits(d.point, point)


def test_sampler_determinism_and_salt(ctx):
    p = Problem()
    x = p.var(0.0, 10.0)
    p.set_objective(x * x)
    p.add_constraint(x, C.LE, 5.0)
    eng = cmi.EngineProblem.from_problem(p)
    tape = cmi.Tape(ctx, eng)
    role = cmi.Role(tape, C.ROLE_SAMPLER, [C.SLOT_CONTINUOUS], 4, samples=32)
    box = [[0.0, 5.0]]
    a1 = role.sample(box, [0], [4], 7)
    a2 = role.sample(box, [0], [4], 7)
    b = role.sample(box, [0], [4], 8)
    assert a1.found and same_bits([a1.value], [a2.value]) and same_bits(a1.point, a2.point)
    assert a1.value != b.value
    # interval bounds do not depend on the salt
    bnd = cmi.Role(tape, C.ROLE_BOUNDER, [C.SLOT_CONTINUOUS], 4)
    f1, l1 = bnd.bound([[0.0, 10.0]], [0], [4])
    f2, l2 = bnd.bound([[0.0, 10.0]], [0], [4])
    # x <= 5 excludes only (7.5, 10]; [5, 7.5] touches the bound and stays
    assert same_bits(l1, l2) and np.array_equal(f1, [1, 1, 1, 0]) and np.array_equal(f1, f2)


def test_infeasible_domain_has_nan_witness(ctx):
    p = Problem()
    x = p.var(0.0, 10.0)
    p.set_objective(x * x)
    p.add_constraint(x, C.LE, -100.0)
    eng = cmi.EngineProblem.from_problem(p)
    role = cmi.Role(cmi.Tape(ctx, eng), C.ROLE_SAMPLER, [C.SLOT_CONTINUOUS], 4, samples=16)
    r = role.sample([[0.0, 10.0]], [0], [4], 1)
    assert not r.found
    assert r.value == np.finfo(np.float64).max
    assert np.all(np.isnan(r.point))


def test_lattice_draws(ctx):
    """Integer/binary samples sit on the lattice (reference sample_discrete_test.cu:173-202)."""
    p = Problem()
    x = p.var(0.0, 10.0)
    k = p.int_var(3.0, 9.0)
    b = p.bin_var()
    p.set_objective(x + k + b)
    eng = cmi.EngineProblem.from_problem(p)
    tape = cmi.Tape(ctx, eng)
    role = cmi.Role(tape, C.ROLE_SAMPLER, [C.SLOT_CONTINUOUS], 4, samples=50)
    for salt in range(20):
        r = role.sample(eng.arrays().box, [0], [4], salt)
        assert r.found
        assert r.point[1] == np.floor(r.point[1]) and 3 <= r.point[1] <= 9
        assert r.point[2] in (0.0, 1.0)
        assert 0.0 <= r.point[0] <= 10.0


def test_enumerator_matches_oracle_and_brute_force(ctx):
    """All-integer problem: the enumerator's value is the exact minimum over the lattice."""
    p = Problem()
    xs = [p.int_var(0, 4), p.int_var(-2, 3), p.int_var(1, 5)]
    obj = 3.0 * xs[0] * xs[1] - 2.0 * (xs[2] * xs[2]) + xs[0] * xs[2] - 0.5 * xs[1]
    p.set_objective(obj)
    p.add_constraint(xs[0] + xs[1] + xs[2], C.LE, 6.0)
    p.add_constraint(xs[0] * xs[1], C.EQ, 2.0)
    eng = cmi.EngineProblem.from_problem(p)
    a = eng.arrays()
    kinds = [C.SLOT_INTEGER_ENUMERATE] * 3
    widths = [5, 6, 5]
    role = cmi.Role(cmi.Tape(ctx, eng), C.ROLE_ENUMERATOR, kinds, 150)
    d = role.enumerate(a.box, [0, 1, 2], widths)
    found, value, index, point = orc.candidate(a, a.box, [0, 1, 2], widths, kinds, 0, 0)
    assert d.found == found and same_bits([d.value], [value]) and d.index == index and same_bits(d.point, point)
    best = min(3.0 * i * j - 2.0 * (k * k) + i * k - 0.5 * j
               for i in range(0, 5) for j in range(-2, 4) for k in range(1, 6) if i + j + k <= 6 and i * j == 2)
    assert d.value == best


@pytest.mark.parametrize("order", ["x_first", "y_first"])
def test_argmin_takes_the_lowest_index_among_equal_minima(ctx, order):
    """CUB ArgMin's rule (reference replay.cuh:439-474): among equal minima the
    lowest index wins. The objective (x - 150)^2 does not depend on y, so every
    lattice point with x = 150 is a minimum; brute force names the lowest
    index, across tiles and the last-tile reduction."""
    p = Problem()
    x, y = p.int_var(0, 300), p.int_var(0, 199)
    p.set_objective((x - 150.0) * (x - 150.0))
    eng = cmi.EngineProblem.from_problem(p)
    a = eng.arrays()
    kinds = [C.SLOT_INTEGER_ENUMERATE] * 2
    vids, widths = ([0, 1], [301, 200]) if order == "x_first" else ([1, 0], [200, 301])
    role = cmi.Role(cmi.Tape(ctx, eng), C.ROLE_ENUMERATOR, kinds, 301 * 200)
    d = role.enumerate(a.box, vids, widths)
    values = {}
    for r in range(301 * 200):
        digits = (r % widths[0], r // widths[0])
        point = [0.0, 0.0]
        for j in range(2):
            point[vids[j]] = float(digits[j])
        values[r] = (point[0] - 150.0) ** 2
    best = min(values.values())
    lowest = min(r for r, v in values.items() if v == best)
    assert d.found and d.value == 0.0 and d.index == lowest
    assert list(d.point) == [150.0, 0.0]
    found, value, index, point = orc.candidate(a, a.box, vids, widths, kinds, 0, 0)
    assert index == lowest and value == 0.0


def test_over_budget_is_reported_before_allocation(ctx):
    eng = P.c1(0)
    tape = cmi.Tape(ctx, eng)
    with pytest.raises(C.CmiError) as ei:
        cmi.Role(tape, C.ROLE_BOUNDER, [C.SLOT_CONTINUOUS] * 2, 400, budget_bytes=16)
    assert ei.value.status == C.OVER_BUDGET


def test_shape_errors(ctx):
    eng = P.c1(0)
    tape = cmi.Tape(ctx, eng)
    role = cmi.Role(tape, C.ROLE_BOUNDER, [C.SLOT_CONTINUOUS], 4)
    with pytest.raises(C.CmiError) as ei:
        role.bound(eng.arrays().box[:3], [0], [4])
    assert ei.value.status == C.SHAPE_MISMATCH
    with pytest.raises(C.CmiError) as ei:
        role.bound(eng.arrays().box, [0], [5])
    assert ei.value.status == C.SHAPE_MISMATCH
    with pytest.raises(C.CmiError) as ei:
        cmi.Role(tape, C.ROLE_ENUMERATOR, [C.SLOT_CONTINUOUS], 4)
    assert ei.value.status == C.INVALID_CONFIG


@pytest.mark.parametrize("groups", [1, 3, 8])
def test_roles_in_function_groups_match_oracle(ctx, groups):
    """Decode-mode items run the tape in its function groups, one wave per
    group: the bounder, sampler and enumerator give the oracle's results for
    every grouping of a 600-node DAG (groups of about 30 tape words)."""
    eng = P.c2()
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    assert tape.set_box_groups(groups) == groups
    kinds, vids, widths = [C.SLOT_CONTINUOUS] * 3, [3, 17, 29], [4, 3, 5]
    n = int(np.prod(widths))
    bnd = cmi.Role(tape, C.ROLE_BOUNDER, kinds, n)
    f_d, lb_d = bnd.bound(a.box, vids, widths)
    f_o, lb_o = orc.bound(a, a.box, vids, widths, kinds)
    assert np.array_equal(f_d, f_o) and same_bits(lb_d, lb_o)
    smp = cmi.Role(tape, C.ROLE_SAMPLER, kinds, n, samples=33)
    for salt in (1, 7):
        d = smp.sample(a.box, vids, widths, salt)
        found, value, index, point = orc.candidate(a, a.box, vids, widths, kinds, 33, salt)
        assert d.found == found and same_bits([d.value], [value]) and d.index == index and same_bits(d.point, point)
    # the enumerator needs integer slots: a QCQP whose first 8 variables are integers in [0, 3]
    ieng = cmi.EngineProblem.qcqp(17, n=24, m=12, n_int=8, int_lb=0.0, int_ub=3.0, cont_lb=-5.0, cont_ub=5.0,
                                  pair_density=0.3)
    ia = ieng.arrays()
    itape = cmi.Tape(ctx, ieng)
    itape.set_box_groups(groups)
    ekinds, evids, ewidths = [C.SLOT_INTEGER_ENUMERATE] * 3, [0, 3, 6], [4, 4, 4]
    enu = cmi.Role(itape, C.ROLE_ENUMERATOR, ekinds, 64)
    d = enu.enumerate(ia.box, evids, ewidths)
    found, value, index, point = orc.candidate(ia, ia.box, evids, ewidths, ekinds, 0, 0)
    assert d.found == found and same_bits([d.value], [value]) and d.index == index and same_bits(d.point, point)


def test_regrouping_after_role_build(ctx):
    """A role built before the tape is regrouped runs the new groups' code."""
    eng = P.c2()
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    tape.set_box_groups(8)
    kinds, vids, widths = [C.SLOT_CONTINUOUS] * 2, [0, 5], [8, 8]
    bnd = cmi.Role(tape, C.ROLE_BOUNDER, kinds, 64)
    f_o, lb_o = orc.bound(a, a.box, vids, widths, kinds)
    for g in (2, 8, 1):
        tape.set_box_groups(g)
        f_d, lb_d = bnd.bound(a.box, vids, widths)
        assert np.array_equal(f_d, f_o) and same_bits(lb_d, lb_o)


@pytest.mark.parametrize("n,m", [(40, 20), (64, 32)])
def test_sliced_items_on_a_long_tape_match_oracle(ctx, n, m):
    """A launch of a few tiles spreads each tile's function groups over
    several workgroups (one wave per SIMD each); a tape of 4,096 words or more
    with more functions than a workgroup holds waves runs one group per
    function. The slices' verdicts and objective meet across workgroups, and
    the bounder, sampler and enumerator still give the oracle's results."""
    eng = cmi.EngineProblem.qcqp(50 + n, n=n, m=m, n_int=8, int_lb=0.0, int_ub=3.0, cont_lb=-5.0, cont_ub=5.0,
                                 pair_density=0.2)
    info, _ = eng.tape_info()
    assert info.box_instr >= 4096 and m + 1 > 16
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    kinds, vids, widths = [C.SLOT_CONTINUOUS] * 2, [9, 30], [4, 4]
    bnd = cmi.Role(tape, C.ROLE_BOUNDER, kinds, 16)
    for _ in range(2):  # the per-tile scratch is left zeroed for the next launch
        f_d, lb_d = bnd.bound(a.box, vids, widths)
        f_o, lb_o = orc.bound(a, a.box, vids, widths, kinds)
        assert np.array_equal(f_d, f_o) and same_bits(lb_d, lb_o)
    smp = cmi.Role(tape, C.ROLE_SAMPLER, kinds, 16, samples=16)
    for salt in (3, 4):
        d = smp.sample(a.box, vids, widths, salt)
        found, value, index, point = orc.candidate(a, a.box, vids, widths, kinds, 16, salt)
        assert d.found == found and same_bits([d.value], [value]) and d.index == index and same_bits(d.point, point)
    ekinds, evids, ewidths = [C.SLOT_INTEGER_ENUMERATE] * 3, [0, 3, 6], [4, 4, 4]
    enu = cmi.Role(tape, C.ROLE_ENUMERATOR, ekinds, 64)
    d = enu.enumerate(a.box, evids, ewidths)
    found, value, index, point = orc.candidate(a, a.box, evids, ewidths, ekinds, 0, 0)
    assert d.found == found and same_bits([d.value], [value]) and d.index == index and same_bits(d.point, point)
