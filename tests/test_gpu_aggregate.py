"""The device's k-child aggregate (cmi_bound_children) against the oracle and
against a host reduction over the device bounder's own per-subregion verdicts
(the reference's central aggregate test, aggregate_backend_test.cu:116-162),
batched, over budget and shape errors."""
import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from oracle import binding as orc

import aggregate_cases as A
import problems as P

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


@pytest.mark.parametrize("n", [16, 256, 4096])
def test_device_aggregate_equals_host_reduction_bitwise(ctx, n):
    eng = A.fixture()
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    kinds, vids, widths = A.partition(a.box, 4, n)
    agg = cmi.AggregateBounder(tape, kinds, n, 4)
    lb, ub, fe = agg.bound_children(a.box, vids, widths)
    # the device bounder's per-subregion verdicts, reduced on the host
    bnd = cmi.Role(tape, C.ROLE_BOUNDER, kinds, n)
    f_d, lb_d = bnd.bound(a.box, vids, widths)
    m = n // 4
    for c in range(4):
        f = f_d[c * m:(c + 1) * m].astype(bool)
        assert fe[0, c] == int(f.any())
        if f.any():
            assert lb[0, c] == lb_d[c * m:(c + 1) * m][f].min()
    lb_o, ub_o, fe_o = orc.aggregate(a, a.box, vids, widths, kinds, 4)
    assert same_bits(lb[0], lb_o) and same_bits(ub[0], ub_o) and np.array_equal(fe[0], fe_o)


def test_enclosure_of_the_objective_over_each_child(ctx):
    """Reference aggregate_backend_test.cu:208-253: every feasible point of a
    child has its objective inside [lb, hull_ub]."""
    eng = A.fixture()
    a = eng.arrays()
    kinds, vids, widths = A.partition(a.box, 4, 256)
    agg = cmi.AggregateBounder(cmi.Tape(ctx, eng), kinds, 256, 4)
    lb, ub, fe = agg.bound_children(a.box, vids, widths)
    rng = np.random.default_rng(0)
    xs = rng.uniform(-2.0, 3.0, 4000)
    ys = rng.uniform(-1.0, 4.0, 4000)
    # the branch variable is x (both have width 5; ties go to the lower index)
    assert vids[-1] == 0
    for x, y in zip(xs, ys):
        if x * x + y > 6.0:
            continue
        c = min(int((x + 2.0) / 5.0 * 4), 3)
        obj = x * x + y + 0.5 * y * y - x
        assert fe[0, c] == 1 and lb[0, c] <= obj + 1e-12 and obj <= ub[0, c] + 1e-12


def test_batched_regions_equal_one_at_a_time(ctx):
    eng = P.c1(2)
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    rng = np.random.default_rng(5)
    boxes = []
    for _ in range(37):
        lo = rng.uniform(-10, 5, 4)
        hi = lo + rng.uniform(0.01, 5, 4)
        boxes.append(np.stack([lo, hi], 1))
    boxes = np.array(boxes)
    n, k = 512, 8
    parts = [A.partition(b, k, n) for b in boxes]
    kinds = parts[0][0]
    assert all(p[0] == kinds for p in parts)
    vids = np.array([p[1] for p in parts])
    widths = np.array([p[2] for p in parts])
    batch = cmi.AggregateBounder(tape, kinds, n, k, max_batch=64)
    lb, ub, fe = batch.bound_children(boxes, vids, widths)
    single = cmi.AggregateBounder(tape, kinds, n, k, max_batch=1)
    for i in range(len(boxes)):
        l1, u1, f1 = single.bound_children(boxes[i], vids[i], widths[i])
        assert same_bits(lb[i], l1[0]) and same_bits(ub[i], u1[0]) and np.array_equal(fe[i], f1[0])
        lo_o, ub_o, fe_o = orc.aggregate(a, boxes[i], vids[i], widths[i], kinds, k)
        assert same_bits(lb[i], lo_o) and same_bits(ub[i], ub_o) and np.array_equal(fe[i], fe_o)


def test_large_n_and_ragged_tiles(ctx):
    """N not a multiple of 64 and M below and above one tile."""
    eng = A.fixture()
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    for n, k in ((6 * 50, 6), (3 * 1000, 3), (2 * 40000, 2), (96, 32)):
        kinds, vids, widths = A.partition(a.box, k, n)
        agg = cmi.AggregateBounder(tape, kinds, n, k)
        lb, ub, fe = agg.bound_children(a.box, vids, widths)
        lb_o, ub_o, fe_o = orc.aggregate(a, a.box, vids, widths, kinds, k)
        assert same_bits(lb[0], lb_o) and same_bits(ub[0], ub_o) and np.array_equal(fe[0], fe_o), (n, k)


def test_over_budget_and_shape_errors(ctx):
    eng = A.fixture()
    tape = cmi.Tape(ctx, eng)
    with pytest.raises(C.CmiError) as ei:
        cmi.AggregateBounder(tape, [C.SLOT_CONTINUOUS, C.SLOT_CONTINUOUS], 64, 4, budget_bytes=16)
    assert ei.value.status == C.OVER_BUDGET
    with pytest.raises(C.CmiError) as ei:
        cmi.AggregateBounder(tape, [C.SLOT_CONTINUOUS], 10, 4)  # k does not divide N
    assert ei.value.status == C.INVALID_CONFIG
    agg = cmi.AggregateBounder(tape, [C.SLOT_CONTINUOUS, C.SLOT_CONTINUOUS], 64, 4, max_batch=2)
    a = eng.arrays()
    with pytest.raises(C.CmiError) as ei:
        agg.bound_children(a.box, [1, 0], [16, 8])  # product 128 != 64
    assert ei.value.status == C.SHAPE_MISMATCH
    with pytest.raises(C.CmiError) as ei:
        agg.bound_children(np.stack([a.box] * 3), [1, 0], [16, 4])  # 3 regions > max_batch 2
    assert ei.value.status == C.SHAPE_MISMATCH


@pytest.mark.parametrize("groups", [1, 4, 8])
def test_aggregate_in_function_groups_equals_oracle(ctx, groups):
    """The aggregate tail over a 600-node DAG run in 1, 4 or 8 function groups
    (one wave per group per tile): every child's lower bound, hull upper bound
    and feasibility equal the oracle's bit for bit."""
    eng = P.c2()
    a = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    assert tape.set_box_groups(groups) == groups
    kinds, vids, widths = A.partition(a.box, 8, 1024)
    agg = cmi.AggregateBounder(tape, kinds, 1024, 8)
    lb, ub, fe = agg.bound_children(a.box, vids, widths)
    lb_o, ub_o, fe_o = orc.aggregate(a, a.box, vids, widths, kinds, 8)
    assert same_bits(lb[0], lb_o) and same_bits(ub[0], ub_o) and np.array_equal(fe[0], fe_o)
