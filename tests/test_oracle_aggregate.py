"""The oracle's k-child aggregate against its own per-subregion verdicts
(the reduction of reference backend/aggregate/replay.cuh:303-322 restated in
numpy) and against the oracle bounder, on the CPU."""
import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from oracle import binding as orc

import aggregate_cases as A
import problems as P


def host_reduction(feas, lo, hi, k):
    m = feas.size // k
    out = []
    for c in range(k):
        f = feas[c * m:(c + 1) * m].astype(bool)
        if not f.any():
            out.append((np.inf, np.inf, 0))
        else:
            out.append((lo[c * m:(c + 1) * m][f].min(), hi[c * m:(c + 1) * m][f].max(), 1))
    return out


@pytest.mark.parametrize("n", [16, 256, 4096])
def test_aggregate_equals_reduction_of_bounder_verdicts(n):
    eng = A.fixture()
    a = eng.arrays()
    kinds, vids, widths = A.partition(a.box, 4, n)
    lb, ub, fe, (rf, rl, ru) = orc.aggregate(a, a.box, vids, widths, kinds, 4, per_region=True)
    f_b, lb_b = orc.bound(a, a.box, vids, widths, kinds)
    assert np.array_equal(rf, f_b) and np.array_equal(rl.view(np.uint64), lb_b.view(np.uint64))
    for c, (l, u, f) in enumerate(host_reduction(rf, rl, ru, 4)):
        assert fe[c] == f and lb[c] == l and ub[c] == u
    assert fe.sum() >= 1
    if n >= 256:
        assert (rf == 0).any()  # the constraint excludes part of the box


def test_all_excluded_child_is_infeasible_and_infinite():
    eng = A.fixture()
    a = eng.arrays()
    # x in [2.9, 3]: x*x + y >= 8.41 - 1 > 6 everywhere, so every child is excluded
    box = np.array([[2.9, 3.0], [-1.0, 4.0]])
    kinds, vids, widths = A.partition(box, 2, 8)
    lb, ub, fe = orc.aggregate(a, box, vids, widths, kinds, 2)
    assert list(fe) == [0, 0] and np.all(np.isinf(lb)) and np.all(np.isinf(ub)) and np.all(lb > 0)


def test_refinement_never_loosens_the_children():
    """Reference aggregate_backend_test.cu:164-206: more subregions per child
    can only raise the min of their lower bounds."""
    eng = A.fixture()
    a = eng.arrays()
    prev = None
    for m in (2, 16, 256):
        kinds, vids, widths = A.partition(a.box, 4, 4 * m)
        lb, ub, fe = orc.aggregate(a, a.box, vids, widths, kinds, 4)
        assert np.all(fe == 1) and np.all(ub >= lb)
        if prev is not None:
            assert np.all(lb >= prev)
        prev = lb
