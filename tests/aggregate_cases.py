"""Problems and partitions for the k-child aggregate tests."""
from __future__ import annotations

import numpy as np

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd.model import Problem


def fixture() -> cmi.EngineProblem:
    """The reference's aggregate fixture (aggregate_backend_test.cu:52-73): a
    shared subexpression and a constraint that excludes part of the box."""
    p = Problem()
    x = p.var(-2.0, 3.0)
    y = p.var(-1.0, 4.0)
    shared = x * x + y
    p.set_objective(shared + 0.5 * y * y - x)
    p.add_constraint(shared, C.LE, 6.0)
    return cmi.EngineProblem.from_problem(p)


def partition(box, k: int, n: int, var_kinds=None):
    """A branch/refine partition in the layout the aggregate relies on: branch
    k ways on the widest live variable (slot last, most significant), refine
    M = n / k ways on the next widest (one slot), nested on the branch
    variable when it is the only live one. Returns (kinds, var_ids, widths)."""
    box = np.asarray(box, np.float64).reshape(-1, 2)
    live = [i for i in range(len(box)) if box[i, 1] > box[i, 0]]
    widths = box[:, 1] - box[:, 0]
    order = sorted(live, key=lambda i: (-widths[i], i))
    kind_of = (lambda v: C.SLOT_CONTINUOUS if var_kinds is None or var_kinds[v] == C.CONTINUOUS
               else C.SLOT_INTEGER_PARTITION)
    b = order[0]
    if len(order) == 1:
        return [kind_of(b)], [b], [n]
    r = order[1]
    return [kind_of(r), kind_of(b)], [r, b], [n // k, k]
