"""The CPU baseline's evaluator (orc_eval_boxes_ru: round-up set once per call,
lower bounds carried negated) equals the oracle's per-operation directed
rounding bit for bit: feasibility, objective lower and upper bounds."""
import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from oracle import binding as orc

import problems as P


def check(arrays, lb, ub):
    f0, l0, u0, _ = orc.eval_boxes(arrays, lb, ub)
    f1, l1, u1 = orc.eval_boxes_ru(arrays, lb, ub)
    assert np.array_equal(f0, f1)
    assert np.array_equal(l0.view(np.uint64), l1.view(np.uint64))
    assert np.array_equal(u0.view(np.uint64), u1.view(np.uint64))


def test_600_node_qcqp_over_random_boxes():
    a = P.c2().arrays()
    lb, ub = P.random_boxes(5, a.n_vars, 3000, -5.0, 5.0)
    check(a, lb, ub)


@pytest.mark.parametrize("seed", range(6))
def test_random_qcqps_with_points_zeros_and_signs(seed):
    a = cmi.EngineProblem.qcqp(900 + seed, n=7, m=4, cont_lb=-3.0, cont_ub=3.0, pair_density=0.6).arrays()
    lb, ub = P.random_boxes(seed, a.n_vars, 2000, -3.0, 3.0, wmin=1e-12, wmax=6.0)
    # degenerate boxes, boxes touching zero and boxes of one sign
    lb[:, :200] = ub[:, :200]
    lb[:, 200:400] = 0.0
    ub[:, 400:600] = 0.0
    lb[:, 400:600] = np.minimum(lb[:, 400:600], -1e-3)
    check(a, lb, ub)


@pytest.mark.parametrize("seed", range(4))
def test_random_dags_of_the_qcqp_op_set(seed):
    a = P.random_dag(seed, n_vars=5, n_ops=80).arrays()
    lb, ub = P.random_boxes(seed + 10, a.n_vars, 1500, -4.0, 4.0, wmin=1e-6, wmax=8.0)
    check(a, lb, ub)


def test_an_extended_op_is_refused():
    a = P.random_ext_dag(1).arrays()
    lb, ub = P.random_boxes(1, a.n_vars, 10, 0.5, 2.0)
    with pytest.raises(RuntimeError, match="QCQP op set"):
        orc.eval_boxes_ru(a, lb, ub)
