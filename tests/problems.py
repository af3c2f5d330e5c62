"""Problem builders shared by the tests: reference fixtures and seeded generators."""
from __future__ import annotations

import numpy as np

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd.model import Problem


def ex2_1_2() -> Problem:
    """MINLPLib ex2_1_2 as the reference's fixed example builds it (source/fixed_examples/qps.cu:22-43).
    Published optimum -213 (MINLP_STATUS.md:528)."""
    p = Problem()
    x = [p.var(0, 1) for _ in range(5)]
    x.append(p.var(0, 20))
    obj = -0.5 * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3] + x[4] * x[4])
    obj = obj - 10.5 * x[0] - 7.5 * x[1] - 3.5 * x[2] - 2.5 * x[3] - 1.5 * x[4] - 10 * x[5]
    p.set_objective(obj)
    p.add_constraint(6 * x[0] + 3 * x[1] + 3 * x[2] + 2 * x[3] + x[4], C.LE, 6.5)
    p.add_constraint(10 * x[0] + 10 * x[2] + x[5], C.LE, 20)
    return p


def ex2_1_9() -> Problem:
    """MINLPLib ex2_1_9 as qps.cu:46-69 builds it (one equality constraint)."""
    p = Problem()
    x = [p.var(0, 1) for _ in range(10)]
    pairs = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 6), (6, 7), (7, 8), (8, 9), (0, 2), (1, 3), (2, 4), (3, 5),
             (4, 6), (5, 7), (6, 8), (7, 9), (0, 8), (0, 9), (1, 9), (0, 4), (3, 6)]
    obj = x[pairs[0][0]] * x[pairs[0][1]]
    for i, j in pairs[1:]:
        obj = obj + x[i] * x[j]
    p.set_objective(-obj)
    c1 = x[0]
    for k in range(1, 10):
        c1 = c1 + x[k]
    p.add_constraint(c1, C.EQ, 1)
    return p


def random_dag(seed: int, n_vars: int = 5, n_ops: int = 60, n_con: int = 3, with_eq: bool = True,
               lo: float = -3.0, hi: float = 3.0, int_vars: int = 0) -> Problem:
    """A random DAG over the device op set: every binary op with node, Const-left,
    Const-right and Const-Const operands, Sqr/Neg, shared subexpressions, LE and EQ
    constraints."""
    rng = np.random.default_rng(seed)
    p = Problem()
    xs = []
    for i in range(n_vars):
        if i < int_vars:
            xs.append(p.int_var(0, 6))
        else:
            a = float(rng.uniform(lo, 0.5 * (lo + hi)))
            b = float(rng.uniform(a + 0.1, hi))
            xs.append(p.var(a, b))
    pool = list(xs)
    for _ in range(n_ops):
        kind = rng.integers(0, 10)
        a = pool[rng.integers(0, len(pool))]
        b = pool[rng.integers(0, len(pool))]
        c = float(np.round(rng.normal(), 3))
        if kind == 0:
            e = a + b
        elif kind == 1:
            e = a - b
        elif kind == 2:
            e = a * b if a.id != b.id else a * a
        elif kind == 3:
            e = a * a
        elif kind == 4:
            e = -a
        elif kind == 5:
            e = c * a if rng.random() < 0.5 else a * c
        elif kind == 6:
            e = c + a if rng.random() < 0.5 else a + c
        elif kind == 7:
            e = c - a if rng.random() < 0.5 else a - c
        elif kind == 8:
            e = p.fixed(c) * p.fixed(0.5 + abs(c)) + a if rng.random() < 0.5 else p.fixed(c) - p.fixed(1.25) + a
        else:
            e = a * 0.0 + b
        pool.append(e)
    tail = pool[n_vars:]
    p.set_objective(tail[-1])
    for k in range(n_con):
        e = tail[rng.integers(0, len(tail))]
        if with_eq and k % 3 == 2:
            p.add_constraint(e, C.EQ, float(np.round(rng.normal(), 2)))
        else:
            p.add_constraint(e, C.LE, float(np.round(rng.uniform(0, 5), 2)))
    return p


def random_ext_dag(seed: int, n_vars: int = 4, n_ops: int = 40, n_con: int = 3, lo: float = -3.0,
                   hi: float = 3.0) -> Problem:
    """A random DAG mixing the extended op set (division, powers, sqrt, exp,
    log, sin, cos, tanh, abs, min, max, x**y) with the QCQP ops. Magnitudes are
    kept in range by squashing: exp and powers take tanh'd or sin'd operands."""
    from cuminlpsearch_amd import model as M
    rng = np.random.default_rng(seed)
    p = Problem()
    xs = []
    for _ in range(n_vars):
        a = float(rng.uniform(lo, 0.5 * (lo + hi)))
        b = float(rng.uniform(a + 0.1, hi))
        xs.append(p.var(a, b))
    pool = list(xs)
    for _ in range(n_ops):
        kind = rng.integers(0, 16)
        a = pool[rng.integers(0, len(pool))]
        b = pool[rng.integers(0, len(pool))]
        c = float(np.round(rng.normal(), 3))
        if kind == 0:
            e = a + b
        elif kind == 1:
            e = a * b if a.id != b.id else a * a
        elif kind == 2:
            e = a / b if rng.random() < 0.7 else (a / (c if c != 0 else 1.5) if rng.random() < 0.5 else c / b)
        elif kind == 3:
            e = M.pown(M.tanh(a), int(rng.integers(-3, 6)))
        elif kind == 4:
            e = M.sqrt(a)
        elif kind == 5:
            e = M.exp(M.sin(a) * 2.0)
        elif kind == 6:
            e = M.log(a)
        elif kind == 7:
            e = M.sin(a * 1.7)
        elif kind == 8:
            e = M.cos(a - b)
        elif kind == 9:
            e = M.tanh(a)
        elif kind == 10:
            e = M.fabs(a) - 0.5
        elif kind == 11:
            e = M.minimum(a, b) if rng.random() < 0.5 else M.minimum(c, a)
        elif kind == 12:
            e = M.maximum(a, b) if rng.random() < 0.5 else M.maximum(a, c)
        elif kind == 13:
            e = M.power(M.fabs(a) + 0.5, M.sin(b))
        elif kind == 14:
            e = a - b
        else:
            e = c * a + b
        pool.append(e)
    tail = pool[n_vars:]
    p.set_objective(tail[-1])
    for k in range(n_con):
        e = tail[rng.integers(0, len(tail))]
        if k % 3 == 2:
            p.add_constraint(e, C.EQ, float(np.round(rng.normal(), 2)))
        else:
            p.add_constraint(e, C.LE, float(np.round(rng.uniform(0, 5), 2)))
    return p


C1_SPEC = dict(n=4, m=2, cont_lb=-10.0, cont_ub=10.0, symmetric_pairs=True)
C2_SPEC = dict(n=32, m=16, cont_lb=-5.0, cont_ub=5.0, pair_density=0.2, target_nodes=600)
C2_SEED = 2026
# MINLP B&B: 64 variables, the first 16 integer in [0, 10], the rest continuous
# in [-5, 5]; 32 constraints; 10 % bilinear pairs plus squares; about 1,500 DAG nodes
C3_SPEC = dict(n=64, m=32, n_int=16, int_lb=0.0, int_ub=10.0, cont_lb=-5.0, cont_ub=5.0, pair_density=0.1,
               target_nodes=1500)


def c1(seed: int) -> cmi.EngineProblem:
    return cmi.EngineProblem.qcqp(seed, **C1_SPEC)


def c2() -> cmi.EngineProblem:
    return cmi.EngineProblem.qcqp(C2_SEED, **C2_SPEC)


def c3(seed: int) -> cmi.EngineProblem:
    return cmi.EngineProblem.qcqp(seed, **C3_SPEC)


def random_boxes(seed: int, n_vars: int, n: int, lo: float, hi: float, wmin: float = 1e-4, wmax: float = 10.0):
    """Sub-boxes with log-uniform widths in [wmin, min(wmax, hi-lo)], SoA (n_vars, n)."""
    rng = np.random.default_rng(seed)
    w = np.exp(rng.uniform(np.log(wmin), np.log(min(wmax, hi - lo)), size=(n_vars, n)))
    lb = rng.uniform(lo, hi - w)
    ub = np.minimum(lb + w, hi)
    return lb, ub
