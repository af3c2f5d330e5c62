"""Device tape evaluation against the CPU oracle: bit-exact interval roots,
objective bounds and feasibility on identical boxes, and enclosure of the
exact (rational) values of the DAG at points inside each box."""
import json
import math
import pathlib
from fractions import Fraction

import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd.model import Problem
from oracle import binding as orc

import op_cases
import problems as P

pytestmark = pytest.mark.gpu

ROWS = json.loads((pathlib.Path(__file__).parent / "golden" / "directed_rounding.json").read_text())["rows"]


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def binary_problem(op):
    p = Problem()
    x = p.var(-1e308, 1e308)
    y = p.var(-1e308, 1e308)
    p.set_objective(x + y if op == "add" else x - y if op == "sub" else x * y)
    return p


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
def test_device_directed_primitives_match_golden(ctx, op):
    """[a,a] op [b,b] on the device gives [RD(a op b), RU(a op b)] bit for bit
    (a zero bound is reported as +0)."""
    rows = [r for r in ROWS if all(math.isfinite(float.fromhex(r[k])) for k in ("a", "b", f"{op}_rd", f"{op}_ru"))]
    a = np.array([float.fromhex(r["a"]) for r in rows])
    b = np.array([float.fromhex(r["b"]) for r in rows])
    lo = np.array([float.fromhex(r[f"{op}_rd"]) for r in rows])
    hi = np.array([float.fromhex(r[f"{op}_ru"]) for r in rows])
    lo = np.where(lo == 0.0, 0.0, lo)
    hi = np.where(hi == 0.0, 0.0, hi)
    prob = cmi.EngineProblem.from_problem(binary_problem(op))
    tape = cmi.Tape(ctx, prob)
    xs = np.stack([a, b])
    feas, olb, oub, _ = tape.eval_boxes(xs, xs)
    assert same_bits(olb, lo), np.flatnonzero(olb.view(np.uint64) != lo.view(np.uint64))[:5]
    assert same_bits(oub, hi)


def compare_with_oracle(ctx, eng: cmi.EngineProblem, lb, ub, engine=C.BOX_INTERPRETED):
    arrays = eng.arrays()
    tape = cmi.Tape(ctx, eng)
    tape.set_box_engine(engine)
    f_d, lo_d, hi_d, r_d = tape.eval_boxes(lb, ub, roots=True)
    f_o, lo_o, hi_o, r_o = orc.eval_boxes(arrays, lb, ub, roots=True)
    assert np.array_equal(f_d, f_o)
    assert same_bits(lo_d, lo_o)
    assert same_bits(hi_d, hi_o)
    assert same_bits(r_d, r_o)
    return f_d, lo_d, hi_d, r_d


@pytest.mark.parametrize("name", sorted(op_cases.CASES))
def test_op_case_bitexact(ctx, name):
    """One instruction form at a time (every tape op code over all cases)."""
    eng = cmi.EngineProblem.from_problem(op_cases.CASES[name]())
    lb, ub = P.random_boxes(5, 3, 400, -3.0, 4.0, 1e-6, 6.0)
    compare_with_oracle(ctx, eng, lb, ub)


@pytest.mark.parametrize("engine", [C.BOX_INTERPRETED, C.BOX_COMPILED, C.BOX_AUTO])
def test_c2_kernel_alone_bitexact(ctx, engine):
    """The 600-node QCQP (n=32, m=16) at full size: 65,536 boxes with
    log-uniform widths in [1e-4, 10], each box kernel."""
    eng = P.c2()
    assert eng.info().reachable_nodes == 600
    lb, ub = P.random_boxes(2026, 32, 65536, -5.0, 5.0)
    f, lo, hi, roots = compare_with_oracle(ctx, eng, lb, ub, engine)
    assert 0 < f.sum() < f.size
    assert np.all(lo <= hi)


def test_compiled_op_cases_bitexact(ctx):
    """Every instruction form through the kernel compiled from the tape."""
    lb, ub = P.random_boxes(5, 3, 400, -3.0, 4.0, 1e-6, 6.0)
    for name in sorted(op_cases.CASES):
        eng = cmi.EngineProblem.from_problem(op_cases.CASES[name]())
        compare_with_oracle(ctx, eng, lb, ub, C.BOX_COMPILED)


@pytest.mark.parametrize("seed", range(6))
def test_compiled_random_dags_bitexact(ctx, seed):
    p = P.random_dag(seed, n_vars=4 + seed % 5, n_ops=40 + 10 * (seed % 4), n_con=4)
    eng = cmi.EngineProblem.from_problem(p)
    lb, ub = P.random_boxes(100 + seed, eng.info().n_vars, 777, -3.0, 3.0, 1e-6, 6.0)
    compare_with_oracle(ctx, eng, lb, ub, C.BOX_COMPILED)


def test_compiled_ragged_counts_and_auto_switch(ctx):
    eng = P.c1(3)
    tape = cmi.Tape(ctx, eng)
    tape.set_box_engine(C.BOX_COMPILED)
    info = tape.info()
    assert info.box_compiled == 1 and info.compile_seconds > 0
    for n in (1, 63, 64, 65, 255, 256, 257, 1000):
        lb, ub = P.random_boxes(n, 4, n, -10.0, 10.0)
        f_d, lo_d, hi_d, r_d = tape.eval_boxes(lb, ub, roots=True)
        f_o, lo_o, hi_o, r_o = orc.eval_boxes(eng.arrays(), lb, ub, roots=True)
        assert np.array_equal(f_d, f_o) and same_bits(lo_d, lo_o) and same_bits(hi_d, hi_o) and same_bits(r_d, r_o)
    # Auto: small calls stay on the interpreter, a large call compiles
    t2 = cmi.Tape(ctx, P.c1(4))
    lb, ub = P.random_boxes(1, 4, 100, -10.0, 10.0)
    t2.eval_boxes(lb, ub)
    assert t2.info().box_compiled == 0
    lb, ub = P.random_boxes(2, 4, 40000, -10.0, 10.0)
    f_d, lo_d, hi_d, _ = t2.eval_boxes(lb, ub)
    assert t2.info().box_compiled == 1
    f_o, lo_o, hi_o, _ = orc.eval_boxes(P.c1(4).arrays(), lb, ub)
    assert np.array_equal(f_d, f_o) and same_bits(lo_d, lo_o) and same_bits(hi_d, hi_o)


@pytest.mark.parametrize("seed", range(12))
def test_random_dags_bitexact(ctx, seed):
    p = P.random_dag(seed, n_vars=4 + seed % 5, n_ops=40 + 10 * (seed % 4), n_con=4)
    eng = cmi.EngineProblem.from_problem(p)
    lb, ub = P.random_boxes(100 + seed, eng.info().n_vars, 777, -3.0, 3.0, 1e-6, 6.0)
    compare_with_oracle(ctx, eng, lb, ub)


def test_function_groups_do_not_change_results(ctx):
    """Every split of the functions over the waves of a workgroup gives the
    oracle's results bit for bit, on C2 and on a random DAG with many constraints."""
    cases = [(P.c2(), P.random_boxes(7, 32, 3000, -5.0, 5.0))]
    p = P.random_dag(77, n_vars=6, n_ops=120, n_con=11)
    e2 = cmi.EngineProblem.from_problem(p)
    cases.append((e2, P.random_boxes(8, 6, 1001, -3.0, 3.0, 1e-5, 6.0)))
    for eng, (lb, ub) in cases:
        arrays = eng.arrays()
        f_o, lo_o, hi_o, r_o = orc.eval_boxes(arrays, lb, ub, roots=True)
        tape = cmi.Tape(ctx, eng)
        seen = set()
        for k in range(1, 9):
            got = tape.set_box_groups(k)
            seen.add(got)
            f_d, lo_d, hi_d, r_d = tape.eval_boxes(lb, ub, roots=True)
            assert np.array_equal(f_d, f_o), k
            assert same_bits(lo_d, lo_o), k
            assert same_bits(hi_d, hi_o), k
            assert same_bits(r_d, r_o), k
        assert len(seen) >= 4


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129, 1000])
def test_ragged_box_counts(ctx, n):
    eng = P.c1(3)
    lb, ub = P.random_boxes(n, 4, n, -10.0, 10.0)
    compare_with_oracle(ctx, eng, lb, ub)


def exact_eval(a, x):
    """Exact rational value of every node at the rational point x."""
    n = a.op.shape[0]
    v = [None] * n
    for i in range(n):
        op = int(a.op[i])
        if op == C.OP_VAR:
            v[i] = Fraction(x[int(a.var_index[i])])
        elif op == C.OP_CONST:
            v[i] = Fraction(float(a.constant[i]))
        else:
            u = v[int(a.in0[i])]
            w = v[int(a.in1[i])] if op in (C.OP_ADD, C.OP_SUB, C.OP_MUL) else None
            v[i] = {C.OP_ADD: lambda: u + w, C.OP_SUB: lambda: u - w, C.OP_MUL: lambda: u * w,
                    C.OP_SQR: lambda: u * u, C.OP_NEG: lambda: -u}[op]()
    return v


@pytest.mark.parametrize("seed", range(4))
def test_enclosure_of_exact_values(ctx, seed):
    p = P.random_dag(50 + seed, n_vars=4, n_ops=30, n_con=3)
    eng = cmi.EngineProblem.from_problem(p)
    a = eng.arrays()
    lb, ub = P.random_boxes(seed, 4, 64, -3.0, 3.0, 1e-3, 6.0)
    tape = cmi.Tape(ctx, eng)
    _, _, _, roots = tape.eval_boxes(lb, ub, roots=True)
    root_ids = [a.objective_root] + [int(r) for r in a.con_root]
    rng = np.random.default_rng(seed)
    for e in range(64):
        for _ in range(4):
            x = [float(rng.uniform(lb[v, e], ub[v, e])) for v in range(4)]
            vals = exact_eval(a, x)
            for k, rid in enumerate(root_ids):
                assert Fraction(roots[k, e, 0]) <= vals[rid] <= Fraction(roots[k, e, 1])


@pytest.mark.parametrize("seed", range(6))
def test_point_evaluation_bitexact(ctx, seed):
    p = P.random_dag(seed, n_vars=5, n_ops=50, n_con=4)
    eng = cmi.EngineProblem.from_problem(p)
    a = eng.arrays()
    rng = np.random.default_rng(seed)
    x = rng.uniform(-3, 3, size=(5, 300))
    f_d, obj_d, r_d = cmi.Tape(ctx, eng).eval_points(x, roots=True)
    f_o, obj_o, r_o = orc.eval_points(a, x, roots=True)
    assert np.array_equal(f_d, f_o)
    assert same_bits(obj_d, obj_o)
    assert same_bits(r_d, r_o)
    for e in range(0, 300, 37):
        vals = cmi.evaluate_point(a, x[:, e])
        assert vals[a.objective_root] == obj_d[e]


def test_zero_boxes_is_a_noop(ctx):
    eng = P.c1(0)
    tape = cmi.Tape(ctx, eng)
    f, lo, hi, _ = tape.eval_boxes(np.zeros((4, 0)), np.zeros((4, 0)))
    assert f.size == 0


def test_autotune_keeps_results_and_picks_the_fastest(ctx):
    eng = P.c2()
    tape = cmi.Tape(ctx, eng)
    table = tape.autotune(8192)
    assert len(table) >= 4
    assert {c["engine"] for c in table} == {"interpreted", "compiled"}
    best = min(table, key=lambda c: c["kernel_ms"])
    info = tape.info()
    assert (info.box_compiled == 1) == (best["engine"] == "compiled")
    lb, ub = P.random_boxes(99, 32, 5000, -5.0, 5.0)
    f_d, lo_d, hi_d, r_d = tape.eval_boxes(lb, ub, roots=True)
    f_o, lo_o, hi_o, r_o = orc.eval_boxes(eng.arrays(), lb, ub, roots=True)
    assert np.array_equal(f_d, f_o) and same_bits(lo_d, lo_o) and same_bits(hi_d, hi_o) and same_bits(r_d, r_o)


@pytest.mark.parametrize("groups,waves", [(1, 1), (4, 1), (8, 2), (8, 4)])
@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("load_batch", [0, 3, 8])
def test_compiled_codegen_bitexact(ctx, groups, waves, schedule, load_batch):
    """Every scheduler and load batching of the compiled kernel gives the
    oracle's results bit for bit, both the kernel that writes root intervals
    and the one that computes only what feasibility and the objective read."""
    eng = P.c2()
    tape = cmi.Tape(ctx, eng)
    tape.set_box_groups(groups)
    tape.set_box_engine(C.BOX_COMPILED)
    tape.set_compiled_shape(False, waves, 1)
    tape.set_compiled_codegen(schedule, load_batch)
    info = tape.info()
    assert info.compiled_schedule == schedule and info.compiled_waves == waves
    assert info.compiled_load_batch == load_batch
    n = 64 * 700 + 5
    lb, ub = P.random_boxes(321 + load_batch, 32, n, -5.0, 5.0)
    f_o, lo_o, hi_o, r_o = orc.eval_boxes(eng.arrays(), lb, ub, roots=True)
    f_n, lo_n, hi_n, _ = tape.eval_boxes(lb, ub)
    assert np.array_equal(f_n, f_o) and same_bits(lo_n, lo_o) and same_bits(hi_n, hi_o)
    f_d, lo_d, hi_d, r_d = tape.eval_boxes(lb, ub, roots=True)
    assert np.array_equal(f_d, f_o) and same_bits(lo_d, lo_o) and same_bits(hi_d, hi_o) and same_bits(r_d, r_o)


@pytest.mark.parametrize("waves", [1, 2, 8])
@pytest.mark.parametrize("persistent", [False, True])
def test_compiled_shapes_bitexact(ctx, waves, persistent):
    """Every compiled shape (groups interleaved in a wave, persistent tiles) gives
    the oracle's results, including a ragged last tile and several tiles per
    persistent workgroup."""
    eng = P.c2()
    tape = cmi.Tape(ctx, eng)
    tape.set_box_groups(8)
    tape.set_box_engine(C.BOX_COMPILED)
    tape.set_compiled_shape(persistent, waves, 1)
    info = tape.info()
    assert info.compiled_waves == waves and info.compiled_persistent == int(persistent)
    n = 256 * 64 * 3 + 17  # more tiles than one workgroup per CU: each persistent workgroup walks several
    lb, ub = P.random_boxes(123, 32, n, -5.0, 5.0)
    f_d, lo_d, hi_d, r_d = tape.eval_boxes(lb, ub, roots=True)
    f_o, lo_o, hi_o, r_o = orc.eval_boxes(eng.arrays(), lb, ub, roots=True)
    assert np.array_equal(f_d, f_o) and same_bits(lo_d, lo_o) and same_bits(hi_d, hi_o) and same_bits(r_d, r_o)
