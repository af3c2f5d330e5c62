"""The extended op set on the device (Div, PowN, Sqrt, Exp, Log, Sin, Cos,
Tanh, Abs, Min, Max, x**y) against the CPU oracle, bit for bit: every word
form through the tape interpreter and through the kernel compiled from the
tape, random DAGs mixing all ops, point evaluation, and directed division
and square root against the x86 correctly rounded ones. The oracle's
enclosures are pinned by mpmath (test_oracle_extended.py)."""
import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd import model as M
from cuminlpsearch_amd.model import Problem
from oracle import binding as orc

import op_cases
import problems as P

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_diff(a, b):
    a = np.ascontiguousarray(a, np.float64).ravel()
    b = np.ascontiguousarray(b, np.float64).ravel()
    return np.flatnonzero(a.view(np.uint64) != b.view(np.uint64))[:5]


def compare(ctx, eng, lb, ub, engine):
    tape = cmi.Tape(ctx, eng)
    tape.set_box_engine(engine)
    f_d, lo_d, hi_d, r_d = tape.eval_boxes(lb, ub, roots=True)
    f_o, lo_o, hi_o, r_o = orc.eval_boxes(eng.arrays(), lb, ub, roots=True)
    assert np.array_equal(f_d, f_o), np.flatnonzero(f_d != f_o)[:5]
    assert same_bits(lo_d, lo_o), first_diff(lo_d, lo_o)
    assert same_bits(hi_d, hi_o), first_diff(hi_d, hi_o)
    assert same_bits(r_d, r_o), first_diff(r_d, r_o)
    return f_d, lo_d, hi_d


def boxes(seed, n_vars, n):
    """Random boxes, plus point boxes and boxes with a zero or domain edge at an end."""
    lb, ub = P.random_boxes(seed, n_vars, n, -3.0, 4.0, 1e-9, 6.0)
    rng = np.random.default_rng(seed)
    k = n // 8
    pts = rng.uniform(-3.0, 4.0, (n_vars, k))
    lb[:, :k] = pts
    ub[:, :k] = pts
    lb[:, k:2 * k] = 0.0
    ub[:, k:2 * k] = np.maximum(ub[:, k:2 * k], 1e-3)
    ub[:, 2 * k:3 * k] = 0.0
    lb[:, 2 * k:3 * k] = np.minimum(lb[:, 2 * k:3 * k], -1e-3)
    assert np.all(lb <= ub)
    return lb, ub


@pytest.mark.parametrize("engine", [C.BOX_INTERPRETED, C.BOX_COMPILED])
@pytest.mark.parametrize("name", sorted(op_cases.EXT_CASES))
def test_ext_op_case_bitexact(ctx, name, engine):
    eng = cmi.EngineProblem.from_problem(op_cases.EXT_CASES[name]())
    print(f"case {name} engine {engine}", flush=True)
    lb, ub = boxes(5, 3, 512)
    compare(ctx, eng, lb, ub, engine)


@pytest.mark.parametrize("engine", [C.BOX_INTERPRETED, C.BOX_COMPILED])
@pytest.mark.parametrize("seed", range(8))
def test_random_extended_dags_bitexact(ctx, seed, engine):
    p = P.random_ext_dag(seed, n_vars=3 + seed % 4, n_ops=30 + 5 * (seed % 3), n_con=5)
    eng = cmi.EngineProblem.from_problem(p)
    print(f"seed {seed} engine {engine}", flush=True)
    lb, ub = P.random_boxes(200 + seed, eng.info().n_vars, 1000, -3.0, 3.0, 1e-9, 6.0)
    compare(ctx, eng, lb, ub, engine)


def _pair_problem(fn):
    p = Problem()
    x = p.var(-1e308, 1e308)
    y = p.var(-1e308, 1e308)
    p.set_objective(fn(x, y))
    return cmi.EngineProblem.from_problem(p)


def _doubles(rng, n, lo=-1070, hi=1020):
    sign = rng.choice([-1.0, 1.0], n)
    return sign * rng.uniform(1.0, 2.0, n) * np.exp2(rng.integers(lo, hi, n).astype(np.float64))


def test_directed_division_is_correctly_rounded(ctx):
    """[a, a] / [b, b] = [RD(a/b), RU(a/b)] over the whole exponent range,
    subnormal and overflowing quotients included."""
    rng = np.random.default_rng(77)
    n = 20000
    a = _doubles(rng, n)
    b = _doubles(rng, n)
    a[:64] = np.array([1.0, 3.0, -7.0, 1e-310, 5e-324, 1.7976931348623157e308] * 11)[:64]
    b[:64] = np.array([3.0, 7.0, 3.0, 3.0, 2.0, 0.5] * 11)[:64]
    eng = _pair_problem(lambda x, y: x / y)
    tape = cmi.Tape(ctx, eng)
    xs = np.stack([a, b])
    _, lo, hi, _ = tape.eval_boxes(xs, xs)
    want_lo = np.array([orc.lib().orc_div_rd(u, v) for u, v in zip(a, b)])
    want_hi = np.array([orc.lib().orc_div_ru(u, v) for u, v in zip(a, b)])
    want_lo = np.where(want_lo == 0.0, 0.0, want_lo)
    want_hi = np.where(want_hi == 0.0, 0.0, want_hi)
    assert same_bits(lo, want_lo), first_diff(lo, want_lo)
    assert same_bits(hi, want_hi), first_diff(hi, want_hi)


def test_directed_sqrt_is_correctly_rounded(ctx):
    rng = np.random.default_rng(78)
    n = 20000
    v = np.abs(_doubles(rng, n, -1074, 1023))
    v[:8] = [4.0, 2.0, 1e-320, 5e-324, 1.7976931348623157e308, 0.25, 9.0, 2.0 ** -1022]
    p = Problem()
    x = p.var(0.0, 1e308)
    p.set_objective(M.sqrt(x))
    tape = cmi.Tape(ctx, cmi.EngineProblem.from_problem(p))
    xs = v[None, :]
    _, lo, hi, _ = tape.eval_boxes(xs, xs)
    want_lo = np.array([orc.lib().orc_sqrt_rd(u) for u in v])
    want_hi = np.array([orc.lib().orc_sqrt_ru(u) for u in v])
    assert same_bits(lo, want_lo), first_diff(lo, want_lo)
    assert same_bits(hi, want_hi), first_diff(hi, want_hi)


@pytest.mark.parametrize("fn,lo,hi", [
    ("exp", -750.0, 712.0), ("log", 1e-300, 1e300), ("sin", -1e6, 1e6), ("cos", -1e6, 1e6), ("tanh", -45.0, 45.0),
])
def test_wide_range_enclosures_bitexact(ctx, fn, lo, hi):
    """Each transcendental enclosure over its whole range of arguments, boxes of
    every width from a point to a whole period."""
    rng = np.random.default_rng(hash(fn) % 1000)
    n = 8192
    if fn == "log":
        a = 10.0 ** rng.uniform(-300, 300, n)
        w = a * 10.0 ** rng.uniform(-17, 1, n)
    else:
        a = rng.uniform(lo, hi, n)
        w = np.where(rng.random(n) < 0.25, 0.0, 10.0 ** rng.uniform(-15, 1, n))
    p = Problem()
    x = p.var(-1e308, 1e308)
    p.set_objective(getattr(M, fn)(x))
    eng = cmi.EngineProblem.from_problem(p)
    for engine in (C.BOX_INTERPRETED, C.BOX_COMPILED):
        compare(ctx, eng, a[None, :], (a + w)[None, :], engine)


@pytest.mark.parametrize("name", sorted(op_cases.EXT_CASES))
def test_ext_point_evaluation_bitexact(ctx, name):
    eng = cmi.EngineProblem.from_problem(op_cases.EXT_CASES[name]())
    rng = np.random.default_rng(9)
    x = rng.uniform(-3.0, 4.0, (3, 1000))
    x[:, :50] = 0.0
    tape = cmi.Tape(ctx, eng)
    f_d, o_d, r_d = tape.eval_points(x, roots=True)
    f_o, o_o, r_o = orc.eval_points(eng.arrays(), x, roots=True)
    assert np.array_equal(f_d, f_o)
    assert same_bits(o_d, o_o), first_diff(o_d, o_o)
    assert same_bits(r_d, r_o), first_diff(r_d, r_o)
