"""Host-side checks that need no GPU: the library loads and exports every
symbol include/cmi.h declares, problem validation maps the reference's error
classes onto status codes, the generator reproduces the benchmark shapes, and
the tape compiler's output is well formed."""
import pathlib
import re

import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd.model import Problem

import problems as P

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "cmi.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cmi_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = C.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes binding covers exactly the declared set
    assert sorted(C.EXPORTED) == names


def test_validation_errors_map_to_status_codes():
    p = Problem()
    x = p.var(0.0, 1.0)
    p.set_objective(x * 2.0)
    a = p.arrays()
    a.in0[-1] = len(a.op) - 1  # operand not strictly earlier
    with pytest.raises(C.CmiError) as e:
        cmi.EngineProblem.from_arrays(a)
    assert e.value.status == C.INVALID_DAG

    p = Problem()
    x = p.var(2.0, 1.0)
    p.set_objective(x * x)
    with pytest.raises(C.CmiError) as e:
        cmi.EngineProblem.from_problem(p)
    assert e.value.status == C.INVALID_PROBLEM

    p = Problem()
    x = p.int_var(0.5, 3.0)
    p.set_objective(x * x)
    with pytest.raises(C.CmiError) as e:
        cmi.EngineProblem.from_problem(p)
    assert e.value.status == C.INVALID_PROBLEM

    p = Problem()
    p.var(0.0, 1.0)
    with pytest.raises(C.CmiError) as e:  # objective never set
        cmi.EngineProblem.from_problem(p)
    assert e.value.status == C.INVALID_PROBLEM

    p = Problem()
    x = p.var(0.0, 1.0)
    p.set_objective(p.fixed(3.0))
    with pytest.raises(C.CmiError) as e:
        cmi.EngineProblem.from_problem(p)
    assert e.value.status == C.INVALID_DAG


def test_extended_ops_compile_to_an_extended_tape():
    p = Problem()
    x = p.var(1.0, 2.0)
    p.set_objective(x / (x + 1.0))
    info, _ = cmi.EngineProblem.from_problem(p).tape_info()
    assert info.extended == 1
    q = Problem()
    y = q.var(1.0, 2.0)
    q.set_objective(y * y + y)
    assert cmi.EngineProblem.from_problem(q).tape_info()[0].extended == 0


def test_extended_unary_on_const_is_rejected():
    p = Problem()
    x = p.var(1.0, 2.0)
    c = p.fixed(2.0)
    e = cmi.Expr(p, p._emit(C.OP_EXP, (c.id,)))
    p.set_objective(x + e)
    eng = cmi.EngineProblem.from_problem(p)
    with pytest.raises(C.CmiError) as err:
        eng.tape_info()
    assert err.value.status == C.INVALID_DAG


def test_unary_on_const_is_rejected():
    p = Problem()
    x = p.var(1.0, 2.0)
    c = p.fixed(2.0)
    neg = cmi.Expr(p, p._emit(C.OP_NEG, (c.id,)))  # the Expr API folds this; a flat DAG can still carry it
    p.set_objective(x + neg)
    eng = cmi.EngineProblem.from_problem(p)
    with pytest.raises(C.CmiError) as e:
        eng.tape_info()
    assert e.value.status == C.INVALID_DAG


def test_generator_shapes_and_determinism():
    c1 = P.c1(0)
    assert c1.info().reachable_nodes == 95  # 4 vars, 10 products, 3 x (14 terms + 13 sums)
    c2 = P.c2()
    assert c2.info().reachable_nodes == 600
    a = P.c2().arrays()
    b = P.c2().arrays()
    assert np.array_equal(a.op, b.op) and np.array_equal(a.constant.view(np.uint64), b.constant.view(np.uint64))
    assert a.n_vars == 32 and a.n_constraints == 16
    assert np.all(a.box[:, 0] == -5.0) and np.all(a.box[:, 1] == 5.0)


def test_export_create_roundtrip():
    a = P.c1(5).arrays()
    again = cmi.EngineProblem.from_arrays(a).arrays()
    for f in ("op", "in0", "in1", "var_index", "con_root", "con_cmp"):
        assert np.array_equal(getattr(a, f), getattr(again, f)), f
    assert np.array_equal(a.constant.view(np.uint64), again.constant.view(np.uint64))


def test_tape_shapes():
    ti, listing = P.c2().tape_info(listing=True)
    assert ti.n_vars == 32
    assert ti.n_slots <= 40  # variables plus a few temporaries
    assert ti.n_outputs == 17
    lines = listing.strip().splitlines()
    assert len(lines) == ti.n_instr
    assert sum(1 for l in lines if re.search(r" (acc\.)?chkle ", l)) == 16
    assert sum(1 for l in lines if re.search(r" (acc\.)?obj ", l)) == 1
    # every slot a tape instruction names is below n_slots
    for l in lines:
        d = int(re.search(r"d=(\d+)", l).group(1))
        assert d < ti.n_slots
