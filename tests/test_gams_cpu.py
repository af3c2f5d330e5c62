"""GAMS scalar-format frontend: lexing, expression parsing, objective
elimination, lowering, diagnostics and the RESULT line.

Follows the reference's own frontend tests case by case
(test/source/gams_test.cpp:71-939); every MINLPLib model the reference tests
with is parsed from tests/golden/gams/, and its variable, constraint and DAG
node counts are checked against the counts the reference frontend recorded
for it (tests/golden/gams_corpus.json). Host only.
"""
from __future__ import annotations

import json
import math
import pathlib

import numpy as np
import pytest

from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd import gams
from cuminlpsearch_amd.model import evaluate_objective, evaluate_point

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
CORPUS = json.loads((GOLDEN / "gams_corpus.json").read_text())


def data_file(name: str) -> pathlib.Path:
    return GOLDEN / "gams" / name


def model_with(expression: str) -> str:
    """`expression` as the objective of a model in GAMS Convert's shape
    (reference gams_test.cpp:26-34)."""
    return ("Variables x1,x2,objvar;\n"
            "Equations e1;\n"
            "e1.. -(" + expression + ") + objvar =E= 0;\n"
            "x1.lo = -10; x1.up = 10;\n"
            "x2.lo = -10; x2.up = 10;\n"
            "Solve m using NLP minimizing objvar;\n")


def parsed_eval(expression: str, x1: float = 0.0, x2: float = 0.0) -> float:
    """The objective of model_with(expression) at (x1, x2); `+ 0*x1` keeps a
    constant expression a variable objective (0*x is never folded)."""
    pm = gams.parse(model_with("(" + expression + ") + 0*x1"))
    return evaluate_objective(pm.problem.arrays(), [x1, x2])


def close(a: float, b: float, tol: float = 1e-9) -> bool:
    return abs(a - b) <= tol * max(1.0, abs(a), abs(b))


def ops_of(pm) -> set:
    return {int(o) for o in pm.problem.arrays().op}


# -- lexing and expressions ----------------------------------------------------


def test_arithmetic_and_precedence():
    assert close(parsed_eval("1 + 2*3"), 7.0)
    assert close(parsed_eval("(1 + 2)*3"), 9.0)
    assert close(parsed_eval("10 - 2 - 3"), 5.0)
    assert close(parsed_eval("100/10/2"), 5.0)
    assert close(parsed_eval("2*x1 + 3*x2", 5.0, 7.0), 31.0)


def test_unary_minus_binds_looser_than_power():
    assert close(parsed_eval("-x1**2", 3.0), -9.0)
    assert close(parsed_eval("(-x1)**2", 3.0), 9.0)
    assert close(parsed_eval("2**3**2"), 512.0)  # right-associative
    assert close(parsed_eval("x1**-1", 4.0), 0.25)


def test_functions_map_onto_dag_ops():
    assert close(parsed_eval("sqr(x1)", 3.0), 9.0)
    assert close(parsed_eval("sqrt(x1)", 9.0), 3.0)
    assert close(parsed_eval("exp(x1)", 1.0), math.exp(1.0))
    assert close(parsed_eval("log(x1)", 2.0), math.log(2.0))
    assert close(parsed_eval("abs(x1)", -3.0), 3.0)
    assert close(parsed_eval("sin(x1)", 1.0), math.sin(1.0))
    assert close(parsed_eval("cos(x1)", 1.0), math.cos(1.0))
    assert close(parsed_eval("tanh(x1)", 1.0), math.tanh(1.0))
    assert close(parsed_eval("power(x1,3)", 2.0), 8.0)
    assert close(parsed_eval("rpower(x1,3)", 2.0), 8.0)
    assert close(parsed_eval("min(x1,x2)", 3.0, 5.0), 3.0)
    assert close(parsed_eval("max(x1,x2,7)", 3.0, 5.0), 7.0)
    assert close(parsed_eval("log10(x1)", 1000.0), 3.0)
    assert close(parsed_eval("log2(x1)", 8.0), 3.0)


def test_function_ops_are_the_named_ones():
    cases = {"sqr(x1)": C.OP_SQR, "sqrt(x1)": C.OP_SQRT, "exp(x1)": C.OP_EXP, "log(x1)": C.OP_LOG,
             "abs(x1)": C.OP_ABS, "sin(x1)": C.OP_SIN, "cos(x1)": C.OP_COS, "tanh(x1)": C.OP_TANH,
             "min(x1,x2)": C.OP_MIN, "max(x1,x2)": C.OP_MAX, "x1**3": C.OP_POWN}
    for text, op in cases.items():
        assert op in ops_of(gams.parse(model_with(text))), text
    # log10 and log2 are Log times the constant 1/log(base)
    pm = gams.parse(model_with("log10(x1)"))
    a = pm.problem.arrays()
    assert C.OP_LOG in set(a.op.tolist()) and C.OP_MUL in set(a.op.tolist())
    assert 1.0 / math.log(10.0) in a.constant[a.op == C.OP_CONST].tolist()


def test_half_power_is_sqrt():
    pm = gams.parse(model_with("x1**0.5"))
    ops = ops_of(pm)
    assert C.OP_SQRT in ops and C.OP_EXP not in ops and C.OP_LOG not in ops
    assert close(parsed_eval("(sqr(x1) + sqr(x2))**0.5", 3.0, 4.0), 5.0)


def test_real_constant_exponent_is_exp_of_log():
    ops = ops_of(gams.parse(model_with("x1**1.5")))
    assert {C.OP_EXP, C.OP_LOG} <= ops and C.OP_POWN not in ops
    assert close(parsed_eval("x1**1.5", 4.0), 4.0 ** 1.5)
    assert close(parsed_eval("x1**0.2", 3.0), 3.0 ** 0.2)
    assert close(parsed_eval("x1**-0.5", 4.0), 4.0 ** -0.5)


def test_variable_exponent_is_exp_of_log():
    ops = ops_of(gams.parse(model_with("x1**x2")))
    assert {C.OP_EXP, C.OP_LOG} <= ops
    assert close(parsed_eval("x1**x2", 2.0, 3.0), 8.0)
    assert close(parsed_eval("10**(x1 + 0.5*x2)", 2.0, 3.0), 10.0 ** 3.5)


def test_integer_exponent_payload():
    a = gams.parse(model_with("x1**7")).problem.arrays()
    assert a.int_exp[a.op == C.OP_POWN].tolist() == [7]
    a = gams.parse(model_with("x1**(-2)")).problem.arrays()
    assert a.int_exp[a.op == C.OP_POWN].tolist() == [-2]


def test_case_insensitive():
    assert close(parsed_eval("POWER(X1,2) + SQR(x1)", 3.0), 18.0)


def test_column_one_star_is_a_comment_but_indented_star_multiplies():
    pm = gams.parse("Variables x1,objvar;\n"
                    "Equations e1;\n"
                    "* this whole line is a comment\n"
                    "e1.. -(2\n"
                    "     *x1) + objvar =E= 0;\n"
                    "x1.lo = 0; x1.up = 10;\n"
                    "Solve m using NLP minimizing objvar;\n")
    assert close(evaluate_objective(pm.problem.arrays(), [3.0]), 6.0)


def test_dollar_control_lines_are_skipped():
    pm = gams.parse("$offlisting\n"
                    "$ontext\n"
                    "this text is not GAMS at all ((( ;;;\n"
                    "$offtext\n"
                    "Variables x1,objvar;\n"
                    "Equations e1;\n"
                    "e1.. -(x1) + objvar =E= 0;\n"
                    "x1.lo = 0; x1.up = 1;\n"
                    "$if NOT '%gams.u1%' == '' $include '%gams.u1%'\n"
                    "Solve m using %NLP% minimizing objvar;\n")
    assert close(evaluate_objective(pm.problem.arrays(), [0.5]), 0.5)


def test_line_numbers_survive_comment_blocks():
    src = ("$ontext\nnot gams\n$offtext\n"
           "Variables x1,objvar;\n"
           "Equations e1;\n"
           "e1.. -(arctan(x1)) + objvar =E= 0;\n")
    with pytest.raises(gams.ParseError) as e:
        gams.parse(src)
    assert e.value.line == 6


def test_special_values():
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. -(x1 + eps) + objvar =E= 0;\n"
                    "x1.lo = -inf; x1.up = 3;\nSolve m using NLP minimizing objvar;\n")
    a = pm.problem.arrays()
    assert a.box[0].tolist() == [-1e6, 3.0]  # -inf became the default bound
    with pytest.raises(gams.ParseError) as e:
        gams.parse(model_with("x1 + na"))
    assert e.value.kind == gams.UNREPRESENTABLE


# -- objective elimination -------------------------------------------------------


def try_eliminate(equation: str):
    """(eliminated, objective at 3, number of variables), reference gams_test.cpp:196-217."""
    pm = gams.parse("Variables x1,objvar;\n"
                    "Equations e1;\n" + equation + "\n"
                    "x1.lo = -10; x1.up = 10;\n"
                    "objvar.lo = -1e6; objvar.up = 1e6;\n"
                    "Solve m using NLP minimizing objvar;\n")
    a = pm.problem.arrays()
    return ("objvar" not in pm.var_names, evaluate_objective(a, [3.0] * a.n_vars), a.n_vars)


def try_eliminate_ineq(equations: str, sense: str = "minimizing", declare: str = "Equations e1;\n"):
    pm = gams.parse("Variables x1,objvar;\n" + declare + equations + "\n"
                    "x1.lo = -10; x1.up = 10;\n"
                    "Solve m using NLP " + sense + " objvar;\n")
    a = pm.problem.arrays()
    return ("objvar" not in pm.var_names, evaluate_objective(a, [3.0] * a.n_vars), a.n_vars)


@pytest.mark.parametrize("equation,value", [
    ("e1.. -(sqr(x1) + 2*x1) + objvar =E= 0;", 15.0),   # canonical -(P) + objvar
    ("e1.. objvar - sqr(x1) =E= 0;", 9.0),
    ("e1.. -(sqr(x1)) + objvar =E= -45;", -36.0),        # non-zero right-hand side
    ("e1.. 2*objvar - sqr(x1) =E= 0;", 4.5),             # coefficient other than 1
    ("e1.. sqr(x1) =E= objvar;", 9.0),                   # objvar on the right
    ("e1.. objvar/4 - sqr(x1) =E= 0;", 36.0),            # divided by a constant
])
def test_objective_rearrangement_shapes(equation, value):
    eliminated, got, n = try_eliminate(equation)
    assert eliminated and n == 1
    assert close(got, value)


def test_canonical_shape_leaves_no_residue():
    # -(P) + objvar =E= 0 rearranges to exactly P: no x-0, no Neg(Neg(x)).
    a = gams.parse(model_with("sqr(x1)")).problem.arrays()
    assert a.op.tolist() == [C.OP_VAR, C.OP_VAR, C.OP_SQR]
    assert int(a.objective_root) == 2


@pytest.mark.parametrize("equation", [
    "e1.. sqr(objvar) - x1 =E= 0;",      # under a function
    "e1.. objvar*x1 - 1 =E= 0;",         # times a variable
    "e1.. x1/objvar - 1 =E= 0;",         # in a denominator
    "e1.. objvar - objvar + x1 =E= 0;",  # cancels
    "e1.. objvar - sqr(x1) =L= 0;",      # inequality pointing the wrong way
])
def test_nonlinear_objvar_falls_back(equation):
    eliminated, _, n = try_eliminate(equation)
    assert not eliminated and n == 2


def test_fallback_is_reported():
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. sqr(objvar) - x1 =E= 0;\n"
                    "x1.lo = 0; x1.up = 1; objvar.lo = 0; objvar.up = 1;\nSolve m using NLP minimizing objvar;\n")
    assert pm.objvar_kept
    assert any("keeping it as a variable" in w.message for w in pm.warnings)


def test_inequality_tight_at_optimum_eliminates():
    r = try_eliminate_ineq("e1.. sqr(x1) - objvar =L= 0;")
    assert r[0] and r[2] == 1 and close(r[1], 9.0)
    r = try_eliminate_ineq("e1.. -(sqr(x1)) + objvar =G= 0;")
    assert r[0] and close(r[1], 9.0)
    r = try_eliminate_ineq("e1.. 2*objvar - sqr(x1) =G= -8;")
    assert r[0] and close(r[1], 0.5)
    r = try_eliminate_ineq("e1.. sqr(x1) - objvar =G= 0;", "maximizing")
    assert r[0] and close(r[1], -9.0)


def test_inequality_wrong_direction_or_second_mention_falls_back():
    r = try_eliminate_ineq("e1.. sqr(x1) - objvar =L= 0;", "maximizing")
    assert not r[0] and r[2] == 2
    r = try_eliminate_ineq("e1.. sqr(x1) - objvar =L= 0;\ne2.. objvar + x1 =L= 4;", "minimizing",
                           "Equations e1,e2;\n")
    assert not r[0] and r[2] == 2


def test_equality_preferred_over_inequality():
    pm = gams.parse("Variables x1,objvar;\nEquations e1,e2;\n"
                    "e1.. -(sqr(x1)) + objvar =E= 0;\ne2.. 2*x1 - objvar =L= 0;\n"
                    "x1.lo = -10; x1.up = 10;\nSolve m using NLP minimizing objvar;\n")
    a = pm.problem.arrays()
    assert pm.var_names == ["x1"] and close(evaluate_objective(a, [3.0]), 9.0)
    # e2 stays a constraint, its objvar reading the objective's value: 2*3 - 9
    assert pm.constraint_names == ["e2"]
    assert close(evaluate_point(a, [3.0])[int(a.con_root[0])], 6.0 - 9.0)


def test_several_definitions_use_the_first():
    pm = gams.parse("Variables x1,objvar;\nEquations e1,e2;\n"
                    "e1.. objvar - sqr(x1) =E= 0;\ne2.. objvar - 2*x1 =E= 0;\n"
                    "x1.lo = -10; x1.up = 10;\nSolve m using NLP minimizing objvar;\n")
    a = pm.problem.arrays()
    assert close(evaluate_objective(a, [3.0]), 9.0)
    assert any("several equations" in w.message for w in pm.warnings)
    assert pm.constraint_names == ["e2"]


def test_bound_on_eliminated_objvar_survives():
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. -(sqr(x1)) + objvar =E= 0;\n"
                    "x1.lo = -10; x1.up = 10;\nobjvar.up = 25;\nSolve m using NLP minimizing objvar;\n")
    a = pm.problem.arrays()
    assert pm.var_names == ["x1"]
    assert a.n_constraints == 1 and int(a.con_cmp[0]) == C.LE and a.con_rhs[0] == 25.0
    assert pm.constraint_names == ["objvar.up"]


def test_surviving_bound_in_the_files_own_sense():
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. -(sqr(x1)) + objvar =E= 0;\n"
                    "x1.lo = -10; x1.up = 10;\nobjvar.up = 25;\nSolve m using NLP maximizing objvar;\n")
    a = pm.problem.arrays()
    assert a.n_constraints == 1 and int(a.con_cmp[0]) == C.LE and a.con_rhs[0] == 25.0
    assert close(evaluate_point(a, [3.0])[int(a.con_root[0])], 9.0)  # f, not -f


def bounded(bounds: str, sense: str):
    return gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. sqr(x1) - objvar =L= 0;\n"
                      "x1.lo = -10; x1.up = 10;\n" + bounds + "Solve m using NLP " + sense + " objvar;\n")


def test_inequality_elimination_keeps_only_the_binding_bound():
    a = bounded("objvar.up = 25;\n", "minimizing").problem.arrays()
    assert a.n_constraints == 1 and int(a.con_cmp[0]) == C.LE and a.con_rhs[0] == 25.0
    pm = bounded("objvar.lo = 5;\n", "minimizing")
    assert pm.problem.arrays().n_constraints == 0
    assert any("objvar.lo" in w.message for w in pm.warnings)
    # maximising mirrors it: objvar.up is dropped
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. sqr(x1) - objvar =G= 0;\n"
                    "x1.lo = -10; x1.up = 10;\nobjvar.lo = 1; objvar.up = 50;\nSolve m using NLP maximizing objvar;\n")
    assert pm.constraint_names == ["objvar.lo"]
    assert any("objvar.up" in w.message for w in pm.warnings)
    # the =E= path keeps both
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. -(sqr(x1)) + objvar =E= 0;\n"
                    "x1.lo = -10; x1.up = 10;\nobjvar.lo = 5; objvar.up = 25;\nSolve m using NLP minimizing objvar;\n")
    a = pm.problem.arrays()
    assert a.n_constraints == 2 and pm.constraint_names == ["objvar.up", "objvar.lo"]
    # objvar.lo = 5 is -f(x) <= -5
    assert close(evaluate_point(a, [3.0])[int(a.con_root[1])], -9.0) and a.con_rhs[1] == -5.0


def test_maximise_negates_and_reports_sense():
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. -(sqr(x1)) + objvar =E= 0;\n"
                    "x1.lo = -10; x1.up = 10;\nSolve m using NLP maximizing objvar;\n")
    assert pm.sense == gams.MAXIMISE
    assert close(evaluate_objective(pm.problem.arrays(), [3.0]), -9.0)


# -- agreement with closed forms of the MINLPLib models ------------------------

EX4_1_2_COEFFS = [
    1.666666666, 1.25, 1.0, 0.8333333, 0.714285714, 0.625, 0.555555555, 1.0, -43.6363636, 0.41666666,
    0.384615384, 0.357142857, 0.3333333, 0.3125, 0.294117647, 0.277777777, 0.263157894, 0.25, 0.238095238,
    0.227272727, 0.217391304, 0.208333333, 0.2, 0.192307692, 0.185185185, 0.178571428, 0.344827586, 0.6666666,
    -15.48387097, 0.15625, 0.1515151, 0.14705882, 0.14285712, 0.138888888, 0.135135135, 0.131578947,
    0.128205128, 0.125, 0.121951219, 0.119047619, 0.116279069, 0.113636363, 0.1111111, 0.108695652,
    0.106382978, 0.208333333, 0.408163265, 0.8]


def test_ex4_1_2_agrees_with_its_closed_form():
    # 2.5 x^2 - 500 x + sum_i c_i x^(i+3) on [1, 2] (the reference's hand-built
    # ex4_1_2, gams_test.cpp:529-561)
    pm = gams.parse_file(data_file("ex4_1_2.gms"))
    a = pm.problem.arrays()
    assert pm.var_names == ["x1"] and a.n_constraints == 0
    assert a.box[0].tolist() == [1.0, 2.0]
    rng = np.random.default_rng(12345)
    for x in rng.uniform(1.0, 2.0, 64):
        want = 2.5 * x * x - 500.0 * x + sum(c * x ** (i + 3) for i, c in enumerate(EX4_1_2_COEFFS))
        assert close(evaluate_objective(a, [x]), want, 1e-12)


def test_ex8_6_2_agrees_with_its_closed_form():
    # -45 + sum_{i<j} (1 - exp(3 (1 - r_ij)))^2 over 10 points in R^3 with six
    # coordinates fixed at 0 (the reference's hand-built ex8_6_2, gams_test.cpp:563-598)
    pm = gams.parse_file(data_file("ex8_6_2.gms"))
    a = pm.problem.arrays()
    assert len(pm.var_names) == 24 and pm.var_names[0] == "x2" and pm.var_names[-1] == "x30"
    assert a.n_constraints == 0
    fixed = {0, 10, 11, 20, 21, 22}
    free = [i for i in range(30) if i not in fixed]
    assert [f"x{i + 1}" for i in free] == pm.var_names
    rng = np.random.default_rng(12345)
    for _ in range(64):
        y = rng.uniform(-5.0, 5.0, 24)
        x = np.zeros(30)
        x[free] = y
        want = 0.0
        for i in range(10):
            for j in range(i + 1, 10):
                r = math.sqrt((x[i] - x[j]) ** 2 + (x[i + 10] - x[j + 10]) ** 2 + (x[i + 20] - x[j + 20]) ** 2)
                want += (1.0 - math.exp(3.0 * (1.0 - r))) ** 2
        assert close(evaluate_objective(a, y), -45.0 + want, 1e-12)


def test_fixed_variables_become_constants():
    pm = gams.parse_file(data_file("ex8_6_2.gms"))
    assert all(lo < up for lo, up in pm.problem.arrays().box)
    kept = gams.parse_file(data_file("ex8_6_2.gms"), gams.ParseOptions(fold_fixed_to_const=False))
    assert len(kept.var_names) == 30
    box = kept.problem.arrays().box
    assert sum(1 for lo, up in box if lo == up == 0.0) == 6


def test_nvs04_agrees_with_its_closed_form():
    pm = gams.parse_file(data_file("nvs04.gms"))
    a = pm.problem.arrays()
    assert a.var_kind.tolist() == [C.INTEGER, C.INTEGER] and a.n_constraints == 0
    rng = np.random.default_rng(20260730)
    for i1, i2 in rng.integers(0, 201, (32, 2)):
        want = 100.0 * (0.5 - (0.6 + i1) ** 2 + i2) ** 2 + (0.4 - i1) ** 2
        assert close(evaluate_objective(a, [float(i1), float(i2)]), want)


def test_autocorr_bern20_03_agrees_with_its_closed_form():
    # sum_i 8 b_i b_{i+2} - sum_j c_j b_j, c_j = 4 at both ends and 8 inside;
    # its objvar is defined by an =L= and is eliminated through it.
    pm = gams.parse_file(data_file("autocorr_bern20-03.gms"))
    a = pm.problem.arrays()
    assert a.var_kind.tolist() == [C.BINARY] * 20 and a.n_constraints == 0
    c = np.full(20, 8.0)
    c[[0, 1, 18, 19]] = 4.0
    rng = np.random.default_rng(20260730)
    for b in rng.integers(0, 2, (32, 20)).astype(float):
        want = sum(8.0 * b[i] * b[i + 2] for i in range(18)) - float(c @ b)
        assert close(evaluate_objective(a, b), want)


def test_integer_variables_parse_as_integer():
    pm = gams.parse_file(data_file("nvs01.gms"))
    a = pm.problem.arrays()
    assert a.var_kind.tolist() == [C.INTEGER, C.INTEGER, C.CONTINUOUS]
    assert a.box[0].tolist() == [0.0, 200.0] and a.box[1].tolist() == [0.0, 200.0]


def test_discrete_bounds_snap_inward_and_empty_integer_boxes_fail():
    pm = gams.parse("Variables i1,objvar;\nInteger Variables i1;\nEquations e1;\ne1.. -(i1) + objvar =E= 0;\n"
                    "i1.lo = 2.3; i1.up = 7.8;\nSolve m using MINLP minimizing objvar;\n")
    assert pm.problem.arrays().box[0].tolist() == [3.0, 7.0]
    with pytest.raises(gams.ParseError):
        gams.parse("Variables i1,objvar;\nInteger Variables i1;\nEquations e1;\ne1.. -(i1) + objvar =E= 0;\n"
                   "i1.lo = 0.3; i1.up = 0.7;\nSolve m using MINLP minimizing objvar;\n")


def test_narrowed_binary_becomes_integer_and_integrality_is_sticky():
    pm = gams.parse("Variables b1,objvar;\nBinary Variables b1;\nEquations e1;\ne1.. -(b1) + objvar =E= 0;\n"
                    "b1.up = 0;\nSolve m using MINLP minimizing objvar;\n")
    a = pm.problem.arrays()
    assert a.var_kind.tolist() == [C.INTEGER] and a.box[0].tolist() == [0.0, 0.0]
    pm = gams.parse("Integer Variables i1;\nPositive Variables i1;\nVariables objvar;\nEquations e1;\n"
                    "e1.. -(i1) + objvar =E= 0;\ni1.up = 10;\nSolve m using MINLP minimizing objvar;\n")
    assert pm.problem.arrays().var_kind.tolist() == [C.INTEGER]


def test_reject_discrete():
    src = ("Variables i1,objvar;\nInteger Variables i1;\nEquations e1;\ne1.. -(i1) + objvar =E= 0;\n"
           "i1.up = 5;\nSolve m using MINLP minimizing objvar;\n")
    with pytest.raises(gams.ParseError) as e:
        gams.parse(src, gams.ParseOptions(reject_discrete=True))
    assert e.value.kind == gams.DISCRETE and e.value.line == 2


@pytest.mark.parametrize("name", sorted(CORPUS["models"]))
def test_every_model_parses_with_the_reference_counts(name):
    pm = gams.parse_file(data_file(name + ".gms"))
    a = pm.problem.arrays()
    row = CORPUS["table"].get(name)
    if row is None:
        pytest.skip(f"{name} has no row in the reference's corpus table")
    assert ("max" if pm.sense == gams.MAXIMISE else "min") == row["sense"]
    assert a.n_vars == row["vars"]
    assert a.n_constraints == row["constraints"]
    assert len(a.op) == row["nodes"]


def test_initial_levels():
    pm = gams.parse("Variables x1,x2,objvar;\nEquations e1;\ne1.. -(x1 + x2) + objvar =E= 0;\n"
                    "x1.lo = 1; x1.up = 4; x2.lo = -3; x2.up = -1;\nx1.l = 9;\n"
                    "Solve m using NLP minimizing objvar;\n")
    # a level is clamped into the box; a variable without one starts at clamp(0)
    assert pm.initial_point.tolist() == [4.0, -1.0]
    pm = gams.parse(model_with("x1"))
    assert pm.initial_point.size == 0


# -- diagnostics -------------------------------------------------------------------


@pytest.mark.parametrize("source,line,kind", [
    (model_with("arctan(x1)"), 3, gams.UNSUPPORTED_FUNCTION),
    (model_with("wibble(x1)"), 3, gams.UNSUPPORTED_FUNCTION),
    ("Set i /1*10/;\n", 1, gams.NON_SCALAR),
    (model_with("zzz"), 3, gams.SYNTAX),
    ("Variables x1;\n", 0, gams.UNREPRESENTABLE),
    ("Variables x1 'a label';\n", 1, gams.NON_SCALAR),
    ("SOS1 Variables s1;\n", 1, gams.DISCRETE),
    ("Variables x1;\nEquations e1;\ne1.. x1 =E= 1;\nSolve m using CNS;\n", 4, gams.UNREPRESENTABLE),
    ("Variables x1,objvar;\nx1.lo = x1;\n", 2, gams.SYNTAX),
    ("Variables x1,objvar;\nEquations e1;\ne1.. x1 + objvar 3;\n", 3, gams.SYNTAX),
    ("Variables x1 ? ;\n", 1, gams.SYNTAX),
    ("Variables x1,objvar;\nEquations e1;\ne1.. x1 =X= objvar;\n", 3, gams.SYNTAX),
    (model_with("power(x1)"), 3, gams.SYNTAX),
])
def test_rejections_carry_line_and_kind(source, line, kind):
    with pytest.raises(gams.ParseError) as e:
        gams.parse(source)
    assert e.value.line == line, str(e.value)
    assert e.value.kind == kind, str(e.value)
    assert str(e.value).startswith(f"line {line}: ")


def test_constant_objective_is_unrepresentable():
    with pytest.raises(gams.ParseError) as e:
        gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. objvar =E= 3;\nx1.lo=0; x1.up=1;\n"
                   "Solve m using NLP minimizing objvar;\n")
    assert e.value.kind == gams.UNREPRESENTABLE


def test_missing_file_is_io():
    with pytest.raises(gams.ParseError) as e:
        gams.parse_file(GOLDEN / "gams" / "no_such_model.gms")
    assert e.value.kind == gams.IO


def test_free_variables_get_a_default_box_and_say_so():
    pm = gams.parse("Variables x1,objvar;\nEquations e1;\ne1.. -(x1) + objvar =E= 0;\n"
                    "Solve m using NLP minimizing objvar;\n", gams.ParseOptions(default_bound=500.0))
    assert pm.problem.arrays().box[0].tolist() == [-500.0, 500.0]
    assert len(pm.warnings) == 2


def test_unbounded_integer_warning_names_the_domain_size():
    pm = gams.parse("Variables i1,objvar;\nInteger Variables i1;\nEquations e1;\ne1.. -(i1) + objvar =E= 0;\n"
                    "Solve m using MINLP minimizing objvar;\n", gams.ParseOptions(default_bound=10.0))
    assert pm.problem.arrays().box[0].tolist() == [0.0, 10.0]
    assert any("integer domain size 11" in w.message for w in pm.warnings)


def test_relations_canonicalise_into_le_and_eq():
    pm = gams.parse("Variables x1,objvar;\nEquations e1,e2,e3;\ne1.. -(x1) + objvar =E= 0;\n"
                    "e2.. 2*x1 =L= 8;\ne3.. 3*x1 =G= 6;\nx1.lo = 0; x1.up = 10;\n"
                    "Solve m using NLP minimizing objvar;\n")
    a = pm.problem.arrays()
    assert a.n_constraints == 2
    assert int(a.con_cmp[0]) == C.LE and a.con_rhs[0] == 8.0
    # 3*x1 >= 6 is stored as 6 <= 3*x1, i.e. 6 - 3*x1 <= 0
    assert int(a.con_cmp[1]) == C.LE and a.con_rhs[1] == 0.0
    assert close(evaluate_point(a, [5.0])[int(a.con_root[1])], 6.0 - 15.0)


def test_constant_and_n_equations_are_dropped_with_warnings():
    pm = gams.parse("Variables x1,objvar;\nEquations e1,e2,e3,e4;\ne1.. -(x1) + objvar =E= 0;\n"
                    "e2.. 1 =L= 2;\ne3.. 3 =E= 4;\ne4.. x1 =N= 0;\nx1.lo = 0; x1.up = 1;\n"
                    "Solve m using NLP minimizing objvar;\n")
    msgs = [w.message for w in pm.warnings]
    assert pm.problem.arrays().n_constraints == 0
    assert any("trivially satisfied" in m for m in msgs)
    assert any("VIOLATED" in m for m in msgs)
    assert any("=N=" in m for m in msgs)


def test_equality_constraint_and_names():
    pm = gams.parse("Variables x1,x2,objvar;\nEquations obj,c1;\nobj.. objvar =E= x1 + x2;\n"
                    "c1.. x1*x1 =E= x2;\nx1.lo=0; x1.up=2; x2.lo=0; x2.up=4;\nSolve m using NLP maximizing objvar;\n")
    a = pm.problem.arrays()
    assert pm.constraint_names == ["c1"] and int(a.con_cmp[0]) == C.EQ
    assert C.OP_SQR in set(a.op.tolist())  # x*x of one variable is Sqr
    assert pm.sense == gams.MAXIMISE


# -- the RESULT line ---------------------------------------------------------------


def test_result_line():
    big = np.finfo(np.float64).max
    # 17 significant digits, as C's %.17g (Python's own formatter is the check)
    lo, up = -17.000000001, -16.9999986035
    assert gams.result_line(gams.MINIMISE, lo, up) == f"RESULT\tsense=min\tprimal={up:.17g}\tdual={lo:.17g}"
    # maximise: both bounds are negated back into the file's sense
    assert gams.result_line(gams.MAXIMISE, -2.5, -2.0) == "RESULT\tsense=max\tprimal=2\tdual=2.5"
    assert gams.result_line(gams.MINIMISE, -big, big) == "RESULT\tsense=min\tprimal=none\tdual=none"
    assert gams.result_line(gams.MINIMISE, -math.inf, math.inf) == "RESULT\tsense=min\tprimal=none\tdual=none"
    assert gams.result_line(gams.MINIMISE, 0.1, 0.1 + 0.2) == \
        "RESULT\tsense=min\tprimal=0.30000000000000004\tdual=0.10000000000000001"
    for v in (1.0 / 3.0, -1e-300, 6.02214076e23, 12.46966882):
        line = gams.result_line(gams.MINIMISE, v, v)
        assert float(line.split("primal=")[1].split("\t")[0]) == v  # 17 digits round-trip


def test_gams_solve_cli_reports_parse_errors(tmp_path, capsys):
    from cuminlpsearch_amd import gams_solve
    bad = tmp_path / "bad.gms"
    bad.write_text("Set i /1*3/;\n")
    assert gams_solve.main([str(bad), "10"]) == 1
    assert "parse error: line 1:" in capsys.readouterr().err
