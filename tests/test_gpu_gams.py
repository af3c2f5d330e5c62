"""MINLPLib models read from their .gms files and solved on the device.

Each model the device op set covers is solved on the device and by the
oracle's restated search with the same settings: the iterates (counts,
bracket, incumbent) must agree bit for bit, and the bracket must be consistent
with MINLPLib's published bounds (tests/golden/gams_corpus.json). The small
ones are solved to the published optimum; the whole set also runs as one
batch of uneven instances. Reference: source/gams/solve.cu:82-122 (parse,
solve), :836-867 (RESULT line)."""
from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import gams
from oracle import binding as orc

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
TABLE = json.loads((GOLDEN / "gams_corpus.json").read_text())["table"]

# models whose DAGs use only Add/Sub/Mul/Sqr/Neg
QCQP_MODELS = ["ex2_1_1", "st_e01", "nvs04", "autocorr_bern20-03", "alkyl", "house", "ex3_1_1", "ex5_2_2_case1",
               "st_bpaf1a", "ex14_1_2", "circle"]
# models that need the extended op set (division, powers, sqrt, exp, log, sin, cos)
EXT_MODELS = ["ex8_1_1", "ex4_1_2", "ex4_1_5", "ex14_1_1", "nvs01", "hs62", "chem", "ex6_1_1", "ex7_2_1", "ex8_6_2"]
OPTS = dict(partition_num=4, enumerate_cap=4, max_cycle_size=2, sample_points=32, iter_limit=300)

FIELDS = ("iterations", "fathomed", "bounded", "sampled", "enqueued", "pruned_infeasible", "pruned_dominated",
          "pending", "stop_reason", "proven_optimal", "infeasible", "found_incumbent")


def same_outcome(d, o):
    for f in FIELDS:
        assert getattr(d, f) == getattr(o, f), f
    assert np.float64(d.lower_bound).tobytes() == np.float64(o.lower_bound).tobytes()
    assert np.float64(d.upper_bound).tobytes() == np.float64(o.upper_bound).tobytes()


def parsed(name):
    return gams.parse_file(GOLDEN / "gams" / f"{name}.gms")


def consistent_with_published(name, out, tol=1e-6):
    """lower <= optimum <= upper, with the optimum bracketed by MINLPLib's
    published primal/dual bounds (all these models minimise)."""
    row = TABLE.get(name)
    if row is None:
        return
    if row["ref_primal"] is not None:
        assert out.lower_bound <= row["ref_primal"] + tol * max(1.0, abs(row["ref_primal"]))
    if out.found_incumbent and row["ref_dual"] is not None:
        assert out.upper_bound >= row["ref_dual"] - tol * max(1.0, abs(row["ref_dual"]))


@pytest.mark.parametrize("name", QCQP_MODELS + EXT_MODELS)
def test_solve_matches_oracle(ctx, name):
    pm = parsed(name)
    print(f"solving {name}", flush=True)
    d, best_d = cmi.solve(ctx, pm.problem, **OPTS)
    o, best_o = orc.solve(pm.problem.arrays(), **OPTS)
    same_outcome(d, o)
    assert np.array_equal(best_d.view(np.uint64), best_o.view(np.uint64))
    consistent_with_published(name, d)


@pytest.mark.parametrize("name,opts,optimum", [
    ("nvs04", OPTS, 0.72),
    ("st_e01", OPTS, -6.666666667),
    ("circle", dict(partition_num=4, enumerate_cap=4, max_cycle_size=3, sample_points=32, iter_limit=2000), 4.57424899),
    # all 20 binaries enumerated in one launch of 2^20 regions
    ("autocorr_bern20-03", dict(partition_num=2, enumerate_cap=2, max_cycle_size=20, sample_points=1, iter_limit=10),
     -72.0),
    # sin, cos and division
    ("ex8_1_1", dict(partition_num=4, enumerate_cap=4, max_cycle_size=2, sample_points=32, iter_limit=3000),
     -2.021806783),
    # a degree-50 polynomial in PowN terms
    ("ex4_1_2", dict(partition_num=8, enumerate_cap=8, max_cycle_size=1, sample_points=32, iter_limit=2000),
     -663.5000966),
    # integers, sqrt and division
    ("nvs01", OPTS, 12.46966882),
])
def test_solved_to_the_published_optimum(ctx, name, opts, optimum):
    pm = parsed(name)
    print(f"solving {name}", flush=True)
    d, best = cmi.solve(ctx, pm.problem, **opts)
    assert d.proven_optimal or d.stop_reason == 0
    assert abs(d.upper_bound - optimum) <= 1e-6 * max(1.0, abs(optimum))
    assert d.lower_bound <= d.upper_bound
    # the incumbent is a point of the model with that objective value
    from cuminlpsearch_amd.model import evaluate_objective
    assert abs(evaluate_objective(pm.problem.arrays(), best) - d.upper_bound) <= 1e-9 * max(1.0, abs(optimum))
    line = gams.result_line(pm.sense, d.lower_bound, d.upper_bound)
    assert line.startswith("RESULT\tsense=min\tprimal=")
    assert float(line.split("primal=")[1].split("\t")[0]) == d.upper_bound


def test_ex2_1_1_brackets_its_optimum(ctx):
    opts = dict(partition_num=2, enumerate_cap=2, max_cycle_size=5, sample_points=64, iter_limit=4000)
    pm = parsed("ex2_1_1")
    d, _ = cmi.solve(ctx, pm.problem, **opts)
    o, _ = orc.solve(pm.problem.arrays(), **opts)
    same_outcome(d, o)
    assert d.lower_bound <= -17.0 <= d.upper_bound
    assert d.upper_bound - d.lower_bound < 0.5
    # to gap 1e-6 with the k-child search, on the published optimum -17
    a, _, _ = cmi.aggregate_solve(ctx, pm.problem, partition_budget=1 << 16, sample_fan_out=1 << 10, sample_points=4,
                                  iter_limit=3000)
    assert a.proven_optimal
    assert abs(a.upper_bound - (-17.0)) <= 1e-6 * 17.0 and a.lower_bound <= a.upper_bound


def test_corpus_batch_equals_plain(ctx):
    """The whole set as one batch of uneven instances (1 to 24 variables, 0 to
    14 constraints, with and without the extended op set in the same launch)."""
    names = QCQP_MODELS + EXT_MODELS
    models = [parsed(n).problem for n in names]
    outs, bests = cmi.solve_batch(ctx, models, **OPTS)
    for name, p, d, b in zip(names, models, outs, bests):
        s, sb = cmi.solve(ctx, p, **OPTS)
        same_outcome(d, s)
        assert np.array_equal(b.view(np.uint64), sb.view(np.uint64)), name


def test_maximise_result_line_reports_the_files_sense(ctx):
    src = ("Variables x1,x2,objvar;\nEquations e1,e2;\n"
           "e1.. objvar =E= x1*x2 + x1;\ne2.. x1 + x2 =L= 3;\n"
           "x1.lo = 0; x1.up = 2; x2.lo = 0; x2.up = 2;\nSolve m using NLP maximizing objvar;\n")
    pm = gams.parse(src)
    d, best = cmi.solve(ctx, pm.problem, partition_num=4, enumerate_cap=4, max_cycle_size=2, sample_points=64,
                        iter_limit=3000, tolerance=1e-7)
    # max x1*x2 + x1 on x1 + x2 <= 3 is 4 at (2, 1)
    assert -d.upper_bound <= 4.0 + 1e-9 and -d.lower_bound >= 4.0 - 1e-9
    line = gams.result_line(pm.sense, d.lower_bound, d.upper_bound)
    primal = float(line.split("primal=")[1].split("\t")[0])
    dual = float(line.split("dual=")[1])
    assert line.startswith("RESULT\tsense=max\t")
    assert primal <= 4.0 + 1e-9 <= dual + 2e-9
    assert abs(primal - 4.0) < 1e-3


def test_gams_solve_cli(capsys):
    from cuminlpsearch_amd import gams_solve
    rc = gams_solve.main([str(GOLDEN / "gams" / "st_e01.gms"), "300", "--partition-num", "4", "--enumerate-cap", "4",
                          "--max-slots", "2", "--sample-points", "32"])
    out = capsys.readouterr().out
    assert rc == 0
    line = [ln for ln in out.splitlines() if ln.startswith("RESULT\t")][0]
    assert line.startswith("RESULT\tsense=min\tprimal=")
    assert abs(float(line.split("primal=")[1].split("\t")[0]) + 6.666666667) < 1e-5  # relative gap 1e-6
    assert "  x1 = " in out and "  x2 = " in out


def test_autocorr_all_binary_run_record(ctx):
    """The reference's run record for autocorr_bern20-03 under --policy=all-binary
    (MINLP_STATUS.md:217): converged at iteration 1 with primal = dual = -72.
    The all-binary row pins its shape whatever the device memory (partition 2,
    enumerate cap 2, 1 sample, 20 slots: catalogue.hpp:92-99), so the record
    is reproducible: the 20 binaries are enumerated in the first iteration."""
    from cuminlpsearch_amd import resolve as R
    pm = parsed("autocorr_bern20-03")
    shape = R.resolve_shape(pm.problem, "all-binary", free_device_bytes=8 << 30)
    opts = shape.solve_options()
    assert opts == dict(partition_num=2, enumerate_cap=2, sample_points=1, max_cycle_size=20)
    d, _ = cmi.solve(ctx, pm.problem, iter_limit=1000, **opts)
    o, _ = orc.solve(pm.problem.arrays(), iter_limit=1000, **opts)
    same_outcome(d, o)
    assert d.iterations == 1 and d.proven_optimal
    assert d.upper_bound == -72.0 and d.lower_bound == -72.0


def test_ex4_1_2_mixed_all_small_run_record(ctx):
    """MINLP_STATUS.md:543: --policy=mixed-all-small converged ex4_1_2 with
    primal = dual = -663.50009661 at record 409. The policy fits one
    variable's partition up to a ceiling that, when the record was made, sat
    at 64 (include/cuminlp/config/catalogue.hpp:74-85 tells of it being raised
    to 2,000,000); this one-variable model reaches the ceiling on any device,
    so the record's shape is partition 64, enumerate cap 1, 10 samples, one
    slot. Under that shape the search converges at iteration 410 (the record
    plus the dequeue that finds the gap closed) on the device and in the
    oracle alike. Under today's ceiling it reaches the same bracket."""
    from cuminlpsearch_amd import resolve as R
    pm = parsed("ex4_1_2")
    shape = R.resolve_shape(pm.problem, "mixed-all-small", free_device_bytes=8 << 30)
    opts = shape.solve_options()
    assert opts == dict(partition_num=2_000_000, enumerate_cap=1, sample_points=10, max_cycle_size=1)
    record = dict(opts, partition_num=64)
    d, best_d = cmi.solve(ctx, pm.problem, iter_limit=1000, **record)
    o, best_o = orc.solve(pm.problem.arrays(), iter_limit=1000, **record)
    same_outcome(d, o)
    assert np.array_equal(best_d.view(np.uint64), best_o.view(np.uint64))
    assert d.proven_optimal and d.iterations == 410
    assert f"{d.upper_bound:.11g}" == "-663.50009661" and f"{d.lower_bound:.11g}" == "-663.50009661"
    now, _ = cmi.solve(ctx, pm.problem, iter_limit=1000, **opts)
    print(f"partition 2,000,000: iterations {now.iterations} primal {now.upper_bound!r}", flush=True)
    assert now.proven_optimal
    assert f"{now.upper_bound:.11g}" == "-663.50009661" and f"{now.lower_bound:.11g}" == "-663.50009661"


def test_nvs04_discrete_run_record(ctx):
    """MINLP_STATUS.md:1021: --policy=discrete converged nvs04
    with primal = dual = 0.72 at record 6. Today the policy's
    coverage fit gives a partition of about 20,000 on a large device, which
    does not converge within 1,000 iterations; when the record was made the
    fit's ceiling was 64 (include/cuminlp/config/catalogue.hpp:74-85), and
    under that shape (enumerate cap 16, 5 samples, 2 slots) the frontier
    empties after iteration 6 on the device and in the oracle alike."""
    from cuminlpsearch_amd import resolve as R
    pm = parsed("nvs04")
    shape = R.resolve_shape(pm.problem, "discrete", free_device_bytes=8 << 30)
    opts = shape.solve_options()
    assert (opts["enumerate_cap"], opts["sample_points"], opts["max_cycle_size"]) == (16, 5, 2)
    assert opts["partition_num"] > 64
    record = dict(opts, partition_num=64)
    d, best_d = cmi.solve(ctx, pm.problem, iter_limit=1000, **record)
    o, best_o = orc.solve(pm.problem.arrays(), iter_limit=1000, **record)
    same_outcome(d, o)
    assert np.array_equal(best_d.view(np.uint64), best_o.view(np.uint64))
    assert d.proven_optimal and d.iterations == 6
    assert f"{d.upper_bound:.12g}" == "0.72" and f"{d.lower_bound:.12g}" == "0.72"


@pytest.mark.parametrize("name", ["ex2_1_1", "nvs04", "nvs01", "autocorr_bern20-03", "st_e01", "ex8_1_1", "circle"])
def test_width_first_solve_matches_oracle(ctx, name):
    """The WidthFirst policy (policy/width_first.hpp) on the device and in the
    oracle: equal iterates, counters and incumbents."""
    from cuminlpsearch_amd import capi as C
    pm = parsed(name)
    opts = dict(policy=C.POLICY_WIDTH_FIRST, partition_num=8, enumerate_cap=4, max_cycle_size=2, sample_points=16,
                iter_limit=200)
    print(f"solving {name} width-first", flush=True)
    d, best_d = cmi.solve(ctx, pm.problem, **opts)
    o, best_o = orc.solve(pm.problem.arrays(), **opts)
    same_outcome(d, o)
    assert np.array_equal(best_d.view(np.uint64), best_o.view(np.uint64))


def test_width_first_roster_row_on_the_device(ctx):
    """--policy width-first: one slot per iteration at the fitted partition."""
    from cuminlpsearch_amd import resolve as R
    pm = parsed("circle")
    shape = R.resolve_shape(pm.problem, "width-first", free_device_bytes=8 << 30)
    opts = shape.solve_options()
    assert opts["max_cycle_size"] == 1 and opts["partition_num"] > 2
    d, _ = cmi.solve(ctx, pm.problem, iter_limit=5, **opts)
    assert d.lower_bound <= 4.57424899 + 1e-6
    consistent_with_published("circle", d)
    small = dict(opts, partition_num=64)
    d, _ = cmi.solve(ctx, pm.problem, iter_limit=100, **small)
    o, _ = orc.solve(pm.problem.arrays(), iter_limit=100, **small)
    same_outcome(d, o)
