"""Resolved shapes on the device: MINLPLib models solved at the shape the
memory-fit resolver picks for a bounded budget, bit-identical to the oracle's
restated search at that shape and without a role build over the budget; the
CLI's automatic path; and continuous-batching admission priced with the same
footprint (reference source/gams/solve.cu:640-800, config/resolve.hpp:95-218)."""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd import gams
from cuminlpsearch_amd import resolve as R
from oracle import binding as orc

import problems as P

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
FIELDS = ("iterations", "fathomed", "bounded", "sampled", "enqueued", "pruned_infeasible", "pruned_dominated",
          "pending", "stop_reason", "proven_optimal", "infeasible", "found_incumbent")
MIB = 1 << 20


def same_outcome(d, o):
    for f in FIELDS:
        assert getattr(d, f) == getattr(o, f), f
    assert np.float64(d.lower_bound).tobytes() == np.float64(o.lower_bound).tobytes()
    assert np.float64(d.upper_bound).tobytes() == np.float64(o.upper_bound).tobytes()


def test_device_memory_is_reported(ctx):
    free, total = R.device_memory(ctx)
    assert 0 < free <= total


# (model, free bytes to fit against, rules the resolver must select)
CASES = [
    ("st_e01", 1 * MIB, "mixed-all-small"),
    ("nvs04", 1 * MIB, "discrete"),
    ("nvs01", 1 * MIB, "mixed-all-small"),
    ("ex8_1_1", 1 * MIB, "mixed-all-small"),
    ("autocorr_bern20-03", 2 * MIB, "all-binary"),
    ("house", 1 * MIB, "mixed-all-small"),
]


@pytest.mark.parametrize("name,free,rules", CASES)
def test_resolved_shape_solves_like_the_oracle_within_its_budget(ctx, name, free, rules):
    pm = gams.parse_file(GOLDEN / "gams" / f"{name}.gms")
    shape = R.resolve_shape(pm.problem, objvar_kept=pm.objvar_kept, free_device_bytes=free)
    print(f"{name}: {shape.params_line()}", flush=True)
    assert shape.rules == rules and shape.source == "auto"
    assert shape.footprint_bytes <= shape.budget_bytes
    opts = dict(shape.solve_options(), iter_limit=25)
    # every role build of the search fits the fitted budget
    d, best_d = cmi.solve(ctx, pm.problem, budget_bytes=shape.budget_bytes, **opts)
    assert C.STOP_NAMES[d.stop_reason] != "device-memory"
    o, best_o = orc.solve(pm.problem.arrays(), **opts)
    same_outcome(d, o)
    assert np.array_equal(best_d.view(np.uint64), best_o.view(np.uint64))


def test_cli_resolves_the_shape_itself(capsys):
    from cuminlpsearch_amd import gams_solve
    rc = gams_solve.main([str(GOLDEN / "gams" / "st_e01.gms"), "300", "--device-budget", str(4 * MIB)])
    out = capsys.readouterr().out
    assert rc == 0
    params = [ln for ln in out.splitlines() if ln.startswith("PARAMS\t")][0]
    assert "\tpolicy=mixed-all-small\tsource=auto\t" in params
    line = [ln for ln in out.splitlines() if ln.startswith("RESULT\t")][0]
    assert abs(float(line.split("primal=")[1].split("\t")[0]) + 6.666666667) < 1e-5


def test_cli_named_policy_and_overrides(capsys):
    from cuminlpsearch_amd import gams_solve
    rc = gams_solve.main([str(GOLDEN / "gams" / "nvs04.gms"), "300", "--policy", "mixed-all-small",
                          "--device-budget", str(4 * MIB), "--sample-points", "16"])
    out = capsys.readouterr().out
    assert rc == 0
    params = [ln for ln in out.splitlines() if ln.startswith("PARAMS\t")][0]
    assert "\tpolicy=mixed-all-small\tsource=overridden\t" in params and "\tsample_points=16\t" in params
    line = [ln for ln in out.splitlines() if ln.startswith("RESULT\t")][0]
    assert abs(float(line.split("primal=")[1].split("\t")[0]) - 0.72) < 1e-5


def test_cli_inapplicable_policy_is_a_configuration_error(capsys):
    from cuminlpsearch_amd import gams_solve
    rc = gams_solve.main([str(GOLDEN / "gams" / "st_e01.gms"), "10", "--policy", "all-binary"])
    assert rc == 2
    assert "do not apply" in capsys.readouterr().err


OPTS = dict(partition_num=4, enumerate_cap=4, max_cycle_size=2, sample_points=32, iter_limit=200)


def test_admission_holds_the_resident_footprint_to_the_budget(ctx):
    probs = [P.c1(0)] * 6  # equal admission charges
    arrival = np.zeros(len(probs))
    # one request alone: its admission charge
    _, _, _, one = cmi.serve_stream(ctx, probs[:1], arrival[:1], max_active=8, **OPTS)
    each = one.peak_resident_bytes
    pp = R.profile(probs[0])
    fp = R.worst_footprint(probs[0], min(OPTS["max_cycle_size"], pp.num_binary + pp.num_integer + pp.num_continuous),
                           OPTS["partition_num"], OPTS["enumerate_cap"], OPTS["sample_points"])
    assert each > fp  # the footprint plus the tape
    plain = [cmi.solve(ctx, p, **OPTS)[0] for p in probs]
    for budget, most in ((0, 6), (int(1.5 * each), 1), (int(2.5 * each), 2)):
        outs, _, _, st = cmi.serve_stream(ctx, probs, arrival, max_active=8, hbm_budget_bytes=budget, **OPTS)
        print(f"budget {budget}: peak_active {st.peak_active}, peak bytes {st.peak_resident_bytes}", flush=True)
        assert st.peak_active == most
        if budget:
            assert st.peak_resident_bytes <= budget
        for d, s in zip(outs, plain):
            same_outcome(d, s)


def test_a_build_over_budget_mid_search_is_reported_with_its_facts(ctx):
    """A role budget below the roles' fixed part: the first composition's build
    fails, the search stops with device-memory and the outcome carries the
    facts (reference search/driver.hpp:408-416)."""
    pm = gams.parse_file(GOLDEN / "gams" / "st_e01.gms")
    d, _ = cmi.solve(ctx, pm.problem, partition_num=4, enumerate_cap=4, max_cycle_size=2, sample_points=32,
                     iter_limit=50, budget_bytes=2000)
    assert C.STOP_NAMES[d.stop_reason] == "device-memory"
    assert d.has_over_budget == 1
    f = d.over_budget
    assert f.budget_bytes == 2000 and f.needed_bytes > 2000 and f.n_regions == 16
    assert f.role.decode() and f.composition.decode()
    assert "not even one slot" in R.explain_over_budget(pm.problem, f, 4, 4, 32)
    assert d.as_dict()["over_budget"]["needed_bytes"] == f.needed_bytes
    # a solve that fits reports no facts
    ok, _ = cmi.solve(ctx, pm.problem, partition_num=4, enumerate_cap=4, max_cycle_size=2, sample_points=32,
                      iter_limit=5)
    assert ok.has_over_budget == 0 and ok.as_dict()["over_budget"] is None
