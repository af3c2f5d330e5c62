"""The memory-fit resolver through the C-ABI (host arithmetic, no device).

Restates the reference's own resolver tests case by case
(test/source/policy_catalogue_test.cpp:58-330: roster lookup and constants,
nvs09's profile, CoverDomains, the two-phase fit re-deriving nvs09's tuned
shape, the no-budget fallbacks, override composition, select_policy's
boundaries, the objvar discount) and pins the footprint against what its
definition fixes: costs affine in the region count, slot order integers then
continuous then binaries, the all-discrete prefix as a separate maximum,
monotonicity in the cap, and the launch-capacity limit of the device roles.
"""
from __future__ import annotations

import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd import gams
from cuminlpsearch_amd import resolve as R
from cuminlpsearch_amd.model import Problem

GIB = 1 << 30


def build(n_bin=0, n_int=0, n_cont=0, int_lb=3.0, int_ub=9.0, cont_ub=1.0) -> cmi.EngineProblem:
    """A problem with the given variable kinds; the objective touches every
    variable (sum of squares), so the tape and its cost depend on the counts."""
    p = Problem()
    xs = [p.var(0, 1, C.BINARY) for _ in range(n_bin)]
    xs += [p.var(int_lb, int_ub, C.INTEGER) for _ in range(n_int)]
    xs += [p.var(0.0, cont_ub) for _ in range(n_cont)]
    obj = xs[0] * xs[0]
    for x in xs[1:]:
        obj = obj + x * x
    p.set_objective(obj)
    return cmi.EngineProblem.from_problem(p)


def nvs09() -> cmi.EngineProblem:
    """nvs09's variable set: 10 integers on [3, 9] (reference
    test/source/nvs09_problem.hpp), with an objective over all of them."""
    return build(n_int=10)


# -- roster ----------------------------------------------------------------------

APPLICABLE = {
    "all-binary": dict(n_bin=20),
    "discrete": dict(n_int=10),
    "mixed-binary": dict(n_bin=3, n_cont=2),
    "mixed-all-small": dict(n_int=2, n_cont=2),
    "mixed-all-large": dict(n_int=2, n_cont=70),
    "width-first": dict(n_int=2, n_cont=2),
}


@pytest.mark.parametrize("name", R.ROSTER)
def test_every_roster_row_resolves_by_name(name):
    s = R.resolve_shape(build(**APPLICABLE[name]), rules=name)
    assert s.rules == name and s.source == "named"


@pytest.mark.parametrize("name", ["quadratic", "", "mixed", "bisection-budget"])
def test_unknown_rules_are_a_configuration_error(name):
    with pytest.raises(C.CmiError) as e:
        R.resolve_shape(build(n_bin=2), rules=name)
    assert e.value.status == C.INVALID_CONFIG
    assert "roster" in str(e.value)


@pytest.mark.parametrize("name,expect", [
    # (partition_num, enumerate_cap, sample_points, max_cycle_size) with no budget:
    # the rows' pinned constants (reference catalogue.hpp:87-143)
    ("all-binary", (2, 2, 1, 20)),
    ("discrete", (2, 7, 5, 7)),          # CoverDomains: nvs-like [3, 9] -> 7
    ("mixed-binary", (2, 2, 5, 4)),      # FollowPartition at the floor
    ("mixed-all-small", (2, 7, 10, 4)),
    ("mixed-all-large", (2, 7, 3, 4)),
    ("width-first", (2, 7, 10, 1)),      # named only; one slot per iteration (catalogue.hpp:132-149)
])
def test_roster_constants_without_a_budget(name, expect):
    s = R.resolve_shape(build(**APPLICABLE[name]), rules=name)
    assert (s.partition_num, s.enumerate_cap, s.sample_points, s.max_cycle_size) == expect
    assert s.budget_bytes == 0
    assert s.solve_options().get("policy", C.POLICY_GREEDY_ENUMERATE) == (
        C.POLICY_WIDTH_FIRST if name == "width-first" else C.POLICY_GREEDY_ENUMERATE)


def test_width_first_is_never_selected_automatically():
    """select_policy classifies into the first five rows only (catalogue.hpp:132-141)."""
    for kinds in APPLICABLE.values():
        assert R.resolve_shape(build(**kinds)).rules != "width-first"
    s = R.resolve_shape(build(n_int=2, n_cont=2), rules="width-first", free_device_bytes=8 << 30)
    assert s.max_cycle_size == 1 and s.partition_num > 2


# -- profile and the resolver ----------------------------------------------------------


def test_profile_reads_nvs09s_ten_integers():
    pp = R.profile(nvs09())
    assert (pp.num_binary, pp.num_integer, pp.num_continuous) == (0, 10, 0)
    assert pp.largest_integer_domain == 7  # [3, 9] has 7 integer points
    assert not pp.objvar_kept


@pytest.mark.parametrize("lb,ub,domain", [(3, 9, 7), (0, 0, 1), (-4, 4, 9), (0, 1000, 1001)])
def test_profile_counts_integer_points(lb, ub, domain):
    assert R.profile(build(n_int=2, int_lb=lb, int_ub=ub)).largest_integer_domain == domain


def test_cover_domains_clamps_enumerate_cap_to_the_domain():
    s = R.resolve_shape(nvs09(), rules="discrete", free_device_bytes=64 * GIB)
    assert s.enumerate_cap == 7
    wide = R.resolve_shape(build(n_int=3, int_lb=0, int_ub=100), rules="discrete", free_device_bytes=64 * GIB)
    assert wide.enumerate_cap == 16  # discrete's ceiling


def test_two_phase_fit_rederives_nvs09s_tuned_shape():
    """Every domain is 7, so for q <= 7 an integer slot costs max(q, 7) = 7
    regions, the same as at q = 2; a budget of exactly the footprint of cap 7
    lets phase 1 buy 7 of the 10 slots and phase 2 widen q from 2 to 7 for
    free (reference policy_catalogue_test.cpp:142-187)."""
    p = nvs09()
    at7 = R.worst_footprint(p, 7, 2, 7, 5)
    at8 = R.worst_footprint(p, 8, 2, 7, 5)
    assert at8 > at7
    free = int(at7 / R.BUDGET_FRACTION + 1024.0)
    s = R.resolve_shape(p, rules="discrete", free_device_bytes=free)
    assert (s.max_cycle_size, s.partition_num, s.enumerate_cap) == (7, 7, 7)
    assert s.max_cycle_size <= R.MAX_SLOTS
    assert s.footprint_bytes == R.worst_footprint(p, 7, 7, 7, 5) <= s.budget_bytes
    # one partition wider would cost coverage
    assert R.auto_max_cycle_size(p, 8, 7, 5, s.budget_bytes, 10) < 7


def test_no_budget_falls_back_to_pinned_cycle_sizes():
    s = R.resolve_shape(nvs09(), rules="discrete")
    assert (s.partition_num, s.enumerate_cap, s.max_cycle_size) == (2, 7, 7)
    s = R.resolve_shape(build(n_bin=20), rules="all-binary")
    assert (s.partition_num, s.enumerate_cap, s.max_cycle_size) == (2, 2, 20)


def test_a_fitted_zero_becomes_one():
    s = R.resolve_shape(nvs09(), rules="discrete", free_device_bytes=16)
    assert s.max_cycle_size == 1 and s.partition_num == 2
    assert s.footprint_bytes > s.budget_bytes  # the role build reports the overflow


def test_budget_is_the_stated_fraction_of_free_memory():
    free = 10 * GIB + 12345
    s = R.resolve_shape(nvs09(), free_device_bytes=free)
    assert s.budget_bytes == int(free * 0.67)
    assert s.footprint_bytes <= s.budget_bytes


def test_pinned_partition_still_fits_the_cycle_with_a_budget():
    """all-binary pins partition and enumerate at 2 but fits the cycle: 30
    binaries in a huge budget stop at the launch-capacity limit, 2^30 regions
    (2^31 would not index)."""
    s = R.resolve_shape(build(n_bin=40), free_device_bytes=1 << 50)
    assert s.rules == "all-binary" and s.source == "auto"
    assert (s.partition_num, s.enumerate_cap, s.max_cycle_size) == (2, 2, 30)


# -- overrides (reference run_spec.hpp:122-166) --------------------------------------------


def test_partition_override_carries_enumerate_cap_along():
    s = R.resolve_shape(nvs09(), rules="discrete", free_device_bytes=64 * GIB, partition_num=5)
    assert (s.partition_num, s.enumerate_cap, s.source) == (5, 5, "overridden")


def test_enumerate_override_alone_keeps_the_fitted_partition():
    base = R.resolve_shape(nvs09(), rules="discrete", free_device_bytes=64 * GIB)
    s = R.resolve_shape(nvs09(), rules="discrete", free_device_bytes=64 * GIB, enumerate_cap=3)
    assert (s.partition_num, s.enumerate_cap, s.max_cycle_size) == (base.partition_num, 3, base.max_cycle_size)


def test_fields_override_independently():
    base = R.resolve_shape(nvs09(), free_device_bytes=64 * GIB)
    s = R.resolve_shape(nvs09(), free_device_bytes=64 * GIB, sample_points=9, max_cycle_size=2)
    assert (s.sample_points, s.max_cycle_size) == (9, 2)
    assert (s.partition_num, s.enumerate_cap) == (base.partition_num, base.enumerate_cap)
    assert s.footprint_bytes == R.worst_footprint(nvs09(), 2, s.partition_num, s.enumerate_cap, 9)


def test_override_below_the_fan_out_floor_is_rejected():
    with pytest.raises(C.CmiError) as e:
        R.resolve_shape(nvs09(), partition_num=1, enumerate_cap=1)
    assert e.value.status == C.INVALID_CONFIG


# -- selection and applicability (reference catalogue.hpp:176-261) ------------------------


@pytest.mark.parametrize("kinds,name", [
    (dict(n_bin=5), "all-binary"),                    # rule 1
    (dict(n_int=1), "discrete"),                      # rule 2
    (dict(n_bin=5, n_int=1), "discrete"),             # rule 2, binaries included
    (dict(n_cont=1, n_bin=1), "mixed-binary"),        # rule 3
    (dict(n_cont=64), "mixed-all-small"),             # 64 live slots: at the cap
    (dict(n_cont=65), "mixed-all-large"),             # one over
    (dict(n_cont=19), "mixed-all-small"),             # continuous only (ex8_6_2's shape)
    (dict(n_cont=2, n_int=1, n_bin=1), "mixed-all-small"),
])
def test_selection_at_its_boundaries(kinds, name):
    s = R.resolve_shape(build(**kinds))
    assert s.rules == name and s.source == "auto"


@pytest.mark.parametrize("n_cont,kept,name", [
    (1, True, "all-binary"),     # the kept objective variable is no degree of freedom
    (1, False, "mixed-binary"),  # an ordinary continuous variable is
    (2, True, "mixed-binary"),   # one discounted, one genuine
])
def test_objvar_kept_discounts_exactly_one_continuous(n_cont, kept, name):
    assert R.resolve_shape(build(n_bin=20, n_cont=n_cont), objvar_kept=kept).rules == name


@pytest.mark.parametrize("name,kinds,ok", [
    ("all-binary", dict(n_bin=3), True),
    ("all-binary", dict(n_bin=3, n_int=1), False),
    ("all-binary", dict(n_bin=3, n_cont=1), False),
    ("discrete", dict(n_bin=3), True),
    ("discrete", dict(n_int=2, n_cont=1), False),
    ("mixed-binary", dict(n_bin=3, n_cont=2), True),
    ("mixed-binary", dict(n_int=1, n_cont=2), False),
    ("mixed-all-small", dict(n_bin=3), True),
    ("mixed-all-large", dict(n_cont=3), True),
])
def test_applicability(name, kinds, ok):
    if ok:
        assert R.resolve_shape(build(**kinds), rules=name).rules == name
    else:
        with pytest.raises(C.CmiError) as e:
            R.resolve_shape(build(**kinds), rules=name)
        assert e.value.status == C.INVALID_CONFIG and "do not apply" in str(e.value)


def test_gams_kept_objective_variable_classifies_as_all_binary():
    src = ("Variables objvar; Binary Variables b1,b2,b3;\nEquations e1;\n"
           "e1.. sqr(objvar) - b1 - b2 - b3 =E= 0;\nobjvar.lo = 0; objvar.up = 2;\n"
           "Solve m using MINLP minimizing objvar;\n")
    pm = gams.parse(src)
    assert pm.objvar_kept
    assert R.resolve_shape(pm.problem, objvar_kept=pm.objvar_kept).rules == "all-binary"
    assert R.resolve_shape(pm.problem).rules == "mixed-binary"


# -- the footprint, pinned by its definition ---------------------------------------------


def prof(**kw) -> R.ProblemProfile:
    return R.ProblemProfile(**kw)


def test_footprint_is_affine_in_the_region_count():
    """All-binary: cap k reaches 2^k regions, so f(k) = A + B 2^k."""
    p = build(n_bin=12)
    f = [R.worst_footprint(p, k, 2, 2, 1) for k in range(1, 8)]
    for k in range(2, len(f)):
        assert f[k] - f[k - 1] == 2 * (f[k - 1] - f[k - 2])


def bundle(p, n, enumerable, samples=5):
    """The roles' bytes for one composition of n regions, measured through the
    footprint of a one-slot profile: an integer slot of width n is enumerable,
    a continuous slot of width n is not."""
    if enumerable:
        return R.worst_footprint(p, 1, 2, n, samples, prof(num_integer=1))
    return R.worst_footprint(p, 1, n, 1, samples, prof(num_continuous=1))


def test_slot_order_is_integers_continuous_binaries():
    p = build(n_int=1, n_cont=1, n_bin=1)
    pp = prof(num_integer=1, num_continuous=1, num_binary=1)
    # cap 1: the integer slot first, at max(q, e) = 5, enumerable
    assert R.worst_footprint(p, 1, 3, 5, 5, pp) == bundle(p, 5, True)
    # cap 2: integer x continuous = 15 regions, not enumerable; the
    # all-discrete prefix integer x binary = 10 regions is
    assert R.worst_footprint(p, 2, 3, 5, 5, pp) == max(bundle(p, 5, True), bundle(p, 15, False),
                                                          bundle(p, 10, True))
    # cap 3: integer x continuous x binary = 30
    assert R.worst_footprint(p, 3, 3, 5, 5, pp) == max(bundle(p, 5, True), bundle(p, 15, False),
                                                          bundle(p, 30, False), bundle(p, 10, True))


def test_all_discrete_prefix_is_a_separate_maximum():
    """1 continuous + 3 binaries at q = 2: the widest prefix has 16 regions but
    no enumerator; the binaries alone have 8 and do, and the enumerator's
    fixed part makes them the larger bundle."""
    p = build(n_cont=1, n_bin=3)
    pp = prof(num_continuous=1, num_binary=3)
    f4 = R.worst_footprint(p, 4, 2, 2, 5, pp)
    assert bundle(p, 8, True) > bundle(p, 16, False)
    assert f4 == bundle(p, 8, True)


def test_footprint_is_monotone_in_the_cap():
    p = build(n_bin=2, n_int=3, n_cont=4)
    f = [R.worst_footprint(p, k, 3, 4, 10) for k in range(0, 10)]
    assert f == sorted(f)


def test_auto_max_cycle_size_is_the_widest_fitting_cap():
    p = build(n_bin=2, n_int=3, n_cont=4)
    for budget in (10_000, 100_000, 1_000_000, 50_000_000):
        k = R.auto_max_cycle_size(p, 3, 4, 10, budget, 64)
        if k > 0:
            assert R.worst_footprint(p, k, 3, 4, 10) <= budget
        if k < 9:
            assert R.worst_footprint(p, k + 1, 3, 4, 10) > budget
    assert R.auto_max_cycle_size(p, 3, 4, 10, 1, 64) == 0
    assert R.auto_max_cycle_size(p, 3, 4, 10, 1 << 50, 64) == 9  # all variables
    assert R.auto_max_cycle_size(p, 3, 4, 10, 1 << 50, 4) == 4   # the ceiling


def test_launch_capacity_bounds_the_cap():
    """Sampler launches index regions x samples in 32 bits: at 1000 samples a
    composition holds at most (2^31 - 1) // 1000 regions whatever the budget."""
    p = build(n_bin=40)
    k = R.auto_max_cycle_size(p, 2, 2, 1000, 1 << 60, 64)
    assert 2 ** k <= (2 ** 31 - 1) // 1000 < 2 ** (k + 1)


# -- the CLI's shape flags (no device: these end before a context is made) ---------------


def test_cli_lists_the_roster(capsys):
    from cuminlpsearch_amd import gams_solve
    assert gams_solve.main(["--list-policies"]) == 0
    out = capsys.readouterr().out
    for name in R.ROSTER:
        assert f"  {name} " in out


def test_cli_rejects_an_unknown_policy(capsys):
    from cuminlpsearch_amd import gams_solve
    assert gams_solve.main(["model.gms", "10", "--policy", "quadratic"]) == 2
    err = capsys.readouterr().err
    assert "unknown policy 'quadratic'" in err and "mixed-all-small" in err


def test_cli_needs_a_model_and_an_iteration_limit(capsys):
    from cuminlpsearch_amd import gams_solve
    assert gams_solve.main([]) == 2


# -- explain_over_budget (reference config/resolve.hpp:247-330) ----------------------------


def facts(role="point sampler", needed=5_000_000, budget=1_000_000, n_regions=250_000, per=20,
          comp=(C.SLOT_CONTINUOUS, C.SLOT_CONTINUOUS)):
    f = C.OverBudget()
    f.role = role.encode()
    f.needed_bytes, f.budget_bytes, f.n_regions, f.bytes_per_element = needed, budget, n_regions, per
    for j, k in enumerate(comp):
        f.composition_kinds[j] = k
    f.composition_size = len(comp)
    return f


def test_explanation_names_the_failure_and_the_cap_that_fits():
    p = build(n_int=2, n_cont=3)
    msg = R.explain_over_budget(p, facts(budget=100_000), 8, 8, 10)
    k = R.auto_max_cycle_size(p, 8, 8, 10, 100_000, 64)
    assert 0 < k < 5
    assert "point sampler" in msg and "composition CC " in msg and "5000000 bytes" in msg and "100000 bytes" in msg
    assert f"at most {k} slot(s) per iteration fit" in msg


def test_explanation_when_not_even_one_slot_fits():
    msg = R.explain_over_budget(build(n_cont=3), facts(budget=100), 8, 8, 10)
    assert "not even one slot" in msg
