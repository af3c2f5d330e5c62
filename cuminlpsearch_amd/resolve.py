"""The memory-fit resolver: the search shape a problem gets for the device
memory it has.

Binds the library's resolver (host arithmetic over the device roles' cost
model; nothing runs on a device). Mirrors the reference's config layer:
``profile`` is profile_problem (config/problem_profile.hpp:43-70),
``worst_footprint`` is worst_composition_footprint (config/footprint.hpp:61-97),
``auto_max_cycle_size`` is config/footprint.hpp:99-149, ``resolve_shape`` is
select_policy + resolve_shape + resolve (config/catalogue.hpp:182-245,
config/resolve.hpp:95-218, config/run_spec.hpp:122-166). ``ROSTER`` is the
rows of policy_roster (config/catalogue.hpp:87-149).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import capi as _c

ROSTER = ("all-binary", "discrete", "mixed-binary", "mixed-all-small", "mixed-all-large", "width-first")
BUDGET_FRACTION = 0.67
MAX_SLOTS = 64


@dataclass(frozen=True)
class ProblemProfile:
    num_binary: int = 0
    num_integer: int = 0
    num_continuous: int = 0
    largest_integer_domain: int = 0
    objvar_kept: bool = False

    def _c(self) -> _c.ProblemProfile:
        return _c.ProblemProfile(self.num_binary, self.num_integer, self.num_continuous, self.largest_integer_domain,
                                 int(self.objvar_kept))


@dataclass(frozen=True)
class ResolvedShape:
    partition_num: int
    enumerate_cap: int
    sample_points: int
    max_cycle_size: int
    budget_bytes: int
    footprint_bytes: int
    source: str  # "auto", "named" or "overridden"
    rules: str

    def solve_options(self) -> dict:
        """Keyword arguments of engine.solve for this shape (the width-first
        row runs the WidthFirst policy; every other row GreedyEnumerate)."""
        opts = dict(partition_num=self.partition_num, enumerate_cap=self.enumerate_cap,
                    sample_points=self.sample_points, max_cycle_size=self.max_cycle_size)
        if self.rules == "width-first":
            opts["policy"] = _c.POLICY_WIDTH_FIRST
        return opts

    def params_line(self) -> str:
        """The machine-readable shape line (reference source/gams/solve.cu:780-800)."""
        return (f"PARAMS\tpolicy={self.rules}\tsource={self.source}\tpartition_num={self.partition_num}"
                f"\tenumerate_cap={self.enumerate_cap}\tsample_points={self.sample_points}"
                f"\tmax_slots={self.max_cycle_size}")


def bisection_budget_fit(problem, sample_points: int, free_device_bytes: int) -> int:
    """B for the bisection-budget policies when none is given: the largest
    B whose 2^B regions fit the budget fraction of the free bytes at the
    reference graph backend's per-region cost -- sampler (8 b + 25) B per
    sample, bounder 16 b + 25, enumerator 8 b + 25 for b buffer nodes
    (backend/graph/cost.hpp:70-137, source/gams/solve.cu:703-711); at least 1."""
    b = problem.info().reachable_nodes
    per_region = sample_points * (8 * b + 25) + (16 * b + 25) + (8 * b + 25)
    budget = int(free_device_bytes * BUDGET_FRACTION)
    if budget < per_region:
        return 1
    return max(1, (budget // per_region).bit_length() - 1)


def profile(problem, objvar_kept: bool = False) -> ProblemProfile:
    """Variable-kind counts and the widest integer domain of an EngineProblem."""
    p = _c.ProblemProfile()
    _c.check(_c.lib().cmi_problem_profile_get(problem.handle, C.byref(p)))
    return ProblemProfile(p.num_binary, p.num_integer, p.num_continuous, p.largest_integer_domain, objvar_kept)


def device_memory(ctx) -> tuple[int, int]:
    """(free, total) bytes of the context's device."""
    f, t = C.c_size_t(), C.c_size_t()
    _c.check(_c.lib().cmi_ctx_memory(ctx.handle, C.byref(f), C.byref(t)))
    return f.value, t.value


def _prof_ptr(pp: ProblemProfile | None):
    return None if pp is None else C.byref(pp._c())


def worst_footprint(problem, cap: int, partition_num: int, enumerate_cap: int, sample_points: int,
                    pp: ProblemProfile | None = None) -> int:
    out = C.c_size_t()
    _c.check(_c.lib().cmi_worst_footprint(problem.handle, _prof_ptr(pp), cap, partition_num, enumerate_cap,
                                          sample_points, C.byref(out)))
    return out.value


def auto_max_cycle_size(problem, partition_num: int, enumerate_cap: int, sample_points: int, budget_bytes: int,
                        ceiling: int = MAX_SLOTS, pp: ProblemProfile | None = None) -> int:
    out = C.c_size_t()
    _c.check(_c.lib().cmi_auto_max_cycle_size(problem.handle, _prof_ptr(pp), partition_num, enumerate_cap,
                                              sample_points, budget_bytes, ceiling, C.byref(out)))
    return out.value


def resolve_shape(problem, rules: str | None = None, objvar_kept: bool = False, free_device_bytes: int = 0,
                  partition_num: int = 0, enumerate_cap: int = 0, sample_points: int = 0,
                  max_cycle_size: int = 0) -> ResolvedShape:
    """The shape for ``problem``: rules by name or selected by its variable
    kinds, fitted to ``free_device_bytes`` (0 = the rules' fallbacks), with
    any non-zero override laid over the result."""
    req = _c.ResolveRequest(rules.encode() if rules is not None else None, int(objvar_kept), free_device_bytes,
                            partition_num, enumerate_cap, sample_points, max_cycle_size)
    out = _c.ResolvedShape()
    _c.check(_c.lib().cmi_resolve_shape(problem.handle, C.byref(req), C.byref(out)))
    return ResolvedShape(out.partition_num, out.enumerate_cap, out.sample_points, out.max_cycle_size,
                         out.budget_bytes, out.footprint_bytes, _c.SOURCE_NAMES[out.source], out.rules.decode())


def explain_over_budget(problem, facts: _c.OverBudget, partition_num: int, enumerate_cap: int, sample_points: int,
                        objvar_kept: bool = False) -> str:
    """A role build that did not fit, explained with the widest slot cap that
    fits its budget (reference config/resolve.hpp:247-330)."""
    n = C.c_size_t()
    buf = C.create_string_buffer(1024)
    _c.check(_c.lib().cmi_explain_over_budget(problem.handle, int(objvar_kept), C.byref(facts), partition_num,
                                              enumerate_cap, sample_points, buf, len(buf), C.byref(n)))
    return buf.value.decode()
