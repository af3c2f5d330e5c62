"""Solve a GAMS scalar-format model on the GPU and print the RESULT line.

    python -m cuminlpsearch_amd.gams_solve MODEL.gms ITERATIONS [options]
    python -m cuminlpsearch_amd.gams_solve --list-policies

The search shape comes from the memory-fit resolver, as in the reference
solver (source/gams/solve.cu:640-800): rules selected by the model's variable
kinds (or named with --policy), fitted to the device's free memory (or to
--device-budget), with any of --partition-num, --enumerate-cap, --max-slots
and --sample-points laid over the result. --policy bisection-budget (or
relative-bisection-budget) with --bisection-budget B runs 2^B regions per
node instead (solve.cu:643-712). The chosen shape is printed as a
``policy:`` line and a machine-readable ``PARAMS`` line.

The output keeps the reference solver's machine-readable contract
(source/gams/solve.cu:836-867): one line

    RESULT<TAB>sense=min|max<TAB>primal=<v|none><TAB>dual=<v|none>

in the model's own sense, then the incumbent as ``  name = value`` lines. Exit
codes follow solve.cu:868-897: 0 solved (any stop reason), 1 parse error,
2 configuration error, 3 out of device memory.
"""
from __future__ import annotations

import argparse
import sys

from . import capi as _c
from . import gams
from . import resolve

ROSTER_NOTES = {
    "all-binary": "binaries only: fan-out 2, 1 sample, up to 20 slots",
    "discrete": "integers (and binaries): enumerate whole integer domains up to 16 wide",
    "mixed-binary": "continuous and binaries: enumerate width follows the partition",
    "mixed-all-small": "integers and continuous, at most 64 variables: 10 samples",
    "mixed-all-large": "integers and continuous, more than 64 variables: 3 samples",
    "width-first": "named only: the single widest variable per iteration, 10 samples",
}
BISECTION_POLICIES = {"bisection-budget": _c.POLICY_BISECTION_BUDGET,
                      "relative-bisection-budget": _c.POLICY_RELATIVE_BISECTION_BUDGET}


def print_roster(out) -> None:
    for name in resolve.ROSTER:
        print(f"  {name:16s} {ROSTER_NOTES[name]}", file=out)
    print("  bisection-budget, relative-bisection-budget: 2^B regions per node (--bisection-budget B; 0 or "
          "absent fits B to the device)", file=out)


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m cuminlpsearch_amd.gams_solve", description=__doc__.split("\n\n")[0])
    ap.add_argument("model", nargs="?")
    ap.add_argument("iterations", type=int, nargs="?")
    ap.add_argument("--policy", default=None, help="roster rules by name (default: selected by variable kinds)")
    ap.add_argument("--list-policies", action="store_true")
    ap.add_argument("--bisection-budget", type=int, default=None,
                    help="B for --policy bisection-budget / relative-bisection-budget (0 = fit to the device)")
    ap.add_argument("--partition-num", type=int, default=None)
    ap.add_argument("--enumerate-cap", type=int, default=None)
    ap.add_argument("--max-slots", type=int, default=None, help="slots (variables) split per iteration")
    ap.add_argument("--sample-points", type=int, default=None)
    ap.add_argument("--device-budget", type=int, default=None,
                    help="fit the shape to this many free device bytes instead of what the device reports")
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--known-primal-bound", type=float, default=None,
                    help="objective value of a known feasible point, in the model's own sense")
    ap.add_argument("--default-bound", type=float, default=1e6)
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)

    if a.list_policies:
        print("policies:")
        print_roster(sys.stdout)
        return 0
    if a.model is None or a.iterations is None:
        ap.print_usage(sys.stderr)
        print("a model file and an iteration limit are required", file=sys.stderr)
        return 2
    bisection = a.policy in BISECTION_POLICIES
    if a.bisection_budget is not None and not bisection:
        print("--bisection-budget only applies under --policy bisection-budget or relative-bisection-budget",
              file=sys.stderr)
        return 2
    if a.policy is not None and a.policy not in resolve.ROSTER and not bisection:
        print(f"unknown policy '{a.policy}'; available policies:", file=sys.stderr)
        print_roster(sys.stderr)
        return 2

    try:
        pm = gams.parse_file(a.model, gams.ParseOptions(default_bound=a.default_bound))
    except gams.ParseError as e:
        print(f"parse error: {e}", file=sys.stderr)
        return 1
    for w in pm.warnings:
        print(f"warning: line {w.line}: {w.message}", file=sys.stderr)

    from . import engine

    known = a.known_primal_bound
    if known is not None and pm.sense == gams.MAXIMISE:
        known = -known
    try:
        ctx = engine.Context(a.device)
        free = a.device_budget if a.device_budget is not None else resolve.device_memory(ctx)[0]
        if bisection:
            # source/gams/solve.cu:684-712: 10 samples unless given, every slot
            # available, the enumerate cap as the enumeration ceiling, B given
            # or fitted to the device
            samples = a.sample_points or 10
            b = a.bisection_budget or resolve.bisection_budget_fit(pm.problem, samples, free)
            print(f"policy: {a.policy} (named), bisection_budget: {b}, sample_points: {samples}")
            print(f"PARAMS\tpolicy={a.policy}\tsource=named\tbisection_budget={b}\tsample_points={samples}")
            sys.stdout.flush()
            try:
                out, best = engine.solve(ctx, pm.problem, policy=BISECTION_POLICIES[a.policy], bisection_budget=b,
                                         enumerate_cap=a.enumerate_cap or 0, max_cycle_size=0,
                                         sample_points=samples, iter_limit=a.iterations, tolerance=a.tolerance,
                                         known_primal=known)
            except _c.CmiError as e:
                print(f"{'out of device memory' if e.status == _c.OVER_BUDGET else 'configuration error'}: {e}",
                      file=sys.stderr)
                return 3 if e.status == _c.OVER_BUDGET else 2
            return report(pm, out, best, None)
        shape = resolve.resolve_shape(pm.problem, rules=a.policy, objvar_kept=pm.objvar_kept, free_device_bytes=free,
                                      partition_num=a.partition_num or 0, enumerate_cap=a.enumerate_cap or 0,
                                      sample_points=a.sample_points or 0, max_cycle_size=a.max_slots or 0)
    except _c.CmiError as e:
        print(f"configuration error: {e}", file=sys.stderr)
        return 2
    print(f"policy: {shape.rules} ({shape.source}), partition_num: {shape.partition_num}, "
          f"enumerate_cap: {shape.enumerate_cap}, sample_points: {shape.sample_points}, "
          f"max_slots: {shape.max_cycle_size}")
    print(shape.params_line())
    sys.stdout.flush()
    try:
        out, best = engine.solve(ctx, pm.problem, **shape.solve_options(), iter_limit=a.iterations,
                                 tolerance=a.tolerance, known_primal=known)
    except _c.CmiError as e:
        print(f"{'out of device memory' if e.status == _c.OVER_BUDGET else 'configuration error'}: {e}",
              file=sys.stderr)
        return 3 if e.status == _c.OVER_BUDGET else 2

    return report(pm, out, best, shape)


def report(pm, out, best, shape) -> int:
    if out.has_over_budget and shape is not None:
        # a composition first reached mid-search did not fit; the bracket so
        # far is still reported (reference search/driver.hpp:408-416)
        print("out of device memory: " + resolve.explain_over_budget(
            pm.problem, out.over_budget, shape.partition_num, shape.enumerate_cap, shape.sample_points,
            pm.objvar_kept), file=sys.stderr)
    print(f"Stop reason: {_c.STOP_NAMES[out.stop_reason]}, iterations {out.iterations}")
    print(gams.result_line(pm.sense, out.lower_bound, out.upper_bound))
    if out.found_incumbent:
        for name, v in zip(pm.var_names, best):
            print(f"  {name} = {v:.17g}")
    else:
        print("  (no feasible sample was found; no solution point to report)")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
