"""Solve a GAMS scalar-format model on the GPU and print the RESULT line.

    python -m cuminlpsearch_amd.gams_solve MODEL.gms ITERATIONS [options]

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


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m cuminlpsearch_amd.gams_solve", description=__doc__.split("\n\n")[0])
    ap.add_argument("model")
    ap.add_argument("iterations", type=int)
    ap.add_argument("--partition-num", type=int, default=4)
    ap.add_argument("--enumerate-cap", type=int, default=4)
    ap.add_argument("--max-slots", type=int, default=2, help="slots (variables) split per iteration")
    ap.add_argument("--sample-points", type=int, default=32)
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--known-primal-bound", type=float, default=None,
                    help="objective value of a known feasible point, in the model's own sense")
    ap.add_argument("--default-bound", type=float, default=1e6)
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)

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
        out, best = engine.solve(ctx, pm.problem, partition_num=a.partition_num, enumerate_cap=a.enumerate_cap,
                                 max_cycle_size=a.max_slots, sample_points=a.sample_points, iter_limit=a.iterations,
                                 tolerance=a.tolerance, known_primal=known)
    except _c.CmiError as e:
        print(f"{'out of device memory' if e.status == _c.OVER_BUDGET else 'configuration error'}: {e}",
              file=sys.stderr)
        return 3 if e.status == _c.OVER_BUDGET else 2

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
