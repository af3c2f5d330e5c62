"""The k-child bound launch timed on its own: one aggregate bounder over a
problem's root box split k ways on its widest variable, each child covered by
N / k subregions along its next-widest variable, run on the compiled or the
interpreted decode-mode kernel and timed over repeated calls (each call is one
launch, the 2k child bounds copied back and a wait).

Run as a module to profile that kernel alone:

    python -m cuminlpsearch_amd.kernel_timing --model ex2_1_1 --log2n 24 --calls 50
"""
from __future__ import annotations

import argparse
import json
import pathlib
import time

import numpy as np

from . import capi as _c
from .engine import AggregateBounder, Context, EngineProblem, Tape

FP64_PEAK_TFLOPS = 78.6
GAMS_DIR = pathlib.Path(__file__).resolve().parent.parent / "tests" / "golden" / "gams"


def root_partition(box, kinds, k: int, n: int):
    """Slot kinds, variables and widths of a launch over the root box: the
    widest live variable split k ways (most significant), the next widest
    n / k ways; one slot of width n if only one variable is live."""
    box = np.asarray(box, np.float64).reshape(-1, 2)
    w = box[:, 1] - box[:, 0]
    live = sorted((i for i in range(len(box)) if w[i] > 0), key=lambda i: (-w[i], i))
    kind = (lambda v: _c.SLOT_CONTINUOUS if kinds[v] == _c.CONTINUOUS else _c.SLOT_INTEGER_PARTITION)
    if len(live) == 1:
        return [kind(live[0])], [live[0]], [n]
    return [kind(live[1]), kind(live[0])], [live[1], live[0]], [n // k, k]


def aggregate_launch(ctx: Context, problem: EngineProblem, n: int, k: int = 4, engine: int = _c.BOX_COMPILED,
                     calls: int = 50) -> dict:
    a = problem.arrays()
    tape = Tape(ctx, problem)
    tape.set_box_engine(engine)
    kinds, vids, widths = root_partition(a.box, a.var_kind, k, n)
    agg = AggregateBounder(tape, kinds, n, k)
    agg.bound_children(a.box, vids, widths)  # first launch: module load, caches
    t0 = time.perf_counter()
    for _ in range(calls):
        agg.bound_children(a.box, vids, widths)
    per_call = (time.perf_counter() - t0) / calls
    flops = int(tape.info().flops_per_box)
    tflops = flops * n / per_call / 1e12
    return {
        "subregions": n, "children": k, "engine": "compiled" if engine == _c.BOX_COMPILED else "interpreted",
        "ms_per_call": per_call * 1e3, "subregions_per_s": n / per_call, "flops_per_subregion": flops,
        "fp64_tflops": tflops, "fp64_frac": tflops / FP64_PEAK_TFLOPS,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--model", default="ex2_1_1", help="a model under tests/golden/gams, or c1 (the 4-variable QCQP, seed 0)")
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--calls", type=int, default=50)
    ap.add_argument("--interpreted", action="store_true")
    a = ap.parse_args(argv)
    from . import gams, workloads
    ctx = Context(0)
    p = workloads.c1(0) if a.model == "c1" else gams.parse_file(GAMS_DIR / f"{a.model}.gms").problem
    r = aggregate_launch(ctx, p, 1 << a.log2n, a.k, _c.BOX_INTERPRETED if a.interpreted else _c.BOX_COMPILED, a.calls)
    r["model"] = a.model
    print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
