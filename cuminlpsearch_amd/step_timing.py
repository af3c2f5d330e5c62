"""The batched and served forms alone, for profiling their step launches:
the 512-instance QCQP batch (n uniform in [4, 48], m = n/2, density 0.15,
100-iteration cap) solved in shared launches, and the 32-request stream of
QCQPs with n log-uniform in [4, 96] served at 200 req/s. Prints the run's
stats (steps, launches, wall time, the decode-mode launches' device time and
FP64 work) as one JSON line.

    python -m cuminlpsearch_amd.step_timing --batch 512
    python -m cuminlpsearch_amd.step_timing --serve-wide
"""
from __future__ import annotations

import argparse
import json

from .engine import Context, serve_stream, solve_batch_stats
from . import workloads as W

OPTS = dict(partition_num=2, enumerate_cap=2, max_cycle_size=4, sample_points=16, tolerance=1e-6)


def _stats(st) -> dict:
    return {"steps": int(st.steps), "launches": int(st.launches), "wall_seconds": st.wall_seconds,
            "kernel_seconds": st.kernel_seconds, "kernel_fp64_ops": st.kernel_fp64_ops,
            "kernel_fp64_tflops": st.kernel_fp64_ops / st.kernel_seconds / 1e12 if st.kernel_seconds > 0 else 0.0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--batch", type=int, default=0, help="solve this many of the batch's QCQPs in shared launches")
    ap.add_argument("--serve-wide", action="store_true", help="serve the 32-request stream of up to 96 variables")
    a = ap.parse_args(argv)
    ctx = Context(0)
    out = {}
    if a.batch:
        b = W.c4_batch(a.batch)
        _, _, st = solve_batch_stats(ctx, b.problems, iter_limit=100, **OPTS)
        out["batch"] = {"workload": b.description, **_stats(st)}
    if a.serve_wide:
        s = W.c5_stream(32, seed=8, n_max=96)
        _, _, _, st = serve_stream(ctx, s.problems, s.arrivals(200.0), max_active=256, iter_limit=100, **OPTS)
        out["serve_wide"] = {"workload": s.description, **_stats(st)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
