"""cuminlpsearch_amd -- MI355X-native interval branch and bound.

The engine (HIP kernels for gfx950 + C++ search host + C-ABI) is
``lib/libcmi.so``; this package binds it with ctypes. There is no CPU
fallback: ``capi.lib()`` raises if the library has not been built.
"""
from . import capi
from .engine import (AggregateBounder, Candidate, Context, EngineProblem, Role, Tape, aggregate_fit, aggregate_solve,
                     qcqp_spec, serve_plain, serve_stream, solve, solve_batch, solve_batch_stats, solve_options)
from .model import Expr, Problem, ProblemArrays, evaluate_point, satisfies, sqr

__all__ = [
    "capi", "aggregate_fit", "aggregate_solve", "Candidate", "Context", "EngineProblem", "Role", "Tape", "qcqp_spec", "serve_plain", "serve_stream",
    "solve", "solve_batch", "solve_batch_stats", "solve_options", "Expr", "Problem", "ProblemArrays", "evaluate_point", "satisfies", "sqr",
]
