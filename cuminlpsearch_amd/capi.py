"""ctypes binding of include/cmi.h (the engine's C-ABI).

Loading fails loudly: there is no CPU fallback for any entry point. Build the
library first with ``python -m cuminlpsearch_amd.build``.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libcmi.so"

# status codes (cmi_status)
BOX_AUTO, BOX_INTERPRETED, BOX_COMPILED = range(3)
SCHED_DEFAULT, SCHED_MAX_ILP = range(2)
(OK, SHAPE_MISMATCH, INVALID_CONFIG, OVER_BUDGET, DEVICE_ERROR, INVALID_HANDLE, INVALID_DAG, INVALID_PROBLEM, ERROR,
 PARSE_ERROR) = range(10)
STATUS_NAMES = {
    OK: "CMI_OK",
    SHAPE_MISMATCH: "CMI_SHAPE_MISMATCH",
    INVALID_CONFIG: "CMI_INVALID_CONFIG",
    OVER_BUDGET: "CMI_OVER_BUDGET",
    DEVICE_ERROR: "CMI_DEVICE_ERROR",
    INVALID_HANDLE: "CMI_INVALID_HANDLE",
    INVALID_DAG: "CMI_INVALID_DAG",
    INVALID_PROBLEM: "CMI_INVALID_PROBLEM",
    ERROR: "CMI_ERROR",
    PARSE_ERROR: "CMI_PARSE_ERROR",
}

# model::Op
(OP_VAR, OP_CONST, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_SQR, OP_NEG, OP_EXP, OP_LOG, OP_SQRT, OP_SIN, OP_COS,
 OP_TANH, OP_POWN, OP_POW, OP_ABS, OP_MIN, OP_MAX) = range(19)
CONTINUOUS, INTEGER, BINARY = range(3)
LE, EQ = range(2)
SLOT_CONTINUOUS, SLOT_INTEGER_PARTITION, SLOT_INTEGER_ENUMERATE, SLOT_BINARY_ENUMERATE = range(4)
ROLE_SAMPLER, ROLE_BOUNDER, ROLE_ENUMERATOR = range(3)
POLICY_GREEDY_ENUMERATE, POLICY_BISECTION_BUDGET, POLICY_RELATIVE_BISECTION_BUDGET, POLICY_WIDTH_FIRST = range(4)
SELECT_LEAST_LOWER_BOUND, SELECT_REJECT_INDEX, SELECT_REJECT_INDEX_STALE = range(3)
STOP_NAMES = ["converged", "exhausted", "iteration-limit", "host-memory", "allocation-failure", "device-memory"]


class Ivl(C.Structure):
    _fields_ = [("lb", C.c_double), ("ub", C.c_double)]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_size_t),
        ("op", C.POINTER(C.c_uint8)),
        ("in0", C.POINTER(C.c_uint64)),
        ("in1", C.POINTER(C.c_uint64)),
        ("constant", C.POINTER(C.c_double)),
        ("var_index", C.POINTER(C.c_uint64)),
        ("int_exp", C.POINTER(C.c_int32)),
        ("n_vars", C.c_size_t),
        ("box", C.POINTER(Ivl)),
        ("var_kind", C.POINTER(C.c_uint8)),
        ("objective_root", C.c_uint64),
        ("n_constraints", C.c_size_t),
        ("con_root", C.POINTER(C.c_uint64)),
        ("con_cmp", C.POINTER(C.c_uint8)),
        ("con_rhs", C.POINTER(C.c_double)),
    ]


class QcqpSpec(C.Structure):
    _fields_ = [
        ("n", C.c_size_t), ("m", C.c_size_t), ("n_int", C.c_size_t),
        ("cont_lb", C.c_double), ("cont_ub", C.c_double), ("int_lb", C.c_double), ("int_ub", C.c_double),
        ("pair_density", C.c_double), ("term_prob", C.c_double), ("lin_prob", C.c_double),
        ("symmetric_pairs", C.c_int),
        ("rhs_lo", C.c_double), ("rhs_hi", C.c_double),
        ("target_nodes", C.c_size_t),
    ]


class ProblemInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_size_t), ("n_vars", C.c_size_t), ("n_constraints", C.c_size_t),
                ("reachable_nodes", C.c_size_t)]


class TapeInfo(C.Structure):
    _fields_ = [("n_instr", C.c_size_t), ("n_slots", C.c_size_t), ("n_vars", C.c_size_t), ("n_outputs", C.c_size_t),
                ("reachable_nodes", C.c_size_t), ("flops_per_box", C.c_uint64), ("bytes_per_box", C.c_size_t),
                ("box_groups", C.c_size_t), ("box_instr", C.c_size_t), ("box_slots", C.c_size_t),
                ("box_engine", C.c_int), ("box_compiled", C.c_int), ("compile_seconds", C.c_double),
                ("compiled_persistent", C.c_int), ("compiled_wgs_per_cu", C.c_size_t),
                ("compiled_waves", C.c_size_t), ("extended", C.c_int), ("compiled_schedule", C.c_int),
                ("compiled_load_batch", C.c_size_t)]


class BoxCandidate(C.Structure):
    _fields_ = [("engine", C.c_int), ("groups", C.c_size_t), ("waves", C.c_size_t), ("persistent", C.c_int),
                ("wgs_per_cu", C.c_size_t), ("schedule", C.c_int), ("load_batch", C.c_size_t),
                ("kernel_ms", C.c_double)]


class CompiledShape(C.Structure):
    _fields_ = [("groups", C.c_size_t), ("persistent", C.c_int), ("waves", C.c_size_t), ("schedule", C.c_int),
                ("load_batch", C.c_size_t), ("roots", C.c_int)]


class AggregateBound(C.Structure):
    _fields_ = [("lb", C.c_double), ("hull_ub", C.c_double), ("feasible", C.c_int)]


class CostModel(C.Structure):
    _fields_ = [("sampler_fixed", C.c_size_t), ("sampler_per_region", C.c_size_t),
                ("bounder_fixed", C.c_size_t), ("bounder_per_region", C.c_size_t),
                ("enumerator_fixed", C.c_size_t), ("enumerator_per_region", C.c_size_t),
                ("max_regions", C.c_size_t)]


class OverBudget(C.Structure):
    _fields_ = [("role", C.c_char * 32), ("needed_bytes", C.c_size_t), ("budget_bytes", C.c_size_t),
                ("n_regions", C.c_size_t), ("elements_per_region", C.c_size_t), ("bytes_per_element", C.c_size_t),
                ("element_breakdown", C.c_char * 96), ("element_inventory", C.c_char * 64), ("cost", CostModel),
                ("composition_kinds", C.c_uint8 * 64), ("composition_size", C.c_size_t),
                ("composition", C.c_char * 68), ("has_fan_out", C.c_int), ("fan_out_partition_num", C.c_size_t),
                ("fan_out_enumerate_cap", C.c_size_t), ("solve_samples_per_region", C.c_size_t),
                ("sampler_bytes", C.c_size_t)]


def over_budget_dict(ob: OverBudget) -> dict:
    """The facts as plain Python values (composition as its kinds list)."""
    out = {}
    for name, _ in ob._fields_:
        v = getattr(ob, name)
        if isinstance(v, bytes):
            v = v.decode()
        elif name == "cost":
            v = {n: getattr(v, n) for n, _ in v._fields_}
        elif name == "composition_kinds":
            v = list(v)[:ob.composition_size]
        out[name] = v
    return out


class BoundView(C.Structure):
    _fields_ = [("feasible", C.POINTER(C.c_uint8)), ("obj_lb", C.POINTER(C.c_double)), ("n_regions", C.c_size_t)]


class CandidateView(C.Structure):
    _fields_ = [("found", C.c_int), ("value", C.c_double), ("point", C.POINTER(C.c_double)), ("index", C.c_int64)]


class SolveOptions(C.Structure):
    _fields_ = [
        ("partition_num", C.c_size_t), ("enumerate_cap", C.c_size_t), ("max_cycle_size", C.c_size_t),
        ("sample_points", C.c_size_t), ("iter_limit", C.c_uint32), ("tolerance", C.c_double),
        ("budget_bytes", C.c_size_t), ("host_budget_bytes", C.c_size_t), ("compact_frontier", C.c_int),
        ("has_known_primal", C.c_int), ("known_primal", C.c_double), ("policy", C.c_int),
        ("bisection_budget", C.c_size_t),
    ]


class AggregateOptions(C.Structure):
    _fields_ = [
        ("branch_fan_out", C.c_size_t), ("partition_budget", C.c_size_t), ("sample_fan_out", C.c_size_t),
        ("sample_points", C.c_size_t), ("iter_limit", C.c_uint32), ("tolerance", C.c_double), ("selection", C.c_int),
        ("relative_width", C.c_int), ("budget_bytes", C.c_size_t), ("host_budget_bytes", C.c_size_t),
        ("has_known_primal", C.c_int), ("known_primal", C.c_double), ("kernel", C.c_int),
    ]


class SolveOutcome(C.Structure):
    _fields_ = [
        ("lower_bound", C.c_double), ("upper_bound", C.c_double), ("stop_reason", C.c_int),
        ("proven_optimal", C.c_int), ("infeasible", C.c_int), ("found_incumbent", C.c_int),
        ("iterations", C.c_uint64), ("fathomed", C.c_uint64), ("bounded", C.c_uint64), ("sampled", C.c_uint64),
        ("enqueued", C.c_uint64), ("pruned_infeasible", C.c_uint64), ("pruned_dominated", C.c_uint64),
        ("pending", C.c_uint64), ("wall_seconds", C.c_double), ("has_over_budget", C.c_int),
        ("over_budget", OverBudget), ("fathomed_points", C.c_uint64), ("witness_rejected", C.c_uint64),
        ("pruned_duplicate", C.c_uint64), ("viable", C.c_uint64), ("dropped_viable", C.c_uint64),
        ("dropped_dominated", C.c_uint64), ("dropped_floor", C.c_double), ("history_freed", C.c_uint64),
        ("history_peak_live", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["stop_reason"] = STOP_NAMES[self.stop_reason]
        ob = self.over_budget
        d["over_budget"] = over_budget_dict(ob) if self.has_over_budget else None
        return d


class ServerOptions(C.Structure):
    _fields_ = [("max_active", C.c_size_t), ("hbm_budget_bytes", C.c_size_t), ("realtime", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("launches", C.c_uint64), ("peak_active", C.c_size_t),
                ("wall_seconds", C.c_double), ("peak_resident_bytes", C.c_size_t), ("kernel_seconds", C.c_double),
                ("kernel_fp64_ops", C.c_double)]


class GamsOptions(C.Structure):
    _fields_ = [("default_bound", C.c_double), ("fold_fixed_to_const", C.c_int), ("reject_discrete", C.c_int)]


class GamsError(C.Structure):
    _fields_ = [("line", C.c_int), ("kind", C.c_int)]


class GamsInfo(C.Structure):
    _fields_ = [("sense", C.c_int), ("n_vars", C.c_size_t), ("n_constraints", C.c_size_t),
                ("n_warnings", C.c_size_t), ("objvar_kept", C.c_int), ("has_initial_point", C.c_int)]


class ProblemProfile(C.Structure):
    _fields_ = [("num_binary", C.c_size_t), ("num_integer", C.c_size_t), ("num_continuous", C.c_size_t),
                ("largest_integer_domain", C.c_size_t), ("objvar_kept", C.c_int)]


SOURCE_AUTO, SOURCE_NAMED, SOURCE_OVERRIDDEN = 0, 1, 2
SOURCE_NAMES = {SOURCE_AUTO: "auto", SOURCE_NAMED: "named", SOURCE_OVERRIDDEN: "overridden"}


class ResolveRequest(C.Structure):
    _fields_ = [("rules", C.c_char_p), ("objvar_kept", C.c_int), ("free_device_bytes", C.c_size_t),
                ("partition_num", C.c_size_t), ("enumerate_cap", C.c_size_t), ("sample_points", C.c_size_t),
                ("max_cycle_size", C.c_size_t)]


class ResolvedShape(C.Structure):
    _fields_ = [("partition_num", C.c_size_t), ("enumerate_cap", C.c_size_t), ("sample_points", C.c_size_t),
                ("max_cycle_size", C.c_size_t), ("budget_bytes", C.c_size_t), ("footprint_bytes", C.c_size_t),
                ("source", C.c_int), ("rules", C.c_char * 32)]


class CmiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_P = C.c_void_p
_SIGNATURES = {
    "cmi_last_error": (C.c_char_p, []),
    "cmi_version": (C.c_char_p, []),
    "cmi_problem_create": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(_P)]),
    "cmi_problem_qcqp": (C.c_int, [C.POINTER(QcqpSpec), C.c_uint64, C.POINTER(_P)]),
    "cmi_problem_info_get": (C.c_int, [_P, C.POINTER(ProblemInfo)]),
    "cmi_problem_export": (C.c_int, [_P] + [_P] * 12),
    "cmi_problem_destroy": (None, [_P]),
    "cmi_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "cmi_ctx_destroy": (None, [_P]),
    "cmi_tape_create": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "cmi_tape_info_get": (C.c_int, [_P, C.POINTER(TapeInfo)]),
    "cmi_tape_destroy": (None, [_P]),
    "cmi_tape_set_box_groups": (C.c_int, [_P, _P, C.c_size_t]),
    "cmi_tape_set_box_engine": (C.c_int, [_P, C.c_int]),
    "cmi_aggregate_build": (C.c_int, [_P, _P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                     C.POINTER(_P), C.POINTER(OverBudget)]),
    "cmi_bound_children": (C.c_int, [_P, C.c_size_t, _P, C.c_size_t, _P, _P, C.c_size_t,
                                    C.POINTER(AggregateBound)]),
    "cmi_tape_autotune": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(BoxCandidate), C.c_size_t,
                                   C.POINTER(C.c_size_t)]),
    "cmi_tape_set_compiled_shape": (C.c_int, [_P, C.c_int, C.c_size_t, C.c_size_t]),
    "cmi_tape_compiled_source": (C.c_int, [_P, C.POINTER(CompiledShape), C.c_char_p, C.c_size_t,
                                           C.POINTER(C.c_size_t)]),
    "cmi_tape_compiled_object": (C.c_int, [_P, C.POINTER(CompiledShape), _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "cmi_tape_set_compiled_codegen": (C.c_int, [_P, C.c_int, C.c_size_t]),
    "cmi_tape_compile_words": (C.c_int, [_P, C.c_size_t, _P, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(TapeInfo)]),
    "cmi_tape_compile_info": (C.c_int, [_P, C.POINTER(TapeInfo), C.c_char_p, C.c_size_t]),
    "cmi_eval_boxes": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_size_t, _P, _P, _P, _P]),
    "cmi_eval_boxes_device": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_size_t, _P, _P, _P, _P, _P]),
    "cmi_eval_points": (C.c_int, [_P, _P, C.c_size_t, C.c_size_t, _P, _P, _P]),
    "cmi_cost_model_get": (C.c_int, [_P, C.c_size_t, C.POINTER(CostModel)]),
    "cmi_role_build": (C.c_int, [_P, C.c_int, _P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(_P),
                                 C.POINTER(OverBudget)]),
    "cmi_bound": (C.c_int, [_P, _P, C.c_size_t, _P, _P, C.c_size_t, C.POINTER(BoundView)]),
    "cmi_sample": (C.c_int, [_P, _P, C.c_size_t, _P, _P, C.c_size_t, C.c_uint64, C.POINTER(CandidateView)]),
    "cmi_enumerate": (C.c_int, [_P, _P, C.c_size_t, _P, _P, C.c_size_t, C.POINTER(CandidateView)]),
    "cmi_role_size": (C.c_size_t, [_P]),
    "cmi_role_destroy": (None, [_P]),
    "cmi_solve_options_default": (None, [C.POINTER(SolveOptions)]),
    "cmi_solve": (C.c_int, [_P, _P, C.POINTER(SolveOptions), C.POINTER(SolveOutcome), _P]),
    "cmi_aggregate_options_default": (None, [C.POINTER(AggregateOptions)]),
    "cmi_aggregate_fit": (C.c_int, [_P, C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(C.c_size_t),
                                    C.POINTER(C.c_size_t)]),
    "cmi_aggregate_solve": (C.c_int, [_P, _P, C.POINTER(AggregateOptions), C.POINTER(SolveOutcome), _P,
                                      C.POINTER(C.c_int)]),
    "cmi_solve_batch": (C.c_int, [_P, C.c_size_t, C.POINTER(_P), C.POINTER(SolveOptions), C.POINTER(SolveOutcome),
                                  C.POINTER(_P)]),
    "cmi_solve_batch_stats": (C.c_int, [_P, C.c_size_t, C.POINTER(_P), C.POINTER(SolveOptions),
                                        C.POINTER(SolveOutcome), C.POINTER(_P), C.POINTER(ServeStats)]),
    "cmi_serve_stream": (C.c_int, [_P, C.c_size_t, C.POINTER(_P), _P, C.POINTER(SolveOptions),
                                   C.POINTER(ServerOptions), C.POINTER(SolveOutcome), _P, _P,
                                   C.POINTER(ServeStats)]),
    "cmi_gams_options_default": (None, [C.POINTER(GamsOptions)]),
    "cmi_gams_parse": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(GamsOptions), C.POINTER(_P), C.POINTER(_P),
                                 C.POINTER(GamsError)]),
    "cmi_gams_parse_file": (C.c_int, [C.c_char_p, C.POINTER(GamsOptions), C.POINTER(_P), C.POINTER(_P),
                                      C.POINTER(GamsError)]),
    "cmi_gams_info_get": (C.c_int, [_P, C.POINTER(GamsInfo)]),
    "cmi_gams_var_name": (C.c_char_p, [_P, C.c_size_t]),
    "cmi_gams_constraint_name": (C.c_char_p, [_P, C.c_size_t]),
    "cmi_gams_warning": (C.c_char_p, [_P, C.c_size_t, C.POINTER(C.c_int)]),
    "cmi_gams_initial_point": (C.c_int, [_P, _P, C.c_size_t]),
    "cmi_gams_destroy": (None, [_P]),
    "cmi_gams_result_line": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_char_p, C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    "cmi_problem_profile_get": (C.c_int, [_P, C.POINTER(ProblemProfile)]),
    "cmi_ctx_memory": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "cmi_worst_footprint": (C.c_int, [_P, C.POINTER(ProblemProfile), C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                      C.POINTER(C.c_size_t)]),
    "cmi_auto_max_cycle_size": (C.c_int, [_P, C.POINTER(ProblemProfile), C.c_size_t, C.c_size_t, C.c_size_t,
                                          C.c_size_t, C.c_size_t, C.POINTER(C.c_size_t)]),
    "cmi_sample_noalloc": (C.c_int, [_P, _P, C.c_size_t, _P, C.c_size_t, _P, _P, C.c_size_t, C.c_uint64, _P, _P, _P,
                                      _P]),
    "cmi_explain_over_budget": (C.c_int, [_P, C.c_int, C.POINTER(OverBudget), C.c_size_t, C.c_size_t, C.c_size_t,
                                          C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "cmi_resolve_shape": (C.c_int, [_P, C.POINTER(ResolveRequest), C.POINTER(ResolvedShape)]),
    "cmi_serve_plain": (C.c_int, [_P, C.c_size_t, C.POINTER(_P), _P, C.POINTER(SolveOptions),
                                  C.POINTER(SolveOutcome), _P, _P, C.POINTER(ServeStats)]),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None

# Engine objects (contexts, tapes, roles, problems) in creation order, so that
# whatever is still alive when the interpreter exits is released then, newest
# first, while the HIP runtime (and a profiler attached to it) is still up,
# rather than by finalisers that may run after either has shut down.
_owners: list = []


def track(owner) -> None:
    """Register an object with a `_release()` method for release at exit."""
    if len(_owners) % 1024 == 1023:
        _owners[:] = [r for r in _owners if r() is not None]
    _owners.append(weakref.ref(owner))


@atexit.register
def _release_at_exit() -> None:
    if _lib is None:
        return
    for r in reversed(_owners):
        o = r()
        if o is not None:
            o._release()
    _owners.clear()


def lib() -> C.CDLL:
    """The loaded engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cuminlpsearch_amd.build`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != OK:
        raise CmiError(status, lib().cmi_last_error().decode())


def ptr(a: np.ndarray | None):
    """Raw data pointer of a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)
