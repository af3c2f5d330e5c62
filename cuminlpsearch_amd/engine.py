"""Python handles over the C-ABI: problems, device contexts, tapes, roles and solves.

Every call goes to the gfx950 library; nothing here computes a result itself.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi as _c
from .model import Problem, ProblemArrays


def _handle() -> C.c_void_p:
    return C.c_void_p()


# One tape instruction as the device reads it (16 bytes).
TAPE_WORD = np.dtype([("op", "u1"), ("pad", "u1"), ("dst", "<u2"), ("a", "<u2"), ("b", "<u2"), ("imm", "<f8")])
assert TAPE_WORD.itemsize == 16


class EngineProblem:
    """A validated Problem held by the engine (cmi_problem)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self._info = None
        _c.track(self)

    @classmethod
    def from_arrays(cls, a: ProblemArrays) -> "EngineProblem":
        d = a.desc()
        h = _handle()
        _c.check(_c.lib().cmi_problem_create(C.byref(d), C.byref(h)))
        return cls(h)

    @classmethod
    def from_problem(cls, p: Problem) -> "EngineProblem":
        return cls.from_arrays(p.arrays())

    @classmethod
    def qcqp(cls, seed: int, **spec) -> "EngineProblem":
        s = qcqp_spec(**spec)
        h = _handle()
        _c.check(_c.lib().cmi_problem_qcqp(C.byref(s), C.c_uint64(seed), C.byref(h)))
        return cls(h)

    def info(self) -> _c.ProblemInfo:
        """Node, variable and constraint counts (a problem does not change
        once created, so the first answer is kept: the reachable-node count
        sweeps the DAG)."""
        if self._info is None:
            out = _c.ProblemInfo()
            _c.check(_c.lib().cmi_problem_info_get(self._h, C.byref(out)))
            self._info = out
        return self._info

    def arrays(self) -> ProblemArrays:
        inf = self.info()
        n, nv, m = inf.n_nodes, inf.n_vars, inf.n_constraints
        a = ProblemArrays(
            op=np.zeros(n, np.uint8), in0=np.zeros(n, np.uint64), in1=np.zeros(n, np.uint64),
            constant=np.zeros(n, np.float64), var_index=np.zeros(n, np.uint64), int_exp=np.zeros(n, np.int32),
            box=np.zeros((nv, 2), np.float64), var_kind=np.zeros(nv, np.uint8), objective_root=0,
            con_root=np.zeros(m, np.uint64), con_cmp=np.zeros(m, np.uint8), con_rhs=np.zeros(m, np.float64))
        root = np.zeros(1, np.uint64)
        p = _c.ptr
        _c.check(_c.lib().cmi_problem_export(self._h, p(a.op), p(a.in0), p(a.in1), p(a.constant), p(a.var_index),
                                             p(a.int_exp), p(a.box), p(a.var_kind), p(root), p(a.con_root),
                                             p(a.con_cmp), p(a.con_rhs)))
        a.objective_root = int(root[0])
        return a

    def tape_info(self, listing: bool = False):
        """Host-only tape compile: (TapeInfo, listing text or None)."""
        out = _c.TapeInfo()
        buf = C.create_string_buffer(1 << 22) if listing else None
        _c.check(_c.lib().cmi_tape_compile_info(self._h, C.byref(out), buf, (1 << 22) if listing else 0))
        return out, (buf.value.decode() if listing else None)

    @staticmethod
    def _shape(persistent, waves, groups, schedule, load_batch, roots):
        return _c.CompiledShape(groups, persistent, waves, schedule, load_batch, int(roots))

    def compiled_source(self, persistent: int = -1, waves: int = 0, groups: int = 0, schedule: int = 0,
                        load_batch: int = 0, roots: bool = True) -> str:
        """Host-only: HIP source of the compiled explicit-box kernel (persistent: 1
        persistent form, 0 direct form, -1 the default; groups: function groups,
        0 = automatic; load_batch: variables per load batch, 0 = all at once;
        roots: the kernel that writes root intervals)."""
        sh = self._shape(persistent, waves, groups, schedule, load_batch, roots)
        n = C.c_size_t()
        _c.check(_c.lib().cmi_tape_compiled_source(self._h, C.byref(sh), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        _c.check(_c.lib().cmi_tape_compiled_source(self._h, C.byref(sh), buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def compiled_object(self, persistent: int = -1, waves: int = 0, schedule: int = 0, groups: int = 0,
                        load_batch: int = 0, roots: bool = True) -> bytes:
        """Host-only: the gfx950 code object of the compiled explicit-box kernel
        (schedule: capi.SCHED_DEFAULT or capi.SCHED_MAX_ILP; other arguments as
        for compiled_source)."""
        sh = self._shape(persistent, waves, groups, schedule, load_batch, roots)
        n = C.c_size_t()
        _c.check(_c.lib().cmi_tape_compiled_object(self._h, C.byref(sh), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        _c.check(_c.lib().cmi_tape_compiled_object(self._h, C.byref(sh), buf, n.value, C.byref(n)))
        return buf.raw[:n.value]

    def tape_words(self, groups: int = 1):
        """Host-only compile into `groups` function groups: (words, group_begin,
        TapeInfo); words is a structured array with the 16-byte tape layout."""
        info = _c.TapeInfo()
        _c.check(_c.lib().cmi_tape_compile_words(self._h, groups, None, 0, None, C.byref(info)))
        words = np.zeros(info.box_instr, TAPE_WORD)
        begin = np.zeros(info.box_groups + 1, np.uint32)
        _c.check(_c.lib().cmi_tape_compile_words(self._h, groups, _c.ptr(words), words.size,
                                                 begin.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(info)))
        return words, begin, info

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def _release(self):
        if getattr(self, "_h", None) and _c._lib is not None:
            _c.lib().cmi_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()


def qcqp_spec(n=4, m=2, n_int=0, cont_lb=-10.0, cont_ub=10.0, int_lb=0.0, int_ub=10.0, pair_density=1.0,
              term_prob=1.0, lin_prob=1.0, symmetric_pairs=False, rhs_lo=1.0, rhs_hi=10.0,
              target_nodes=0) -> _c.QcqpSpec:
    return _c.QcqpSpec(n, m, n_int, cont_lb, cont_ub, int_lb, int_ub, pair_density, term_prob, lin_prob,
                       int(symmetric_pairs), rhs_lo, rhs_hi, target_nodes)


class Context:
    """One device and its stream (cmi_ctx)."""

    def __init__(self, device: int = 0):
        self._h = _handle()
        _c.track(self)
        _c.check(_c.lib().cmi_ctx_create(device, C.byref(self._h)))

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def _release(self):
        if getattr(self, "_h", None) and _c._lib is not None:
            _c.lib().cmi_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()


class Tape:
    """A problem compiled to a device tape (cmi_tape)."""

    def __init__(self, ctx: Context, problem: EngineProblem):
        self.ctx = ctx
        self.problem = problem
        self._h = _handle()
        _c.track(self)
        _c.check(_c.lib().cmi_tape_create(ctx.handle, problem.handle, C.byref(self._h)))

    def info(self) -> _c.TapeInfo:
        out = _c.TapeInfo()
        _c.check(_c.lib().cmi_tape_info_get(self._h, C.byref(out)))
        return out

    def set_box_engine(self, engine: int) -> None:
        """capi.BOX_AUTO, BOX_INTERPRETED or BOX_COMPILED (compiles now)."""
        _c.check(_c.lib().cmi_tape_set_box_engine(self._h, engine))

    def set_compiled_shape(self, persistent: bool = True, waves: int = 0, wgs_per_cu: int = 1) -> None:
        """Shape of the compiled explicit-box kernel (persistent only when its LDS tiles fit;
        waves per workgroup, 0 = one per function group)."""
        _c.check(_c.lib().cmi_tape_set_compiled_shape(self._h, int(persistent), waves, wgs_per_cu))

    def set_compiled_codegen(self, schedule: int = 0, load_batch: int = 8) -> None:
        """Scheduler (capi.SCHED_DEFAULT or capi.SCHED_MAX_ILP) and load batching
        (variables per batch, 0 = all at once) of the compiled kernel; results
        are the same."""
        _c.check(_c.lib().cmi_tape_set_compiled_codegen(self._h, schedule, load_batch))

    def autotune(self, n_boxes: int = 65536) -> list[dict]:
        """Time the explicit-box kernel configurations on n_boxes synthetic boxes,
        keep the fastest; returns every candidate with its median kernel time."""
        cap = 64
        out = (_c.BoxCandidate * cap)()
        n = C.c_size_t()
        _c.check(_c.lib().cmi_tape_autotune(self._h, self.problem.handle, n_boxes, out, cap, C.byref(n)))
        names = {_c.BOX_INTERPRETED: "interpreted", _c.BOX_COMPILED: "compiled"}
        return [dict(engine=names.get(r.engine, r.engine), groups=r.groups, waves=r.waves,
                     persistent=bool(r.persistent), wgs_per_cu=r.wgs_per_cu,
                     schedule="max-ilp" if r.schedule == _c.SCHED_MAX_ILP else "default", load_batch=r.load_batch,
                     kernel_ms=r.kernel_ms)
                for r in out[:min(n.value, cap)]]

    def set_box_groups(self, k: int = 0) -> int:
        """Waves (function groups) per explicit-box workgroup; 0 = automatic.
        Returns the count in effect after clamping."""
        _c.check(_c.lib().cmi_tape_set_box_groups(self._h, self.problem.handle, k))
        return int(self.info().box_groups)

    def eval_boxes(self, lb: np.ndarray, ub: np.ndarray, roots: bool = False):
        """lb, ub: (n_vars, n_boxes) float64. Returns (feasible, obj_lb, obj_ub, roots or None)."""
        lb = np.ascontiguousarray(lb, np.float64)
        ub = np.ascontiguousarray(ub, np.float64)
        nv, n = lb.shape
        feas = np.zeros(n, np.uint8)
        olb = np.zeros(n)
        oub = np.zeros(n)
        r = np.zeros((self.info().n_outputs, n, 2)) if roots else None
        _c.check(_c.lib().cmi_eval_boxes(self._h, _c.ptr(lb), _c.ptr(ub), n, n, _c.ptr(feas), _c.ptr(olb),
                                         _c.ptr(oub), _c.ptr(r)))
        return feas, olb, oub, r

    def eval_points(self, x: np.ndarray, roots: bool = False):
        x = np.ascontiguousarray(x, np.float64)
        nv, n = x.shape
        feas = np.zeros(n, np.uint8)
        obj = np.zeros(n)
        r = np.zeros((self.info().n_outputs, n)) if roots else None
        _c.check(_c.lib().cmi_eval_points(self._h, _c.ptr(x), n, n, _c.ptr(feas), _c.ptr(obj), _c.ptr(r)))
        return feas, obj, r

    def cost_model(self, samples: int = 1) -> _c.CostModel:
        out = _c.CostModel()
        _c.check(_c.lib().cmi_cost_model_get(self._h, samples, C.byref(out)))
        return out

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def _release(self):
        if getattr(self, "_h", None) and _c._lib is not None:
            _c.lib().cmi_tape_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()


class AggregateBounder:
    """The k-child aggregate bounder (cmi_aggregate_build / cmi_bound_children)."""

    def __init__(self, tape: Tape, slot_kinds, n_regions: int, k: int, max_batch: int = 1, budget_bytes: int = 0):
        self.tape = tape
        self.k = k
        self.slot_kinds = np.ascontiguousarray(slot_kinds, np.uint8)
        self._h = _handle()
        _c.track(self)
        self.facts = _c.OverBudget()
        _c.check(_c.lib().cmi_aggregate_build(tape.handle, _c.ptr(self.slot_kinds), len(self.slot_kinds), n_regions,
                                              k, max_batch, budget_bytes, C.byref(self._h), C.byref(self.facts)))

    def bound_children(self, boxes, var_ids, widths):
        """boxes (R, n_vars, 2), var_ids and widths (R, slot_count) -> (lb, hull_ub, feasible), each (R, k)."""
        boxes = np.ascontiguousarray(boxes, np.float64)
        if boxes.ndim == 2:
            boxes = boxes[None]
        r, nv = boxes.shape[0], boxes.shape[1]
        slots = len(self.slot_kinds)
        vid = np.ascontiguousarray(np.broadcast_to(np.asarray(var_ids, np.uint64).reshape(-1, slots), (r, slots)))
        wid = np.ascontiguousarray(np.broadcast_to(np.asarray(widths, np.uint64).reshape(-1, slots), (r, slots)))
        out = (_c.AggregateBound * (r * self.k))()
        _c.check(_c.lib().cmi_bound_children(self._h, r, _c.ptr(boxes), nv, _c.ptr(vid), _c.ptr(wid),
                                             len(self.slot_kinds), out))
        lb = np.array([o.lb for o in out]).reshape(r, self.k)
        ub = np.array([o.hull_ub for o in out]).reshape(r, self.k)
        fe = np.array([o.feasible for o in out], np.int32).reshape(r, self.k)
        return lb, ub, fe

    def size(self) -> int:
        return int(_c.lib().cmi_role_size(self._h))

    def _release(self):
        if getattr(self, "_h", None) and _c._lib is not None:
            _c.lib().cmi_role_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()


@dataclass
class Candidate:
    found: bool
    value: float
    point: np.ndarray
    index: int


class Role:
    """One backend role over a tape (cmi_role): bounder, sampler or enumerator."""

    def __init__(self, tape: Tape, kind: int, slot_kinds, n_regions: int, samples: int = 1, budget_bytes: int = 0):
        self.tape = tape
        self.kind = kind
        self.slot_kinds = np.ascontiguousarray(slot_kinds, np.uint8)
        self._h = _handle()
        _c.track(self)
        self.facts = _c.OverBudget()
        _c.check(_c.lib().cmi_role_build(tape.handle, kind, _c.ptr(self.slot_kinds), len(self.slot_kinds),
                                         n_regions, samples, budget_bytes, C.byref(self._h), C.byref(self.facts)))

    def _args(self, box, var_ids, widths):
        box = np.ascontiguousarray(box, np.float64).reshape(-1, 2)
        vid = np.ascontiguousarray(var_ids, np.uint64)
        wid = np.ascontiguousarray(widths, np.uint64)
        return box, vid, wid

    def bound(self, box, var_ids, widths):
        box, vid, wid = self._args(box, var_ids, widths)
        v = _c.BoundView()
        _c.check(_c.lib().cmi_bound(self._h, _c.ptr(box), box.shape[0], _c.ptr(vid), _c.ptr(wid), len(vid),
                                    C.byref(v)))
        n = v.n_regions
        feas = np.ctypeslib.as_array(v.feasible, shape=(n,)).copy()
        olb = np.ctypeslib.as_array(v.obj_lb, shape=(n,)).copy()
        return feas, olb

    def sample(self, box, var_ids, widths, salt: int) -> Candidate:
        box, vid, wid = self._args(box, var_ids, widths)
        v = _c.CandidateView()
        _c.check(_c.lib().cmi_sample(self._h, _c.ptr(box), box.shape[0], _c.ptr(vid), _c.ptr(wid), len(vid),
                                     C.c_uint64(salt), C.byref(v)))
        return Candidate(bool(v.found), v.value, np.ctypeslib.as_array(v.point, shape=(box.shape[0],)).copy(),
                         v.index)

    def enumerate(self, box, var_ids, widths) -> Candidate:
        box, vid, wid = self._args(box, var_ids, widths)
        v = _c.CandidateView()
        _c.check(_c.lib().cmi_enumerate(self._h, _c.ptr(box), box.shape[0], _c.ptr(vid), _c.ptr(wid), len(vid),
                                        C.byref(v)))
        return Candidate(bool(v.found), v.value, np.ctypeslib.as_array(v.point, shape=(box.shape[0],)).copy(),
                         v.index)

    def size(self) -> int:
        return _c.lib().cmi_role_size(self._h)

    def _release(self):
        if getattr(self, "_h", None) and _c._lib is not None:
            _c.lib().cmi_role_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()


def solve_options(partition_num=2, enumerate_cap=2, max_cycle_size=0, sample_points=1, iter_limit=1000000,
                  tolerance=1e-6, budget_bytes=0, host_budget_bytes=0, compact_frontier=False,
                  known_primal=None, policy=0, bisection_budget=0) -> _c.SolveOptions:
    """Solve options; policy is capi.POLICY_* (for the bisection-budget
    policies enumerate_cap is the enumeration ceiling, 0 = none)."""
    o = _c.SolveOptions()
    _c.lib().cmi_solve_options_default(C.byref(o))
    o.policy = policy
    o.bisection_budget = bisection_budget
    o.partition_num = partition_num
    o.enumerate_cap = enumerate_cap
    o.max_cycle_size = max_cycle_size
    o.sample_points = sample_points
    o.iter_limit = iter_limit
    o.tolerance = tolerance
    o.budget_bytes = budget_bytes
    o.host_budget_bytes = host_budget_bytes
    o.compact_frontier = int(compact_frontier)
    if known_primal is not None:
        o.has_known_primal = 1
        o.known_primal = known_primal
    return o


def aggregate_options(branch_fan_out=4, partition_budget=1024, sample_fan_out=8, sample_points=1,
                      iter_limit=1000000, tolerance=1e-6, selection=0, relative_width=False, budget_bytes=0,
                      host_budget_bytes=0, known_primal=None, kernel=0) -> _c.AggregateOptions:
    """Settings of the k-child search; selection is capi.SELECT_*, kernel a
    capi.BOX_* engine for the bound launches."""
    o = _c.AggregateOptions()
    _c.lib().cmi_aggregate_options_default(C.byref(o))
    o.branch_fan_out = branch_fan_out
    o.partition_budget = partition_budget
    o.sample_fan_out = sample_fan_out
    o.sample_points = sample_points
    o.iter_limit = iter_limit
    o.tolerance = tolerance
    o.selection = selection
    o.relative_width = int(relative_width)
    o.budget_bytes = budget_bytes
    o.host_budget_bytes = host_budget_bytes
    o.kernel = kernel
    if known_primal is not None:
        o.has_known_primal = 1
        o.known_primal = known_primal
    return o


def aggregate_fit(problem: EngineProblem, free_device_bytes: int, branch_fan_out=4, sample_points=1):
    """The reference's device fit: (partition_budget, sample_fan_out)."""
    n, s = C.c_size_t(), C.c_size_t()
    _c.check(_c.lib().cmi_aggregate_fit(problem.handle, free_device_bytes, branch_fan_out, sample_points,
                                        C.byref(n), C.byref(s)))
    return n.value, s.value


def aggregate_solve(ctx: Context, problem: EngineProblem, **opts):
    """The k-child search: (outcome, best point, whether the selection rule
    ever left its lb fallback)."""
    o = aggregate_options(**opts)
    out = _c.SolveOutcome()
    best = np.full(problem.info().n_vars, np.nan)
    engaged = C.c_int(0)
    _c.check(_c.lib().cmi_aggregate_solve(ctx.handle, problem.handle, C.byref(o), C.byref(out), _c.ptr(best),
                                          C.byref(engaged)))
    return out, best, bool(engaged.value)


def solve(ctx: Context, problem: EngineProblem, **opts):
    o = solve_options(**opts)
    out = _c.SolveOutcome()
    n = problem.info().n_vars
    best = np.zeros(n)
    _c.check(_c.lib().cmi_solve(ctx.handle, problem.handle, C.byref(o), C.byref(out), _c.ptr(best)))
    return out, best


def solve_batch(ctx: Context, problems: list, **opts):
    outs, bests, _ = solve_batch_stats(ctx, problems, **opts)
    return outs, bests


def solve_batch_stats(ctx: Context, problems: list, **opts):
    """solve_batch, also returning its ServeStats (steps, launches, wall time,
    device time and FP64 work of the decode-mode launches)."""
    o = solve_options(**opts)
    n = len(problems)
    outs = (_c.SolveOutcome * n)()
    hs = (C.c_void_p * n)(*[p.handle.value for p in problems])
    bests = [np.zeros(p.info().n_vars) for p in problems]
    bp = (C.c_void_p * n)(*[b.ctypes.data for b in bests])
    st = _c.ServeStats()
    _c.check(_c.lib().cmi_solve_batch_stats(ctx.handle, n, hs, C.byref(o), outs, bp, C.byref(st)))
    return list(outs), bests, st


def serve_stream(ctx: Context, problems: list, arrival, max_active: int = 256, hbm_budget_bytes: int = 0,
                 realtime: bool = True, **opts):
    o = solve_options(**opts)
    n = len(problems)
    outs = (_c.SolveOutcome * n)()
    hs = (C.c_void_p * n)(*[p.handle.value for p in problems])
    arr = np.ascontiguousarray(arrival, np.float64)
    fin = np.zeros(n)
    lat = np.zeros(n)
    so = _c.ServerOptions(max_active, hbm_budget_bytes, int(realtime))
    st = _c.ServeStats()
    _c.check(_c.lib().cmi_serve_stream(ctx.handle, n, hs, _c.ptr(arr), C.byref(o), C.byref(so), outs, _c.ptr(fin),
                                       _c.ptr(lat), C.byref(st)))
    return list(outs), fin, lat, st


def serve_plain(ctx: Context, problems: list, arrival, **opts):
    o = solve_options(**opts)
    n = len(problems)
    outs = (_c.SolveOutcome * n)()
    hs = (C.c_void_p * n)(*[p.handle.value for p in problems])
    arr = np.ascontiguousarray(arrival, np.float64)
    fin = np.zeros(n)
    lat = np.zeros(n)
    st = _c.ServeStats()
    _c.check(_c.lib().cmi_serve_plain(ctx.handle, n, hs, _c.ptr(arr), C.byref(o), outs, _c.ptr(fin), _c.ptr(lat),
                                      C.byref(st)))
    return list(outs), fin, lat, st
