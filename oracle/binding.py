"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

SO = pathlib.Path(__file__).resolve().parent / "liboracle.so"

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not SO.exists():
            raise RuntimeError(f"{SO} missing: run `make -C oracle`")
        h = C.CDLL(str(SO))
        d = C.c_double
        P = C.c_void_p
        for name in ("orc_add_rd", "orc_add_ru", "orc_sub_rd", "orc_sub_ru", "orc_mul_rd", "orc_mul_ru"):
            getattr(h, name).restype = d
            getattr(h, name).argtypes = [d, d]
        for name in ("orc_add_ru_n", "orc_mul_ru_n"):
            getattr(h, name).restype = None
            getattr(h, name).argtypes = [P, P, P, C.c_size_t]
        for name in ("orc_iv_mul", "orc_iv_mul_case"):
            getattr(h, name).restype = None
            getattr(h, name).argtypes = [P, P, P]
        h.orc_iv_mul_s.restype = None
        h.orc_iv_mul_s.argtypes = [P, d, P]
        h.orc_iv_sqr.restype = None
        h.orc_iv_sqr.argtypes = [P, P]
        h.orc_slot_bounds.restype = None
        h.orc_slot_bounds.argtypes = [C.c_int, P, C.c_size_t, C.c_size_t, P]
        h.orc_sample_hash.restype = C.c_uint64
        h.orc_sample_hash.argtypes = [C.c_uint64] * 4
        h.orc_draw.restype = d
        h.orc_draw.argtypes = [C.c_int, d, d, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
        h.orc_last_error.restype = C.c_char_p
        h.orc_eval_boxes.restype = C.c_int
        h.orc_eval_boxes.argtypes = [P, P, P, C.c_size_t, C.c_size_t, P, P, P, P]
        h.orc_eval_boxes_ru.restype = C.c_int
        h.orc_eval_boxes_ru.argtypes = [P, P, P, C.c_size_t, C.c_size_t, P, P, P]
        h.orc_eval_points.restype = C.c_int
        h.orc_eval_points.argtypes = [P, P, C.c_size_t, C.c_size_t, P, P, P]
        h.orc_bound.restype = C.c_int
        h.orc_bound.argtypes = [P, P, P, P, P, C.c_size_t, P, P]
        h.orc_aggregate.restype = C.c_int
        h.orc_aggregate.argtypes = [P, P, P, P, P, C.c_size_t, C.c_size_t, P, P, P, P, P, P]
        h.orc_candidate.restype = C.c_int
        h.orc_candidate.argtypes = [P, P, P, P, P, C.c_size_t, C.c_size_t, C.c_uint64, P, P, P, P]
        for name in ("orc_fma_ru", "orc_fma_rd"):
            getattr(h, name).restype = d
            getattr(h, name).argtypes = [d, d, d]
        for name in ("orc_div_ru", "orc_div_rd"):
            getattr(h, name).restype = d
            getattr(h, name).argtypes = [d, d]
        for name in ("orc_sqrt_ru", "orc_sqrt_rd"):
            getattr(h, name).restype = d
            getattr(h, name).argtypes = [d]
        h.orc_iv_ext.restype = C.c_int
        h.orc_iv_ext.argtypes = [C.c_int, P, P, P, P, C.c_int, C.c_size_t, P, P, P]
        h.orc_pt_ext.restype = C.c_int
        h.orc_pt_ext.argtypes = [C.c_int, P, P, C.c_int, C.c_size_t, P]
        h.orc_solve.restype = C.c_int
        h.orc_solve.argtypes = [P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint32, d, P, P]
        h.orc_solve_opts.restype = C.c_int
        h.orc_solve_opts.argtypes = [P, P, P, P]
        h.orc_aggregate_solve.restype = C.c_int
        h.orc_aggregate_solve.argtypes = [P, P, P, P, P]
        _lib = h
    return _lib


def _ok(rc: int) -> None:
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def iv(fn: str, *args):
    out = np.zeros(2)
    arrs = [np.ascontiguousarray(a, np.float64) if isinstance(a, (list, tuple, np.ndarray)) else a for a in args]
    cargs = [_p(a) if isinstance(a, np.ndarray) else a for a in arrs]
    getattr(lib(), fn)(*cargs, _p(out))
    return out


def eval_boxes(arrays, lb, ub, roots=False):
    d = arrays.desc()
    lb = np.ascontiguousarray(lb, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    nv, n = lb.shape
    feas = np.zeros(n, np.uint8)
    olb = np.zeros(n)
    oub = np.zeros(n)
    r = np.zeros((1 + arrays.n_constraints, n, 2)) if roots else None
    _ok(lib().orc_eval_boxes(C.byref(d), _p(lb), _p(ub), n, n, _p(feas), _p(olb), _p(oub), _p(r)))
    return feas, olb, oub, r


def eval_boxes_ru(arrays, lb, ub):
    """orc_eval_boxes with the rounding mode set once per call (the CPU
    baseline's evaluator; QCQP op set only). lb, ub: (n_vars, n) views of one
    array may be strided: ld is their row stride."""
    d = arrays.desc()
    lb = np.ascontiguousarray(lb, np.float64)
    ub = np.ascontiguousarray(ub, np.float64)
    nv, n = lb.shape
    feas = np.zeros(n, np.uint8)
    olb = np.zeros(n)
    oub = np.zeros(n)
    _ok(lib().orc_eval_boxes_ru(C.byref(d), _p(lb), _p(ub), n, n, _p(feas), _p(olb), _p(oub)))
    return feas, olb, oub


def eval_points(arrays, x, roots=False):
    d = arrays.desc()
    x = np.ascontiguousarray(x, np.float64)
    nv, n = x.shape
    feas = np.zeros(n, np.uint8)
    obj = np.zeros(n)
    r = np.zeros((1 + arrays.n_constraints, n)) if roots else None
    _ok(lib().orc_eval_points(C.byref(d), _p(x), n, n, _p(feas), _p(obj), _p(r)))
    return feas, obj, r


def bound(arrays, box, var_ids, widths, kinds):
    d = arrays.desc()
    box = np.ascontiguousarray(box, np.float64).reshape(-1, 2)
    vid = np.ascontiguousarray(var_ids, np.uint64)
    wid = np.ascontiguousarray(widths, np.uint64)
    kd = np.ascontiguousarray(kinds, np.uint8)
    n = int(np.prod(wid)) if len(wid) else 1
    feas = np.zeros(n, np.uint8)
    olb = np.zeros(n)
    _ok(lib().orc_bound(C.byref(d), _p(box), _p(vid), _p(wid), _p(kd), len(vid), _p(feas), _p(olb)))
    return feas, olb


def candidate(arrays, box, var_ids, widths, kinds, samples, salt):
    """samples == 0: the enumerator; otherwise the sampler with that many points per region."""
    d = arrays.desc()
    box = np.ascontiguousarray(box, np.float64).reshape(-1, 2)
    vid = np.ascontiguousarray(var_ids, np.uint64)
    wid = np.ascontiguousarray(widths, np.uint64)
    kd = np.ascontiguousarray(kinds, np.uint8)
    found = C.c_int()
    value = C.c_double()
    index = C.c_int64()
    point = np.zeros(arrays.n_vars)
    _ok(lib().orc_candidate(C.byref(d), _p(box), _p(vid), _p(wid), _p(kd), len(vid), samples, salt,
                            C.byref(found), C.byref(value), C.byref(index), _p(point)))
    return bool(found.value), value.value, index.value, point


def solve(arrays, partition_num=2, enumerate_cap=2, max_cycle_size=0, sample_points=1, iter_limit=1000000,
          tolerance=1e-6, host_budget_bytes=0, compact_frontier=False, known_primal=None, policy=0,
          bisection_budget=0):
    """The oracle's best-first search with the engine's solve options. A host
    budget of 0 means none here (the engine reads 0 as half of MemAvailable).
    `policy` is a cmi_policy value; `bisection_budget` is the B of the
    bisection-budget policies."""
    from cuminlpsearch_amd import capi as _c

    d = arrays.desc()
    o = _c.SolveOptions()
    o.partition_num = partition_num
    o.enumerate_cap = enumerate_cap
    o.max_cycle_size = max_cycle_size
    o.sample_points = sample_points
    o.iter_limit = iter_limit
    o.tolerance = tolerance
    o.host_budget_bytes = host_budget_bytes
    o.compact_frontier = int(compact_frontier)
    o.policy = policy
    o.bisection_budget = bisection_budget
    if known_primal is not None:
        o.has_known_primal = 1
        o.known_primal = known_primal
    out = _c.SolveOutcome()
    best = np.zeros(arrays.n_vars)
    _ok(lib().orc_solve_opts(C.byref(d), C.byref(o), C.byref(out), _p(best)))
    return out, best


def aggregate_solve(arrays, **opts):
    """The oracle's k-child search with the engine's aggregate options:
    (outcome, best point, whether the selection rule left its lb fallback)."""
    from cuminlpsearch_amd import engine as _e
    o = _e.aggregate_options(**opts)
    from cuminlpsearch_amd import capi as _c
    d = arrays.desc()
    out = _c.SolveOutcome()
    best = np.zeros(arrays.n_vars)
    engaged = C.c_int(0)
    _ok(lib().orc_aggregate_solve(C.byref(d), C.byref(o), C.byref(out), _p(best), C.byref(engaged)))
    return out, best, bool(engaged.value)


def _vec(fn, a, b):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), np.broadcast_shapes(np.shape(a), np.shape(b))))
    b = np.ascontiguousarray(np.broadcast_to(np.asarray(b, np.float64), a.shape))
    out = np.empty(a.shape)
    fn(a.ctypes.data, b.ctypes.data, out.ctypes.data, a.size)
    return out


def add_ru_n(a, b) -> np.ndarray:
    """Element-wise RU(a + b)."""
    return _vec(lib().orc_add_ru_n, a, b)


def mul_ru_n(a, b) -> np.ndarray:
    """Element-wise RU(a * b)."""
    return _vec(lib().orc_mul_ru_n, a, b)


def aggregate(arrays, box, var_ids, widths, kinds, k, per_region=False):
    """k-child aggregate of one launch: (lb, hull_ub, feasible) per child, and
    with per_region the per-subregion (feasible, obj_lb, obj_ub)."""
    d = arrays.desc()
    box = np.ascontiguousarray(box, np.float64).reshape(-1, 2)
    vid = np.ascontiguousarray(var_ids, np.uint64)
    wid = np.ascontiguousarray(widths, np.uint64)
    kin = np.ascontiguousarray(kinds, np.uint8)
    n = int(np.prod(wid)) if wid.size else 1
    lb = np.zeros(k)
    ub = np.zeros(k)
    fe = np.zeros(k, np.int32)
    rf = np.zeros(n, np.uint8) if per_region else None
    rl = np.zeros(n) if per_region else None
    ru = np.zeros(n) if per_region else None
    _ok(lib().orc_aggregate(C.byref(d), _p(box), _p(vid), _p(wid), _p(kin), vid.size, k, _p(lb), _p(ub), _p(fe),
                            _p(rf), _p(rl), _p(ru)))
    if per_region:
        return lb, ub, fe, (rf, rl, ru)
    return lb, ub, fe


def iv_ext(op: int, a_lb, a_ub, b_lb=None, b_ub=None, n_exp: int = 0):
    """Extended interval op n at a time: (lb, ub, ok) arrays."""
    a_lb = np.ascontiguousarray(a_lb, np.float64)
    a_ub = np.ascontiguousarray(a_ub, np.float64)
    n = a_lb.shape[0]
    if b_lb is not None:
        b_lb = np.ascontiguousarray(b_lb, np.float64)
        b_ub = np.ascontiguousarray(b_ub, np.float64)
    lo, hi, ok = np.zeros(n), np.zeros(n), np.zeros(n, np.uint8)
    rc = lib().orc_iv_ext(op, _p(a_lb), _p(a_ub), _p(b_lb), _p(b_ub), int(n_exp), n, _p(lo), _p(hi), _p(ok))
    if rc != 0:
        raise ValueError(f"oracle: op {op} has no extended interval form")
    return lo, hi, ok.astype(bool)


def pt_ext(op: int, a, b=None, n_exp: int = 0) -> np.ndarray:
    """Extended point op (round to nearest), n at a time."""
    a = np.ascontiguousarray(a, np.float64)
    if b is not None:
        b = np.ascontiguousarray(b, np.float64)
    out = np.zeros(a.shape[0])
    rc = lib().orc_pt_ext(op, _p(a), _p(b), int(n_exp), a.shape[0], _p(out))
    if rc != 0:
        raise ValueError(f"oracle: op {op} has no extended point form")
    return out
