"""Host-side problem builder mirroring the reference's Expr/Problem API.

Same construction rules as the C++ model (and the reference's
model/problem.hpp:35-358): ``x * x`` on the same node emits Sqr, a Python
number operand becomes a Const node consumed by value, unary ops on a Const
fold. ``Problem.arrays()`` spells the DAG as the flat arrays of
``cmi_problem_desc`` (include/cmi.h), which is what crosses the C-ABI.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import capi as _c


@dataclass
class _Node:
    op: int
    ins: tuple
    constant: float = 0.0
    var_index: int = 0
    int_exp: int = 0


class Expr:
    __slots__ = ("p", "id")

    def __init__(self, p: "Problem", node_id: int):
        self.p = p
        self.id = node_id

    # -- helpers -----------------------------------------------------------
    def _wrap(self, other) -> "Expr":
        if isinstance(other, Expr):
            return other
        return self.p.fixed(float(other))

    def _is_const(self) -> bool:
        return self.p.nodes[self.id].op == _c.OP_CONST

    def _binary(self, op: int, other: "Expr") -> "Expr":
        return Expr(self.p, self.p._emit(op, (self.id, other.id)))

    # -- operators ---------------------------------------------------------
    def __add__(self, o):
        return self._binary(_c.OP_ADD, self._wrap(o))

    def __radd__(self, o):
        return self._wrap(o)._binary(_c.OP_ADD, self)

    def __sub__(self, o):
        return self._binary(_c.OP_SUB, self._wrap(o))

    def __rsub__(self, o):
        return self._wrap(o)._binary(_c.OP_SUB, self)

    def __mul__(self, o):
        o = self._wrap(o)
        if o.id == self.id:
            if self._is_const():
                c = self.p.nodes[self.id].constant
                return self.p.fixed(c * c)
            return Expr(self.p, self.p._emit(_c.OP_SQR, (self.id,)))
        return self._binary(_c.OP_MUL, o)

    def __rmul__(self, o):
        return self._wrap(o).__mul__(self)

    def __truediv__(self, o):
        return self._binary(_c.OP_DIV, self._wrap(o))

    def __rtruediv__(self, o):
        return self._wrap(o)._binary(_c.OP_DIV, self)

    def __neg__(self):
        if self._is_const():
            return self.p.fixed(-self.p.nodes[self.id].constant)
        return Expr(self.p, self.p._emit(_c.OP_NEG, (self.id,)))


def sqr(e: Expr) -> Expr:
    return e * e


def _unary(op: int, fold, e: Expr) -> Expr:
    """A unary op, folded to a Const on a Const operand (model/problem.hpp)."""
    if e._is_const():
        return e.p.fixed(fold(e.p.nodes[e.id].constant))
    return Expr(e.p, e.p._emit(op, (e.id,)))


def _fold(fn):
    def safe(v):
        with np.errstate(all="ignore"):
            return float(fn(np.float64(v)))
    return safe


def exp(e: Expr) -> Expr:
    return _unary(_c.OP_EXP, _fold(np.exp), e)


def log(e: Expr) -> Expr:
    return _unary(_c.OP_LOG, _fold(np.log), e)


def sqrt(e: Expr) -> Expr:
    return _unary(_c.OP_SQRT, _fold(np.sqrt), e)


def sin(e: Expr) -> Expr:
    return _unary(_c.OP_SIN, _fold(np.sin), e)


def cos(e: Expr) -> Expr:
    return _unary(_c.OP_COS, _fold(np.cos), e)


def tanh(e: Expr) -> Expr:
    return _unary(_c.OP_TANH, _fold(np.tanh), e)


def fabs(e: Expr) -> Expr:
    return _unary(_c.OP_ABS, _fold(np.fabs), e)


def pown(e: Expr, n: int) -> Expr:
    """e**n for an integer n (Op::PowN)."""
    if e._is_const():
        return e.p.fixed(_fold(lambda v: np.power(v, float(n)))(e.p.nodes[e.id].constant))
    return Expr(e.p, e.p._emit(_c.OP_POWN, (e.id,), int_exp=int(n)))


def power(a: Expr, b) -> Expr:
    """a**b for an expression exponent (Op::Pow)."""
    b = a._wrap(b)
    if a._is_const() and b._is_const():
        return a.p.fixed(_fold(lambda v: np.power(v, a.p.nodes[b.id].constant))(a.p.nodes[a.id].constant))
    return a._binary(_c.OP_POW, b)


def _pair(a, b):
    if not isinstance(a, Expr):
        a = b._wrap(a)
    return a, a._wrap(b)


def minimum(a, b) -> Expr:
    a, b = _pair(a, b)
    return a._binary(_c.OP_MIN, b)


def maximum(a, b) -> Expr:
    a, b = _pair(a, b)
    return a._binary(_c.OP_MAX, b)


@dataclass
class Problem:
    nodes: list = field(default_factory=list)
    box: list = field(default_factory=list)
    kinds: list = field(default_factory=list)
    objective_root: int = -1
    constraints: list = field(default_factory=list)  # (root, cmp, rhs)

    def _emit(self, op: int, ins: tuple, **payload) -> int:
        self.nodes.append(_Node(op, tuple(ins), **payload))
        return len(self.nodes) - 1

    def var(self, lb: float, ub: float, kind: int = _c.CONTINUOUS) -> Expr:
        idx = len(self.box)
        self.box.append((float(lb), float(ub)))
        self.kinds.append(kind)
        return Expr(self, self._emit(_c.OP_VAR, (), var_index=idx))

    def int_var(self, lb: float, ub: float) -> Expr:
        return self.var(lb, ub, _c.INTEGER)

    def bin_var(self) -> Expr:
        return self.var(0.0, 1.0, _c.BINARY)

    def fixed(self, value: float) -> Expr:
        return Expr(self, self._emit(_c.OP_CONST, (), constant=float(value)))

    def set_objective(self, e: Expr) -> None:
        self.objective_root = e.id

    def add_constraint(self, lhs: Expr, cmp: int, rhs: float) -> None:
        self.constraints.append((lhs.id, cmp, float(rhs)))

    @property
    def n_vars(self) -> int:
        return len(self.box)

    def arrays(self) -> "ProblemArrays":
        n = len(self.nodes)
        a = ProblemArrays(
            op=np.array([nd.op for nd in self.nodes], dtype=np.uint8),
            in0=np.array([nd.ins[0] if len(nd.ins) > 0 else 0 for nd in self.nodes], dtype=np.uint64),
            in1=np.array([nd.ins[1] if len(nd.ins) > 1 else 0 for nd in self.nodes], dtype=np.uint64),
            constant=np.array([nd.constant for nd in self.nodes], dtype=np.float64),
            var_index=np.array([nd.var_index for nd in self.nodes], dtype=np.uint64),
            int_exp=np.array([nd.int_exp for nd in self.nodes], dtype=np.int32),
            box=np.array(self.box, dtype=np.float64).reshape(-1, 2),
            var_kind=np.array(self.kinds, dtype=np.uint8),
            objective_root=self.objective_root if self.objective_root >= 0 else 2**64 - 1,
            con_root=np.array([c[0] for c in self.constraints], dtype=np.uint64),
            con_cmp=np.array([c[1] for c in self.constraints], dtype=np.uint8),
            con_rhs=np.array([c[2] for c in self.constraints], dtype=np.float64),
        )
        assert a.op.shape[0] == n
        return a


@dataclass
class ProblemArrays:
    """A Problem as the flat arrays of ``cmi_problem_desc``."""

    op: np.ndarray
    in0: np.ndarray
    in1: np.ndarray
    constant: np.ndarray
    var_index: np.ndarray
    int_exp: np.ndarray
    box: np.ndarray  # (n_vars, 2)
    var_kind: np.ndarray
    objective_root: int
    con_root: np.ndarray
    con_cmp: np.ndarray
    con_rhs: np.ndarray

    @property
    def n_vars(self) -> int:
        return int(self.box.shape[0])

    @property
    def n_constraints(self) -> int:
        return int(self.con_root.shape[0])

    def desc(self) -> _c.ProblemDesc:
        """A cmi_problem_desc viewing these arrays (keep self alive while it is used)."""
        import ctypes as C

        for name in ("op", "in0", "in1", "constant", "var_index", "int_exp", "box", "var_kind", "con_root",
                     "con_cmp", "con_rhs"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name)))
        d = _c.ProblemDesc()
        d.n_nodes = int(self.op.shape[0])
        d.op = self.op.ctypes.data_as(C.POINTER(C.c_uint8))
        d.in0 = self.in0.ctypes.data_as(C.POINTER(C.c_uint64))
        d.in1 = self.in1.ctypes.data_as(C.POINTER(C.c_uint64))
        d.constant = self.constant.ctypes.data_as(C.POINTER(C.c_double))
        d.var_index = self.var_index.ctypes.data_as(C.POINTER(C.c_uint64))
        d.int_exp = self.int_exp.ctypes.data_as(C.POINTER(C.c_int32))
        d.n_vars = self.n_vars
        d.box = self.box.ctypes.data_as(C.POINTER(_c.Ivl))
        d.var_kind = self.var_kind.ctypes.data_as(C.POINTER(C.c_uint8))
        d.objective_root = int(self.objective_root)
        d.n_constraints = self.n_constraints
        d.con_root = self.con_root.ctypes.data_as(C.POINTER(C.c_uint64))
        d.con_cmp = self.con_cmp.ctypes.data_as(C.POINTER(C.c_uint8))
        d.con_rhs = self.con_rhs.ctypes.data_as(C.POINTER(C.c_double))
        return d


_UNARY = {
    _c.OP_EXP: np.exp,
    _c.OP_LOG: np.log,
    _c.OP_SQRT: np.sqrt,
    _c.OP_SIN: np.sin,
    _c.OP_COS: np.cos,
    _c.OP_TANH: np.tanh,
    _c.OP_ABS: np.fabs,
}
_BINARY = {
    _c.OP_ADD: np.add,
    _c.OP_SUB: np.subtract,
    _c.OP_MUL: np.multiply,
    _c.OP_DIV: np.divide,
    _c.OP_MIN: np.fmin,
    _c.OP_MAX: np.fmax,
    _c.OP_POW: np.power,
}


def evaluate_point(a: ProblemArrays, x) -> np.ndarray:
    """Round-to-nearest value of every node at point x, with C semantics for
    out-of-domain arguments (log(0) = -inf, sqrt(-1) = nan); a value, not a
    bound (reference model/eval.hpp:38-116)."""
    n = a.op.shape[0]
    val = np.zeros(n)
    with np.errstate(all="ignore"):
        for i in range(n):
            op = int(a.op[i])
            if op == _c.OP_VAR:
                val[i] = x[int(a.var_index[i])]
            elif op == _c.OP_CONST:
                val[i] = a.constant[i]
            elif op in _BINARY:
                val[i] = _BINARY[op](val[int(a.in0[i])], val[int(a.in1[i])])
            elif op in _UNARY:
                val[i] = _UNARY[op](val[int(a.in0[i])])
            elif op == _c.OP_SQR:
                u = val[int(a.in0[i])]
                val[i] = u * u
            elif op == _c.OP_NEG:
                val[i] = -val[int(a.in0[i])]
            elif op == _c.OP_POWN:
                val[i] = np.power(val[int(a.in0[i])], float(a.int_exp[i]))
            else:
                raise ValueError(f"evaluate_point: unknown op {op}")
    return val


def evaluate_objective(a: ProblemArrays, x) -> float:
    """Value of the objective at x (reference model/eval.hpp:165-171)."""
    return float(evaluate_point(a, x)[int(a.objective_root)])


def satisfies(lhs: float, cmp: int, rhs: float) -> bool:
    if cmp == _c.LE:
        return lhs <= rhs
    diff = abs(lhs - rhs)
    return diff <= 1e-8 or diff <= 1e-6 * max(abs(lhs), abs(rhs))


__all__ = ["Expr", "Problem", "ProblemArrays", "sqr", "exp", "log", "sqrt", "sin", "cos", "tanh", "fabs", "pown",
           "power", "minimum", "maximum", "evaluate_point", "evaluate_objective", "satisfies", "math"]
