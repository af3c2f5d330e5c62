"""One small problem per tape instruction form, so that a wrong device
operation shows up as the case that names it. Shared by the CPU replay test
(which also checks that together the cases reach every tape op code) and the
device test."""
from __future__ import annotations

from cuminlpsearch_amd import capi as C
from cuminlpsearch_amd import model as M
from cuminlpsearch_amd.model import Problem


def _xyz():
    p = Problem()
    return p, p.var(-3.0, 2.0), p.var(-1.5, 4.0), p.var(0.5, 3.0)


def _case(expr_fn, cmp=C.LE, rhs=0.75, second=None):
    def build():
        p, x, y, z = _xyz()
        e = expr_fn(p, x, y, z)
        p.set_objective(e)
        p.add_constraint(e, cmp, rhs)
        if second is not None:
            e2 = second(p, x, y, z)
            p.add_constraint(e2, C.EQ, 0.5)
        return p
    return build


# A node that stays in a slot (read twice, not a leaf) makes the compiler
# store it and run the generic form of the op that consumes it.
def _shared(p, x, y, z):
    return (x * y) + z


CASES = {
    "add": _case(lambda p, x, y, z: x + y),
    "sub": _case(lambda p, x, y, z: x - y),
    "mul": _case(lambda p, x, y, z: x * y),
    "sqr": _case(lambda p, x, y, z: x * x),
    "neg": _case(lambda p, x, y, z: -x),
    "add_imm": _case(lambda p, x, y, z: x + 1.25),
    "imm_add": _case(lambda p, x, y, z: 1.25 + x),
    "sub_imm": _case(lambda p, x, y, z: x - 0.625),
    "imm_sub": _case(lambda p, x, y, z: 0.625 - x),
    "mul_imm_pos": _case(lambda p, x, y, z: 1.5 * x),
    "mul_imm_neg": _case(lambda p, x, y, z: x * -2.5),
    "mul_imm_zero": _case(lambda p, x, y, z: x * 0.0 + y),
    "sum3": _case(lambda p, x, y, z: x + y + z),
    "diff3": _case(lambda p, x, y, z: x - y - z),
    "bilinear_scaled": _case(lambda p, x, y, z: -0.75 * (x * y) + 2.0 * (y * z)),
    "square_scaled": _case(lambda p, x, y, z: 3.0 * (x * x) - 0.5 * (z * z)),
    "acc_add_imm": _case(lambda p, x, y, z: (x + y) + 0.5),
    "acc_sub_imm": _case(lambda p, x, y, z: (x + y) - 0.5),
    "acc_imm_sub": _case(lambda p, x, y, z: 0.5 - (x + y)),
    "acc_scale": _case(lambda p, x, y, z: (x + y) * -1.75),
    "acc_neg": _case(lambda p, x, y, z: -(x + y)),
    "eq_leaf": _case(lambda p, x, y, z: x - y, cmp=C.EQ, rhs=0.5),
    "eq_acc": _case(lambda p, x, y, z: x + y + z, cmp=C.EQ, rhs=2.0),
    "const_root": _case(lambda p, x, y, z: p.fixed(1.5) * p.fixed(2.0) + x),
    "shared_add": _case(lambda p, x, y, z: _shared(p, x, y, z) * _shared(p, x, y, z) + _shared(p, x, y, z)),
    "shared_generic": _case(lambda p, x, y, z: (lambda s: (s - y) * (s + 1.0) - (2.0 - s) * -s)(_shared(p, x, y, z)),
                            second=lambda p, x, y, z: (lambda s: s * s - s)(x * z + y)),
    "shared_scaled": _case(lambda p, x, y, z: (lambda s: s * 0.5 + (s - 0.25) + (0.75 - s) + s * s)(x * y - z)),
    "var_root": _case(lambda p, x, y, z: x, second=lambda p, x, y, z: y),
    "add_scaled": _case(lambda p, x, y, z: (x + y) + 2.0 * z),
    "acc_forms": lambda: _separate(lambda p, x, y, z: x + 2.0 * y - 1.5 * z + 3.0 * (y * y) - x * x
                                   - 0.5 * (x * z) + y * z - z * z,
                                   lambda p, x, y, z: 2.0 * x - y * 0.5 + (x * y) * 1.5 - 2.0 * (z * z)),
}


# The extended op set: one case per word form, over boxes that reach every
# domain case (partly or wholly outside the domain, zero inside or at an end).
EXT_CASES = {
    "div": _case(lambda p, x, y, z: x / y),
    "div_pos": _case(lambda p, x, y, z: (x + y) / z),
    "div_imm": _case(lambda p, x, y, z: x / 1.5),
    "imm_div": _case(lambda p, x, y, z: 2.0 / y),
    "pown_even": _case(lambda p, x, y, z: M.pown(x, 4) - y),
    "pown_odd": _case(lambda p, x, y, z: M.pown(y, 3) + x),
    "pown_neg_even": _case(lambda p, x, y, z: M.pown(y, -2)),
    "pown_neg_odd": _case(lambda p, x, y, z: M.pown(x, -3) + z),
    "pown_big": _case(lambda p, x, y, z: M.pown(z, 17) * 1e-6),
    "sqrt": _case(lambda p, x, y, z: M.sqrt(x) + M.sqrt(z)),
    "exp": _case(lambda p, x, y, z: M.exp(x * y) - z),
    "log": _case(lambda p, x, y, z: M.log(y) + M.log(z)),
    "sin": _case(lambda p, x, y, z: M.sin(3.0 * x) + y),
    "cos": _case(lambda p, x, y, z: M.cos(x * z) - 0.5),
    "tanh": _case(lambda p, x, y, z: M.tanh(x * y), cmp=C.EQ, rhs=0.25),
    "abs": _case(lambda p, x, y, z: M.fabs(x) - M.fabs(y)),
    "min": _case(lambda p, x, y, z: M.minimum(x, y)),
    "min_imm": _case(lambda p, x, y, z: M.minimum(x, 0.5) + M.minimum(0.25, z)),
    "max": _case(lambda p, x, y, z: M.maximum(x, y) * z),
    "max_imm": _case(lambda p, x, y, z: M.maximum(y, -0.5) + M.maximum(1.0, x)),
    "pow": _case(lambda p, x, y, z: M.power(z, x)),
    "pow_const_exp": _case(lambda p, x, y, z: M.power(y, 2.5)),
    "pow_const_base": _case(lambda p, x, y, z: M.power(p.fixed(1.5), x * z)),
    "ext_mixed": _case(lambda p, x, y, z: M.exp(M.sin(x) / z) + M.log(z + M.fabs(y)) * M.tanh(x),
                       second=lambda p, x, y, z: M.sqrt(z * z + x * x) - M.pown(y, 2)),
}


def _separate(obj_fn, con_fn):
    """Objective and constraint in different accumulator chains."""
    p, x, y, z = _xyz()
    p.set_objective(obj_fn(p, x, y, z))
    p.add_constraint(con_fn(p, x, y, z), C.LE, 1.0)
    return p
