"""Host replay of device tapes -- test infrastructure.

Runs the exact 16-byte words the device runs (EngineProblem.tape_words) over
many boxes at once, with the interval kernel's representation: an interval
[lb, ub] is held as (-lb, ub) and every bound is one round-toward-+infinity
operation (the oracle's RU primitives); lower bounds come out by exact
negation. Outputs report a zero bound as +0, as the device does.

Comparing this replay with the oracle's node-by-node evaluation checks the tape
compiler (term fusion, leaf rematerialisation, slot allocation, function
groups) and the (-lb, ub) formulas on the CPU, without a GPU.
"""
from __future__ import annotations

import numpy as np

from cuminlpsearch_amd import capi as _c
from oracle import binding as orc

OPS = dict(Nop=0, LoadConst=1, Add=2, AddImm=3, Sub=4, SubImm=5, ImmSub=6, Mul=7, MulImm=8, Sqr=9, Neg=10,
           CheckLE=11, CheckEQ=12, Objective=13, Output=14, AccStore=16, AccAddImm=17, AccSubImm=18,
           AccImmSub=19, AccScale=20, AccNeg=21, AccCheckLE=22, AccCheckEQ=23, AccObjective=24, AccOutput=25,
           Div=32, DivImm=33, ImmDiv=34, PowN=35, Sqrt=36, Exp=37, Log=38, Sin=39, Cos=40, Tanh=41, Abs=42,
           Min=43, MinImm=44, Max=45, MaxImm=46)
EXT_BASE, EXT_END = 32, 47
TERM_BASE = 64
TERM_S, TERM_SC, TERM_SS, TERM_Q, TERM_SSC, TERM_QC = range(6)
MODE_SET, MODE_ADD, MODE_SUB = range(3)

ru_add = orc.add_ru_n
ru_mul = orc.mul_ru_n


def add(a, b):
    return ru_add(a[0], b[0]), ru_add(a[1], b[1])


def sub(a, b):
    return ru_add(a[0], b[1]), ru_add(a[1], b[0])


def add_imm(a, c):
    return ru_add(a[0], -c), ru_add(a[1], c)


def sub_imm(a, c):
    return ru_add(a[0], c), ru_add(a[1], -c)


def imm_sub(c, a):
    return ru_add(a[1], -c), ru_add(c, a[0])


def neg(a):
    return a[1], a[0]


def mul(x, y):
    xn, xu = x
    yn, yu = y
    # max skipping NaN products (0 * inf), as the device's v_max_f64
    nl = np.fmax.reduce([ru_mul(-xn, yn), ru_mul(xn, yu), ru_mul(xu, yn), ru_mul(-xu, yu)])
    ub = np.fmax.reduce([ru_mul(xn, yn), ru_mul(-xn, yu), ru_mul(-xu, yn), ru_mul(xu, yu)])
    return nl, ub


def mul_imm(a, c):
    if c > 0:
        return ru_mul(a[0], c), ru_mul(a[1], c)
    if c < 0:
        return ru_mul(-a[1], c), ru_mul(-a[0], c)
    z = np.zeros_like(a[0])
    return z, z.copy()


def sqr(a):
    s = np.fmax(-a[0], np.fmin(a[1], 0.0))
    m = np.fmax(np.abs(a[0]), np.abs(a[1]))
    return ru_mul(-s, s), ru_mul(m, m)


def output(a):
    """[lb, ub] with +0 for a zero bound and the infinite bound for a NaN one."""
    nl = np.where(np.isnan(a[0]), np.inf, a[0])
    ub = np.where(np.isnan(a[1]), np.inf, a[1])
    return ru_add(-nl, 0.0), ru_add(ub, 0.0)


def check_le(a, imm):
    return ~(-a[0] > imm)


def check_eq(a, imm):
    return ~(-a[0] > imm) & ~(imm > a[1])


_EXT = {OPS["Div"]: _c.OP_DIV, OPS["DivImm"]: _c.OP_DIV, OPS["ImmDiv"]: _c.OP_DIV, OPS["PowN"]: _c.OP_POWN,
        OPS["Sqrt"]: _c.OP_SQRT, OPS["Exp"]: _c.OP_EXP, OPS["Log"]: _c.OP_LOG, OPS["Sin"]: _c.OP_SIN,
        OPS["Cos"]: _c.OP_COS, OPS["Tanh"]: _c.OP_TANH, OPS["Abs"]: _c.OP_ABS, OPS["Min"]: _c.OP_MIN,
        OPS["MinImm"]: _c.OP_MIN, OPS["Max"]: _c.OP_MAX, OPS["MaxImm"]: _c.OP_MAX}


def ext(op, a, b, imm):
    """An extended-op word through the oracle's interval op: ((nl, ub), ok)."""
    n = a[0].shape[0]
    point = (np.full(n, imm), np.full(n, imm))
    x = (-a[0], a[1])
    y = None
    if op in (OPS["Div"], OPS["Min"], OPS["Max"]):
        y = (-b[0], b[1])
    elif op in (OPS["DivImm"], OPS["MinImm"], OPS["MaxImm"]):
        y = point
    elif op == OPS["ImmDiv"]:
        x, y = point, (-a[0], a[1])
    lo, hi, ok = orc.iv_ext(_EXT[op], x[0], x[1], None if y is None else y[0], None if y is None else y[1],
                            n_exp=int(imm) if op == OPS["PowN"] else 0)
    return (-lo, hi), ok


def run_group(words, slots, n, n_outputs, roots):
    """One group's words over the slot table; returns (feasible, objective or None)."""
    feas = np.ones(n, bool)
    obj = None
    acc = (np.zeros(n), np.zeros(n))
    for w in words:
        op = int(w["op"])
        a_i, b_i, d_i, imm = int(w["a"]), int(w["b"]), int(w["dst"]), float(w["imm"])
        if op >= TERM_BASE:
            term, mode = op & 7, (op - TERM_BASE) >> 3
            a = slots[a_i]
            x = {TERM_S: lambda: a,
                 TERM_SC: lambda: mul_imm(a, imm),
                 TERM_SS: lambda: mul(a, slots[b_i]),
                 TERM_Q: lambda: sqr(a),
                 TERM_SSC: lambda: mul_imm(mul(a, slots[b_i]), imm),
                 TERM_QC: lambda: mul_imm(sqr(a), imm)}[term]()
            acc = x if mode == MODE_SET else add(acc, x) if mode == MODE_ADD else sub(acc, x)
            continue
        if OPS["AccStore"] <= op <= OPS["AccOutput"]:
            if op == OPS["AccStore"]:
                slots[d_i] = acc
            elif op == OPS["AccAddImm"]:
                acc = add_imm(acc, imm)
            elif op == OPS["AccSubImm"]:
                acc = sub_imm(acc, imm)
            elif op == OPS["AccImmSub"]:
                acc = imm_sub(imm, acc)
            elif op == OPS["AccScale"]:
                acc = mul_imm(acc, imm)
            elif op == OPS["AccNeg"]:
                acc = neg(acc)
            elif op == OPS["AccCheckLE"]:
                feas &= check_le(acc, imm)
            elif op == OPS["AccCheckEQ"]:
                feas &= check_eq(acc, imm)
            elif op == OPS["AccObjective"]:
                obj = acc
            elif op == OPS["AccOutput"]:
                roots[b_i] = output(acc)
            continue
        if op == OPS["Nop"]:
            continue
        a = slots[a_i]
        if EXT_BASE <= op < EXT_END:
            r, ok = ext(op, a, slots[b_i] if op in (OPS["Div"], OPS["Min"], OPS["Max"]) else None, imm)
            feas &= ok
            slots[d_i] = r
            continue
        if op == OPS["LoadConst"]:
            r = (np.full(n, -imm), np.full(n, imm))
        elif op == OPS["Add"]:
            r = add(a, slots[b_i])
        elif op == OPS["AddImm"]:
            r = add_imm(a, imm)
        elif op == OPS["Sub"]:
            r = sub(a, slots[b_i])
        elif op == OPS["SubImm"]:
            r = sub_imm(a, imm)
        elif op == OPS["ImmSub"]:
            r = imm_sub(imm, a)
        elif op == OPS["Mul"]:
            r = mul(a, slots[b_i])
        elif op == OPS["MulImm"]:
            r = mul_imm(a, imm)
        elif op == OPS["Sqr"]:
            r = sqr(a)
        elif op == OPS["Neg"]:
            r = neg(a)
        elif op == OPS["CheckLE"]:
            feas &= check_le(a, imm)
            continue
        elif op == OPS["CheckEQ"]:
            feas &= check_eq(a, imm)
            continue
        elif op == OPS["Objective"]:
            obj = a
            continue
        elif op == OPS["Output"]:
            roots[b_i] = output(a)
            continue
        else:
            raise ValueError(f"unknown tape op {op}")
        slots[d_i] = r
    return feas, obj


def eval_boxes(eng, lb, ub, groups: int = 1):
    """Replay of the explicit-box kernel: (feasible, obj_lb, obj_ub, roots[k, e, 2])."""
    words, begin, info = eng.tape_words(groups)
    lb = np.asarray(lb, np.float64)
    ub = np.asarray(ub, np.float64)
    n = lb.shape[1]
    n_slots = max(int(info.box_slots), 1)
    slots = [None] * n_slots
    for v in range(int(info.n_vars)):
        slots[v] = (-lb[v], ub[v].copy())
    roots = [None] * int(info.n_outputs)
    feas = np.ones(n, bool)
    obj = (np.zeros(n), np.zeros(n))
    for g in range(int(info.box_groups)):
        f, o = run_group(words[begin[g]:begin[g + 1]], slots, n, info.n_outputs, roots)
        feas &= f
        if o is not None:
            obj = o
    olb, oub = output(obj)
    r = np.zeros((int(info.n_outputs), n, 2))
    for k, v in enumerate(roots):
        if v is not None:
            r[k, :, 0], r[k, :, 1] = v
    return feas.astype(np.uint8), olb, oub, r
