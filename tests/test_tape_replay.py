"""CPU checks of the tape compiler: the words the device runs, replayed on the
host in the device's (-lb, ub) round-toward-+infinity form, give the oracle's
node-by-node results bit for bit, for every instruction form and for every
split of the functions into groups."""
import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from oracle import binding as orc

import op_cases
import problems as P
import tape_emulator as T


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def check(eng, lb, ub, groups):
    f_o, lo_o, hi_o, r_o = orc.eval_boxes(eng.arrays(), lb, ub, roots=True)
    f_e, lo_e, hi_e, r_e = T.eval_boxes(eng, lb, ub, groups)
    assert np.array_equal(f_e, f_o)
    assert same_bits(lo_e, lo_o)
    assert same_bits(hi_e, hi_o)
    assert same_bits(r_e, r_o)


ALL_CASES = {**op_cases.CASES, **op_cases.EXT_CASES}


@pytest.mark.parametrize("name", sorted(ALL_CASES))
def test_op_case_replay_matches_oracle(name):
    eng = cmi.EngineProblem.from_problem(ALL_CASES[name]())
    lb, ub = P.random_boxes(5, 3, 400, -3.0, 4.0, 1e-6, 6.0)
    check(eng, lb, ub, 1)


def test_op_cases_reach_every_op_code():
    seen = set()
    for build in ALL_CASES.values():
        words, _, _ = cmi.EngineProblem.from_problem(build()).tape_words(1)
        seen |= {int(o) for o in words["op"]}
    generic = set(range(T.OPS["LoadConst"], T.OPS["Output"] + 1))
    acc = set(range(T.OPS["AccStore"], T.OPS["AccOutput"] + 1))
    terms = {T.TERM_BASE + 8 * m + t for m in range(3) for t in range(6)}
    ext = set(range(T.EXT_BASE, T.EXT_END))
    assert generic | acc | terms | ext <= seen


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("groups", [1, 2, 3, 8])
def test_random_dag_replay_matches_oracle(seed, groups):
    p = P.random_dag(seed, n_vars=4 + seed % 5, n_ops=40 + 10 * (seed % 4), n_con=6)
    eng = cmi.EngineProblem.from_problem(p)
    lb, ub = P.random_boxes(100 + seed, eng.info().n_vars, 300, -3.0, 3.0, 1e-6, 6.0)
    check(eng, lb, ub, groups)


def test_c2_grouped_replay_matches_oracle():
    eng = P.c2()
    lb, ub = P.random_boxes(11, 32, 64, -5.0, 5.0)
    for groups in (1, 8):
        check(eng, lb, ub, groups)


def test_group_split_shape():
    """Groups partition the functions: each root's check or read-out appears in
    exactly one group; temporaries of different groups do not overlap."""
    eng = P.c2()
    words, begin, info = eng.tape_words(8)
    assert info.box_groups == 8 and begin[0] == 0 and begin[-1] == words.size
    out_ops = {T.OPS["Output"], T.OPS["AccOutput"]}
    outs = [int(w["b"]) for w in words if int(w["op"]) in out_ops]
    assert sorted(outs) == list(range(info.n_outputs))
    n_vars = int(info.n_vars)
    temps = []
    for g in range(8):
        seg = words[begin[g]:begin[g + 1]]
        written = {int(w["dst"]) for w in seg if int(w["op"]) in (T.OPS["AccStore"],) or
                   T.OPS["LoadConst"] <= int(w["op"]) <= T.OPS["Neg"]}
        temps.append({s for s in written if s >= n_vars})
    for g in range(8):
        for h in range(g + 1, 8):
            assert not (temps[g] & temps[h])
    assert max((max(t) for t in temps if t), default=-1) < info.box_slots


def test_compile_words_rejects_small_buffer():
    import ctypes as C
    from cuminlpsearch_amd import capi
    eng = P.c1(0)
    info = capi.TapeInfo()
    from cuminlpsearch_amd.engine import TAPE_WORD
    buf = np.zeros(2, TAPE_WORD)
    rc = capi.lib().cmi_tape_compile_words(eng.handle, 1, capi.ptr(buf), 2, None, C.byref(info))
    assert rc == capi.SHAPE_MISMATCH


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("groups", [1, 3])
def test_random_extended_dag_replay_matches_oracle(seed, groups):
    p = P.random_ext_dag(seed, n_vars=3 + seed % 4, n_ops=30 + 5 * (seed % 3), n_con=5)
    eng = cmi.EngineProblem.from_problem(p)
    lb, ub = P.random_boxes(200 + seed, eng.info().n_vars, 300, -3.0, 3.0, 1e-9, 6.0)
    check(eng, lb, ub, groups)
