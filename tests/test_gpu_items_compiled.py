"""The decode-mode kernels compiled from the tape (jit::items_source) against
the interpreted ones and the oracle: per-region bounder verdicts and
objective lower bounds, and the k-child aggregate, bit for bit, for QCQPs,
MINLPLib models (the extended op set included), integer slots and partial
tiles; and the aggregate search driven by the compiled kernel equals the
oracle's."""
import pathlib

import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as C, gams
from cuminlpsearch_amd import workloads as W
from oracle import binding as orc

import aggregate_cases as A

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def problems():
    yield "c1-3", W.c1(3)
    yield "qcqp-n32", cmi.EngineProblem.qcqp(11, n=32, m=16, cont_lb=-5.0, cont_ub=5.0, pair_density=0.2)
    for name in ("ex2_1_1", "ex8_1_1", "hs62", "nvs01", "ex4_1_2", "st_bpaf1a"):
        yield name, gams.parse_file(GOLDEN / "gams" / f"{name}.gms").problem


def tapes(ctx, p):
    ti = cmi.Tape(ctx, p)
    ti.set_box_engine(C.BOX_INTERPRETED)
    tc = cmi.Tape(ctx, p)
    tc.set_box_engine(C.BOX_COMPILED)
    return ti, tc


@pytest.mark.parametrize("n_regions", [4096, 1000])
def test_compiled_bounder_equals_interpreted_and_oracle(ctx, n_regions):
    for name, p in problems():
        a = p.arrays()
        kinds_v = list(a.var_kind)
        # 1000 regions: 100 x 10 on two variables, a partial last tile
        kinds, vids, widths = A.partition(a.box, 4 if n_regions == 4096 else 10, n_regions, kinds_v)
        print(f"{name}: slots {kinds} widths {widths}", flush=True)
        ti, tc = tapes(ctx, p)
        fi, li = cmi.Role(ti, C.ROLE_BOUNDER, kinds, n_regions).bound(a.box, vids, widths)
        fc, lc = cmi.Role(tc, C.ROLE_BOUNDER, kinds, n_regions).bound(a.box, vids, widths)
        assert np.array_equal(fi, fc), name
        assert same_bits(li, lc), name
        fo, lo = orc.bound(a, a.box, vids, widths, kinds)
        assert np.array_equal(fc, fo) and same_bits(lc, lo), name


def test_compiled_aggregate_equals_interpreted_and_oracle(ctx):
    for name, p in problems():
        a = p.arrays()
        kinds, vids, widths = A.partition(a.box, 8, 8192, list(a.var_kind))
        ti, tc = tapes(ctx, p)
        ri = cmi.AggregateBounder(ti, kinds, 8192, 8).bound_children(a.box, vids, widths)
        rc = cmi.AggregateBounder(tc, kinds, 8192, 8).bound_children(a.box, vids, widths)
        for x, y in zip(ri, rc):
            assert same_bits(x, y), name
        lo, uo, fo = orc.aggregate(a, a.box, vids, widths, kinds, 8)
        assert same_bits(rc[0][0], lo) and same_bits(rc[1][0], uo) and np.array_equal(rc[2][0], fo), name


def test_compiled_aggregate_with_integer_and_enumerate_slots(ctx):
    """nvs01 (integers): an enumerate slot in the refine and an integer
    partition branch."""
    p = gams.parse_file(GOLDEN / "gams" / "nvs01.gms").problem
    a = p.arrays()
    kinds_v = list(a.var_kind)
    ints = [i for i, k in enumerate(kinds_v) if k != C.CONTINUOUS]
    cont = [i for i, k in enumerate(kinds_v) if k == C.CONTINUOUS]
    kinds = [C.SLOT_INTEGER_ENUMERATE, C.SLOT_CONTINUOUS, C.SLOT_INTEGER_PARTITION]
    vids = [ints[0], cont[0], ints[1]]
    widths = [16, 256, 4]
    ti, tc = tapes(ctx, p)
    ri = cmi.AggregateBounder(ti, kinds, 16384, 4).bound_children(a.box, vids, widths)
    rc = cmi.AggregateBounder(tc, kinds, 16384, 4).bound_children(a.box, vids, widths)
    for x, y in zip(ri, rc):
        assert same_bits(x, y)
    lo, uo, fo = orc.aggregate(a, a.box, vids, widths, kinds, 4)
    assert same_bits(rc[0][0], lo) and same_bits(rc[1][0], uo) and np.array_equal(rc[2][0], fo)


def test_aggregate_search_on_the_compiled_kernel_equals_the_oracle(ctx):
    """The search with its bound launches on the compiled kernel from the
    first iteration, against the oracle; and under the automatic choice,
    which switches from the interpreter to the compiled kernel mid-solve,
    against the interpreter throughout."""
    from test_gpu_aggregate_search import same
    for name in ("ex8_1_1", "st_e01", "nvs01"):
        pm = gams.parse_file(GOLDEN / "gams" / f"{name}.gms")
        opts = dict(partition_budget=1 << 15, sample_fan_out=64, sample_points=2, iter_limit=25)
        d, bd, _ = cmi.aggregate_solve(ctx, pm.problem, kernel=C.BOX_COMPILED, **opts)
        o, bo, _ = orc.aggregate_solve(pm.problem.arrays(), **opts)
        same(d, o)
        assert np.array_equal(bd.view(np.uint64), bo.view(np.uint64))
    pm = gams.parse_file(GOLDEN / "gams" / "ex2_1_1.gms")
    opts = dict(partition_budget=1 << 24, sample_fan_out=1 << 12, sample_points=1, iter_limit=150)
    auto, ba, _ = cmi.aggregate_solve(ctx, pm.problem, **opts)
    interp, bi, _ = cmi.aggregate_solve(ctx, pm.problem, kernel=C.BOX_INTERPRETED, **opts)
    same(auto, interp)
    assert np.array_equal(ba.view(np.uint64), bi.view(np.uint64))
    assert auto.proven_optimal
