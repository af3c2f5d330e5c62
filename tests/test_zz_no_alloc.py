"""cmi_sample_noalloc (the sampler over the static scratch array, allocating
nothing) against cmi_sample on a sampler role built for the same arguments:
the same candidate value, index and witness bits, on consecutive calls (the
second smaller than the first, so stale scratch from the first must not leak
into it), and a size beyond the array rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import capi as K
from oracle import binding as orc

import problems as P

pytestmark = pytest.mark.gpu


def noalloc(tape, kinds, box, vids, widths, samples, salt):
    kinds = np.ascontiguousarray(kinds, np.uint8)
    box = np.ascontiguousarray(box, np.float64).reshape(-1, 2)
    vids = np.ascontiguousarray(vids, np.uint64)
    widths = np.ascontiguousarray(widths, np.uint64)
    found = C.c_int(-1)
    value = C.c_double(0.0)
    index = C.c_int64(-7)
    point = np.full(box.shape[0], -1.0)
    st = K.lib().cmi_sample_noalloc(tape.handle, K.ptr(kinds), len(kinds), K.ptr(box), box.shape[0], K.ptr(vids),
                                    K.ptr(widths), samples, C.c_uint64(salt), C.byref(found), C.byref(value),
                                    C.byref(index), K.ptr(point))
    return st, bool(found.value), value.value, index.value, point


def with_role(tape, kinds, box, vids, widths, samples, salt):
    n_regions = int(np.prod(widths))
    role = cmi.Role(tape, K.ROLE_SAMPLER, kinds, n_regions, samples=samples)
    return role.sample(box, vids, widths, salt)


def same(a, b_found, b_value, b_index, b_point):
    assert a.found == b_found
    assert np.float64(a.value).tobytes() == np.float64(b_value).tobytes()
    assert a.index == b_index
    assert np.array_equal(np.asarray(a.point).view(np.uint64), np.asarray(b_point).view(np.uint64))


CASES = [
    # (problem, slot kinds, var_ids, widths, samples, salt): the first large, the next smaller
    ("c1", [K.SLOT_CONTINUOUS] * 3, [0, 1, 3], [16, 16, 8], 32, 11),
    ("c1", [K.SLOT_CONTINUOUS] * 2, [2, 0], [2, 3], 4, 12),
    ("n24", [K.SLOT_CONTINUOUS] * 4, [0, 5, 9, 17], [8, 8, 4, 4], 16, 13),
    ("n24", [K.SLOT_CONTINUOUS], [3], [5], 1, 14),
]


def test_noalloc_sampler_equals_the_role_on_consecutive_calls(ctx):
    probs = {"c1": P.c1(4),
             "n24": cmi.EngineProblem.qcqp(77, n=24, m=12, cont_lb=-5.0, cont_ub=5.0, pair_density=0.2)}
    tapes = {k: cmi.Tape(ctx, p) for k, p in probs.items()}
    for name, kinds, vids, widths, samples, salt in CASES:
        print(f"case {name} widths={widths} samples={samples}", flush=True)
        arrays = probs[name].arrays()
        box = arrays.box.copy()
        box[:, 1] = np.minimum(box[:, 1], box[:, 0] + 3.0)  # a sub-box of the root
        st, found, value, index, point = noalloc(tapes[name], kinds, box, vids, widths, samples, salt)
        assert st == K.OK, K.lib().cmi_last_error()
        ref = with_role(tapes[name], kinds, box, vids, widths, samples, salt)
        same(ref, found, value, index, point)
        o = orc.candidate(arrays, box, vids, widths, kinds, samples, salt)
        assert o[0] == found and np.float64(o[1]).tobytes() == np.float64(value).tobytes() and o[2] == index


def test_noalloc_sampler_rejects_a_launch_beyond_the_static_array(ctx):
    tape = cmi.Tape(ctx, P.c1(0))
    box = P.c1(0).arrays().box
    # 64^3 regions x 64 samples = 262,144 tiles: more tile partials than 4 MiB holds
    st, *_ = noalloc(tape, [K.SLOT_CONTINUOUS] * 3, box, [0, 1, 2], [64, 64, 64], 64, 1)
    assert st == K.INVALID_CONFIG
    assert "static scratch" in K.lib().cmi_last_error().decode()
    # and the array still serves the next call
    st, found, value, index, point = noalloc(tape, [K.SLOT_CONTINUOUS] * 2, box, [0, 1], [4, 4], 8, 2)
    assert st == K.OK
    same(with_role(tape, [K.SLOT_CONTINUOUS] * 2, box, [0, 1], [4, 4], 8, 2), found, value, index, point)
