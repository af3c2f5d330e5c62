"""The batching and serving workloads themselves, at their real sizes, with
iteration caps that keep them within a test's time: every batched and every
served solve equals the same solve run plainly, and a sample equals the
oracle's restated search."""
import numpy as np
import pytest

import cuminlpsearch_amd as cmi
from cuminlpsearch_amd import workloads as W
from oracle import binding as orc

pytestmark = pytest.mark.gpu

FIELDS = ("iterations", "fathomed", "bounded", "sampled", "enqueued", "pruned_infeasible", "pruned_dominated",
          "pending", "stop_reason", "proven_optimal", "infeasible", "found_incumbent", "viable", "history_freed",
          "history_peak_live")


def same_outcome(d, o):
    for f in FIELDS:
        assert getattr(d, f) == getattr(o, f), f
    for f in ("lower_bound", "upper_bound"):
        assert np.float64(getattr(d, f)).tobytes() == np.float64(getattr(o, f)).tobytes(), f


OPTS = dict(partition_num=2, enumerate_cap=2, max_cycle_size=4, sample_points=16, tolerance=1e-6)


def test_c4_batch_of_512_equals_plain_and_the_oracle(ctx):
    """512 seeded QCQPs (n uniform in [4, 48], m = n/2, density 0.15) in one
    lock-step batch, 12 iterations each."""
    b = W.c4_batch(512)
    assert len(b.problems) == 512 and b.sizes.min() >= 4 and b.sizes.max() <= 48
    opts = dict(OPTS, iter_limit=12)
    outs, bests = cmi.solve_batch(ctx, b.problems, **opts)
    print("batched: done", flush=True)
    for i, (p, d, best) in enumerate(zip(b.problems, outs, bests)):
        s, sb = cmi.solve(ctx, p, **opts)
        same_outcome(d, s)
        assert np.array_equal(best.view(np.uint64), sb.view(np.uint64)), i
    for i in (0, int(np.argmax(b.sizes)), 255):
        o, _ = orc.solve(b.problems[i].arrays(), **opts)
        same_outcome(outs[i], o)


def test_c5_stream_under_an_hbm_budget_equals_plain(ctx):
    """A Poisson stream at 2,000 req/s of 96 QCQPs with n log-uniform in
    [4, 96] (density U[0.1, 0.3]), admitted under an HBM budget of a third of
    the resident bytes the stream reaches without one, so requests are held
    back; 10 iterations each. Every served result equals the plain solve of
    the same request."""
    s = W.c5_stream(96, seed=7)
    assert s.sizes.min() >= 4 and s.sizes.max() <= 96 and s.sizes.max() >= 48
    opts = dict(OPTS, iter_limit=10)
    arrival = s.arrivals(2000.0)
    _, _, _, free = cmi.serve_stream(ctx, s.problems, arrival, max_active=32, **opts)
    budget = free.peak_resident_bytes // 3
    outs, fin, lat, st = cmi.serve_stream(ctx, s.problems, arrival, max_active=32, hbm_budget_bytes=budget, **opts)
    print(f"served: peak_active {free.peak_active} unbudgeted, {st.peak_active} under {budget} bytes", flush=True)
    assert 1 < st.peak_active < free.peak_active  # the budget held requests back
    assert np.all(lat >= 0) and np.all(fin >= arrival)
    for p, d in zip(s.problems, outs):
        plain, _ = cmi.solve(ctx, p, **opts)
        same_outcome(d, plain)
    i = int(np.argmax(s.sizes))
    o, _ = orc.solve(s.problems[i].arrays(), **opts)
    same_outcome(outs[i], o)
