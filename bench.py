"""Benchmark of the MI355X interval B&B engine; prints one JSON line.

Default workload (one GPU): the batched DAG interval-evaluation kernel alone --
a seeded QCQP with n=32 variables in [-5,5], m=16 constraints, squares plus 20%
bilinear pairs, 600 reachable DAG nodes, evaluated over 65,536 sub-boxes
(log-uniform widths 1e-4..10) resident in HBM. One step = one launch over all
65,536 boxes. value = boxes/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solves] [--no-serve] [--no-batch] [--no-cpu]
                    [--no-aggregate]
                    [--no-full-stream] [--full-stream-cap C]

Also reported, on the same line:
  "cpu_baseline"  the CPU port of the evaluation on 1 thread and on every
                  worker thread, with the CPU model;
  "solves"        the 4-variable QCQP solved to gap 1e-6 for 32 seeds: GPU
                  plain and batched, and the CPU oracle's search;
  "serving"       served (continuous batching) vs plain solves/s and p50/p99
                  latency on Poisson streams at 50, 200 and 2,000 req/s, over
                  the full size range n in [4, 96] ("wide"), and the
                  20,000-request stream at 2,000 req/s (served; plain on a
                  1-in-50 sample);
  "batching"      512 seeded QCQPs solved in shared launches vs one after
                  another;
  "minlp"         the 64-variable MINLP (16 integers) under an iteration budget:
                  plain, batched and k-child search, the bracket reached;
  "aggregate"     the k-child search: the 4-variable QCQPs to gap 1e-6 and the
                  MINLPLib models it closes, with the reference's recorded
                  times for those models;
  "autotune"      every explicit-box kernel shape the autotuner timed.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import platform
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_BOXES = 65536
SEED = 2026
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def make_workload():
    import cuminlpsearch_amd as cmi

    eng = cmi.EngineProblem.qcqp(SEED, n=32, m=16, cont_lb=-5.0, cont_ub=5.0, pair_density=0.2, target_nodes=600)
    rng = np.random.default_rng(SEED)
    w = np.exp(rng.uniform(np.log(1e-4), np.log(10.0), size=(32, N_BOXES)))
    lb = rng.uniform(-5.0, 5.0 - w)
    ub = np.minimum(lb + w, 5.0)
    return eng, lb, ub


def cpu_model() -> str:
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_threads() -> int:
    """Worker threads for the CPU legs: the host's cores, at most 16 (one GPU's
    share of a GPU box's CPUs)."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline(arrays, lb, ub, budget_s: float = 8.0):
    """The CPU port of the DAG evaluation (the oracle's node-by-node evaluation
    with round-up set once per call, bit-identical to the oracle) over bounded
    prefixes of the same boxes: on 1 thread, then on every worker thread.
    Returns a dict with both rates and the all-threads results for the parity
    check."""
    from oracle import binding as orc

    threads = host_threads()
    probe = 2000
    t0 = time.perf_counter()
    orc.eval_boxes_ru(arrays, lb[:, :probe], ub[:, :probe])
    per_box = (time.perf_counter() - t0) / probe
    n1 = int(min(N_BOXES, max(probe, 0.25 * budget_s / per_box)))
    t0 = time.perf_counter()
    orc.eval_boxes_ru(arrays, np.ascontiguousarray(lb[:, :n1]), np.ascontiguousarray(ub[:, :n1]))
    one = n1 / (time.perf_counter() - t0)

    n = int(min(N_BOXES, max(threads * 1000, 0.75 * budget_s * threads / per_box)))
    n -= n % threads
    chunk = n // threads
    results = [None] * threads

    def work(k):
        s = slice(k * chunk, (k + 1) * chunk)
        results[k] = orc.eval_boxes_ru(arrays, np.ascontiguousarray(lb[:, s]), np.ascontiguousarray(ub[:, s]))

    t0 = time.perf_counter()
    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    many = n / (time.perf_counter() - t0)
    merged = tuple(np.concatenate([r[i] for r in results]) for i in range(3))
    return dict(one_thread=one, one_sample=n1, all_threads=many, threads=threads, sample=n, results=merged)


CONVERGED_STOPS = (0, 1)  # gap closed, or every node fathomed


def _converged(outs) -> float:
    return float(np.mean([o.stop_reason in CONVERGED_STOPS for o in outs]))


def _same_results(a_outs, b_outs) -> bool:
    return all(a.iterations == b.iterations and a.upper_bound == b.upper_bound and a.lower_bound == b.lower_bound
               and a.stop_reason == b.stop_reason for a, b in zip(a_outs, b_outs))


C1_SOLVE_OPTS = dict(partition_num=8, enumerate_cap=8, max_cycle_size=4, sample_points=16, tolerance=1e-6)


def c1_solves(ctx, seeds=range(32), iter_limit: int = 5000, cpu_iter_cap: int = 300):
    """The 4-variable QCQP (benchmark config 1) solved to gap 1e-6 for 32
    seeds: plain GPU solves one after another, the same 32 as one batch, and
    the CPU oracle's restated search on the converged seeds that close within
    `cpu_iter_cap` iterations (1 thread, then one seed per worker thread)."""
    import cuminlpsearch_amd as cmi
    from cuminlpsearch_amd import workloads as W
    from oracle import binding as orc

    opts = dict(C1_SOLVE_OPTS, iter_limit=iter_limit)
    probs = [W.c1(s) for s in seeds]
    rows = []
    t_all = time.perf_counter()
    for s, p in zip(seeds, probs):
        t0 = time.perf_counter()
        out, _ = cmi.solve(ctx, p, **opts)
        rows.append(dict(seed=int(s), converged=out.stop_reason in CONVERGED_STOPS, stop=int(out.stop_reason),
                         iterations=int(out.iterations), nodes=int(out.enqueued),
                         bounded=int(out.bounded), seconds=time.perf_counter() - t0,
                         lower=float(out.lower_bound), upper=float(out.upper_bound)))
    t_plain = time.perf_counter() - t_all
    conv = [r for r in rows if r["converged"]]
    t0 = time.perf_counter()
    outs_b, _ = cmi.solve_batch(ctx, probs, **opts)
    t_batch = time.perf_counter() - t0
    same = all(o.iterations == r["iterations"] and o.upper_bound == r["upper"] and o.lower_bound == r["lower"]
               for o, r in zip(outs_b, rows))

    cpu_seeds = [r["seed"] for r in conv if r["iterations"] <= cpu_iter_cap]
    cpu = None
    if cpu_seeds:
        arrays = {s: W.c1(s).arrays() for s in cpu_seeds}
        t0 = time.perf_counter()
        cpu_outs = {s: orc.solve(arrays[s], **opts)[0] for s in cpu_seeds}
        t_one = time.perf_counter() - t0
        threads = min(host_threads(), len(cpu_seeds))
        done = {}

        def work(k):
            for s in cpu_seeds[k::threads]:
                done[s] = orc.solve(arrays[s], **opts)[0]

        t0 = time.perf_counter()
        ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        t_many = time.perf_counter() - t0
        gpu_time = sum(r["seconds"] for r in rows if r["seed"] in cpu_outs)
        cpu = {
            "seeds": cpu_seeds,
            "solves_per_s_1_thread": len(cpu_seeds) / t_one,
            "solves_per_s_all_threads": len(cpu_seeds) / t_many,
            "threads": threads,
            "gpu_plain_solves_per_s_same_seeds": len(cpu_seeds) / gpu_time,
            "results_identical_to_gpu": all(cpu_outs[s].iterations == done[s].iterations
                                            and next(r for r in rows if r["seed"] == s)["iterations"]
                                            == cpu_outs[s].iterations for s in cpu_seeds),
        }
    return {
        "workload": f"QCQP n=4 in [-10,10], m=2, Q and A_i symmetric N(0,1), seeds {seeds[0]}..{seeds[-1]}; "
                    f"GreedyEnumerate partition {opts['partition_num']} over all 4 variables "
                    f"({opts['partition_num'] ** 4} regions per iteration), {opts['sample_points']} samples per "
                    f"region, gap 1e-6, cap {iter_limit} iterations",
        "solves": len(rows),
        "converged": len(conv),
        "converged_fraction": len(conv) / len(rows),
        "plain_solves_per_s_converged": (len(conv) / sum(r["seconds"] for r in conv)) if conv else None,
        "plain_seconds_total": t_plain,
        "batched_seconds_total": t_batch,
        "batched_equals_plain": same,
        "median_nodes_converged": float(np.median([r["nodes"] for r in conv])) if conv else None,
        "per_seed": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()} for r in rows],
        "cpu_oracle": cpu,
        "cpu_model": cpu_model(),
    }


# The reference's k-child (aggregate) runs that converged on models of this
# corpus, and its per-region gams_solve beside them: wall seconds as its
# AGGREGATE_STATUS.md table records them (its own GPU, whole process: parse,
# graph builds and search).
REFERENCE_AGGREGATE_SECONDS = {"ex2_1_1": (5.6, 235.2), "ex8_1_1": (0.8, 35.7)}
AGG_MODELS = ("ex2_1_1", "ex8_1_1", "circle", "st_e01", "nvs01", "ex4_1_2")


def c3_solves(ctx, seeds=range(4), iter_limit: int = 2000, agg_iter_limit: int = 300):
    """The 64-variable MINLP (16 integers in [0, 10], 32 constraints, 64
    samples per region) under a fixed iteration budget: plain solves one after
    another, the same seeds as one batch (identical results), and the k-child
    search; the bracket each reaches and the rate of search iterations and
    nodes. Interval bounds on 64 variables do not close a 1e-6 gap within any
    budget a benchmark can spend, so the gap reached is the figure."""
    import cuminlpsearch_amd as cmi
    from cuminlpsearch_amd import workloads as W

    opts = dict(partition_num=2, enumerate_cap=2, max_cycle_size=4, sample_points=64, iter_limit=iter_limit,
                tolerance=1e-6)
    probs = [W.c3(s) for s in seeds]
    cmi.solve(ctx, probs[0], **dict(opts, iter_limit=2))
    rows = []
    t_plain = 0.0
    outs_p = []
    for s, p in zip(seeds, probs):
        t0 = time.perf_counter()
        out, _ = cmi.solve(ctx, p, **opts)
        dt = time.perf_counter() - t0
        t_plain += dt
        outs_p.append(out)
        gap = (out.upper_bound - out.lower_bound) / max(1.0, abs(out.upper_bound))
        rows.append(dict(seed=int(s), converged=out.stop_reason in CONVERGED_STOPS, iterations=int(out.iterations),
                         nodes=int(out.enqueued), bounded=int(out.bounded), seconds=round(dt, 4), lower=float(out.lower_bound),
                         upper=float(out.upper_bound), relative_gap=float(gap)))
    t0 = time.perf_counter()
    outs_b, _ = cmi.solve_batch(ctx, probs, **opts)
    t_batch = time.perf_counter() - t0
    agg_opts = dict(partition_budget=1 << 20, sample_fan_out=1024, sample_points=64, iter_limit=agg_iter_limit)
    agg = []
    t_agg = 0.0
    for s, p in zip(seeds, probs):
        t0 = time.perf_counter()
        out, _, _ = cmi.aggregate_solve(ctx, p, **agg_opts)
        dt = time.perf_counter() - t0
        t_agg += dt
        agg.append(dict(seed=int(s), converged=bool(out.proven_optimal), iterations=int(out.iterations),
                        children=int(out.enqueued), seconds=round(dt, 4), lower=float(out.lower_bound),
                        upper=float(out.upper_bound)))
    iters = sum(r["iterations"] for r in rows)
    return {
        "workload": f"MINLP n=64 (16 integer in [0,10], 48 continuous in [-5,5]), m=32, 10% bilinear density, "
                    f"seeds {seeds[0]}..{seeds[-1]}; GreedyEnumerate partition 2, 64 samples per region, gap 1e-6, "
                    f"cap {iter_limit} iterations; k-child search N=2^20, cap {agg_iter_limit}",
        "converged": sum(r["converged"] for r in rows),
        "plain_iterations_per_s": iters / t_plain,
        "plain_nodes_per_s": sum(r["nodes"] for r in rows) / t_plain,
        "plain_regions_bounded_per_s": sum(r["bounded"] for r in rows) / t_plain,
        "batched_seconds": t_batch,
        "plain_seconds": t_plain,
        "batched_equals_plain": _same_results(outs_b, outs_p),
        "per_seed": rows,
        "aggregate": {"seconds": round(t_agg, 4), "per_seed": agg},
    }


def aggregate_solves(ctx, seeds=range(32), n_c1: int = 1 << 20, n_models: int = 1 << 24, iter_limit: int = 2000):
    """The k-child search (k = 4 children per node, each bounded over N / 4
    subregions in one launch): the 4-variable QCQPs' 32 seeds to gap 1e-6,
    and the MINLPLib models of the corpus it closes, each timed on its first
    solve (which compiles the problem's kernels once the interpreted work
    would have paid for it) and on a repeat (kernels cached)."""
    import cuminlpsearch_amd as cmi
    from cuminlpsearch_amd import gams, workloads as W
    from oracle import binding as orc

    c1_opts = dict(partition_budget=n_c1, sample_fan_out=4096, sample_points=4, iter_limit=iter_limit)
    cmi.aggregate_solve(ctx, W.c1(seeds[0]), **dict(c1_opts, iter_limit=1))
    rows = []
    for s in seeds:
        t0 = time.perf_counter()
        out, _, _ = cmi.aggregate_solve(ctx, W.c1(s), **c1_opts)
        rows.append(dict(seed=int(s), converged=bool(out.proven_optimal), iterations=int(out.iterations),
                         children=int(out.enqueued), seconds=time.perf_counter() - t0,
                         lower=float(out.lower_bound), upper=float(out.upper_bound)))
    conv = [r for r in rows if r["converged"]]
    # the oracle's restated search on the same workload, two iterations
    t0 = time.perf_counter()
    o, _, _ = orc.aggregate_solve(W.c1(1).arrays(), **dict(c1_opts, iter_limit=2))
    cpu_iter = (time.perf_counter() - t0) / max(int(o.iterations), 1)
    gpu_iter = sum(r["seconds"] for r in rows) / max(sum(r["iterations"] for r in rows), 1)
    models = {}
    root = pathlib.Path(ROOT) / "tests" / "golden" / "gams"
    m_opts = dict(partition_budget=n_models, sample_fan_out=1 << 14, sample_points=1, iter_limit=1000)
    for name in AGG_MODELS:
        pm = gams.parse_file(root / f"{name}.gms")
        t0 = time.perf_counter()
        out, _, _ = cmi.aggregate_solve(ctx, pm.problem, **m_opts)
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        again, _, _ = cmi.aggregate_solve(ctx, pm.problem, **m_opts)
        repeat = time.perf_counter() - t0
        assert again.iterations == out.iterations and again.upper_bound == out.upper_bound
        row = dict(iterations=int(out.iterations), converged=bool(out.proven_optimal), seconds=round(first, 4),
                   seconds_repeat=round(repeat, 4), lower=float(out.lower_bound), upper=float(out.upper_bound))
        if name in REFERENCE_AGGREGATE_SECONDS:
            agg_s, plain_s = REFERENCE_AGGREGATE_SECONDS[name]
            row["reference_aggregate_seconds"] = agg_s
            row["reference_gams_solve_seconds"] = plain_s
        models[name] = row
    from cuminlpsearch_amd.kernel_timing import aggregate_launch
    launch = {name: aggregate_launch(ctx, prob, n_models)
              for name, prob in (("ex2_1_1", gams.parse_file(root / "ex2_1_1.gms").problem), ("c1_seed0", W.c1(0)))}
    return {
        "bound_launch": launch,
        "workload": f"k-child search, k=4, N={n_c1} subregions bounded per node (N/4 per child), sampling "
                    f"4096 subregions x 4 points, gap 1e-6, cap {iter_limit} iterations; QCQP n=4 seeds "
                    f"{seeds[0]}..{seeds[-1]}",
        "converged": len(conv),
        "converged_fraction": len(conv) / len(rows),
        "solves_per_s_converged": (len(conv) / sum(r["seconds"] for r in conv)) if conv else None,
        "median_iterations_converged": float(np.median([r["iterations"] for r in conv])) if conv else None,
        "gpu_seconds_per_iteration": gpu_iter,
        "cpu_oracle_seconds_per_iteration": cpu_iter,
        "per_seed": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()} for r in rows],
        "minlplib": {"settings": f"N={n_models}, 16384 sampled subregions x 1 point, gap 1e-6, cap 1000; "
                                 "seconds: first solve of the model in this process, parse excluded",
                     "models": models},
    }


def serve_comparison(ctx, n_req: int = 128, rates=(50.0, 200.0, 2000.0), seed: int = 7, n_max: int = 16,
                     iter_limit: int = 1000, plain: bool = True):
    """Served (continuous batching) vs plain (one solve at a time) on the same
    Poisson streams of seeded QCQPs, one stream per arrival rate: n
    log-uniform in [4, n_max], m = n/2, bilinear density U[0.1, 0.3], gap 1e-6,
    iteration cap per request. The first fields are those of the 200 req/s
    stream (or of the only rate); "rates" holds every stream. Served rates on
    converged requests only are reported beside the all-request rates."""
    import cuminlpsearch_amd as cmi
    from cuminlpsearch_amd import workloads as W

    st = W.c5_stream(n_req, seed=seed, n_max=n_max)
    opts = dict(partition_num=2, enumerate_cap=2, max_cycle_size=4, sample_points=16, iter_limit=iter_limit,
                tolerance=1e-6)
    per_rate = []
    for rate in rates:
        arrival = st.arrivals(rate)
        outs_s, fin_s, lat_s, st_s = cmi.serve_stream(ctx, st.problems, arrival, max_active=256, **opts)
        conv = np.array([o.stop_reason in CONVERGED_STOPS for o in outs_s])
        row = {
            "arrival_rate_per_s": rate,
            "served_solves_per_s": n_req / st_s.wall_seconds,
            "served_converged_solves_per_s": float(conv.sum()) / st_s.wall_seconds,
            "served_p50_ms": float(np.percentile(lat_s, 50) * 1e3),
            "served_p99_ms": float(np.percentile(lat_s, 99) * 1e3),
            "converged_fraction": float(conv.mean()),
            "served_p50_ms_converged": float(np.percentile(lat_s[conv], 50) * 1e3) if conv.any() else None,
            "steps": int(st_s.steps),
            "peak_active": int(st_s.peak_active),
            "step_kernels": kernel_roofline(st_s),
        }
        if plain:
            outs_p, fin_p, lat_p, st_p = cmi.serve_plain(ctx, st.problems, arrival, **opts)
            row.update({
                "plain_solves_per_s": n_req / st_p.wall_seconds,
                "plain_converged_solves_per_s": float(conv.sum()) / st_p.wall_seconds,
                "speedup": st_p.wall_seconds / st_s.wall_seconds,
                "plain_p50_ms": float(np.percentile(lat_p, 50) * 1e3),
                "plain_p99_ms": float(np.percentile(lat_p, 99) * 1e3),
                "results_identical": _same_results(outs_s, outs_p),
            })
        per_rate.append(row)
    main = next((r for r in per_rate if r["arrival_rate_per_s"] == 200.0), per_rate[0])
    return {
        "workload": st.description + f", gap 1e-6, cap {iter_limit} iterations, rates {list(rates)} req/s",
        "requests": n_req,
        **main,
        "results_identical": all(r.get("results_identical", True) for r in per_rate),
        "rates": per_rate,
    }


def full_stream(ctx, iter_limit: int, n_req: int = 20000, rate: float = 2000.0, sample_every: int = 50):
    """The serving configuration at its full size: 20,000 QCQPs (n log-uniform
    in [4, 96], m = n/2, density U[0.1, 0.3], gap 1e-6) arriving as a Poisson
    stream at 2,000 req/s, served with continuous batching under an iteration
    cap per request. The plain form is timed on every 50th request of the same
    stream solved back to back (serving all 20,000 plainly would take the
    better part of an hour); the served results of those requests must equal
    the plain ones."""
    import cuminlpsearch_amd as cmi
    from cuminlpsearch_amd import workloads as W

    st = W.c5_stream(n_req, seed=9, n_max=96)
    opts = dict(partition_num=2, enumerate_cap=2, max_cycle_size=4, sample_points=16, iter_limit=iter_limit,
                tolerance=1e-6)
    arrival = st.arrivals(rate)
    outs_s, fin_s, lat_s, st_s = cmi.serve_stream(ctx, st.problems, arrival, max_active=256, **opts)
    conv = np.array([o.stop_reason in CONVERGED_STOPS for o in outs_s])
    idx = list(range(0, n_req, sample_every))
    sample = [st.problems[i] for i in idx]
    outs_p, _, _, st_p = cmi.serve_plain(ctx, sample, np.zeros(len(sample)), **opts)
    plain_rate = len(sample) / st_p.wall_seconds
    served_rate = n_req / st_s.wall_seconds
    return {
        "workload": st.description + f", gap 1e-6, cap {iter_limit} iterations, {rate:.0f} req/s",
        "requests": n_req,
        "arrival_rate_per_s": rate,
        "served_solves_per_s": served_rate,
        "served_converged_solves_per_s": float(conv.sum()) / st_s.wall_seconds,
        "served_p50_ms": float(np.percentile(lat_s, 50) * 1e3),
        "served_p99_ms": float(np.percentile(lat_s, 99) * 1e3),
        "converged_fraction": float(conv.mean()),
        "steps": int(st_s.steps),
        "peak_active": int(st_s.peak_active),
        "step_kernels": kernel_roofline(st_s),
        "plain_sample": f"every {sample_every}th request ({len(sample)}), solved back to back",
        "plain_solves_per_s": plain_rate,
        "speedup": served_rate / plain_rate,
        "results_identical": _same_results([outs_s[i] for i in idx], outs_p),
    }


def batch_comparison(ctx, n_inst: int = 512, iter_limit: int = 100):
    """Batched (every instance's regions in the same launches) vs sequential
    plain solves of the same seeded QCQPs: n uniform in [4, 48], m = n/2,
    bilinear density 0.15, gap 1e-6, iteration cap per instance."""
    import cuminlpsearch_amd as cmi
    from cuminlpsearch_amd import workloads as W

    b = W.c4_batch(n_inst)
    opts = dict(partition_num=2, enumerate_cap=2, max_cycle_size=4, sample_points=16, iter_limit=iter_limit,
                tolerance=1e-6)
    t0 = time.perf_counter()
    outs_b, _, st_b = cmi.solve_batch_stats(ctx, b.problems, **opts)
    t_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    outs_p = [cmi.solve(ctx, p, **opts)[0] for p in b.problems]
    t_p = time.perf_counter() - t0
    conv = sum(o.stop_reason in CONVERGED_STOPS for o in outs_b)
    return {
        "workload": b.description + f", gap 1e-6, cap {iter_limit} iterations",
        "instances": n_inst,
        "batched_solves_per_s": n_inst / t_b,
        "plain_solves_per_s": n_inst / t_p,
        "speedup": t_p / t_b,
        "converged_fraction": conv / n_inst,
        "batched_converged_solves_per_s": conv / t_b,
        "plain_converged_solves_per_s": conv / t_p,
        "iterations_total": int(sum(o.iterations for o in outs_b)),
        "results_identical": _same_results(outs_b, outs_p),
        "steps": int(st_b.steps),
        "step_kernels": kernel_roofline(st_b),
    }


def kernel_roofline(st) -> dict:
    """The decode-mode launches of a batched or served run, from the engine's
    own HIP events and FP64 accounting: device time, its share of the wall
    time, and the algorithmic FP64 rate against the FP64 vector peak."""
    tflops = st.kernel_fp64_ops / st.kernel_seconds / 1e12 if st.kernel_seconds > 0 else 0.0
    return {
        "bound": "fp64",
        "kernel_seconds": st.kernel_seconds,
        "kernel_share_of_wall": st.kernel_seconds / st.wall_seconds if st.wall_seconds > 0 else None,
        "fp64_ops": st.kernel_fp64_ops,
        "achieved_tflops": tflops,
        "peak_tflops": FP64_PEAK_TFLOPS,
        "frac": tflops / FP64_PEAK_TFLOPS,
        "ms_per_step": st.kernel_seconds * 1e3 / st.steps if st.steps else None,
    }


PROFILES = pathlib.Path(__file__).resolve().parent / "profiles"


def traffic_profile(kernel: str) -> pathlib.Path:
    return PROFILES / f"{kernel}_qcqp-n32-m16-600node-65536box_hbm-traffic.md"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    summary (counter passes scaled by a 512 MiB copy), or None if absent."""
    path = traffic_profile(kernel)
    if not path.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch = (\d+)$", path.read_text(), re.M)
    return int(m.group(1)) if m else None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served-vs-plain Poisson-stream comparison (continuous batching)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-vs-sequential comparison over 512 seeded instances")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-solves", action="store_true",
                    help="skip the 32 solves of the 4-variable QCQP to gap 1e-6 (GPU plain, batched, CPU oracle)")
    ap.add_argument("--no-aggregate", action="store_true",
                    help="skip the k-child (aggregate) search runs")
    ap.add_argument("--no-full-stream", action="store_true",
                    help="skip the 20,000-request stream (served, and plain on a 1-in-50 sample)")
    ap.add_argument("--full-stream-cap", type=int, default=20,
                    help="iteration cap per request of the 20,000-request stream")
    ap.add_argument("--box-groups", type=int, default=0,
                    help="waves (function groups) per 64-box workgroup; 0 = chosen from the tape")
    ap.add_argument("--box-engine", choices=["tuned", "compiled", "interpreted", "auto"], default="tuned",
                    help="explicit-box kernel: the autotuner's pick over interpreter and compiled shapes on this "
                         "workload (default), the tape compiled to a straight-line kernel at run time, the tape "
                         "interpreter, or the engine's automatic choice")
    ap.add_argument("--waves", type=int, default=0,
                    help="compiled kernel: waves per workgroup (0 = one per function group)")
    ap.add_argument("--persistent", action="store_true",
                    help="compiled kernel: persistent double-buffered form instead of one workgroup per tile")
    ap.add_argument("--schedule", choices=["default", "max-ilp"], default="default",
                    help="compiled kernel: instruction scheduler of the run-time compiler")
    ap.add_argument("--load-batch", type=int, default=8,
                    help="compiled direct form: variables per load batch (0 = every load at the start)")
    ap.add_argument("--wgs-per-cu", type=int, default=1,
                    help="compiled kernel, persistent form: resident workgroups per CU (grid = CUs x this)")
    ap.add_argument("--sweep-groups", action="store_true",
                    help="print the interpreter's kernel time for each group count, and the compiled kernel's, "
                         "to stderr before the timed run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device copy (read 512 MiB, write 512 MiB) after the timed region, as "
                         "the reference kernel that scales hardware byte counters")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every run here uses one MI355X).",
              file=sys.stderr)
        return 2

    import ctypes as C

    import torch

    import cuminlpsearch_amd as cmi
    from cuminlpsearch_amd import capi

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 2
    t_start = time.perf_counter()
    dev = torch.device("cuda:0")
    ctx = cmi.Context(0)
    eng, lb, ub = make_workload()
    tape = cmi.Tape(ctx, eng)

    # The boxes are replicated into enough distinct buffers (> the 256 MB
    # Infinity Cache in total) that every step reads boxes that are not
    # cache-resident: each step's inputs come from HBM.
    n_sets = max(1, int(np.ceil(320e6 / (16 * 32 * N_BOXES))))
    h_lb = torch.from_numpy(lb)
    h_ub = torch.from_numpy(ub)
    d_lb = [h_lb.to(dev) for _ in range(n_sets)]
    d_ub = [h_ub.to(dev) for _ in range(n_sets)]
    d_feas = torch.empty(N_BOXES, dtype=torch.uint8, device=dev)
    d_olb = torch.empty(N_BOXES, dtype=torch.float64, device=dev)
    d_oub = torch.empty(N_BOXES, dtype=torch.float64, device=dev)
    lib = capi.lib()

    def launch(k, stream):
        capi.check(lib.cmi_eval_boxes_device(tape.handle, C.c_void_p(d_lb[k % n_sets].data_ptr()),
                                             C.c_void_p(d_ub[k % n_sets].data_ptr()), N_BOXES, N_BOXES,
                                             C.c_void_p(d_feas.data_ptr()), C.c_void_p(d_olb.data_ptr()),
                                             C.c_void_p(d_oub.data_ptr()), None, C.c_void_p(stream.cuda_stream)))

    # A stream of our own: its handle is non-null, so the engine launches on it
    # (a null handle would select the engine context's internal stream) and the
    # HIP events below time exactly these launches.
    stream = torch.cuda.Stream(dev)
    engines = {"compiled": capi.BOX_COMPILED, "interpreted": capi.BOX_INTERPRETED, "auto": capi.BOX_AUTO}
    schedules = {"default": capi.SCHED_DEFAULT, "max-ilp": capi.SCHED_MAX_ILP}
    # The kernel configuration is fixed before any launch (the autotuner runs on
    # its own synthetic boxes), so every launch below is the measured kernel.
    tuning = None
    if not a.sweep_groups:
        if a.box_engine == "tuned":
            tuning = tape.autotune(N_BOXES)
        else:
            tape.set_box_groups(a.box_groups)
            tape.set_compiled_shape(a.persistent, a.waves, a.wgs_per_cu)
            tape.set_compiled_codegen(schedules[a.schedule], a.load_batch)
            tape.set_box_engine(engines[a.box_engine])
    torch.cuda.synchronize()
    for k in range(a.warmup):
        launch(k, stream)
    torch.cuda.synchronize()

    def kernel_ms(steps):
        """Mean kernel duration: HIP events on the launch stream around `steps`
        back-to-back launches (the gap between two launches on one stream is
        well under a microsecond, so this tracks the kernels' own duration)."""
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for k in range(steps):
            launch(k, stream)
        end.record(stream)
        torch.cuda.synchronize()
        return float(start.elapsed_time(end)) / steps

    if a.sweep_groups:
        tape.set_box_engine(capi.BOX_INTERPRETED)
        for g in range(1, 9):
            got = tape.set_box_groups(g)
            if got != g:
                continue
            for k in range(3):
                launch(k, stream)
            i = tape.info()
            print(f"interpreted box_groups={g} instr={i.box_instr} slots={i.box_slots} kernel_ms={kernel_ms(20):.4f}",
                  file=sys.stderr, flush=True)
        tape.set_box_engine(capi.BOX_COMPILED)
        for g in (4, 8):
            if tape.set_box_groups(g) != g:
                continue
            for waves in (1, 2, 4, 8):
                if waves > g:
                    continue
                for persistent, wgs in ((False, 1), (True, 1), (True, 2)):
                    tape.set_compiled_shape(persistent, waves, wgs)
                    for k in range(3):
                        launch(k, stream)
                    i = tape.info()
                    print(f"compiled box_groups={g} waves={i.compiled_waves} persistent={i.compiled_persistent} "
                          f"wgs_per_cu={wgs} kernel_ms={kernel_ms(20):.4f} compile_s={i.compile_seconds:.3f}",
                          file=sys.stderr, flush=True)
        if a.box_engine == "tuned":
            tuning = tape.autotune(N_BOXES)
        else:
            tape.set_box_groups(a.box_groups)
            tape.set_compiled_shape(a.persistent, a.waves, a.wgs_per_cu)
            tape.set_compiled_codegen(schedules[a.schedule], a.load_batch)
            tape.set_box_engine(engines[a.box_engine])
    info = tape.info()
    for k in range(3):
        launch(k, stream)
    torch.cuda.synchronize()
    kern_ms_single = kernel_ms(a.steps)

    # Throughput: the K steps are captured once into a HIP graph and replayed,
    # so the timed region is device work, not Python launch overhead.
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream(dev)
    cap.wait_stream(stream)
    with torch.cuda.stream(cap):
        graph.capture_begin()
        for k in range(a.steps):
            launch(k, cap)
        graph.capture_end()
    torch.cuda.synchronize()
    graph.replay()  # untimed: the first replay uploads the graph
    torch.cuda.synchronize()
    # HIP events on the replay stream bracket the timed region: their span / K
    # is the average duration of one launch as the GPU ran it.
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    replay_stream = torch.cuda.current_stream(dev)
    t0 = time.perf_counter()
    ev0.record(replay_stream)
    graph.replay()
    ev1.record(replay_stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(ev0.elapsed_time(ev1)) / a.steps
    ms_per_step = wall * 1e3 / a.steps
    value = N_BOXES / (ms_per_step / 1e3)

    feas = d_feas.cpu().numpy()
    olb = d_olb.cpu().numpy()
    oub = d_oub.cpu().numpy()

    if a.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device=dev)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    kernel = "cmi_boxes_compiled" if info.box_compiled else "k_boxes"
    bytes_per_box = info.bytes_per_box
    flops_per_box = info.flops_per_box
    achieved_gbs = bytes_per_box * N_BOXES / (kern_ms * 1e-3) / 1e9
    achieved_tflops = flops_per_box * N_BOXES / (kern_ms * 1e-3) / 1e12

    line = {
        "metric": "boxes/s DAG eval (interval bounds of a 600-node QCQP DAG per sub-box, outward-rounded FP64)",
        "value": value,
        "unit": "boxes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded QCQP and sub-boxes)",
        "config": {
            "workload": "QCQP n=32 in [-5,5], m=16, squares + 20% bilinear pairs, 600 reachable nodes, "
                        "65536 sub-boxes, widths log-uniform 1e-4..10, seed 2026",
            "boxes_per_step": N_BOXES,
            "box_sets_rotated": n_sets,
            "timed_region": "one HIP-graph replay of K launches",
            "tape_instructions": info.n_instr,
            "tape_slots": info.n_slots,
            "box_engine": ("compiled" if info.box_compiled else "interpreted"),
            "box_engine_choice": a.box_engine,
            "compiled_waves": info.compiled_waves if info.box_compiled else None,
            "compiled_persistent": bool(info.compiled_persistent) if info.box_compiled else None,
            "compiled_schedule": ("max-ilp" if info.compiled_schedule == capi.SCHED_MAX_ILP else "default")
            if info.box_compiled else None,
            "compiled_load_batch": info.compiled_load_batch if info.box_compiled else None,
            "compile_seconds": info.compile_seconds,
            "box_groups": info.box_groups,
            "box_instructions": info.box_instr,
            "box_slots": info.box_slots,
            "reachable_nodes": info.reachable_nodes,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(kernel),
            "traffic_source": traffic_profile(kernel).name,
            "kernel": kernel,
            "kernel_ms": kern_ms,
            "kernel_ms_eager_launches": kern_ms_single,
            "bytes_per_box": bytes_per_box,
            "flops_per_box": flops_per_box,
            "fp64_achieved_tflops": achieved_tflops,
            "fp64_frac": achieved_tflops / FP64_PEAK_TFLOPS,
        },
    }
    if tuning is not None:
        line["autotune"] = sorted(tuning, key=lambda c: c["kernel_ms"])

    if not a.no_cpu:
        print(f"bench: CPU baseline ({time.perf_counter() - t_start:.0f} s)", file=sys.stderr, flush=True)
        arrays = eng.arrays()
        cpu = cpu_baseline(arrays, lb, ub)
        f_o, lo_o, hi_o = cpu["results"]
        n_sample = cpu["sample"]
        parity = bool(np.array_equal(f_o, feas[:n_sample]) and np.array_equal(lo_o.view(np.uint64),
                                                                              olb[:n_sample].view(np.uint64))
                      and np.array_equal(hi_o.view(np.uint64), oub[:n_sample].view(np.uint64)))
        line["cpu_baseline"] = {
            "value": cpu["all_threads"],
            "unit": "boxes/s",
            "cores": cpu["threads"],
            "kind": "port",
            "value_1_thread": cpu["one_thread"],
            "cpu_model": cpu_model(),
            "sample": f"first {n_sample} of the 65536 boxes on {cpu['threads']} host threads "
                      f"(first {cpu['one_sample']} on 1 thread); the oracle's node-by-node evaluation with "
                      f"round-up set once per call, bit-identical to the oracle",
        }
        line["parity_vs_cpu"] = "bit-exact" if parity else "MISMATCH"

    def progress(what):
        print(f"bench: {what} ({time.perf_counter() - t_start:.0f} s)", file=sys.stderr, flush=True)

    if not a.no_solves:
        progress("solves of the 4-variable QCQP")
        line["solves"] = c1_solves(ctx)
    if not a.no_solves:
        progress("the 64-variable MINLP")
        line["minlp"] = c3_solves(ctx)
    if not a.no_aggregate:
        progress("k-child searches")
        line["aggregate"] = aggregate_solves(ctx)
    if not a.no_serve:
        progress("serving")
        line["serving"] = serve_comparison(ctx)
        # the survey's size range (n up to 96), on a shorter stream and cap
        wide = serve_comparison(ctx, n_req=64, rates=(200.0, 2000.0), seed=8, n_max=96, iter_limit=100)
        line["serving"]["wide"] = wide
        if not a.no_full_stream:
            progress("serving the 20,000-request stream")
            line["serving"]["full_stream"] = full_stream(ctx, a.full_stream_cap)
    if not a.no_batch:
        progress("batching")
        line["batching"] = batch_comparison(ctx)

    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
