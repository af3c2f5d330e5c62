"""The seeded QCQP workloads the benchmark and the tests share: the 4-variable
plain solves, the 512-instance batch and the Poisson request stream.

Every instance comes from the engine's own seeded generator
(cmi_problem_qcqp); the size draws use numpy's seeded generator, and the
seeds are part of each workload's description.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .engine import EngineProblem

# 4 continuous variables in [-10, 10], 2 quadratic constraints, Q and A_i
# symmetric N(0, 1), c and b_i U[-1, 1], r_i U[1, 10]; seeds 0..31.
C1_SPEC = dict(n=4, m=2, cont_lb=-10.0, cont_ub=10.0, symmetric_pairs=True)


def c1(seed: int) -> EngineProblem:
    return EngineProblem.qcqp(seed, **C1_SPEC)


# 64 variables, the first 16 integer in [0, 10], the rest continuous in
# [-5, 5]; 32 constraints; 10 % bilinear pairs plus squares, coefficients
# N(0, 1); about 1,500 DAG nodes.
C3_SPEC = dict(n=64, m=32, n_int=16, int_lb=0.0, int_ub=10.0, cont_lb=-5.0, cont_ub=5.0, pair_density=0.1,
               target_nodes=1500)


def c3(seed: int) -> EngineProblem:
    return EngineProblem.qcqp(seed, **C3_SPEC)


@dataclass
class Batch:
    problems: list
    sizes: np.ndarray
    description: str


def c4_batch(count: int = 512, seed: int = 11, first_problem_seed: int = 5000) -> Batch:
    """`count` seeded QCQPs, n uniform in [4, 48], m = n/2, bilinear density
    0.15, continuous variables in [-5, 5]."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(4, 49, count)
    probs = [EngineProblem.qcqp(first_problem_seed + i, n=int(n), m=int(n) // 2, cont_lb=-5.0, cont_ub=5.0,
                                pair_density=0.15) for i, n in enumerate(sizes)]
    return Batch(probs, sizes, f"{count} QCQPs, n uniform [4,48] (seed {seed}), m=n/2, density 0.15, "
                               f"problem seeds {first_problem_seed}..{first_problem_seed + count - 1}")


@dataclass
class Stream:
    problems: list
    sizes: np.ndarray
    gaps: np.ndarray  # inter-arrival gaps at rate 1 (exponential, mean 1)
    description: str

    def arrivals(self, rate: float) -> np.ndarray:
        """Poisson arrival times at `rate` requests per second, the first at 0."""
        t = np.cumsum(self.gaps / rate)
        return t - t[0]


def c5_stream(count: int, seed: int = 7, n_max: int = 96, first_problem_seed: int = 1000) -> Stream:
    """`count` seeded QCQPs, n = floor(exp(U[ln 4, ln(n_max + 1)])), m = n/2,
    bilinear density U[0.1, 0.3], continuous variables in [-5, 5]; Poisson
    arrivals."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(4), np.log(n_max + 1), count))).astype(int)
    dens = rng.uniform(0.1, 0.3, count)
    probs = [EngineProblem.qcqp(first_problem_seed + i, n=int(n), m=int(n) // 2, cont_lb=-5.0, cont_ub=5.0,
                                pair_density=float(d)) for i, (n, d) in enumerate(zip(sizes, dens))]
    gaps = rng.exponential(1.0, count)
    return Stream(probs, sizes, gaps, f"{count} QCQPs, n log-uniform [4,{n_max}], m=n/2, density U[0.1,0.3], "
                                      f"Poisson arrivals (seed {seed})")
