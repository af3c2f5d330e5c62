"""The reference's host tests of the k-child search (aggregate_policy_test.cpp,
aggregate_frontier_test.cpp), restated in tests/aggregate_pins.cpp against the
engine's cmi/aggregate_search.hpp and the oracle's search_core.hpp, compiled
with g++ and run here."""
from __future__ import annotations

import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_restated_aggregate_cases_hold_for_engine_and_oracle(tmp_path):
    exe = tmp_path / "aggregate_pins"
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'cuminlpsearch_amd' / 'csrc'}",
           f"-I{ROOT / 'include'}", f"-I{ROOT / 'oracle'}", str(ROOT / "tests" / "aggregate_pins.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-4000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout[-4000:]
    assert " 0 failed" in run.stdout
