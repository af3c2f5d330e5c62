"""The reference-side adapter (include/cmi_reference_adapter.hpp) on the device:
tests/adapter/adapter_test.cpp, built against reference-shaped types, builds,
bounds, samples and enumerates through RegionBackendFactory<double> and maps
CMI_OVER_BUDGET to OverBudgetError with every fact intact (backend.hpp:81-287)."""
import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "adapter" / "build" / "adapter_test"


def test_reference_adapter_on_the_device():
    assert EXE.exists(), "the adapter test is built by cuminlpsearch_amd/build.py (build() of __graft_entry__.py)"
    run = subprocess.run([str(EXE), "0"], capture_output=True, text=True, timeout=120)
    print(run.stdout)
    assert run.returncode == 0, run.stdout[-3000:] + run.stderr[-2000:]
    assert " 0 failed" in run.stdout
