"""Static check of the gfx950 code: inside every round-toward-+infinity region
of the interval interpreter, each FP64 instruction that rounds comes from the
engine's own inline asm (which computes lower bounds by negation); none is
compiler-generated, since compiler-generated arithmetic there would silently
round upward. Compiles kernels.hip to assembly on the host (no GPU needed)."""
import pathlib
import subprocess

import pytest

from isa_tools import check_asm_regions

ROOT = pathlib.Path(__file__).resolve().parent.parent
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not pathlib.Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    from cuminlpsearch_amd.build import HIP_FLAGS

    out = tmp_path_factory.mktemp("isa") / "kernels.s"
    cmd = [HIPCC, *HIP_FLAGS, "--cuda-device-only", "-S", "-o", str(out),
           str(ROOT / "cuminlpsearch_amd" / "csrc" / "kernels.hip")]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text().splitlines()


def test_round_up_regions_hold_only_asm_arithmetic(asm):
    # the explicit-box kernel and the decode bounder
    assert check_asm_regions(asm) >= 2
