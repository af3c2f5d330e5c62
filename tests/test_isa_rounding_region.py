"""Static check of the gfx950 code: inside every round-toward-+infinity region
of the interval interpreter, each FP64 instruction that rounds comes from the
engine's own inline asm (which computes lower bounds by negation); none is
compiler-generated, since compiler-generated arithmetic there would silently
round upward. Compiles kernels.hip to assembly on the host (no GPU needed)."""
import pathlib
import re
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HIPCC = "/opt/rocm/bin/hipcc"

ROUNDING = re.compile(r"^\s*(v_(add|mul|fma|sub|subrev|div_fmas|div_scale|rcp|sqrt|ldexp|cvt)[a-z0-9_]*f64)\b")


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


def regions(lines):
    inside = False
    cur = []
    for ln in lines:
        if "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 1" in ln:
            inside = True
            cur = []
            continue
        if "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0" in ln and inside:
            inside = False
            yield cur
            continue
        if inside:
            cur.append(ln)


def test_round_up_regions_hold_only_asm_arithmetic(asm):
    found = 0
    for body in regions(asm):
        found += 1
        in_asm = False
        for ln in body:
            if ";;#ASMSTART" in ln:
                in_asm = True
                continue
            if ";;#ASMEND" in ln:
                in_asm = False
                continue
            m = ROUNDING.match(ln)
            assert not (m and not in_asm), f"compiler-generated rounding op inside the round-up region: {ln.strip()}"
            assert "s_swappc" not in ln and "s_setpc" not in ln, "call inside the round-up region"
            assert "s_setreg" not in ln or in_asm, "stray mode write inside the round-up region"
    assert found >= 2  # the explicit-box kernel and the decode bounder
