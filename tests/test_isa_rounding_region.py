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


# -- the checker itself, on small hand-written listings --------------------------

_HEAD = ["\t.type\tk,@function", "k:"]
_ENTER = ["\t;;#ASMSTART", "\ts_nop 1", "\ts_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 1", "\t;;#ASMEND"]
_LEAVE = ["\t;;#ASMSTART", "\ts_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0 ; s0", "\t;;#ASMEND"]
_ASM_ADD = ["\t;;#ASMSTART", "\tv_add_f64 v[0:1], v[2:3], v[4:5] ; s0", "\t;;#ASMEND"]
_TAIL = ["\ts_endpgm", ".Lfunc_end0:"]


def test_checker_accepts_blocks_laid_out_after_the_exit():
    # a branch inside the region to a block placed after s_endpgm that
    # branches back: still inside the region by control flow
    body = _HEAD + _ENTER + ["\ts_cbranch_execz .LBB0_2", ".LBB0_1:"] + _ASM_ADD + _LEAVE + ["\ts_endpgm",
                                                                                          ".LBB0_2:"] + _ASM_ADD + [
        "\ts_branch .LBB0_1", ".Lfunc_end0:"]
    assert check_asm_regions(body) == 1


@pytest.mark.parametrize("bad", ["compiler_op_inside", "asm_op_outside", "mixed_modes", "call_inside"])
def test_checker_rejects(bad):
    plain_add = ["\tv_add_f64 v[0:1], v[2:3], v[4:5]"]
    if bad == "compiler_op_inside":
        body = _HEAD + _ENTER + plain_add + _LEAVE + _TAIL
    elif bad == "asm_op_outside":
        body = _HEAD + _ENTER + _LEAVE + _ASM_ADD + _TAIL
    elif bad == "mixed_modes":
        # a block reached both before and after the region's entry
        body = (_HEAD + ["\ts_cbranch_execz .LBB0_1"] + _ENTER + [".LBB0_1:"] + _ASM_ADD + _LEAVE + _TAIL)
    else:
        body = _HEAD + _ENTER + ["\ts_swappc_b64 s[30:31], s[16:17]"] + _LEAVE + _TAIL
    with pytest.raises(AssertionError):
        check_asm_regions(body)
