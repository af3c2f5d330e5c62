"""CPU checks of the explicit-box kernels compiled at run time from a tape:
the generated source and the code object the run-time compiler makes from it
(exactly as the engine does on a GPU) keep every rounding FP64 instruction
inside the one round-toward-+infinity region, and every instruction there comes
from the interval operations' inline asm."""
import pathlib
import re
import shutil
import subprocess

import pytest

import cuminlpsearch_amd as cmi
from isa_tools import ENTER, LEAVE, PACKED_F32, ROUNDING, check_asm_regions

import op_cases
import problems as P

HIPCC = "/opt/rocm/bin/hipcc"
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"
# the run-time compiler's options (jit.cpp)
RTC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def problems():
    yield "c2", P.c2()
    for name in ("const_root", "shared_generic", "acc_forms", "eq_acc", "mul_imm_zero"):
        yield name, cmi.EngineProblem.from_problem(op_cases.CASES[name]())
    for name in ("ext_mixed", "div", "pown_neg_odd", "sin", "log", "tanh", "pow"):
        yield name, cmi.EngineProblem.from_problem(op_cases.EXT_CASES[name]())


@pytest.mark.parametrize("form", ["direct", "persistent", "batched_no_roots"])
@pytest.mark.parametrize("case", [name for name, _ in problems()])
def test_generated_source_regions(tmp_path, case, form):
    """Every form of the generated source -- direct, persistent, and the direct
    form with batched loads (scheduling barriers) and no root stores -- keeps
    every rounding instruction inside the one round-up region."""
    if not pathlib.Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    eng = dict(problems())[case]
    src = tmp_path / "k.hip"
    if form == "batched_no_roots":
        text = eng.compiled_source(0, 0, 0, 0, 8, False)
    else:
        text = eng.compiled_source(1 if form == "persistent" else 0)
    # hiprtc includes the HIP device header by itself; hipcc needs it named
    src.write_text("#include <hip/hip_runtime.h>\n" + text)
    out = tmp_path / "k.s"
    subprocess.run([HIPCC, *RTC_FLAGS, "--cuda-device-only", "-S", "-o", str(out), str(src)], check=True,
                   capture_output=True)
    assert check_asm_regions(out.read_text().splitlines()) == 1


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("case", ["c2", "shared_generic"])
def test_runtime_code_object_rounds_only_inside_the_region(tmp_path, case, persistent, schedule):
    """schedule 1: the max-ILP instruction scheduler, which reorders independent
    instructions but keeps the region's entry and exit around every rounding one."""
    if not pathlib.Path(OBJDUMP).exists():
        pytest.skip("llvm-objdump not available")
    eng = dict(problems())[case]
    co = tmp_path / "k.co"
    co.write_bytes(eng.compiled_object(persistent, 0, schedule, load_batch=8 * schedule, roots=bool(schedule)))
    dis = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", str(co)], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    enters = [i for i, ln in enumerate(dis) if ENTER in ln]
    leaves = [i for i, ln in enumerate(dis) if LEAVE in ln]
    assert len(enters) == 1 and len(leaves) == 1 and enters[0] < leaves[0]
    rounding = [i for i, ln in enumerate(dis) if ROUNDING.match(ln.strip())]
    assert rounding and all(enters[0] < i < leaves[0] for i in rounding)
    assert not any(PACKED_F32.match(ln.strip()) for ln in dis)
    assert not any(re.match(r"\s*scratch_", ln) for ln in dis), "register spills in the compiled kernel"


@pytest.mark.parametrize("batch", [1, 3, 8, 64])
@pytest.mark.parametrize("groups,waves", [(1, 1), (8, 2)])
def test_batched_loads_follow_first_use(groups, waves, batch):
    """Direct form with load batching: each wave loads every variable once,
    batches of `batch` in first-use order between scheduling barriers, the
    first two up front, and never reads a variable before its load."""
    eng = P.c2()
    src = eng.compiled_source(0, waves, groups, 0, batch, True)
    kern = src[src.index("__global__"):]
    blocks = re.split(r"\n  (?:else )?if \(wave == \d+u\) \{\n", kern)[1:]
    assert len(blocks) == waves
    for blk in blocks:
        lines = blk.split("\n")
        loads = [re.match(r"\s+const Iv (s\d+) = load_box", ln) for ln in lines]
        names = [m.group(1) for m in loads if m]
        assert len(names) == len(set(names))
        loaded = set()
        first = []
        for ln, m in zip(lines, loads):
            if m:
                loaded.add(m.group(1))
                continue
            for v in re.findall(r"\b(s\d+)\b", ln):
                assert v in loaded, (v, ln)
                if v not in first:
                    first.append(v)
        assert names[:len(first)] == first  # loads issued in first-use order
        barriers = [i for i, ln in enumerate(lines) if "sched_barrier" in ln]
        assert len(barriers) == 2 * ((len(names) + batch - 1) // batch)


def _statements(src):
    """The kernel's statements without loads and barriers, with every held
    product (const Iv pN = ...) written back into the statements that use it."""
    kern = src[src.index("__global__"):]
    held = {}
    out = []
    for ln in kern.split("\n"):
        ln = ln.strip()
        if not ln or "load_box" in ln or "sched_barrier" in ln:
            continue
        m = re.match(r"const Iv (p\d+) = (.*);$", ln)
        if m:
            held[m.group(1)] = m.group(2)
            continue
        out.append(re.sub(r"\b(p\d+)\b", lambda q: held.get(q.group(1), q.group(1)), ln))
    return out


@pytest.mark.parametrize("case", ["c2", "shared_generic", "acc_forms", "ext_mixed"])
def test_batched_form_reorders_whole_functions_only(case):
    """With batched loads the statements are the same multiset as in tape
    order. For a tape whose functions share no temporary slot (c2) whole
    function blocks move, the one reading the fewest variables first; a tape
    with temporaries keeps its order."""
    eng = dict(problems())[case]
    plain = _statements(eng.compiled_source(0, 1, 1, 0, 0, True))
    batched = _statements(eng.compiled_source(0, 1, 1, 0, 8, True))
    assert sorted(plain) == sorted(batched)
    words, _, _ = eng.tape_words(1)
    nv = eng.info().n_vars
    temps = any(int(w["dst"]) >= nv and (int(w["op"]) < 16 or 32 <= int(w["op"]) < 47 or int(w["op"]) == 16)
                for w in words)
    if temps:
        assert plain == batched
    if case == "c2":
        assert plain != batched
        first_block = []
        for ln in batched[batched.index("Iv obj = {0.0, 0.0};") + 1:]:
            if first_block and not ln.startswith("acc0 = iv_add(acc0"):
                break
            first_block.append(ln)
        vars_first = set(re.findall(r"\bs(\d+)\b", " ".join(first_block)))
        assert len(vars_first) == 7  # the constraint of the 600-node DAG that reads the fewest variables


def test_no_roots_kernel_writes_no_roots():
    """The kernel for calls without root intervals has no root stores, and the
    compiler then drops the constraint upper bounds nothing reads: fewer FP64
    instructions than the kernel that writes them."""
    if not pathlib.Path(OBJDUMP).exists():
        pytest.skip("llvm-objdump not available")
    eng = P.c2()
    assert "out_store(rt" in eng.compiled_source(0, 1, 1, 0, 8, True)
    assert "out_store(rt" not in eng.compiled_source(0, 1, 1, 0, 8, False)
    counts = []
    for roots in (True, False):
        dis = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", "/dev/stdin"], input=eng.compiled_object(
            0, 1, 0, groups=1, load_batch=8, roots=roots), check=True, capture_output=True).stdout.decode()
        counts.append(len(re.findall(r"\sv_\w+_f64", dis)))
    assert counts[1] < 0.8 * counts[0], counts


def test_source_is_deterministic_and_exact():
    eng = P.c2()
    a, b = eng.compiled_source(), eng.compiled_source()
    assert a == b
    words, _, _ = eng.tape_words(1)
    # every immediate appears as an exact hexadecimal literal
    for imm in {float(w["imm"]) for w in words if float(w["imm"]) != 0.0}:
        assert f"({_c_hex(imm)})" in a, imm


def _c_hex(x: float) -> str:
    """printf("%a") spelling of a finite nonzero double."""
    m, e = float.hex(x).split("p")
    if "." in m:
        m = m.rstrip("0").rstrip(".")
    return f"{m}p{e}"


def _wave_bodies(src: str):
    body = src[src.index("cmi_boxes_compiled("):]
    parts = re.split(r"(?:else )?if \(wave == \d+u\) \{", body)[1:]
    return parts


def test_shared_products_are_formed_once_per_wave():
    """The QCQP's product nodes are shared between functions: a wave forms
    each distinct product of variables once and reuses it (c2: 145 bilinear
    and 48 square terms over 77 + 32 distinct products)."""
    eng = P.c2()
    one = _wave_bodies(eng.compiled_source(0, 1))
    assert len(one) == 1
    formed = len(re.findall(r"iv_mul\(s\d+, s\d+, t\)", one[0])) + len(re.findall(r"iv_sqr\(s\d+, t\)", one[0]))
    assert formed == 109
    # split over waves, each wave still forms each of its products once
    for body in _wave_bodies(eng.compiled_source(0, 2)):
        calls = re.findall(r"iv_mul\(s\d+, s\d+, t\)|iv_sqr\(s\d+, t\)", body)
        assert len(calls) == len(set(calls))


def test_products_of_rewritten_slots_are_not_shared():
    """A term over a temporary slot is formed where it stands: the tape may
    write that slot again between two such terms."""
    eng = cmi.EngineProblem.from_problem(op_cases.CASES["shared_generic"]())
    src = eng.compiled_source(0, 1)
    n_vars = eng.info().n_vars
    for m in re.finditer(r"const Iv p\d+ = iv_(?:mul\(s(\d+), s(\d+)|sqr\(s(\d+)), t\)", src):
        assert all(int(x) < n_vars for x in m.groups() if x is not None)
