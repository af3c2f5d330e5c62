"""Build the gfx950 engine library and the CPU oracle, in-tree.

    python -m cuminlpsearch_amd.build            # both
    python -m cuminlpsearch_amd.build --engine   # libcmi.so only

Outputs: cuminlpsearch_amd/lib/libcmi.so (HIP kernels + C++ host + C-ABI) and
oracle/liboracle.so (the CPU restatement used by tests and the CPU baseline).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG = ROOT / "cuminlpsearch_amd"
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
OBJ_DIR = ROOT / "_build" / "engine"
ENGINE_SO = LIB_DIR / "libcmi.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no a*b+c fusion anywhere (the decode and the point
# evaluator are specified unfused); packed-fp32-ops disabled for correctness.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{CSRC}",
    f"-I{ROOT / 'include'}",
]

ENGINE_SOURCES = ["kernels.hip", "hip_backend.cpp", "batch.cpp", "capi.cpp", "server.cpp", "jit.cpp",
                  "aggregate_backend.cpp", "aggregate_solve.cpp", "gams.cpp"]
# interval_ops.hpp is also the prelude of every kernel compiled at run time
# from a tape; its text is embedded into the library.
EMBEDDED_TEXT = {"cmi_interval_ops_source": "interval_ops.hpp"}


def _embed_sources() -> list[pathlib.Path]:
    out = []
    for symbol, name in EMBEDDED_TEXT.items():
        text = (CSRC / name).read_text()
        if ")CMI_EMBED" in text:
            raise RuntimeError(f"{name} contains the raw-string delimiter")
        path = OBJ_DIR / f"embed_{symbol}.cpp"
        body = f'extern "C" const char {symbol}[] = R"CMI_EMBED({text})CMI_EMBED";\n'
        if not path.exists() or path.read_text() != body:
            path.write_text(body)
        out.append(path)
    return out


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_engine(jobs: int = 8) -> pathlib.Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    objs = []
    cmds = []
    for src in ENGINE_SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = OBJ_DIR / (src + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", str(path), "-o", str(obj)])
        objs.append(obj)
    for path in _embed_sources():
        obj = path.with_suffix(".o")
        cmds.append(["g++", "-O2", "-fPIC", "-c", str(path), "-o", str(obj)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = ENGINE_SO.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", "-o", str(tmp), *map(str, objs), "-lhiprtc"])
    os.replace(tmp, ENGINE_SO)
    return ENGINE_SO


ADAPTER_DIR = ROOT / "tests" / "adapter"
ADAPTER_TEST = ADAPTER_DIR / "build" / "adapter_test"


def adapter_test_command(out: pathlib.Path) -> list[str]:
    """g++ line of the reference-side adapter's test program (host code only:
    it reaches the device through libcmi.so, and checks against liboracle.so)."""
    return ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{ADAPTER_DIR}", f"-I{ROOT / 'include'}",
            str(ADAPTER_DIR / "adapter_test.cpp"), "-o", str(out), f"-L{LIB_DIR}", f"-L{ORACLE_DIR}", "-lcmi",
            "-loracle", "-Wl,-rpath,$ORIGIN/../../../cuminlpsearch_amd/lib:$ORIGIN/../../../oracle"]


def build_adapter_test() -> pathlib.Path:
    ADAPTER_TEST.parent.mkdir(parents=True, exist_ok=True)
    tmp = ADAPTER_TEST.with_suffix(".tmp")
    _run(adapter_test_command(tmp))
    os.replace(tmp, ADAPTER_TEST)
    return ADAPTER_TEST


def build_oracle() -> pathlib.Path:
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-s", "-C", str(ORACLE_DIR), "liboracle.so"])
    return ORACLE_SO


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--engine", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    both = not (a.engine or a.oracle)
    if a.engine or both:
        print(build_engine(a.j))
    if a.oracle or both:
        print(build_oracle())
    if both:
        print(build_adapter_test())
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
