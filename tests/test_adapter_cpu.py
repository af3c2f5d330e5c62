"""The reference-side adapter compiles against include/cmi.h and reference-shaped
types with every warning an error, and links against libcmi.so and the oracle."""
import subprocess
import sys

from cuminlpsearch_amd import build as B


def test_adapter_compiles_and_links(tmp_path):
    out = tmp_path / "adapter_test"
    run = subprocess.run(B.adapter_test_command(out), capture_output=True, text=True)
    assert run.returncode == 0, run.stderr[-4000:]
    assert out.exists()


def test_adapter_header_is_c_abi_only():
    """The adapter reaches the engine only through the C-ABI: it includes cmi.h
    and no engine-internal header."""
    text = (B.ROOT / "include" / "cmi_reference_adapter.hpp").read_text()
    includes = [ln for ln in text.splitlines() if ln.startswith("#include")]
    assert '#include "cmi.h"' in includes
    assert not any("cmi/" in ln or "hip" in ln for ln in includes), includes
