"""The C++ snippets of INTEGRATION.md compile against this build's headers
(hipcc -fsyntax-only for gfx950), so the integration guide stays true."""
from __future__ import annotations

import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def cxx_blocks() -> list[str]:
    text = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"```c\+\+\n(.*?)```", text, flags=re.S)


def test_integration_has_cxx_snippets():
    assert cxx_blocks(), "INTEGRATION.md lost its C++ snippets"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
@pytest.mark.parametrize("idx", range(len(cxx_blocks())))
def test_integration_snippet_compiles(tmp_path, idx):
    src = cxx_blocks()[idx]
    if "int main" not in src:
        src += "\nint main() { return 0; }\n"
    f = tmp_path / f"snippet{idx}.cpp"
    f.write_text(src)
    cmd = [HIPCC, "-std=c++17", "-fsyntax-only", "--offload-arch=gfx950", "-x", "hip",
           "-I", str(ROOT / "kokkos_amd" / "include"), "-I", str(ROOT / "include"),
           "-Wno-unused-command-line-argument", str(f)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
