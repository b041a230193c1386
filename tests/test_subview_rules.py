"""Compile-time checks of Kokkos::subview's result layouts (the
static_asserts in tests/cpp/subview_rules.cpp), on the CPU: hipcc parses the
file for gfx950 without running anything."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
def test_subview_layout_rules_compile():
    cmd = [HIPCC, "-std=c++17", "--offload-arch=gfx950", "-fsyntax-only", "-x", "hip",
           "-I", str(ROOT / "include"), "-I", str(ROOT / "kokkos_amd" / "include"),
           str(ROOT / "tests" / "cpp" / "subview_rules.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
