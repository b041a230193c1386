"""Host-side atomics on types wider than 8 bytes (CPU only): two threads add
to and exchange / compare-exchange one Kokkos::complex<double>; every
operation on one address must take the same lock, so no update is lost
(tests/cpp/host_wide_atomics.cpp).  Built for the host only and run here."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
def test_wide_host_atomics_share_one_lock(tmp_path):
    exe = tmp_path / "host_wide_atomics"
    cmd = [HIPCC, "-std=c++17", "-O2", "-x", "hip", "--offload-host-only",
           "-I", str(ROOT / "include"), "-I", str(ROOT / "kokkos_amd" / "include"),
           str(ROOT / "tests" / "cpp" / "host_wide_atomics.cpp"), "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    for _ in range(3):
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
