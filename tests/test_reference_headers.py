"""Source compatibility with the reference's own unit tests (CPU only).

Each header under the reference's core/unit_test/ is compiled where it lies,
read-only, against this build's headers with `hipcc -fsyntax-only` for
gfx950, exactly as the reference's test drivers include it: gtest, a category
header naming the execution space (TEST_EXECSPACE = Kokkos::MI355X), then the
test header.  Nothing is built, linked or run from the reference.

PASS lists the headers that must compile unchanged; the test fails if any of
them stops compiling.  The test skips when the reference tree is absent (it
is not on the GPU box).  `python tests/test_reference_headers.py` runs the
sweep over every header and prints the headers that compile and the first
error of each that does not.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
UNIT = REF / "core" / "unit_test"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Headers of core/unit_test/ that compile unchanged.
PASS = sorted("""
KokkosExecutionEnvironmentNeverInitializedFixture.hpp TestAbort.hpp TestArrayOps.hpp TestAtomicOperations.hpp
TestAtomicOperations_complexdouble.hpp TestAtomicOperations_complexfloat.hpp TestAtomicOperations_double.hpp
TestAtomicOperations_float.hpp TestAtomicOperations_int.hpp TestAtomicOperations_longint.hpp
TestAtomicOperations_longlongint.hpp TestAtomicOperations_shared.hpp TestAtomicOperations_unsignedint.hpp
TestAtomicOperations_unsignedlongint.hpp TestAtomicOperations_unsignedlonglongint.hpp TestAtomicViews.hpp
TestAtomics.hpp TestBlockSizeDeduction.hpp TestCXX11.hpp TestCXX11Deduction.hpp TestExecSpacePartitioning.hpp
TestExecSpaceThreadSafety.hpp TestExecutionSpace.hpp TestGraph.hpp TestGraphAtomicLocks.hpp
TestJoinBackwardCompatibility.hpp TestLocalDeepCopy.hpp TestMDRange.hpp TestMDRangePolicyConstructors.hpp
TestMDRangeReduce.hpp TestMDRange_a.hpp TestMDRange_b.hpp TestMDRange_c.hpp TestMDRange_d.hpp
TestMDRange_e.hpp TestMDRange_f.hpp TestMDRange_g.hpp TestMDSpan.hpp TestMathematicalConstants.hpp
TestMultiGPU.hpp TestNonTrivialScalarTypes.hpp TestOccupancyControlTrait.hpp TestParallelScanRangePolicy.hpp
TestPrintf.hpp TestQuadPrecisionMath.hpp TestRange.hpp TestRangePolicy.hpp TestRangePolicyConstructors.hpp
TestRangePolicyRequire.hpp TestRealloc.hpp TestReduce.hpp
TestReduceCombinatorical.hpp TestReductions.hpp TestReductions_DeviceView.hpp TestResize.hpp TestSwap.hpp
TestTaskScheduler.hpp TestTeam.hpp TestTeamBasic.hpp TestTeamCombinedReducers.hpp TestTeamMDRange.hpp
TestTeamPolicyConstructors.hpp TestTeamReductionScan.hpp TestTeamScan.hpp TestTeamScratch.hpp
TestTeamTeamSize.hpp TestTeamVector.hpp TestTeamVectorRange.hpp TestViewAPI.hpp TestViewAPI_a.hpp
TestViewAPI_c.hpp TestViewAPI_d.hpp TestViewBadAlloc.hpp TestViewCopy_a.hpp TestViewCopy_b.hpp
TestViewCopy_c.hpp TestViewCtorDimMatch.hpp TestViewEmptyRuntimeUnmanaged.hpp
TestViewLayoutStrideAssignment.hpp TestViewMemoryAccessViolation.hpp TestViewOfClass.hpp
TestViewSpaceAssign.hpp TestView_64bit.hpp Test_InterOp_Streams.hpp
""".split())

# Whole test programs of core/unit_test/tools/ that compile unchanged (the
# built-in tuners and the tuning interface).
TOOLS_PASS = ["TestBuiltinTuners.cpp", "TestCategoricalTuner.cpp", "TestTuning.cpp"]

SOURCE = """#include <gtest/gtest.h>
#define TEST_CATEGORY mi355x
#define TEST_CATEGORY_NUMBER 6
#define TEST_CATEGORY_DEATH mi355x_DeathTest
#define TEST_EXECSPACE Kokkos::MI355X
#define TEST_CATEGORY_FIXTURE(name) mi355x_##name
#include <Kokkos_Core.hpp>
#include <{header}>
"""


def compile_header(header: str) -> tuple[str, bool, str]:
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "probe.cpp"
        src.write_text(SOURCE.format(header=header))
        cmd = [
            HIPCC, "-std=c++17", "-fsyntax-only", "--offload-arch=gfx950", "-x", "hip",
            "-I", str(ROOT / "kokkos_amd" / "include"), "-I", str(ROOT / "include"),
            "-I", str(UNIT), "-I", str(UNIT / "category_files"), "-I", str(REF / "tpls" / "gtest"),
            "-Wno-unused-command-line-argument", "-ferror-limit=3", str(src),
        ]
        r = subprocess.run(cmd, capture_output=True, text=True)
        first = next((ln for ln in (r.stdout + r.stderr).splitlines() if "error:" in ln), "")
        return header, r.returncode == 0, first


def compile_tools_program(name: str) -> tuple[str, bool, str]:
    cmd = [
        HIPCC, "-std=c++17", "-fsyntax-only", "--offload-arch=gfx950", "-x", "hip",
        "-I", str(ROOT / "kokkos_amd" / "include"), "-I", str(ROOT / "include"),
        "-I", str(UNIT), "-I", str(REF / "tpls" / "gtest"),
        "-Wno-unused-command-line-argument", "-ferror-limit=3", str(UNIT / "tools" / name),
    ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    first = next((ln for ln in (r.stdout + r.stderr).splitlines() if "error:" in ln), "")
    return name, r.returncode == 0, first


def _sweep(headers, jobs):
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(compile_header, headers))


needs_reference = pytest.mark.skipif(not UNIT.is_dir(), reason="reference tree not present")


@needs_reference
def test_pinned_reference_headers_compile():
    missing = [h for h in PASS if not (UNIT / h).is_file()]
    assert not missing, missing
    res = _sweep(PASS, min(8, os.cpu_count() or 1))
    bad = [(h, err) for h, ok, err in res if not ok]
    assert not bad, bad


@needs_reference
def test_pinned_tools_programs_compile():
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        res = list(ex.map(compile_tools_program, TOOLS_PASS))
    bad = [(h, err) for h, ok, err in res if not ok]
    assert not bad, bad


if __name__ == "__main__":
    headers = sorted(p.name for p in UNIT.glob("*.hpp"))
    res = _sweep(headers, min(8, os.cpu_count() or 1))
    ok = [h for h, good, _ in res if good]
    print(f"{len(ok)} of {len(headers)} compile")
    for h, good, err in res:
        print(("PASS " if good else "FAIL ") + h + ("" if good else "  " + err[:200]))
    sys.exit(0)
