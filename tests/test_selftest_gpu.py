"""The reference's own unit tests (core/unit_test: TestGraph, TestReduce,
TestReducers, TestParallelScanRangePolicy, TestMDRange, TestTeam*,
TestAtomics*), restated in C++ against the MI355X headers and run on the GPU.

Each case is compiled user code (tests/cpp/selftest.hip) checking the
closed forms the reference asserts: integer results exactly, the float scan
and reducer cases within the tolerance written next to each check.
"""
import pytest

import kokkos_amd as K
from selftest_lib import run_selftest, selftests

try:
    NAMES = selftests()
except K.KkmxError:  # not built: the GPU run fails loudly in test_selftests_built
    NAMES = []


@pytest.mark.gpu
def test_selftests_built():
    assert len(selftests()) >= 40


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_reference_unit(paths, name):
    print(f"selftest {name}", flush=True)
    rc, msg = run_selftest(name)
    assert rc == 0, f"{name}: status {rc}: {msg}"
