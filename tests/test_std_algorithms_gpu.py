"""Kokkos::Experimental std-algorithm scans and reductions through the C ABI
(algorithms/src/std_algorithms/Kokkos_InclusiveScan.hpp and siblings)
against numpy / the oracle on the same seeded inputs: integers bit-exact,
float64 within 1e-12 relative of a compensated reference.  The iterator
forms (sub-ranges, strided Views, returned iterators) run as C++ self-tests
(std_* in tests/cpp/selftest.hip)."""
import ctypes as C
import math

import numpy as np
import pytest

import kokkos_amd as K

pytestmark = pytest.mark.gpu
SIZES = [1, 1000, 1 << 22]


def run(P, rc):
    K.check(rc, P)
    K.fence()


@pytest.mark.parametrize("n", SIZES)
def test_inclusive_scan_i32(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_i32(n, 51, 0, 7)
    S, D = K.DeviceBuffer.from_array(x), K.DeviceBuffer(n * 4)
    run(P, P.kkmx_std_inclusive_scan_i32(S.ptr, D.ptr, n))
    assert np.array_equal(D.download(np.int32, n), np.cumsum(x, dtype=np.int32))
    # in place
    run(P, P.kkmx_std_inclusive_scan_i32(S.ptr, S.ptr, n))
    assert np.array_equal(S.download(np.int32, n), np.cumsum(x, dtype=np.int32))


@pytest.mark.parametrize("n", SIZES)
def test_exclusive_scan_i64_with_init(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_i64(n, 52, -(1 << 40), 1 << 40)
    S, D = K.DeviceBuffer.from_array(x), K.DeviceBuffer(n * 8)
    run(P, P.kkmx_std_exclusive_scan_i64(S.ptr, D.ptr, n, 42))
    want = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int64) + 42
    assert np.array_equal(D.download(np.int64, n), want)
    run(P, P.kkmx_std_exclusive_scan_i64(S.ptr, S.ptr, n, 42))  # in place
    assert np.array_equal(S.download(np.int64, n), want)


@pytest.mark.parametrize("n", SIZES)
def test_inclusive_max_scan_no_neutral_element(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_i32(n, 53, -(1 << 30), 1 << 30)
    S, D = K.DeviceBuffer.from_array(x), K.DeviceBuffer(n * 4)
    run(P, P.kkmx_std_inclusive_scan_max_i32(S.ptr, D.ptr, n))
    assert np.array_equal(D.download(np.int32, n), np.maximum.accumulate(x))


@pytest.fixture
def scan_body(paths):
    """Pin the scan tile body for one test; the automatic choice afterwards."""
    def pin(body):
        K.check(paths.kkmx_set_scan_body(body), paths)
    yield pin
    K.check(paths.kkmx_set_scan_body(-1), paths)


@pytest.mark.parametrize("n", [1, 255, 256, 4099, 100003, (1 << 21) + 17, (1 << 22) + 5, (1 << 24) + 3])
@pytest.mark.parametrize("body", [0, 1])
def test_max_scan_both_tile_bodies(paths, oracle, scan_body, body, n):
    """The running maximum's value type wraps the int with an emptiness flag,
    so the contribution-keeping body gives each lane 4 consecutive indices
    per round; both bodies, ragged tiles and in-place calls equal numpy."""
    P = paths
    scan_body(body)
    x = oracle.fill_uniform_i32(n, 55 + n % 5, -(1 << 30), 1 << 30)
    want = np.maximum.accumulate(x)
    S, D = K.DeviceBuffer.from_array(x), K.DeviceBuffer(n * 4)
    run(P, P.kkmx_std_inclusive_scan_max_i32(S.ptr, D.ptr, n))
    assert np.array_equal(D.download(np.int32, n), want)
    run(P, P.kkmx_std_inclusive_scan_max_i32(S.ptr, S.ptr, n))  # in place
    assert np.array_equal(S.download(np.int32, n), want)


@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_transform_exclusive_scan_squares_f64(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_f64(n, 54, -1.0, 1.0)
    S, D = K.DeviceBuffer.from_array(x), K.DeviceBuffer(n * 8)
    run(P, P.kkmx_std_transform_exclusive_scan_sq_f64(S.ptr, D.ptr, n, 0.5))
    pre, _ = oracle.scan_excl_f64_compensated(x * x)  # x*x is exact in the functor and here
    want = pre + 0.5
    got = D.download(np.float64, n)
    err = np.abs(got - want) / want
    print(f"n={n} max rel err {err.max():.2e}")
    assert err.max() <= 1e-12


@pytest.mark.parametrize("n", SIZES)
def test_reduce_i64(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_i64(n, 55, -(1 << 40), 1 << 40)
    S = K.DeviceBuffer.from_array(x)
    out = C.c_int64()
    run(P, P.kkmx_std_reduce_i64(S.ptr, n, 7, C.byref(out)))
    assert out.value == int(x.sum()) + 7
    run(P, P.kkmx_std_reduce_max_i64(S.ptr, n, -(1 << 62), C.byref(out)))
    assert out.value == int(x.max())
    run(P, P.kkmx_std_reduce_max_i64(S.ptr, n, 1 << 50, C.byref(out)))  # init wins
    assert out.value == 1 << 50


@pytest.mark.parametrize("n", SIZES)
def test_transform_reduce_dot_f64(paths, oracle, n):
    P = paths
    a = oracle.fill_uniform_f64(n, 56, 0.0, 1.0)
    b = oracle.fill_uniform_f64(n, 57, 0.0, 1.0)
    A, B = K.DeviceBuffer.from_array(a), K.DeviceBuffer.from_array(b)
    out = C.c_double()
    run(P, P.kkmx_std_transform_reduce_dot_f64(A.ptr, B.ptr, n, 1.5, C.byref(out)))
    want = math.fsum((a * b).tolist()) + 1.5  # products rounded as in the kernel
    rel = abs(out.value - want) / want
    print(f"n={n} rel err {rel:.2e}")
    assert rel <= 1e-12


def test_empty_ranges(paths):
    P = paths
    out = C.c_int64()
    run(P, P.kkmx_std_reduce_i64(None, 0, 9, C.byref(out)))
    assert out.value == 9
    run(P, P.kkmx_std_inclusive_scan_i32(None, None, 0))
