"""Kokkos::Experimental::ScatterView (containers/src/Kokkos_ScatterView.hpp)
through the C ABI, against the oracle on the same seeded inputs: the
float64 scatter-add within 1e-10 relative per bin (atomic order varies),
min / max exact, the Zipf histogram bit-exact.  Reset, contribute into a
separate View, prod and the integer forms run as C++ self-tests
(scatter_view_* in tests/cpp/selftest.hip)."""
import numpy as np
import pytest

import kokkos_amd as K

pytestmark = pytest.mark.gpu


def run(P, rc):
    K.check(rc, P)
    K.fence()


@pytest.mark.parametrize("n,nbins", [(1000, 16), (1 << 20, 1 << 10), (1 << 24, 1 << 20)])
def test_scatter_view_sum_f64(paths, oracle, n, nbins):
    P = paths
    vals = oracle.fill_uniform_f64(n, 9, 0.0, 1.0)
    idx = oracle.fill_uniform_i32(n, 9 ^ 0xABCDEF, 0, nbins - 1)
    V, I, B = K.DeviceBuffer.from_array(vals), K.DeviceBuffer.from_array(idx), K.DeviceBuffer(nbins * 8)
    run(P, P.kkmx_deep_fill_f64(B.ptr, nbins, 0.0))
    run(P, P.kkmx_scatter_view_f64(V.ptr, I.ptr, n, B.ptr, nbins, 0))
    got = B.download(np.float64, nbins)
    ref, absb = oracle.scatter_f64(vals, idx, nbins)
    nz = absb > 0
    err = np.max(np.abs(got[nz] - ref[nz]) / absb[nz]) if nz.any() else 0.0
    print(f"n={n} bins={nbins} max rel err {err:.2e}")
    assert err <= 1e-10
    assert np.all(got[~nz] == 0.0)


@pytest.mark.parametrize("op", [1, 2])
def test_scatter_view_min_max_f64(paths, oracle, op):
    P = paths
    n, nbins = 1 << 20, 4096
    vals = oracle.fill_uniform_f64(n, 19, -1.0, 1.0)
    idx = oracle.fill_uniform_i32(n, 23, 0, nbins - 1)
    V, I, B = K.DeviceBuffer.from_array(vals), K.DeviceBuffer.from_array(idx), K.DeviceBuffer(nbins * 8)
    start = 1e300 if op == 1 else -1e300
    run(P, P.kkmx_deep_fill_f64(B.ptr, nbins, start))
    run(P, P.kkmx_scatter_view_f64(V.ptr, I.ptr, n, B.ptr, nbins, op))
    want = np.full(nbins, start)
    (np.minimum if op == 1 else np.maximum).at(want, idx, vals)
    assert np.array_equal(B.download(np.float64, nbins), want)


@pytest.mark.parametrize("duplicated", [0, 1])
def test_scatter_view_histogram(paths, oracle, duplicated):
    P = paths
    n, nb = 1 << 22, 65536
    Kb, Bb = K.DeviceBuffer(n * 4), K.DeviceBuffer(nb * 4)
    run(P, P.kkmx_zipf_keys(Kb.ptr, n, 1.1, nb, 8))
    K.load("kkmx_rt").mx_memset_async(Bb.ptr, 0, nb * 4, P.kkmx_stream())
    run(P, P.kkmx_scatter_view_histogram(Kb.ptr, n, Bb.ptr, nb, duplicated))
    assert np.array_equal(Bb.download(np.int32, nb), oracle.histogram(oracle.zipf_keys(n, 1.1, nb, 8), nb))
