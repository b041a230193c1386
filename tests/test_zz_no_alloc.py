"""The no-allocation variant of the inclusive int32 scan
(kkmx_scan_incl_i32_noalloc) against the ordinary entry point
(kkmx_scan_incl_i32), bit for bit, through the C ABI: two calls in a row,
the second on a smaller input, so the second call runs over the first call's
leftover scratch contents."""
import numpy as np
import pytest

import kokkos_amd as K


def run(P, rc):
    K.check(rc, P)
    K.fence()


@pytest.mark.gpu
def test_noalloc_scan_matches_scan_entry_point(paths, oracle):
    P = paths
    for n, seed in ((3 * (1 << 22) + 17, 41), (100003, 42), (1, 43)):
        print(f"noalloc scan n={n} seed={seed}", flush=True)
        x = oracle.fill_uniform_i32(n, seed, 0, 7)
        X = K.DeviceBuffer.from_array(x)
        A, B = K.DeviceBuffer(n * 4), K.DeviceBuffer(n * 4)
        run(P, P.kkmx_scan_incl_i32(X.ptr, A.ptr, n))
        run(P, P.kkmx_scan_incl_i32_noalloc(X.ptr, B.ptr, n))
        a, b = A.download(np.int32, n), B.download(np.int32, n)
        assert np.array_equal(a, b), n
        assert np.array_equal(b, oracle.scan_incl_i32(x)), n


@pytest.mark.gpu
def test_noalloc_scan_rejects_sizes_beyond_its_scratch(paths):
    P = paths
    X = K.DeviceBuffer(16)
    rc = P.kkmx_scan_incl_i32_noalloc(X.ptr, X.ptr, (1 << 31) + 1)
    assert rc == -1
    assert b"2^31" in P.kkmx_last_error()
    assert P.kkmx_scan_incl_i32_noalloc(X.ptr, X.ptr, 0) == 0
