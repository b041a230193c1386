"""deep_copy between device Views through the C ABI (ViewFill / ViewCopy,
core/src/Kokkos_CopyViews.hpp:58-420): fills, contiguous copies and
layout-changing copies, bit-exact against numpy on the same inputs.  The
strided-window, rank-2/3/5 and stream-ordered cases run as C++ self-tests
(deep_copy_* in tests/cpp/selftest.hip)."""
import numpy as np
import pytest

import kokkos_amd as K

pytestmark = pytest.mark.gpu


def run(P, rc):
    K.check(rc, P)
    K.fence()


@pytest.mark.parametrize("n", [1, 7, 1000003])
def test_fill(paths, n):
    P = paths
    D = K.DeviceBuffer(n * 8)
    run(P, P.kkmx_deep_fill_f64(D.ptr, n, 1.25))
    assert np.all(D.download(np.float64, n) == 1.25)


@pytest.mark.parametrize("n", [1, 1001, 1 << 20])
def test_contiguous_copy(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_f64(n, 41, -1.0, 1.0)
    S, D = K.DeviceBuffer.from_array(x), K.DeviceBuffer(n * 8)
    run(P, P.kkmx_deep_copy_f64(D.ptr, S.ptr, n, 1, 1, 0, 0))
    assert np.array_equal(D.download(np.float64, n), x)


@pytest.mark.parametrize("shape", [(37, 53, 11), (1, 5, 9), (64, 64, 64), (3, 1, 129), (100, 3, 70), (65, 1, 16), (16, 7, 200)])
@pytest.mark.parametrize("dst_layout,src_layout", [(1, 0), (0, 1), (1, 1)])
def test_layout_copy(paths, oracle, shape, dst_layout, src_layout):
    P = paths
    n0, n1, n2 = shape
    n = n0 * n1 * n2
    x = oracle.fill_uniform_f64(n, 42, -1.0, 1.0)
    S, D = K.DeviceBuffer.from_array(x), K.DeviceBuffer(n * 8)
    run(P, P.kkmx_deep_copy_f64(D.ptr, S.ptr, n0, n1, n2, dst_layout, src_layout))
    got = D.download(np.float64, n)
    # index space (i, j, k): LayoutLeft stores at i + n0*(j + n1*k), LayoutRight at (i*n1 + j)*n2 + k
    src = x.reshape(n2, n1, n0).transpose(2, 1, 0) if src_layout == 0 else x.reshape(n0, n1, n2)
    dst = got.reshape(n2, n1, n0).transpose(2, 1, 0) if dst_layout == 0 else got.reshape(n0, n1, n2)
    assert np.array_equal(dst, src)


def test_transpose_512_cubed(paths, oracle):
    """The measured case: 512^3 doubles, LayoutLeft -> LayoutRight."""
    P = paths
    n0 = n1 = n2 = 512
    n = n0 * n1 * n2
    S, D = K.DeviceBuffer(n * 8), K.DeviceBuffer(n * 8)
    run(P, P.kkmx_fill_uniform_f64(S.ptr, n, 43, -1.0, 1.0))
    run(P, P.kkmx_deep_copy_f64(D.ptr, S.ptr, n0, n1, n2, 1, 0))
    x = S.download(np.float64, n)
    got = D.download(np.float64, n)
    assert np.array_equal(got.reshape(n0, n1, n2), x.reshape(n2, n1, n0).transpose(2, 1, 0))
