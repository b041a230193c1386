"""Parity of the RangePolicy hot path on the MI355X against the CPU oracle.

Integers bit-exact; FP64 reductions within 1e-12 relative error of the exact
(correctly rounded) sum, with the measured error printed.  Sizes cover the
edge cases the reference tests (empty, one element, tile boundaries) and the
full configured sizes (2^24, 2^28) through checksums of whole outputs.
"""
import numpy as np
import pytest

import kokkos_amd as K

pytestmark = pytest.mark.gpu

EDGE = [0, 1, 2, 63, 64, 65, 255, 256, 257, 4095, 4096, 4097, 16383, 16384, 16385, 100003, 1 << 20]


def dev(arr):
    return K.DeviceBuffer.from_array(arr) if arr.size else K.DeviceBuffer(8)


def run(P, rc):
    K.check(rc, P)
    K.fence()


@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_fills_match_oracle(paths, oracle, n):
    P = paths
    b = K.DeviceBuffer(n * 8)
    run(P, P.kkmx_fill_uniform_f64(b.ptr, n, 0, 0.0, 1.0))
    assert np.array_equal(b.download(np.float64, n), oracle.fill_uniform_f64(n, 0, 0.0, 1.0))
    run(P, P.kkmx_fill_uniform_i64(b.ptr, n, 3, -(1 << 40), 1 << 40))
    assert np.array_equal(b.download(np.int64, n), oracle.fill_uniform_i64(n, 3, -(1 << 40), 1 << 40))
    run(P, P.kkmx_fill_uniform_i32(b.ptr, n, 4, 0, 7))
    assert np.array_equal(b.download(np.int32, n), oracle.fill_uniform_i32(n, 4, 0, 7))


@pytest.mark.parametrize("n", [0, 1, 257, 100003])
def test_stream_kernels_bit_exact(paths, oracle, n):
    P = paths
    a = oracle.fill_uniform_f64(n, 2, -1.0, 1.0)
    b = oracle.fill_uniform_f64(n, 3, -1.0, 1.0)
    c = oracle.fill_uniform_f64(n, 4, -1.0, 1.0)
    A, B, C = dev(a), dev(b), dev(c)
    run(P, P.kkmx_stream_copy(A.ptr, C.ptr, n))
    assert np.array_equal(C.download(np.float64, n), a)
    run(P, P.kkmx_stream_scale(B.ptr, C.ptr, n, 3.0))
    assert np.array_equal(B.download(np.float64, n), 3.0 * a)
    B.upload(b) if n else None
    run(P, P.kkmx_stream_add(A.ptr, B.ptr, C.ptr, n))
    assert np.array_equal(C.download(np.float64, n), a + b)
    run(P, P.kkmx_stream_triad(A.ptr, B.ptr, C.ptr, n, 3.0))
    got = A.download(np.float64, n)
    # device code contracts b + s*c into one fma: bit-exact against the fused
    # oracle, and within one ulp of the unfused Serial expression
    assert np.array_equal(got, oracle.stream_triad(b, a + b, 3.0, fused=True))
    plain = oracle.stream_triad(b, a + b, 3.0, fused=False)
    assert np.all(np.abs(got - plain) <= np.spacing(np.abs(plain)))
    run(P, P.kkmx_stream_set(A.ptr, n, 2.5))
    assert np.all(A.download(np.float64, n) == 2.5)


@pytest.mark.parametrize("n", EDGE)
def test_reduce_int64_bit_exact(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_i64(n, 3, -(1 << 40), 1 << 40)
    X, out = dev(x), K.DeviceBuffer(32)
    for op in (0, 1, 2):
        run(P, P.kkmx_reduce_i64(X.ptr, n, op, out.ptr))
        assert int(out.download(np.int64, 1)[0]) == oracle.reduce_i64(x, op)
    run(P, P.kkmx_reduce_minmaxsum_i64(X.ptr, n, out.ptr))
    mn, mx, s = out.download(np.int64, 3).tolist()
    assert (mn, mx, s) == (oracle.reduce_i64(x, 1), oracle.reduce_i64(x, 2), oracle.reduce_i64(x, 0))


def test_reduce_int32_and_closed_forms(paths, oracle):
    P = paths
    out = K.DeviceBuffer(8)
    for n in [1, 10, 1000, 1 << 20]:
        i = np.arange(n, dtype=np.int64)
        X = dev(i + 1)
        run(P, P.kkmx_reduce_i64(X.ptr, n, 0, out.ptr))
        assert int(out.download(np.int64, 1)[0]) == n * (n + 1) // 2
        Y = dev((n - i).astype(np.int32))
        run(P, P.kkmx_reduce_i32(Y.ptr, n, 0, out.ptr))
        assert int(out.download(np.int64, 1)[0]) == n * (n + 1) // 2
    x = oracle.fill_uniform_i32(99991, 7, -1000, 1000)
    X = dev(x)
    for op in (0, 1, 2):
        run(P, P.kkmx_reduce_i32(X.ptr, x.size, op, out.ptr))
        assert int(out.download(np.int64, 1)[0]) == oracle.reduce_i32(x, op)


@pytest.mark.parametrize("n,seed", [(1, 0), (1000, 0), (1 << 24, 0)])
def test_reduce_sum_f64_error_vs_exact(paths, oracle, n, seed):
    P = paths
    x = oracle.fill_uniform_f64(n, seed, 0.0, 1.0)
    X, out = dev(x), K.DeviceBuffer(8)
    run(P, P.kkmx_reduce_sum_f64(X.ptr, n, out.ptr))
    got = float(out.download(np.float64, 1)[0])
    exact = oracle.sum_exact(x)
    rel = abs(got - exact) / abs(exact)
    neu = abs(oracle.sum_neumaier(x) - exact) / abs(exact)
    print(f"sum_f64 n={n}: device={got!r} exact={exact!r} rel_err={rel:.3e} (neumaier {neu:.3e})")
    assert rel <= 1e-12
    # the result is reproducible run to run (fixed grid, ordered combine)
    run(P, P.kkmx_reduce_sum_f64(X.ptr, n, out.ptr))
    assert float(out.download(np.float64, 1)[0]) == got


def test_reduce_sum_f64_known_answers(paths):
    P = paths
    out = K.DeviceBuffer(8)
    # integer-valued doubles: exact in any order (TestReduce.hpp:656-667 idea)
    ones = np.ones(3 << 20, np.float64)
    run(P, P.kkmx_reduce_sum_f64(dev(ones).ptr, ones.size, out.ptr))
    assert out.download(np.float64, 1)[0] == ones.size
    d = np.arange(1 << 16, dtype=np.float64) * 2.0**-20
    run(P, P.kkmx_reduce_sum_f64(dev(d).ptr, d.size, out.ptr))
    assert out.download(np.float64, 1)[0] == ((1 << 16) - 1) * (1 << 16) / 2 * 2.0**-20
    run(P, P.kkmx_reduce_sum_f64(K.DeviceBuffer(8).ptr, 0, out.ptr))
    assert out.download(np.float64, 1)[0] == 0.0


@pytest.mark.parametrize("n", EDGE)
def test_scans_int_bit_exact(paths, oracle, n):
    P = paths
    x = oracle.fill_uniform_i32(n, 1, 0, 15)
    X = dev(x)
    o64, tot = K.DeviceBuffer(max(n, 1) * 8), K.DeviceBuffer(8)
    run(P, P.kkmx_scan_excl_i32_i64(X.ptr, o64.ptr, n, tot.ptr))
    ref, rtot = oracle.scan_excl_i32_i64(x)
    assert np.array_equal(o64.download(np.int64, n), ref)
    assert int(tot.download(np.int64, 1)[0]) == rtot
    o32 = K.DeviceBuffer(max(n, 1) * 4)
    run(P, P.kkmx_scan_incl_i32(X.ptr, o32.ptr, n))
    assert np.array_equal(o32.download(np.int32, n), oracle.scan_incl_i32(x))
    y = oracle.fill_uniform_i64(n, 5, -(1 << 40), 1 << 40)
    Y = dev(y)
    run(P, P.kkmx_scan_excl_i64(Y.ptr, o64.ptr, n, tot.ptr))
    ref, rtot = oracle.scan_excl_i64(y)
    assert np.array_equal(o64.download(np.int64, n), ref)
    assert int(tot.download(np.int64, 1)[0]) == rtot


def test_scan_reference_closed_form(paths):
    # TestParallelScanRangePolicy: over 0..N-1 prefix i(i-1)/2, total N(N-1)/2
    P = paths
    for n in [1, 2, 7, 64, 1000, 4097, 1 << 20]:
        i = np.arange(n, dtype=np.int64)
        X = dev(i)
        o, t = K.DeviceBuffer(n * 8), K.DeviceBuffer(8)
        run(P, P.kkmx_scan_excl_i64(X.ptr, o.ptr, n, t.ptr))
        assert np.array_equal(o.download(np.int64, n), i * (i - 1) // 2)
        assert int(t.download(np.int64, 1)[0]) == n * (n - 1) // 2


def test_scan_f64(paths, oracle):
    P = paths
    # integer-valued doubles: exact
    n = 300001
    x = oracle.fill_uniform_i32(n, 8, 0, 100).astype(np.float64)
    X, o, t = dev(x), K.DeviceBuffer(n * 8), K.DeviceBuffer(8)
    run(P, P.kkmx_scan_excl_f64(X.ptr, o.ptr, n, t.ptr))
    ref, rt = oracle.scan_excl_f64(x)
    assert np.array_equal(o.download(np.float64, n), ref) and t.download(np.float64, 1)[0] == rt
    # random values: each prefix within n*eps of the Serial prefix
    y = oracle.fill_uniform_f64(n, 9, 0.0, 1.0)
    Y = dev(y)
    run(P, P.kkmx_scan_excl_f64(Y.ptr, o.ptr, n, t.ptr))
    ref, rt = oracle.scan_excl_f64(y)
    got = o.download(np.float64, n)
    err = np.max(np.abs(got - ref) / np.maximum(ref, 1.0))
    print(f"scan_f64 max rel err vs Serial: {err:.3e}")
    assert err < 1e-12


def test_scan_state_reuse_across_launches(paths, oracle):
    """Tile counters and epochs are self-cleaning: alternate big and small
    scans on the same instance and check every result."""
    P = paths
    for n in [1 << 20, 5, 70000, 1 << 20, 1, 4097]:
        x = oracle.fill_uniform_i32(n, n, 0, 15)
        X, o = dev(x), K.DeviceBuffer(n * 4)
        run(P, P.kkmx_scan_incl_i32(X.ptr, o.ptr, n))
        assert np.array_equal(o.download(np.int32, n), oracle.scan_incl_i32(x)), n


@pytest.fixture
def scan_body(paths):
    """Pin the scan tile body for one test; the automatic choice afterwards."""
    def pin(body):
        K.check(paths.kkmx_set_scan_body(body), paths)
    yield pin
    K.check(paths.kkmx_set_scan_body(-1), paths)


@pytest.mark.parametrize("n", [1, 64, 4097, 100003, (1 << 21) + 17, 1 << 22, (1 << 24) + 5])
def test_scan_bodies_and_view_traits_bit_exact(paths, oracle, scan_body, n):
    """Both register-resident tile bodies (keep the wave scans / keep the
    contributions), with the input read through a plain, a RandomAccess and a
    Restrict View, equal the Serial scan bit for bit; ragged last tiles
    included (n not a multiple of any tile)."""
    P = paths
    x = oracle.fill_uniform_i32(n, 40 + n % 7, -100, 100)
    ref = oracle.scan_incl_i32(x)
    X, o = dev(x), K.DeviceBuffer(n * 4)
    for body in (0, 1):
        scan_body(body)
        for traits in (0, 1, 2):
            o.upload(np.zeros(n, np.int32))
            run(P, P.kkmx_scan_incl_i32_traits(X.ptr, o.ptr, n, traits))
            assert np.array_equal(o.download(np.int32, n), ref), (body, traits)


def test_scan_bodies_f64(paths, oracle, scan_body):
    """Both tile bodies on FP64 exclusive scans with totals.  Integer-valued
    doubles add exactly, so both equal the Serial scan bit for bit; for
    U[0,1) values every prefix is within 1e-12 of the compensated scan.  (The
    look-back combines the nearest predecessor that has published an
    inclusive prefix, which depends on timing, so rounded FP64 scans are not
    bitwise reproducible from run to run under either body.)"""
    P = paths
    n = (1 << 22) + 333
    xi = oracle.fill_uniform_i32(n, 12, 0, 100).astype(np.float64)
    y = oracle.fill_uniform_f64(n, 11, 0.0, 1.0)
    XI, Y, o, t = dev(xi), dev(y), K.DeviceBuffer(n * 8), K.DeviceBuffer(8)
    ref_i, rt_i = oracle.scan_excl_f64(xi)
    ref, rt = oracle.scan_excl_f64_compensated(y)
    for body in (0, 1):
        scan_body(body)
        run(P, P.kkmx_scan_excl_f64(XI.ptr, o.ptr, n, t.ptr))
        assert np.array_equal(o.download(np.float64, n), ref_i) and t.download(np.float64, 1)[0] == rt_i, body
        run(P, P.kkmx_scan_excl_f64(Y.ptr, o.ptr, n, t.ptr))
        got = o.download(np.float64, n)
        err = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0))
        terr = abs(t.download(np.float64, 1)[0] - rt) / rt
        print(f"scan_f64 body {body}: max rel err vs compensated {err:.3e}, total {terr:.3e}")
        assert err < 1e-12 and terr < 1e-12, body


def test_scan_in_place_and_overlap_rules(paths, oracle):
    """kkmx_scan_incl_i32 in place (in == out) reads through a plain View and
    is exact; asking for a RandomAccess / Restrict input that overlaps the
    output is rejected before anything runs."""
    P = paths
    for n in (5, 70001, (1 << 22) + 3):
        x = oracle.fill_uniform_i32(n, 50, 0, 9)
        B = dev(x)
        run(P, P.kkmx_scan_incl_i32(B.ptr, B.ptr, n))
        assert np.array_equal(B.download(np.int32, n), oracle.scan_incl_i32(x)), n
        for traits in (1, 2):
            assert P.kkmx_scan_incl_i32_traits(B.ptr, B.ptr, n, traits) == -1
    assert P.kkmx_scan_incl_i32_traits(B.ptr, B.ptr, 4, 3) == -1
    assert P.kkmx_set_scan_body(2) == -1


def test_scan_automatic_body_choice_exact_every_call(paths, oracle):
    """With the automatic choice the first large launches of a functor type
    alternate the two bodies before one is kept: every call is exact."""
    P = paths
    n = (1 << 21) + 1000
    for k in range(10):
        x = oracle.fill_uniform_i32(n, 60 + k, -5, 50)
        X, o = dev(x), K.DeviceBuffer(n * 4)
        run(P, P.kkmx_scan_incl_i32(X.ptr, o.ptr, n))
        assert np.array_equal(o.download(np.int32, n), oracle.scan_incl_i32(x)), k
        y = oracle.fill_uniform_i64(n, 70 + k, -(1 << 40), 1 << 40)
        Y, p, t = dev(y), K.DeviceBuffer(n * 8), K.DeviceBuffer(8)
        run(P, P.kkmx_scan_excl_i64(Y.ptr, p.ptr, n, t.ptr))
        ref, rt = oracle.scan_excl_i64(y)
        assert np.array_equal(p.download(np.int64, n), ref) and int(t.download(np.int64, 1)[0]) == int(rt), k


def test_full_size_config2(paths, oracle):
    """2^28 elements: int64 reduce max/min/sum (seed 3) and int32 inclusive
    scan (seed 4), generated on the device, compared bit for bit."""
    P = paths
    N = 1 << 28
    X, out = K.DeviceBuffer(N * 8), K.DeviceBuffer(32)
    run(P, P.kkmx_fill_uniform_i64(X.ptr, N, 3, -(1 << 40), 1 << 40))
    h = X.download(np.int64, N)
    assert oracle.checksum(h) == oracle.checksum(oracle.fill_uniform_i64(N, 3, -(1 << 40), 1 << 40))
    for op in (0, 1, 2):
        run(P, P.kkmx_reduce_i64(X.ptr, N, op, out.ptr))
        assert int(out.download(np.int64, 1)[0]) == oracle.reduce_i64(h, op, 8)
    del h, X
    S, O = K.DeviceBuffer(N * 4), K.DeviceBuffer(N * 4)
    run(P, P.kkmx_fill_uniform_i32(S.ptr, N, 4, 0, 7))
    s = S.download(np.int32, N)
    run(P, P.kkmx_scan_incl_i32(S.ptr, O.ptr, N))
    assert oracle.checksum(O.download(np.int32, N)) == oracle.checksum(oracle.scan_incl_i32(s, 8))


def test_full_size_config2_stream(paths, oracle):
    """2^28 doubles per array (seeds 2, 3, 4): one STREAM step in the bench's
    order (copy c=a, scale b=3c, add c=a+b, triad a=b+3c), each output
    compared with the oracle's by a checksum of the whole 2 GiB array."""
    P = paths
    N = 1 << 28
    A, B, C = K.DeviceBuffer(N * 8), K.DeviceBuffer(N * 8), K.DeviceBuffer(N * 8)
    run(P, P.kkmx_fill_uniform_f64(A.ptr, N, 2, -1.0, 1.0))
    run(P, P.kkmx_fill_uniform_f64(B.ptr, N, 3, -1.0, 1.0))
    run(P, P.kkmx_fill_uniform_f64(C.ptr, N, 4, -1.0, 1.0))
    a = oracle.fill_uniform_f64(N, 2, -1.0, 1.0)
    assert oracle.checksum(A.download(np.float64, N)) == oracle.checksum(a)
    run(P, P.kkmx_stream_copy(A.ptr, C.ptr, N))
    c = oracle.stream_copy(a, 8)
    assert oracle.checksum(C.download(np.float64, N)) == oracle.checksum(c), "copy"
    run(P, P.kkmx_stream_scale(B.ptr, C.ptr, N, 3.0))
    b = oracle.stream_scale(c, 3.0, 8)
    assert oracle.checksum(B.download(np.float64, N)) == oracle.checksum(b), "scale"
    run(P, P.kkmx_stream_add(A.ptr, B.ptr, C.ptr, N))
    c = oracle.stream_add(a, b, 8)
    assert oracle.checksum(C.download(np.float64, N)) == oracle.checksum(c), "add"
    run(P, P.kkmx_stream_triad(A.ptr, B.ptr, C.ptr, N, 3.0))
    a = oracle.stream_triad(b, c, 3.0, fused=True, nthreads=8)
    assert oracle.checksum(A.download(np.float64, N)) == oracle.checksum(a), "triad"


def test_full_size_config1(paths, oracle):
    """2^24: f64 sum (seed 0) within 1e-12 of the exact sum; int32 U[0,16)
    exclusive scan into int64 (seed 1) bit-exact."""
    P = paths
    N = 1 << 24
    X, out = K.DeviceBuffer(N * 8), K.DeviceBuffer(8)
    run(P, P.kkmx_fill_uniform_f64(X.ptr, N, 0, 0.0, 1.0))
    run(P, P.kkmx_reduce_sum_f64(X.ptr, N, out.ptr))
    got = float(out.download(np.float64, 1)[0])
    exact = oracle.sum_exact(oracle.fill_uniform_f64(N, 0, 0.0, 1.0))
    rel = abs(got - exact) / exact
    print(f"config1a sum: device={got!r} exact={exact!r} rel_err={rel:.3e}")
    assert rel <= 1e-12
    S, O, T = K.DeviceBuffer(N * 4), K.DeviceBuffer(N * 8), K.DeviceBuffer(8)
    run(P, P.kkmx_fill_uniform_i32(S.ptr, N, 1, 0, 15))
    run(P, P.kkmx_scan_excl_i32_i64(S.ptr, O.ptr, N, T.ptr))
    ref, rt = oracle.scan_excl_i32_i64(oracle.fill_uniform_i32(N, 1, 0, 15), 8)
    assert np.array_equal(O.download(np.int64, N), ref)
    assert int(T.download(np.int64, 1)[0]) == rt
