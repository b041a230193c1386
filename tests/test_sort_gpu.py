"""Kokkos::sort (LSD radix sort on gfx950) and Kokkos::BinSort / BinOp1D
through the C ABI (algorithms/src/sorting/Kokkos_SortPublicAPI.hpp,
Kokkos_BinSortPublicAPI.hpp), checked against numpy on the same seeded
inputs: sorted keys bit-exact; BinSort's permutation equal to numpy's stable
argsort when keys are sorted within bins, and bin-monotone otherwise."""
import numpy as np
import pytest

import kokkos_amd as K

pytestmark = pytest.mark.gpu
TYPES = {0: np.int32, 1: np.int64, 2: np.float32, 3: np.float64, 4: np.uint32, 5: np.uint64}


def run(P, rc):
    K.check(rc, P)
    K.fence()


def keys_for(oracle, code, n, seed):
    dt = TYPES[code]
    if code in (2, 3):
        x = oracle.fill_uniform_f64(n, seed, -1e6, 1e6)
        x[::7] = np.round(x[::7])  # ties
        x[::11] *= -1e-30          # tiny magnitudes of both signs
        x[3::13] = 0.0             # signed zeros, interleaved: they compare
        x[5::26] = -0.0            # equal and keep their input order
        x[::97] = np.inf
        x[1::97] = -np.inf
        return x.astype(dt)
    r = oracle.fill_uniform_i64(n, seed, -(1 << 62), 1 << 62)
    if code in (0, 4):
        r = r >> 32
    if code == 0:
        r[::5] = r[::5] % 50       # many duplicates
    return r.astype(dt)


@pytest.mark.parametrize("code", sorted(TYPES))
@pytest.mark.parametrize("n", [1, 2, 1000, 4097, 1 << 20])
def test_sort_matches_numpy(paths, oracle, code, n):
    P = paths
    x = keys_for(oracle, code, n, 70 + code)
    D = K.DeviceBuffer.from_array(x)
    run(P, P.kkmx_sort(D.ptr, n, code))
    got = D.download(TYPES[code], n)
    want = np.sort(x, kind="stable")
    bits = {4: np.uint32, 8: np.uint64}[x.dtype.itemsize]
    assert np.array_equal(got.view(bits), want.view(bits))  # bitwise: -0.0 != +0.0 here


def test_sort_large_int32(paths, oracle):
    P = paths
    n = 1 << 26
    D = K.DeviceBuffer(n * 4)
    run(P, P.kkmx_fill_uniform_i32(D.ptr, n, 77, -(1 << 30), 1 << 30))
    x = D.download(np.int32, n)
    run(P, P.kkmx_sort(D.ptr, n, 0))
    assert np.array_equal(D.download(np.int32, n), np.sort(x))


@pytest.mark.parametrize("n,nbins", [(1, 4), (1000, 16), (1 << 20, 1 << 12)])
def test_binsort_within_bins_is_stable_argsort(paths, oracle, n, nbins):
    P = paths
    # keys on a 2^-16 grid: equal keys (16 per value at 2^20) must keep index
    # order; bins stay small, since the within-bin sort is an insertion sort
    # per bin (quadratic in the bin size, as in the reference)
    keys = np.floor(oracle.fill_uniform_f64(n, 81, 0.0, 1.0) * 65536.0) / 65536.0
    vals = np.arange(n, dtype=np.int64) * 3
    Kd, Vd, Pd = K.DeviceBuffer.from_array(keys), K.DeviceBuffer.from_array(vals), K.DeviceBuffer(n * 4)
    run(P, P.kkmx_binsort_f64(Kd.ptr, Vd.ptr, n, nbins, 0.0, 1.0, 1, Pd.ptr))
    perm = Pd.download(np.int32, n)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, ref)
    assert np.array_equal(Vd.download(np.int64, n), vals[ref])


def test_binsort_without_bin_sort_is_bin_monotone(paths, oracle):
    P = paths
    n, nbins = 1 << 18, 1000
    keys = oracle.fill_uniform_f64(n, 82, -5.0, 5.0)
    vals = np.arange(n, dtype=np.int64)
    Kd, Vd, Pd = K.DeviceBuffer.from_array(keys), K.DeviceBuffer.from_array(vals), K.DeviceBuffer(n * 4)
    run(P, P.kkmx_binsort_f64(Kd.ptr, Vd.ptr, n, nbins, -5.0, 5.0, 0, Pd.ptr))
    perm = Pd.download(np.int32, n)
    assert np.array_equal(np.sort(perm), np.arange(n))
    bins = (nbins / 10.0 * (keys[perm] + 5.0)).astype(np.int64)
    assert np.all(np.diff(bins) >= 0)
    assert np.array_equal(Vd.download(np.int64, n), vals[perm])


@pytest.mark.parametrize("code", sorted(TYPES))
@pytest.mark.parametrize("n", [1, 2, 4097, 1 << 20])
@pytest.mark.parametrize("vbytes", [4, 8])
def test_sort_by_key_is_stable(paths, oracle, code, n, vbytes):
    """Kokkos::Experimental::sort_by_key (Kokkos_SortByKeyPublicAPI.hpp:27-64):
    keys as numpy's stable sort (bitwise), values moved with their keys, and
    equal keys keeping their input order (values = input positions here)."""
    P = paths
    x = keys_for(oracle, code, n, 90 + code)
    vdt = np.uint32 if vbytes == 4 else np.uint64
    vals = (np.arange(n, dtype=np.uint64) * 2654435761 % (1 << 32)).astype(vdt)
    Dk, Dv = K.DeviceBuffer.from_array(x), K.DeviceBuffer.from_array(vals)
    run(P, P.kkmx_sort_by_key(Dk.ptr, Dv.ptr, n, code, vbytes, 0))
    order = np.argsort(x, kind="stable")
    bits = {4: np.uint32, 8: np.uint64}[x.dtype.itemsize]
    assert np.array_equal(Dk.download(TYPES[code], n).view(bits), x[order].view(bits))
    assert np.array_equal(Dv.download(vdt, n), vals[order])


@pytest.mark.parametrize("code", [0, 3])
def test_sort_by_key_with_comparator(paths, oracle, code):
    """The comparator overload (Kokkos_SortByKeyPublicAPI.hpp:67-108) with a
    greater-than comparator: descending keys, ties in input order."""
    P = paths
    n = 5000
    x = keys_for(oracle, code, n, 95)
    vals = np.arange(n, dtype=np.uint64)
    Dk, Dv = K.DeviceBuffer.from_array(x), K.DeviceBuffer.from_array(vals)
    run(P, P.kkmx_sort_by_key(Dk.ptr, Dv.ptr, n, code, 8, 1))
    # stable descending: sort by negated key; ties keep input order
    order = sorted(range(n), key=lambda i: -float(x[i]))
    assert np.array_equal(Dk.download(TYPES[code], n), x[order])
    assert np.array_equal(Dv.download(np.uint64, n), vals[order])


@pytest.mark.parametrize("code,lo,hi", [(0, 0, 1000), (0, -5, 5), (1, 0, 1 << 20), (4, 7, 8), (5, 0, 1 << 40),
                                        (1, -(1 << 40), -(1 << 39))])
@pytest.mark.parametrize("n", [4097, 1 << 20])
def test_small_range_keys_skip_constant_digits(paths, oracle, code, lo, hi, n):
    """Keys whose high (or all) digits are the same everywhere: those passes
    are skipped, with an odd or even number of passes left (the odd case ends
    with a copy back), or none at all; results equal numpy's stable sort and,
    for sort_by_key, its stable argsort."""
    P = paths
    dt = TYPES[code]
    x = oracle.fill_uniform_i64(n, 300 + code, lo, hi - 1).astype(dt)
    D = K.DeviceBuffer.from_array(x)
    run(P, P.kkmx_sort(D.ptr, n, code))
    assert np.array_equal(D.download(dt, n), np.sort(x, kind="stable"))
    vals = np.arange(n, dtype=np.uint32)
    Dk, Dv = K.DeviceBuffer.from_array(x), K.DeviceBuffer.from_array(vals)
    run(P, P.kkmx_sort_by_key(Dk.ptr, Dv.ptr, n, code, 4, 0))
    order = np.argsort(x, kind="stable")
    assert np.array_equal(Dk.download(dt, n), x[order])
    assert np.array_equal(Dv.download(np.uint32, n), vals[order])


def test_float_keys_with_constant_low_digits_and_signed_zeros(paths, oracle):
    """Doubles that are small integers (their low mantissa bytes are zero in
    every key) and a mix of -0.0 and +0.0 only (equal keys: nothing moves,
    the input order stays)."""
    P = paths
    n = 100000
    x = (oracle.fill_uniform_i64(n, 310, 0, 7).astype(np.float64) - 3.0)
    x[::9] = -0.0
    D = K.DeviceBuffer.from_array(x)
    run(P, P.kkmx_sort(D.ptr, n, 3))
    want = np.sort(x, kind="stable")
    assert np.array_equal(D.download(np.float64, n).view(np.uint64), want.view(np.uint64))
    z = np.where(np.arange(n) % 3 == 0, -0.0, 0.0)
    vals = np.arange(n, dtype=np.uint64)
    Dk, Dv = K.DeviceBuffer.from_array(z), K.DeviceBuffer.from_array(vals)
    run(P, P.kkmx_sort_by_key(Dk.ptr, Dv.ptr, n, 3, 8, 0))
    assert np.array_equal(Dk.download(np.float64, n).view(np.uint64), z.view(np.uint64))
    assert np.array_equal(Dv.download(np.uint64, n), vals)
