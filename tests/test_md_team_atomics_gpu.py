"""Parity of MDRange (7-point stencil), TeamPolicy (CSR SpMV with team
scratch) and the atomics paths (GUPS, histogram, FP64 scatter) against the
CPU oracle on the same seeded inputs.

Tolerances, stated per result:
  * stencil: |device - oracle| <= 1e-12 * max(1, |oracle|) per point (the
    device fuses one multiply-add; every sweep is otherwise the same
    arithmetic), i.e. within the 1e-12 of the configuration;
  * SpMV: normwise per row, |y - y_ref| <= 1e-12 * sum_j |a_ij x_j|;
  * GUPS, histogram: bit-exact (integer atomics commute);
  * FP64 scatter: per-bin relative error <= 1e-10 against compensated sums.
"""
import numpy as np
import pytest

import kokkos_amd as K

pytestmark = pytest.mark.gpu

C0, C1 = 0.4, 0.1  # stencil coefficients of the workload
X_SEED = 606       # seed of the SpMV input vector


def run(P, rc):
    K.check(rc, P)
    K.fence()


def dev(a):
    return K.DeviceBuffer.from_array(a) if a.size else K.DeviceBuffer(8)


# ---------------------------------------------------------------- stencil
@pytest.mark.parametrize("n,sweeps,tile", [(8, 1, (0, 0, 0)), (31, 2, (0, 0, 0)), (64, 3, (0, 0, 0)),
                                           (65, 2, (32, 4, 2)), (100, 1, (16, 8, 8)), (40, 2, (4, 4, 4))])
def test_stencil7_matches_oracle(paths, oracle, n, sweeps, tile):
    P = paths
    m = n + 2
    u0 = oracle.fill_uniform_f64(m**3, 5, 0.0, 1.0)
    U, V = dev(u0), dev(u0)
    run(P, P.kkmx_stencil7_sweeps(U.ptr, V.ptr, n, sweeps, C0, C1, *tile))
    ref_u, ref_v = u0.copy(), u0.copy()
    for s in range(sweeps):
        if s % 2 == 0:
            ref_v = oracle.stencil7(ref_u, n, C0, C1)
        else:
            ref_u = oracle.stencil7(ref_v, n, C0, C1)
    got = (V if sweeps % 2 else U).download(np.float64, m**3)
    ref = ref_v if sweeps % 2 else ref_u
    err = np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref)))
    print(f"stencil n={n} sweeps={sweeps} tile={tile} max rel err {err:.2e}")
    assert err <= 1e-12


def test_stencil7_full_size_config3(paths, oracle):
    """512^3 interior, 10 sweeps, input U[0,1) seed 5."""
    P = paths
    n, sweeps = 512, 10
    m = n + 2
    U, V = K.DeviceBuffer(m**3 * 8), K.DeviceBuffer(m**3 * 8)
    run(P, P.kkmx_fill_uniform_f64(U.ptr, m**3, 5, 0.0, 1.0))
    run(P, P.kkmx_fill_uniform_f64(V.ptr, m**3, 5, 0.0, 1.0))
    run(P, P.kkmx_stencil7_sweeps(U.ptr, V.ptr, n, sweeps, C0, C1, 0, 0, 0))
    got = U.download(np.float64, m**3)
    a = oracle.fill_uniform_f64(m**3, 5, 0.0, 1.0)
    for _ in range(sweeps // 2):
        b = oracle.stencil7(a, n, C0, C1)
        a = oracle.stencil7(b, n, C0, C1)
    err = np.max(np.abs(got - a) / np.maximum(1.0, np.abs(a)))
    print(f"stencil 512^3 x10 max rel err {err:.2e}")
    assert err <= 1e-12


# ---------------------------------------------------------------- SpMV
def gen_csr(P, nrows, seed=6):
    import ctypes as C

    rp = K.DeviceBuffer((nrows + 1) * 8)
    nnz = C.c_int64()
    run(P, P.kkmx_csr_generate(rp.ptr, None, None, nrows, nrows, 4, 64, seed, C.byref(nnz)))
    col, val = K.DeviceBuffer(nnz.value * 4), K.DeviceBuffer(nnz.value * 8)
    run(P, P.kkmx_csr_fill(col.ptr, val.ptr, nnz.value, nrows, seed))
    return rp, col, val, nnz.value


@pytest.mark.parametrize("nrows", [1, 7, 1000, 70001])
def test_csr_generation_matches_oracle(paths, oracle, nrows):
    P = paths
    rp, col, val, nnz = gen_csr(P, nrows)
    orp, ocol, oval = oracle.csr(nrows, nrows, 4, 64, 6)
    assert nnz == orp[-1]
    assert np.array_equal(rp.download(np.int64, nrows + 1), orp)
    assert np.array_equal(col.download(np.int32, nnz), ocol)
    assert np.array_equal(val.download(np.float64, nnz), oval)


@pytest.mark.parametrize("nrows,variant,ts,vl", [(1, 1, 64, 8), (1000, 0, 0, 0), (1000, 1, 64, 8),
                                                 (70001, 1, 32, 16), (70001, 1, 64, 4), (1 << 20, 1, 64, 8)])
def test_spmv_normwise(paths, oracle, nrows, variant, ts, vl):
    P = paths
    rp, col, val, nnz = gen_csr(P, nrows)
    X, Y = K.DeviceBuffer(nrows * 8), K.DeviceBuffer(nrows * 8)
    run(P, P.kkmx_fill_uniform_f64(X.ptr, nrows, X_SEED, -1.0, 1.0))
    run(P, P.kkmx_spmv_csr(rp.ptr, col.ptr, val.ptr, X.ptr, Y.ptr, nrows, nnz, variant, ts, vl))
    orp, ocol, oval = oracle.csr(nrows, nrows, 4, 64, 6)
    x = oracle.fill_uniform_f64(nrows, X_SEED, -1.0, 1.0)
    yref, absrow = oracle.spmv(orp, ocol, oval, x)
    y = Y.download(np.float64, nrows)
    err = np.max(np.abs(y - yref) / absrow)
    print(f"spmv rows={nrows} variant={variant} ts={ts} vl={vl} normwise err {err:.2e}")
    assert err <= 1e-12


# ---------------------------------------------------------------- GUPS
def test_xorshift_stream_matches_sequential_generator(paths, oracle):
    P = paths
    for n, seed in [(1, 7), (4095, 7), (4097, 0), (100003, 123)]:
        B = K.DeviceBuffer(n * 8)
        run(P, P.kkmx_xorshift64_stream(B.ptr, n, seed))
        assert np.array_equal(B.download(np.uint64, n), oracle.xorshift64star(n, seed)), n


@pytest.mark.parametrize("log_table,log_updates", [(10, 12), (16, 18), (20, 22)])
def test_gups_bit_exact(paths, oracle, log_table, log_updates):
    P = paths
    T, U = 1 << log_table, 1 << log_updates
    S, Tb = K.DeviceBuffer(2 * U * 8), K.DeviceBuffer(T * 8)
    run(P, P.kkmx_xorshift64_stream(S.ptr, 2 * U, 7))
    K.load("kkmx_rt").mx_memset_async(Tb.ptr, 0, T * 8, P.kkmx_stream())
    run(P, P.kkmx_gups(Tb.ptr, T, S.ptr, U))
    ref = oracle.gups(T, oracle.xorshift64star(2 * U, 7), U)
    assert np.array_equal(Tb.download(np.int64, T), ref)


def test_gups_full_size_config4(paths, oracle):
    """2^30-entry int64 table, 2^28 updates from the xorshift64 stream, seed 7."""
    P = paths
    T, U = 1 << 30, 1 << 28
    S, Tb = K.DeviceBuffer(2 * U * 8), K.DeviceBuffer(T * 8)
    run(P, P.kkmx_xorshift64_stream(S.ptr, 2 * U, 7))
    stream = oracle.xorshift64star(2 * U, 7)
    assert oracle.checksum(S.download(np.uint64, 2 * U)) == oracle.checksum(stream)
    K.load("kkmx_rt").mx_memset_async(Tb.ptr, 0, T * 8, P.kkmx_stream())
    run(P, P.kkmx_gups(Tb.ptr, T, S.ptr, U))
    ref = oracle.gups(T, stream, U, 16)
    assert oracle.checksum(Tb.download(np.int64, T)) == oracle.checksum(ref)


# ---------------------------------------------------------------- histogram
@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_zipf_keys_match_oracle(paths, oracle, n):
    P = paths
    Kb = K.DeviceBuffer(n * 4)
    run(P, P.kkmx_zipf_keys(Kb.ptr, n, 1.1, 65536, 8))
    assert np.array_equal(Kb.download(np.int32, n), oracle.zipf_keys(n, 1.1, 65536, 8))


@pytest.mark.parametrize("mode,lds_bins", [(0, 0), (1, 32768), (1, 1024), (1, 65536 // 2 + 7), (2, 32768),
                                           (2, 100), (2, 65536 // 2 + 7), (3, 32768), (3, 100), (3, 65536)])
def test_histogram_bit_exact(paths, oracle, mode, lds_bins):
    P = paths
    n, nb = 1 << 22, 65536
    Kb, Bb = K.DeviceBuffer(n * 4), K.DeviceBuffer(nb * 4)
    run(P, P.kkmx_zipf_keys(Kb.ptr, n, 1.1, nb, 8))
    K.load("kkmx_rt").mx_memset_async(Bb.ptr, 0, nb * 4, P.kkmx_stream())
    run(P, P.kkmx_histogram(Kb.ptr, n, Bb.ptr, nb, mode, lds_bins))
    got = Bb.download(np.int32, nb)
    ref = oracle.histogram(oracle.zipf_keys(n, 1.1, nb, 8), nb)
    assert got.sum() == n
    assert np.array_equal(got, ref)


def test_histogram_full_size_config4(paths, oracle):
    """2^28 Zipf(1.1) keys over 65536 bins, seed 8, LDS-privatised."""
    P = paths
    n, nb = 1 << 28, 65536
    Kb, Bb = K.DeviceBuffer(n * 4), K.DeviceBuffer(nb * 4)
    run(P, P.kkmx_zipf_keys(Kb.ptr, n, 1.1, nb, 8))
    keys = oracle.zipf_keys(n, 1.1, nb, 8)
    assert oracle.checksum(Kb.download(np.int32, n)) == oracle.checksum(keys)
    ref = oracle.histogram(keys, nb)
    for mode in (1, 2, 3):
        K.load("kkmx_rt").mx_memset_async(Bb.ptr, 0, nb * 4, P.kkmx_stream())
        run(P, P.kkmx_histogram(Kb.ptr, n, Bb.ptr, nb, mode, 32768))
        assert np.array_equal(Bb.download(np.int32, nb), ref), f"mode {mode}"


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 1023, 4097, 1 << 16])
def test_histogram_wide_ragged(paths, oracle, n):
    """Modes 2 and 3 on extents that are not multiples of 4 keys or of a team's
    share, and on keys at an offset of 4 bytes (not 16-byte aligned, so the
    entry point runs mode 1)."""
    P = paths
    nb = 300
    keys = oracle.fill_uniform_i32(n + 1, 81, 0, nb - 1)
    Kb, Bb = dev(keys), K.DeviceBuffer(nb * 4)
    for mode in (2, 3):
        for off in (0, 1):
            K.load("kkmx_rt").mx_memset_async(Bb.ptr, 0, nb * 4, P.kkmx_stream())
            run(P, P.kkmx_histogram(Kb.ptr + 4 * off, n, Bb.ptr, nb, mode, 128))
            got = Bb.download(np.int32, nb)
            ref = oracle.histogram(np.ascontiguousarray(keys[off:off + n]), nb)
            assert np.array_equal(got, ref), (mode, off)


def test_histogram_rows_adversarial_keys(paths, oracle):
    """Mode 3 with every key in one bin above the LDS range (one team-row
    counter takes a whole team's share) and with all keys in bin 0."""
    P = paths
    n, nb = 1 << 22, 65536
    Bb = K.DeviceBuffer(nb * 4)
    for key in (nb - 1, 0):
        keys = np.full(n, key, np.int32)
        Kb = dev(keys)
        K.load("kkmx_rt").mx_memset_async(Bb.ptr, 0, nb * 4, P.kkmx_stream())
        run(P, P.kkmx_histogram(Kb.ptr, n, Bb.ptr, nb, 3, 32768))
        assert np.array_equal(Bb.download(np.int32, nb), oracle.histogram(keys, nb)), key


@pytest.mark.parametrize("pattern", ["one_packed_bin", "word_pair", "last_bin", "hot_bin"])
def test_histogram_packed_carries(paths, oracle, pattern):
    """Mode 3 with every bin in LDS on keys that drive one 16-bit counter far
    past its range (2^26 keys, about 16 carries of 16384 per team), two
    counters sharing one LDS word, the upper half of the last word, and a
    replicated hot bin."""
    P = paths
    n, nb = 1 << 26, 65536
    if pattern == "one_packed_bin":
        keys = np.full(n, 300, np.int32)
    elif pattern == "word_pair":
        keys = np.tile(np.array([300, 301], np.int32), n // 2)
    elif pattern == "last_bin":
        keys = np.full(n, nb - 1, np.int32)
    else:
        keys = np.full(n, 7, np.int32)
    Kb, Bb = dev(keys), K.DeviceBuffer(nb * 4)
    K.load("kkmx_rt").mx_memset_async(Bb.ptr, 0, nb * 4, P.kkmx_stream())
    run(P, P.kkmx_histogram(Kb.ptr, n, Bb.ptr, nb, 3, 0))
    assert np.array_equal(Bb.download(np.int32, nb), oracle.histogram(keys, nb)), pattern


@pytest.mark.parametrize("nb", [257, 1001, 69000, 70000, 1 << 17])
def test_histogram_mode3_bin_counts(paths, oracle, nb):
    """Mode 3 across bin counts: 257 and 1001 (odd packed tails), 69000 (near
    the LDS limit of the packed layout, 69725 bins), and 70000 and 2^17 (too
    many bins: team rows for the bins above the LDS range)."""
    P = paths
    n = (1 << 22) + 13
    keys = oracle.fill_uniform_i32(n, 90 + nb % 11, 0, nb - 1)
    Kb, Bb = dev(keys), K.DeviceBuffer(nb * 4)
    K.load("kkmx_rt").mx_memset_async(Bb.ptr, 0, nb * 4, P.kkmx_stream())
    run(P, P.kkmx_histogram(Kb.ptr, n, Bb.ptr, nb, 3, 32768))
    assert np.array_equal(Bb.download(np.int32, nb), oracle.histogram(keys, nb)), nb


# ---------------------------------------------------------------- FP64 scatter
@pytest.mark.parametrize("n,nbins", [(1000, 16), (1 << 20, 1 << 10), (1 << 26, 1 << 20)])
def test_scatter_f64_within_1e10(paths, oracle, n, nbins):
    P = paths
    vals = oracle.fill_uniform_f64(n, 9, 0.0, 1.0)
    idx = oracle.fill_uniform_i32(n, 9 ^ 0xABCDEF, 0, nbins - 1)
    V, I, B = dev(vals), dev(idx), K.DeviceBuffer(nbins * 8)
    K.load("kkmx_rt").mx_memset_async(B.ptr, 0, nbins * 8, P.kkmx_stream())
    run(P, P.kkmx_scatter_add_f64(V.ptr, I.ptr, n, B.ptr, nbins))
    got = B.download(np.float64, nbins)
    ref, absb = oracle.scatter_f64(vals, idx, nbins)
    nz = absb > 0
    err = np.max(np.abs(got[nz] - ref[nz]) / absb[nz]) if nz.any() else 0.0
    print(f"scatter n={n} bins={nbins} max rel err {err:.2e}")
    assert err <= 1e-10
    assert np.all(got[~nz] == 0.0)
