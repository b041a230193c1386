"""Pins the CPU oracle (CPU only, no GPU).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, an off-by-one in a scan, a wrong rounding, a transposed operand) fails:
published known answers, the closed forms of the reference's own unit tests,
an independent exact summation (Python's math.fsum), and brute force on tiny
inputs.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLD = json.loads((Path(__file__).parent / "golden" / "known_answers.json").read_text())


def test_splitmix64_known_answers(oracle):
    got = [oracle.lib().or_splitmix64(1234567, i) for i in range(5)]
    assert got == GOLD["splitmix64_seed_1234567"]


def test_fill_uniform_f64_range_and_moments(oracle):
    x = oracle.fill_uniform_f64(1 << 20, 0, 0.0, 1.0)
    assert x.min() >= 0.0 and x.max() < 1.0
    assert abs(x.mean() - 0.5) < 3e-3
    assert abs(x.var() - 1.0 / 12.0) < 3e-3
    y = oracle.fill_uniform_f64(1 << 16, 2, -1.0, 1.0)
    assert y.min() >= -1.0 and y.max() < 1.0 and abs(y.mean()) < 1e-2
    # deterministic in the seed, different across seeds
    assert np.array_equal(y, oracle.fill_uniform_f64(1 << 16, 2, -1.0, 1.0))
    assert not np.array_equal(y, oracle.fill_uniform_f64(1 << 16, 3, -1.0, 1.0))
    # element i is the top 53 bits of splitmix64(seed, i) scaled to [0,1)
    for i in (0, 1, 12345):
        z = oracle.lib().or_splitmix64(0, i)
        assert x[i] == (z >> 11) * 2.0**-53


def test_fill_uniform_int_inclusive_bounds(oracle):
    x = oracle.fill_uniform_i32(1 << 16, 1, 0, 15)
    assert x.min() == 0 and x.max() == 15
    counts = np.bincount(x, minlength=16)
    assert counts.min() > 0.9 * (1 << 16) / 16
    y = oracle.fill_uniform_i64(1 << 16, 3, -(1 << 40), 1 << 40)
    assert y.min() >= -(1 << 40) and y.max() <= (1 << 40)
    assert abs(float(y.mean())) < 0.02 * (1 << 40)
    z = oracle.fill_uniform_i32(1000, 4, 0, 7)
    assert set(np.unique(z)) == set(range(8))


def test_xorshift64star_reference_semantics(oracle):
    # Random_XorShift64 replaces a zero state by 1318319
    assert np.array_equal(oracle.xorshift64star(8, 0), oracle.xorshift64star(8, 1318319))
    # drand() = urand64 / (2^64-1) is uniform: the statistical checks of the
    # reference's algorithms/unit_tests/TestRandom.hpp (mean 1/2, variance 1/12)
    u = oracle.xorshift64star(1 << 20, 7).astype(np.float64) / 2.0**64
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 3e-3
    # indices mod 2^30 cover the table uniformly
    idx = (oracle.xorshift64star(1 << 20, 7) & np.uint64((1 << 30) - 1)) >> np.uint64(26)
    c = np.bincount(idx.astype(np.int64), minlength=16)
    assert c.min() > 0.9 * (1 << 20) / 16


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 1000, 1 << 16])
def test_exact_sum_matches_fsum(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    assert oracle.sum_exact(x) == math.fsum(x)


def test_exact_sum_special_cases(oracle):
    # half-way rounding across partials (the case fsum's correction handles)
    assert oracle.sum_exact(np.array([1e-16, 1.0, 1e16])) == math.fsum([1e-16, 1.0, 1e16])
    # massive cancellation
    assert oracle.sum_exact(np.array([1e100, 1.0, -1e100, 1e-100])) == 1.0 + 1e-100
    # integer-valued doubles below 2^53 are exact for every summation order
    ints = np.arange(1, 100001, dtype=np.float64)
    assert oracle.sum_exact(ints) == 100000 * 100001 / 2
    assert oracle.sum_serial(ints) == 100000 * 100001 / 2
    # dyadic values k * 2^-20 sum exactly
    d = np.arange(4096, dtype=np.float64) * 2.0**-20
    assert oracle.sum_exact(d) == (4095 * 4096 / 2) * 2.0**-20


def test_neumaier_and_serial_error_bounds(oracle):
    x = oracle.fill_uniform_f64(1 << 20, 0, 0.0, 1.0)
    ex = oracle.sum_exact(x)
    assert abs(oracle.sum_neumaier(x) - ex) <= 2 * np.finfo(float).eps * abs(ex)
    n = x.size
    # recursive summation: |err| <= (n-1) eps sum|x|
    assert abs(oracle.sum_serial(x) - ex) <= (n - 1) * np.finfo(float).eps * np.sum(np.abs(x))
    for t in (1, 3, 8):
        assert abs(oracle.sum_openmp(x, t) - ex) <= n * np.finfo(float).eps * ex
    assert oracle.sum_openmp(x, 1) == oracle.sum_serial(x)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_scan_reference_closed_forms(oracle, threads):
    # TestParallelScanRangePolicy: scan of 0..N-1, prefix i(i-1)/2, postfix i(i+1)/2
    for n in [1, 2, 7, 64, 1000, 4097]:
        x = np.arange(n, dtype=np.int32)
        ex, tot = oracle.scan_excl_i32_i64(x, threads)
        i = np.arange(n, dtype=np.int64)
        assert np.array_equal(ex, i * (i - 1) // 2)
        assert tot == GOLD["scan_closed_forms"][str(n)]["total"]
        inc = oracle.scan_incl_i32(x, threads)
        assert np.array_equal(inc.astype(np.int64), i * (i + 1) // 2)


def test_scan_brute_force_and_thread_invariance(oracle):
    rng = np.random.default_rng(5)
    for n in [0, 1, 5, 33]:
        x = rng.integers(-50, 50, n).astype(np.int64)
        ex, tot = oracle.scan_excl_i64(x, 1)
        acc, ref = 0, []
        for v in x:
            ref.append(acc)
            acc += int(v)
        assert ex.tolist() == ref and tot == acc
    x = oracle.fill_uniform_i32(100003, 1, 0, 15)
    a, ta = oracle.scan_excl_i32_i64(x, 1)
    for t in (2, 7, 8):
        b, tb = oracle.scan_excl_i32_i64(x, t)
        assert np.array_equal(a, b) and ta == tb


def test_reduce_closed_forms_and_brute_force(oracle):
    for n in [1, 10, 1000]:
        i = np.arange(n, dtype=np.int64)
        assert oracle.reduce_i64(i + 1, 0) == GOLD["reduce_closed_forms"][str(n)]["sum_i_plus_1"]
        assert oracle.reduce_i64(n - i, 0) == GOLD["reduce_closed_forms"][str(n)]["sum_n_minus_i"]
    rng = np.random.default_rng(3)
    x = rng.integers(-(1 << 40), 1 << 40, 999).astype(np.int64)
    for t in (1, 4):
        assert oracle.reduce_i64(x, 0, t) == int(sum(int(v) for v in x))
        assert oracle.reduce_i64(x, 1, t) == min(int(v) for v in x)
        assert oracle.reduce_i64(x, 2, t) == max(int(v) for v in x)
    # the sum wraps mod 2^64 like the device's unsigned accumulation
    big = np.full(4, np.iinfo(np.int64).max, dtype=np.int64)
    assert oracle.reduce_i64(big, 0) == ((4 * (2**63 - 1) + 2**63) % 2**64) - 2**63
    assert oracle.reduce_i64(np.zeros(0, np.int64), 1) == np.iinfo(np.int64).max
    assert oracle.reduce_i64(np.zeros(0, np.int64), 2) == np.iinfo(np.int64).min


def test_stream_triad_fused_vs_unfused(oracle):
    b = oracle.fill_uniform_f64(1 << 16, 3, -1.0, 1.0)
    c = oracle.fill_uniform_f64(1 << 16, 4, -1.0, 1.0)
    fused = oracle.stream_triad(b, c, 3.0, True)
    plain = oracle.stream_triad(b, c, 3.0, False)
    assert np.array_equal(plain, b + 3.0 * c)
    # one rounding instead of two: within one ulp of the unfused value
    assert np.all(np.abs(fused - plain) <= np.spacing(np.abs(plain)) + 1e-300)
    # the fused form is the single correct rounding of the exact b + 3c
    # (Fraction arithmetic is exact; float() rounds to nearest even)
    from fractions import Fraction

    for i in range(0, 1 << 16, 257):
        assert fused[i] == float(Fraction(b[i]) + 3 * Fraction(c[i]))


def test_checksum_sensitivity(oracle):
    x = oracle.fill_uniform_i64(4096, 9, -100, 100)
    h = oracle.checksum(x)
    y = x.copy()
    y[[10, 20]] = y[[20, 10]]
    if y[10] != y[20]:
        assert oracle.checksum(y) != h
    y = x.copy()
    y[100] += 1
    assert oracle.checksum(y) != h
    assert oracle.checksum(x.copy()) == h


def test_compensated_f64_scan_against_fsum(oracle):
    """Every prefix of the compensated scan is within 2 ulp of math.fsum of
    the same prefix (fsum is correctly rounded)."""
    import math

    x = oracle.fill_uniform_f64(3000, 77, -1.0, 1.0) * np.exp2(np.arange(3000) % 40 - 20)
    pre, tot = oracle.scan_excl_f64_compensated(x)
    for i in range(0, 3000, 7):
        ref = math.fsum(x[:i])
        assert abs(pre[i] - ref) <= 2 * np.spacing(abs(ref)) + 1e-300, i
    assert abs(tot - math.fsum(x)) <= 2 * np.spacing(abs(math.fsum(x)))
    assert pre[0] == 0.0
    # a planted error (dropping the compensation term) fails the same check
    naive = np.concatenate([[0.0], np.cumsum(x)[:-1]])
    worst = max(abs(naive[i] - math.fsum(x[:i])) / (np.spacing(abs(math.fsum(x[:i]))) + 1e-300) for i in range(1, 3000, 7))
    assert worst > 2


# ---------------------------------------------------------------- MDRange / Team / atomics restatements
def test_stencil7_constant_field_and_brute_force(oracle):
    """A constant field maps to (c0 + 6 c1) times itself in the interior and
    the halo is untouched; a direct triple loop agrees on a random grid."""
    n, c0, c1 = 6, 0.4, 0.1
    m = n + 2
    u = np.full(m**3, 2.0)
    v = oracle.stencil7(u, n, c0, c1, 1)
    V = v.reshape(m, m, m)
    assert np.allclose(V[1:-1, 1:-1, 1:-1], 2.0 * (c0 + 6 * c1), rtol=0, atol=1e-15)
    assert np.all(V[0] == 2.0) and np.all(V[:, :, -1] == 2.0)
    u = oracle.fill_uniform_f64(m**3, 5, 0.0, 1.0)
    U = u.reshape(m, m, m)  # index [k][j][i], i fastest
    ref = U.copy()
    for k in range(1, n + 1):
        for j in range(1, n + 1):
            for i in range(1, n + 1):
                nb = U[k, j, i - 1] + U[k, j, i + 1] + U[k, j - 1, i] + U[k, j + 1, i] + U[k - 1, j, i] + U[k + 1, j, i]
                ref[k, j, i] = c0 * U[k, j, i] + c1 * nb
    got = oracle.stencil7(u, n, c0, c1, 4).reshape(m, m, m)
    assert np.max(np.abs(got - ref)) <= 1e-15
    # a transposed neighbour (wrong stride) is caught by the same check
    bad = ref.copy()
    bad[1:-1, 1:-1, 1:-1] += c1 * (U[1:-1, 1:-1, 2:] - U[2:, 1:-1, 1:-1])
    assert np.max(np.abs(got - bad)) > 1e-3


def test_csr_shape_and_spmv_against_dense(oracle):
    nrows, ncols = 37, 53
    rp, col, val = oracle.csr(nrows, ncols, 4, 9, 6)
    nnz_row = np.diff(rp)
    assert rp[0] == 0 and np.all(nnz_row >= 4) and np.all(nnz_row <= 9)
    assert np.all((col >= 0) & (col < ncols)) and np.all((val >= -1) & (val < 1))
    A = np.zeros((nrows, ncols))
    for r in range(nrows):
        for k in range(rp[r], rp[r + 1]):
            A[r, col[k]] += val[k]
    x = oracle.fill_uniform_f64(ncols, 606, -1.0, 1.0)
    y, absrow = oracle.spmv(rp, col, val, x, 3)
    assert np.allclose(y, A @ x, rtol=0, atol=1e-14)
    absref = np.array([sum(abs(val[k] * x[col[k]]) for k in range(rp[r], rp[r + 1])) for r in range(nrows)])
    assert np.allclose(absrow, absref, rtol=1e-14, atol=0)


def test_gups_brute_force(oracle):
    T, U = 64, 1000
    s = oracle.xorshift64star(2 * U, 7)
    acc = [0] * T
    for u in range(U):
        t = int(s[2 * u]) & (T - 1)
        acc[t] = (acc[t] + int(s[2 * u + 1])) % (1 << 64)  # two's-complement wrap
    want = np.array(acc, np.uint64).view(np.int64)
    assert np.array_equal(oracle.gups(T, s, U, 4), want)


def test_histogram_and_zipf_distribution(oracle):
    nb = 64
    k = oracle.zipf_keys(200000, 1.1, nb, 8)
    assert k.min() >= 0 and k.max() < nb
    h = oracle.histogram(k, nb)
    assert np.array_equal(h, np.bincount(k, minlength=nb))
    pmf = np.arange(1, nb + 1, dtype=np.float64) ** -1.1
    pmf /= pmf.sum()
    # every bin within 5 binomial standard deviations of the Zipf(1.1) pmf
    sd = np.sqrt(k.size * pmf * (1 - pmf))
    assert np.all(np.abs(h - k.size * pmf) <= 5 * sd + 1)


def test_scatter_against_sorted_sums(oracle):
    n, nb = 20000, 97
    v = oracle.fill_uniform_f64(n, 9, 0.0, 1.0)
    i = oracle.fill_uniform_i32(n, 9 ^ 0xABCDEF, 0, nb - 1)
    b, a = oracle.scatter_f64(v, i, nb)
    ref = np.array([math.fsum(v[i == k]) for k in range(nb)])
    assert np.max(np.abs(b - ref) / np.maximum(ref, 1e-300)) <= 2e-16
    assert np.allclose(a, ref)  # inputs are non-negative: sum of |v| = sum of v
