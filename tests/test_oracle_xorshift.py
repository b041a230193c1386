"""Pins the oracle's xorshift64* generator (CPU only, no GPU).

The oracle's `or_xorshift64star` restates Random_XorShift64::urand64
(algorithms/src/Kokkos_Random.hpp:788-808): the state s (0 replaced by
1318319) advances by three xor-shifts, s ^= s >> 12; s ^= s << 25;
s ^= s >> 27, and each output is s * M - 1 mod 2^64 with M =
2685821657736338717 (Vigna, "An experimental exploration of Marsaglia's
xorshift generators, scrambled", ACM TOMS 42(4), 2016: xorshift64*, M_32).

Nothing here retypes that loop and compares.  The checks observe the
generator only through its outputs and test what the mathematics fixes:

* State recovery.  M is odd, so s_k = (out_k + 1) * M^-1 mod 2^64 recovers
  every state.  The state map must then be GF(2)-linear: T(a ^ b) = T(a) ^
  T(b).  A dropped "- 1" (or a wrong multiplier) turns the recovered map into
  an affine or scrambled one and fails.
* The matrix of T, read column by column from the outputs for the seeds 2^j,
  must equal (I + R^27)(I + L^25)(I + R^12), built from shift matrices here.
  A wrong shift amount or direction fails.
* Full period.  T must have order exactly 2^64 - 1 in GL(64, GF(2)):
  T^(2^64-1) = I and T^((2^64-1)/p) != I for each prime p of
  2^64 - 1 = 3 * 5 * 17 * 257 * 641 * 65537 * 6700417.  This is the property
  Marsaglia's shift triples are chosen for; it holds for (12, 25, 27).
* Successive outputs follow s_{k+1} = T(s_k) for 1000 steps of the seeds
  0, 7 and 2^63 + 1.

Each check is also run against generators with a planted mistake (restated
in this file, on the CPU) and must reject them.
"""
from __future__ import annotations

import numpy as np
import pytest

MASK = (1 << 64) - 1
M_XS = 2685821657736338717
M_INV = pow(M_XS, -1, 1 << 64)
PERIOD = (1 << 64) - 1
PERIOD_PRIMES = (3, 5, 17, 257, 641, 65537, 6700417)


# ------------------------------------------------------------ GF(2) matrices
# A 64x64 matrix over GF(2) is a uint8 array A with A[i, j] = bit i of the
# image of e_j.
def _vec(x: int) -> np.ndarray:
    return np.array([(x >> i) & 1 for i in range(64)], dtype=np.uint8)


def _int(v: np.ndarray) -> int:
    return sum(int(b) << i for i, b in enumerate(v))


def _mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return ((a.astype(np.int64) @ b.astype(np.int64)) & 1).astype(np.uint8)


def _pow(a: np.ndarray, e: int) -> np.ndarray:
    r = np.eye(64, dtype=np.uint8)
    sq = a
    while e:
        if e & 1:
            r = _mul(r, sq)
        sq = _mul(sq, sq)
        e >>= 1
    return r


def _shift_matrix(k: int, left: bool) -> np.ndarray:
    """Matrix of x -> x << k (left) or x >> k (right) on 64-bit words."""
    s = np.zeros((64, 64), dtype=np.uint8)
    for j in range(64):
        i = j + k if left else j - k
        if 0 <= i < 64:
            s[i, j] = 1
    return s


def _xorshift_matrix(a: int, b: int, c: int) -> np.ndarray:
    eye = np.eye(64, dtype=np.uint8)
    ta = eye ^ _shift_matrix(a, left=False)
    tb = eye ^ _shift_matrix(b, left=True)
    tc = eye ^ _shift_matrix(c, left=False)
    return _mul(tc, _mul(tb, ta))


# ------------------------------------------------------------ observation
def _recover(out: int) -> int:
    return ((out + 1) * M_INV) & MASK


def _observed_matrix(gen) -> np.ndarray:
    """Column j = recovered first state of the stream seeded with 2^j."""
    cols = [_vec(_recover(int(gen(1, 1 << j)[0]))) for j in range(64)]
    return np.stack(cols, axis=1)


def _apply(t: np.ndarray, s: int) -> int:
    return _int(_mul(t, _vec(s)[:, None])[:, 0])


def _check_generator(gen) -> None:
    """Raises AssertionError unless `gen(n, seed)` is xorshift64* with the
    (12, 25, 27) triple, multiplier M and the "- 1" of urand64."""
    rng = np.random.default_rng(2024)
    # GF(2)-linearity of the recovered state map.
    for _ in range(8):
        a = int(rng.integers(1, 1 << 63, dtype=np.int64)) | 1
        b = int(rng.integers(1, 1 << 63, dtype=np.int64)) << 1
        ra, rb, rab = (_recover(int(gen(1, s)[0])) for s in (a, b, a ^ b))
        assert rab == ra ^ rb, "recovered state map is not GF(2)-linear"
    t = _observed_matrix(gen)
    assert np.array_equal(t, _xorshift_matrix(12, 25, 27)), "state map != (12, 25, 27) xorshift"
    assert np.array_equal(_pow(t, PERIOD), np.eye(64, dtype=np.uint8)), "T^(2^64-1) != I"
    for p in PERIOD_PRIMES:
        assert not np.array_equal(_pow(t, PERIOD // p), np.eye(64, dtype=np.uint8)), p
    # Whole streams: s_0 = T(seed'), s_{k+1} = T(s_k).
    for seed in (0, 7, (1 << 63) + 1):
        outs = [int(x) for x in gen(1000, seed)]
        states = [_recover(o) for o in outs]
        s = 1318319 if seed == 0 else seed
        for k, sk in enumerate(states):
            s = _apply(t, s)
            assert sk == s, (seed, k)


# ------------------------------------------------------------ tests
def test_oracle_xorshift64star_pinned(oracle):
    _check_generator(lambda n, seed: oracle.xorshift64star(n, seed))


def test_oracle_zero_seed_substitution(oracle):
    # Random_XorShift64's constructor maps a zero state to 1318319.
    assert np.array_equal(oracle.xorshift64star(16, 0), oracle.xorshift64star(16, 1318319))


def test_shift_triple_has_full_period_and_neighbours_do_not_all():
    # The (12, 25, 27) triple is full period; the check is not vacuous: some
    # neighbouring triples fail it.
    t = _xorshift_matrix(12, 25, 27)
    assert np.array_equal(_pow(t, PERIOD), np.eye(64, dtype=np.uint8))
    failing = 0
    for abc in ((12, 25, 26), (13, 25, 27), (12, 24, 27), (11, 25, 27)):
        u = _xorshift_matrix(*abc)
        full = np.array_equal(_pow(u, PERIOD), np.eye(64, dtype=np.uint8)) and all(
            not np.array_equal(_pow(u, PERIOD // p), np.eye(64, dtype=np.uint8)) for p in PERIOD_PRIMES
        )
        failing += not full
    assert failing >= 1


def _planted(shifts=(12, 25, 27), left_middle=True, minus_one=True, mult=M_XS, zero_seed=1318319):
    a, b, c = shifts

    def gen(n, seed):
        s = zero_seed if seed == 0 else seed
        out = []
        for _ in range(n):
            s ^= s >> a
            s = (s ^ (s << b)) & MASK if left_middle else s ^ (s >> b)
            s ^= s >> c
            out.append(((s * mult) - (1 if minus_one else 0)) & MASK)
        return out

    return gen


def test_planted_reference_model_passes():
    # The same checks accept a correct stand-alone model, so the mutants
    # below fail for their mistake and not for the harness.
    _check_generator(_planted())


@pytest.mark.parametrize(
    "mutant",
    [
        dict(minus_one=False),
        dict(shifts=(12, 25, 26)),
        dict(shifts=(13, 25, 27)),
        dict(shifts=(12, 24, 27)),
        dict(left_middle=False),
        dict(mult=M_XS + 2),
        dict(zero_seed=1318318),
    ],
    ids=["dropped_minus_one", "shift_c", "shift_a", "shift_b", "middle_shift_direction", "multiplier", "zero_seed"],
)
def test_planted_mistakes_fail(mutant):
    with pytest.raises(AssertionError):
        _check_generator(_planted(**mutant))
