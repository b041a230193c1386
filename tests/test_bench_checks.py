"""CPU checks of the helpers bench.py uses to verify its timed outputs.

The triad check in bench.py rounds b + 3c correctly from numpy alone; here it
is compared with the oracle's fused triad (C fma) on random inputs and on
inputs built so that the x87 sum lands exactly half way between two doubles
(the double-rounding case the helper corrects).
"""
import numpy as np

import bench


def test_triad_exact_matches_c_fma(oracle):
    rng = np.random.default_rng(11)
    n = 1 << 18
    b = rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-60, 2, n)
    c = rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-60, 2, n)
    assert np.array_equal(bench.triad_exact(b, c, chunk=1 << 14), oracle.stream_triad(b, c, 3.0, fused=True))
    b = oracle.fill_uniform_f64(n, 3, -1.0, 1.0)
    c = oracle.fill_uniform_f64(n, 4, -1.0, 1.0)
    assert np.array_equal(bench.triad_exact(b, c), oracle.stream_triad(b, c, 3.0, fused=True))


def test_triad_exact_double_rounding_cases(oracle):
    # 3c = 3 + 1.5 * 2^-51 is the midpoint of two doubles; a tiny b pushes
    # the exact sum just above or below it, which the long double sum loses.
    c0 = 1.0 + 2.0**-52
    b = np.array([2.0**-70, -(2.0**-70), 2.0**-64, -(2.0**-64), 0.0, 2.0**-62, -(2.0**-62)])
    c = np.full(b.size, c0)
    want = oracle.stream_triad(b, c, 3.0, fused=True)
    assert want[1] == 3.0 + 2.0**-51 and want[0] == 3.0 + 2.0**-50
    assert np.array_equal(bench.triad_exact(b, c), want)
    # and the unfused expression differs on some of them: the check is strict
    assert not np.array_equal(b + 3.0 * c, want)
