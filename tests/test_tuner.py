"""CPU checks of the tuner's host logic: targets, model file format and the
feature function of the tool library (host code, no GPU needed)."""
import numpy as np

from kokkos_amd import tuner as T


def test_soft_targets_prefer_fastest_and_sum_to_one():
    t = np.array([[1.0] * 16 + [1.0], [2.0] * 8 + [1.0] + [3.0] * 7 + [2.0]], np.float32)
    y = T.soft_targets(t)
    assert np.allclose(y.sum(axis=1), 1.0)
    assert np.allclose(y[0], 1 / 16)
    assert y[1].argmax() == 8 and y[1, 8] > 0.99


def test_regret_against_exhaustive_best():
    t = np.array([[2.0] * 15 + [1.0, 3.0]], np.float32)
    assert T.regret(t, np.array([15]))[0] == 0.0
    assert T.regret(t, np.array([16]))[0] == 2.0


def test_model_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    p = rng.standard_normal(12 * 64 + 64 + 64 * 64 + 64 + 16 * 64 + 16).astype(np.float32)
    m, s = rng.standard_normal(12).astype(np.float32), rng.random(12).astype(np.float32)
    T.save_model(p, m, s, {"x": 1}, tmp_path / "a.bin", tmp_path / "a.json")
    p2, m2, s2 = T.read_model(tmp_path / "a.bin")
    assert np.array_equal(p, p2) and np.array_equal(m, m2) and np.array_equal(s, s2)


def test_tool_features_match_their_definition():
    fam = dict(pattern=1, value_bytes=8, floating=1, functor_bytes=24, bytes=8.0, flops=1.0, kind=2)
    f = T.raw_features(fam, 2.0 ** 20, 256)
    want = [20 / 32, 0, 1, 0, np.log2(9) / 4, np.log2(25) / 10, 1, np.log2(9) / 6, 1 / 8,
            np.log2(2.0 ** 20 / (256 * 2048) + 2.0 ** -20) / 20, np.log2(2.0 ** 23 + 1) / 40, 1.0]
    assert np.allclose(f, want, rtol=1e-6, atol=1e-7)
