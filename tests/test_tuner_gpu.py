"""The learned tuner on the GPU: the HIP training kernels learn a task with a
known answer, the sweep produces well-formed timing groups, and the tuning
tool answers the runtime's requests without changing any result."""
import ctypes as C

import numpy as np
import pytest

import kokkos_amd as K
from kokkos_amd import tuner as T

pytestmark = pytest.mark.gpu


def test_mlp_training_learns_a_separable_task(paths):
    """Labels are the arg-max of a fixed linear map of the features: the MLP
    must fit them (train accuracy >= 90%) and the GPU forward kernel must
    agree with a numpy forward pass of the trained weights."""
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2048, T.NF)).astype(np.float32)
    W = rng.standard_normal((T.NF, T.NC)).astype(np.float32)
    lab = (X @ W).argmax(axis=1)
    Y = np.eye(T.NC, dtype=np.float32)[lab]
    params, loss = T.train(X, Y, epochs=150, batch=128, lr=3e-3, weight_decay=0.0)
    print(f"loss {loss[0]:.3f} -> {loss[-1]:.3f}")
    assert loss[-1] < 0.5 * loss[0]
    Z = T.predict(params, X)
    acc = float((Z.argmax(axis=1) == lab).mean())
    print(f"train accuracy {acc:.3f}")
    assert acc >= 0.9
    # numpy restatement of the forward pass
    o = 0
    W1 = params[o:o + 64 * 12].reshape(64, 12); o += 64 * 12
    b1 = params[o:o + 64]; o += 64
    W2 = params[o:o + 64 * 64].reshape(64, 64); o += 64 * 64
    b2 = params[o:o + 64]; o += 64
    W3 = params[o:o + 16 * 64].reshape(16, 64); o += 16 * 64
    b3 = params[o:o + 16]
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    z = h2 @ W3.T + b3
    assert np.allclose(Z, z, rtol=1e-4, atol=1e-4)


def test_sweep_groups_are_well_formed(paths):
    data = T.collect(extents_per_family=2, seed=5, lo=10.0, hi=14.0, reps=1)
    G = len(data["times"])
    assert G >= 2 * len(data["families"]) - 5
    assert data["times"].shape == (G, T.NC + 1)
    assert np.all(data["times"] > 0) and np.all(np.isfinite(data["times"]))
    assert data["feats"].shape == (G, T.NF)


def test_tool_answers_requests_and_results_stay_exact(paths, oracle, tmp_path):
    """Load the tuner as a Kokkos Tools library, give it a model, and run
    reductions and scans through it: every request is answered and every
    integer result is still bit-exact."""
    P = paths
    rng = np.random.default_rng(1)
    X = rng.standard_normal((512, T.NF)).astype(np.float32)
    Y = np.eye(T.NC, dtype=np.float32)[rng.integers(0, T.NC, 512)]
    params, _ = T.train(X, Y, epochs=5, batch=64)
    mean, inv = T.normaliser(X)
    path = tmp_path / "m.bin"
    T.save_model(params, mean, inv, {}, path, tmp_path / "m.json")
    lib = str(K.lib_path("kkmx_tuner"))
    K.check(P.kkmx_tools_load(lib.encode()), P)
    try:
        T.load_into_tool(path)
        before = T.tool_stats()
        for n in (1000, 100003, 1 << 22):
            x = oracle.fill_uniform_i64(n, 31, -(1 << 40), 1 << 40)
            D, R = K.DeviceBuffer.from_array(x), K.DeviceBuffer(8)
            for op in (0, 1, 2):
                K.check(P.kkmx_reduce_i64(D.ptr, n, op, R.ptr), P)
                K.fence()
                assert R.download(np.int64, 1)[0] == oracle.reduce_i64(x, op)
            xi = oracle.fill_uniform_i32(n, 32, 0, 7)
            Di, O = K.DeviceBuffer.from_array(xi), K.DeviceBuffer(4 * n)
            K.check(P.kkmx_scan_incl_i32(Di.ptr, O.ptr, n), P)
            K.fence()
            assert np.array_equal(O.download(np.int32, n), oracle.scan_incl_i32(xi))
        after = T.tool_stats()
        print(before, after)
        assert after["requests"] > before["requests"]
        assert after["answered"] == after["requests"] - before["requests"] + before["answered"]
        assert after["inferences"] >= 1
    finally:
        K.check(P.kkmx_tools_unload(), P)


def test_tool_chooses_batch_tile_shape_and_results_match(paths, oracle):
    """With the tuner loaded, a batch planned without an explicit tile shape
    asks the tool (kokkos.mi355x.batch_shape); the shape is a valid one and
    the results equal those of the default shape bit for bit on integers and
    within 1e-12 on float64."""
    from kokkos_amd import batch as B
    from kokkos_amd import workloads as W

    P = paths
    wl = W.batch_workload(300, 21, 4, 16)
    ref = B.BatchRun(wl, 5)
    ref.run(B.MODE_FUSED, 1)
    K.fence()
    ri, rf = (a.copy() for a in ref.results())
    ref.close()
    lib = str(K.lib_path("kkmx_tuner"))
    K.check(P.kkmx_tools_load(lib.encode()), P)
    try:
        T.load_into_tool(T.MODEL_PATH)
        before = T.tool_stats()
        run = B.BatchRun(wl)
        after = T.tool_stats()
        print(f"tuned tile shape {run.shape}, tool {before} -> {after}")
        assert 0 <= run.shape < 8
        # the input fills are tuned range dispatches too; the plan adds one
        assert after["requests"] >= before["requests"] + 1
        run.run(B.MODE_FUSED_GRAPH, 2)
        K.fence()
        gi, gf = run.results()
        isint = wl["type"] != W.T_F64
        assert np.array_equal(gi[isint], ri[isint])
        rel = np.abs(gf[~isint] - rf[~isint]) / np.maximum(np.abs(rf[~isint]), 1e-300)
        assert rel.size == 0 or rel.max() <= 1e-12
        run.close()
    finally:
        K.check(P.kkmx_tools_unload(), P)
