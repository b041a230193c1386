"""The learned launch-shape tuner: data collection, training and evaluation.

* collect(): the runtime sweeps itself. For each dispatch family
  (libkkmx_sweep.so) and a log-uniform set of extents, it times the 16
  launch-shape candidates and the heuristic with HIP events. One sample is
  one (family, extent, candidate, time) record.
* train(): soft-target cross-entropy on the 16 candidate logits, run by the
  HIP training kernels in libkkmx_tuner.so (kkmx_mlp_train).
* evaluate(): regret of the MLP's arg-max and of the heuristic against the
  exhaustive best, on held-out (family, extent) groups. Inference uses the
  HIP forward kernel (kkmx_mlp_predict).
* save_model() / load: the file format the tuning tool reads
  (kkmx_tuner_load_model, or the KKMX_TUNER_MODEL environment variable).
"""
from __future__ import annotations

import ctypes as C
import json
import struct
from pathlib import Path

import numpy as np

from . import KkmxError, load

NF, NH, NC = 12, 64, 16
MODEL_DIR = Path(__file__).resolve().parent / "tuner"
MODEL_PATH = MODEL_DIR / "mlp_gfx950.bin"
MODEL_META = MODEL_DIR / "mlp_gfx950.json"

f32p = C.POINTER(C.c_float)


def _libs():
    S = load("kkmx_sweep")
    T = load("kkmx_tuner")
    S.kkmx_sweep_family_count.restype = C.c_int
    S.kkmx_sweep_family_info.argtypes = ([C.c_int] + [C.POINTER(C.c_int)] * 4 + [C.POINTER(C.c_double)] * 2 +
                                         [C.POINTER(C.c_int)])
    S.kkmx_sweep_setup.argtypes = [C.c_int64]
    S.kkmx_sweep_time.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, f32p, C.POINTER(C.c_int64)]
    S.kkmx_sweep_last_error.restype = C.c_char_p
    S.kkmx_sweep_last_features.argtypes = [C.POINTER(C.c_double)]
    T.kkmx_mlp_param_count.restype = C.c_int
    T.kkmx_mlp_train.argtypes = [f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, C.c_int,
                                 f32p, f32p]
    T.kkmx_mlp_predict.argtypes = [f32p, f32p, C.c_int, f32p]
    T.kkmx_tuner_features.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                      C.c_int, C.c_int, f32p]
    T.kkmx_tuner_load_model.argtypes = [C.c_char_p]
    T.kkmx_tuner_stats.argtypes = [C.POINTER(C.c_uint64)] * 3
    T.kkmx_tuner_last_error.restype = C.c_char_p
    T.kkmx_tuner_last_answer.argtypes = [C.POINTER(C.c_int)] * 3
    return S, T


def _fp(a: np.ndarray):
    return a.ctypes.data_as(f32p)


def families():
    S, _ = _libs()
    out = []
    for f in range(S.kkmx_sweep_family_count()):
        p, vb, fl, fb, kd = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        by, fo = C.c_double(), C.c_double()
        S.kkmx_sweep_family_info(f, C.byref(p), C.byref(vb), C.byref(fl), C.byref(fb), C.byref(by), C.byref(fo),
                                 C.byref(kd))
        out.append(dict(pattern=p.value, value_bytes=vb.value, floating=fl.value, functor_bytes=fb.value,
                        bytes=by.value, flops=fo.value, kind=kd.value))
    return out


# Raw request inputs of one group: pattern, extent, value bytes, functor
# bytes, floating, bytes/index, flops/index, policy kind.
NRAW = 8


def family_raw(fam: dict, extent: float) -> np.ndarray:
    return np.array([fam["pattern"], extent, fam["value_bytes"], fam["functor_bytes"], fam["floating"],
                     fam["bytes"], fam["flops"], fam.get("kind", 0)], np.float64)


def features_from_raw(raw: np.ndarray, cus: int) -> np.ndarray:
    """The tool's 12 features (kkmx_tuner_features) of raw request inputs."""
    _, T = _libs()
    raw = np.atleast_2d(raw)
    out = np.zeros((len(raw), NF), np.float32)
    for i, r in enumerate(raw):
        T.kkmx_tuner_features(int(r[0]), float(r[1]), float(r[2]), float(r[3]), int(r[4]), float(r[5]), float(r[6]),
                              cus, int(r[7]), _fp(out[i]))
    return out


def raw_features(fam: dict, extent: float, cus: int) -> np.ndarray:
    return features_from_raw(family_raw(fam, extent), cus)[0]


def ncand_of_kind(kind: np.ndarray) -> np.ndarray:
    """Valid candidates per group: 8 tile shapes for batches, else 16."""
    return np.where(np.asarray(kind) == 3, 8, NC)


def collect(extents_per_family: int = 160, seed: int = 2024, lo: float = 8.0, hi: float = 27.0, reps: int = 2,
            cus: int = 256, progress=None) -> dict:
    """Times every family at `extents_per_family` log-uniform extents under the
    16 candidates and the heuristic.  Returns the groups as arrays."""
    S, _ = _libs()
    fams = families()
    max_n = int(2 ** hi) + 1
    if S.kkmx_sweep_setup(max_n) != 0:
        raise KkmxError(S.kkmx_sweep_last_error().decode())
    rng = np.random.default_rng(seed)
    fam_ids, extents, times, feats, raws = [], [], [], [], []
    buf = np.zeros(NC + 1, np.float32)
    neff = C.c_int64()
    lf = (C.c_double * 7)()
    # extent ranges (log2) per policy kind: Range indices, Team leagues,
    # MDRange points, batch element budgets
    ranges = {0: (lo, hi), 1: (6.0, 18.0), 2: (12.0, hi), 3: (14.0, hi)}
    try:
        for f in range(len(fams)):
            a, b = ranges[fams[f]["kind"]]
            ns = np.unique(np.floor(np.exp2(rng.uniform(a, b, extents_per_family))).astype(np.int64))
            for n in ns:
                if S.kkmx_sweep_time(f, int(n), reps, -1, 1, _fp(buf), C.byref(neff)) != 0:
                    raise KkmxError(S.kkmx_sweep_last_error().decode())
                fam_ids.append(f)
                extents.append(int(neff.value))
                times.append(buf.copy())
                if fams[f]["kind"] == 3:
                    S.kkmx_sweep_last_features(lf)
                    raw = np.array(list(lf) + [3.0], np.float64)
                else:
                    raw = family_raw(fams[f], neff.value)
                raws.append(raw)
                feats.append(features_from_raw(raw, cus)[0])
            if progress:
                progress(f"family {f} ({len(ns)} extents) done, {len(times)} groups", )
    finally:
        S.kkmx_sweep_teardown()
    return dict(family=np.array(fam_ids, np.int32), extent=np.array(extents, np.int64),
                times=np.array(times, np.float32), feats=np.array(feats, np.float32), raw=np.array(raws),
                families=fams)


def soft_targets(times: np.ndarray, tau: float = 0.03) -> np.ndarray:
    """Target distribution over the 16 candidates: softmax of -(t/t_best - 1)/tau."""
    t = times[:, :NC].astype(np.float64)
    rel = t / t.min(axis=1, keepdims=True) - 1.0
    z = np.exp(-rel / tau)
    return (z / z.sum(axis=1, keepdims=True)).astype(np.float32)


def normaliser(feats: np.ndarray):
    mean = feats.mean(axis=0)
    std = feats.std(axis=0)
    inv = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 0.0)
    return mean.astype(np.float32), inv.astype(np.float32)


def train(feats, targets, epochs=400, batch=128, lr=3e-3, weight_decay=1e-4, seed=7):
    _, T = _libs()
    X = np.ascontiguousarray(feats, np.float32)
    Y = np.ascontiguousarray(targets, np.float32)
    params = np.zeros(T.kkmx_mlp_param_count(), np.float32)
    loss = np.zeros(epochs, np.float32)
    rc = T.kkmx_mlp_train(_fp(X), _fp(Y), X.shape[0], epochs, batch, lr, weight_decay, seed, 1, _fp(params), _fp(loss))
    if rc != 0:
        raise KkmxError(T.kkmx_tuner_last_error().decode())
    return params, loss


def predict(params, feats) -> np.ndarray:
    _, T = _libs()
    X = np.ascontiguousarray(feats, np.float32)
    Z = np.zeros((X.shape[0], NC), np.float32)
    if T.kkmx_mlp_predict(_fp(params), _fp(X), X.shape[0], _fp(Z)) != 0:
        raise KkmxError(T.kkmx_tuner_last_error().decode())
    return Z


def regret(times: np.ndarray, choice: np.ndarray) -> np.ndarray:
    """t(choice)/t(best) - 1 per group; choice 16 = the heuristic column."""
    best = times[:, :NC].min(axis=1)
    return times[np.arange(len(times)), choice] / best - 1.0


def summary(r: np.ndarray) -> dict:
    return dict(mean=float(r.mean()), median=float(np.median(r)), p90=float(np.quantile(r, 0.9)),
                max=float(r.max()), within_5pct=float((r <= 0.05).mean()))


def save_model(params, mean, inv_std, meta: dict, path: Path = MODEL_PATH, meta_path: Path = MODEL_META):
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(b"KKMXMLP1")
        fh.write(struct.pack("<4i", NF, NH, NH, NC))
        fh.write(np.asarray(mean, "<f4").tobytes())
        fh.write(np.asarray(inv_std, "<f4").tobytes())
        fh.write(np.asarray(params, "<f4").tobytes())
    meta_path.write_text(json.dumps(meta, indent=1, sort_keys=True) + "\n")


def read_model(path: Path = MODEL_PATH):
    raw = path.read_bytes()
    assert raw[:8] == b"KKMXMLP1"
    dims = struct.unpack("<4i", raw[8:24])
    assert dims == (NF, NH, NH, NC)
    off = 24
    mean = np.frombuffer(raw, "<f4", NF, off)
    inv = np.frombuffer(raw, "<f4", NF, off + 4 * NF)
    params = np.frombuffer(raw, "<f4", offset=off + 8 * NF).copy()
    return params, mean.copy(), inv.copy()


def load_into_tool(path: Path = MODEL_PATH) -> None:
    _, T = _libs()
    if T.kkmx_tuner_load_model(str(path).encode()) != 0:
        raise KkmxError(T.kkmx_tuner_last_error().decode())


def last_answer():
    """(variable, heuristic's candidate, model's candidate) of the last
    tuning request the tool answered, or None."""
    _, T = _libs()
    v, h, c = C.c_int(), C.c_int(), C.c_int()
    if T.kkmx_tuner_last_answer(C.byref(v), C.byref(h), C.byref(c)) != 0:
        return None
    return v.value, h.value, c.value


def tool_stats():
    _, T = _libs()
    a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
    T.kkmx_tuner_stats(C.byref(a), C.byref(b), C.byref(c))
    return dict(requests=a.value, answered=b.value, inferences=c.value)


def group_raw(data: dict) -> np.ndarray:
    """Raw request inputs of every group (rebuilt from the family table for
    sample files written before they were stored)."""
    if "raw" in data:
        return np.asarray(data["raw"], np.float64)
    return np.array([family_raw(data["families"][f], e) for f, e in zip(data["family"], data["extent"])])


def hintless_features(data: dict, cus: int = 256) -> np.ndarray:
    """Features of every group as a functor without the optional
    kkmx_bytes_per_index / kkmx_flops_per_index hints would present them
    (batch groups keep theirs: the runtime computes them)."""
    raw = group_raw(data).copy()
    hinted = raw[:, 7] < 3
    raw[hinted, 5] = 0.0
    raw[hinted, 6] = 0.0
    return features_from_raw(raw, cus)


def masked_argmax(Z: np.ndarray, ncand: np.ndarray) -> np.ndarray:
    Z = Z.copy()
    Z[np.arange(NC)[None, :] >= np.asarray(ncand)[:, None]] = -np.inf
    return Z.argmax(axis=1)


def fit_and_evaluate(data: dict, seed: int = 11, holdout: float = 0.2, hint_dropout: bool = True, cus: int = 256,
                     **train_kw) -> dict:
    """Split groups, train on the rest, report regret of MLP vs heuristic on
    the held-out groups (and on all groups).

    hint_dropout: the training set also holds every training group with the
    functor's cost hints removed, so functors that declare none (most user
    code) still get answers learned from their pattern, types and extent."""
    rng = np.random.default_rng(seed)
    G = len(data["times"])
    idx = rng.permutation(G)
    nt = int(G * (1 - holdout))
    tr, te = idx[:nt], idx[nt:]
    Y = soft_targets(data["times"])
    X_tr, Y_tr = data["feats"][tr], Y[tr]
    hl = None
    if hint_dropout:
        hl = hintless_features(data, cus)
        X_tr = np.concatenate([X_tr, hl[tr]])
        Y_tr = np.concatenate([Y_tr, Y[tr]])
    mean, inv = normaliser(X_tr)
    params, loss = train((X_tr - mean) * inv, Y_tr, **train_kw)
    kinds = group_raw(data)[:, 7].astype(int)
    ncand = ncand_of_kind(kinds)
    Z = predict(params, (data["feats"] - mean) * inv)
    choice = masked_argmax(Z, ncand)
    r_mlp = regret(data["times"], choice)
    r_heur = regret(data["times"], np.full(G, NC))
    res = dict(
        groups=G, samples=int(G * NC), train_groups=int(nt), test_groups=int(G - nt),
        loss_first=float(loss[0]), loss_last=float(loss[-1]), hint_dropout=bool(hint_dropout),
        test=dict(mlp=summary(r_mlp[te]), heuristic=summary(r_heur[te])),
        train=dict(mlp=summary(r_mlp[tr]), heuristic=summary(r_heur[tr])),
    )
    if hl is not None:
        c_hl = masked_argmax(predict(params, (hl - mean) * inv), ncand)
        res["test_without_hints"] = dict(mlp=summary(regret(data["times"], c_hl)[te]))
    # held-out regret per policy kind (RangePolicy launch shape, Team shape,
    # MDRange tile)
    is_te = np.zeros(G, bool)
    is_te[te] = True
    res["test_by_kind"] = {
        name: dict(groups=int((is_te & (kinds == k)).sum()), mlp=summary(r_mlp[is_te & (kinds == k)]),
                   heuristic=summary(r_heur[is_te & (kinds == k)]))
        for k, name in enumerate(("range", "team", "mdrange", "batch")) if (is_te & (kinds == k)).any()
    }
    return dict(params=params, mean=mean, inv_std=inv, loss=loss, report=res, choice=choice)
