"""Sweep, train and evaluate the learned launch-shape tuner on the GPU.

    python -m kokkos_amd.tuner_train --out <dir> [--extents 160] [--epochs 400]
                                     [--samples <dir>/samples.npz]

--samples retrains on the groups of an earlier sweep instead of sweeping.

Writes to <dir>: samples.npz (every timed group), mlp_gfx950.bin and
mlp_gfx950.json (the model the tuning tool loads) and report.json (regret
of the MLP and of the heuristic against exhaustive search on held-out groups).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--extents", type=int, default=160)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--samples", help="train on an existing samples.npz instead of sweeping")
    a = ap.parse_args(argv)
    out = Path(a.out)
    out.mkdir(parents=True, exist_ok=True)

    import kokkos_amd as K
    from kokkos_amd import tuner as T

    K.init(0)
    cus = int(K.device_props()["cu_count"])
    t0 = time.time()
    if a.samples:
        z = np.load(a.samples)
        data = {k: z[k] for k in ("family", "extent", "times", "feats", "raw") if k in z}
        data["families"] = T.families()
        data["raw"] = T.group_raw(data)
    else:
        data = T.collect(extents_per_family=a.extents, seed=a.seed, reps=a.reps, cus=cus,
                         progress=lambda m: print(f"[{time.time() - t0:6.1f}s] {m}", flush=True))
    np.savez_compressed(out / "samples.npz", family=data["family"], extent=data["extent"], times=data["times"],
                        feats=data["feats"], raw=data["raw"])
    print(f"collected {len(data['times'])} groups = {len(data['times']) * T.NC} samples "
          f"in {time.time() - t0:.1f}s", flush=True)
    t1 = time.time()
    fit = T.fit_and_evaluate(data, epochs=a.epochs, cus=cus)
    rep = fit["report"]
    rep["train_seconds"] = time.time() - t1
    rep["sweep_seconds"] = t1 - t0
    rep["families"] = data["families"]
    rep["cu_count"] = int(cus)
    print(json.dumps({k: rep.get(k) for k in ("groups", "samples", "loss_first", "loss_last", "test", "test_by_kind",
                                              "test_without_hints")},
                     indent=1), flush=True)
    meta = dict(features=T.NF, hidden=T.NH, configs=T.NC, report=rep,
                candidates={"range": "for/reduce: block=128<<(c/4), items=1<<(c%4); scan: block=128<<(c/4), items=8<<(c%4)",
                            "team": "(team_size, vector_length) = kTeamCandidates[c]",
                            "mdrange": "tile: fastest dim 16<<(c%4), next 1<<(c/4), others 1",
                            "batch": "fused-kernel tile shape Impl::kBatchShape[c], c < 8"})
    T.save_model(fit["params"], fit["mean"], fit["inv_std"], meta, out / "mlp_gfx950.bin", out / "mlp_gfx950.json")
    (out / "report.json").write_text(json.dumps(rep, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
