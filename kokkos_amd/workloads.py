"""Seeded synthetic workloads shared by the tests and the benchmark.

Batched uneven dispatch: 4096 independent ops per step, sizes floor(2^U) with
U uniform in [8, 20], input element types int32 / int64 / float64 in equal
proportion, each op a Sum reduction or an exclusive scan, all chosen from the
SplitMix64 stream of the seed (10 for the full workload).  Input values
follow the small plain-method workload: float64 uniform in [0, 1), integers
uniform in [0, 15].
"""
from __future__ import annotations

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """Output idx of the SplitMix64 stream seeded with `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


OP_SUM, OP_MIN, OP_MAX, OP_SCAN_EXCL, OP_SCAN_INCL, OP_AXPB = range(6)
T_I32, T_I64, T_F64 = range(3)
ELEM_BYTES = {T_I32: 4, T_I64: 8, T_F64: 8}
POOL_SEEDS = {T_I32: 11, T_I64: 12, T_F64: 10}


def batch_workload(nops: int = 4096, seed: int = 10, lo: float = 8.0, hi: float = 20.0) -> dict:
    """Sizes, element types and kinds of the ops, plus each op's offset in
    the input pool of its type and in the output pool of its accumulator."""
    k = np.arange(nops, dtype=np.uint64)
    r0 = splitmix64(seed, 3 * k)
    r1 = splitmix64(seed, 3 * k + np.uint64(1))
    r2 = splitmix64(seed, 3 * k + np.uint64(2))
    u = lo + (hi - lo) * ((r0 >> np.uint64(11)).astype(np.float64) * 2.0**-53)
    n = np.floor(np.exp2(u)).astype(np.int64)
    typ = ((r1 >> np.uint64(32)) * np.uint64(3) >> np.uint64(32)).astype(np.int32)
    op = np.where((r2 >> np.uint64(63)) == 0, OP_SUM, OP_SCAN_EXCL).astype(np.int32)
    in_off = np.zeros(nops, np.int64)
    out_off = np.zeros(nops, np.int64)
    in_tot = {T_I32: 0, T_I64: 0, T_F64: 0}
    out_tot = {"i64": 0, "f64": 0}
    for i in range(nops):
        t = int(typ[i])
        in_off[i] = in_tot[t]
        in_tot[t] += int(n[i])
        if op[i] == OP_SCAN_EXCL:
            key = "f64" if t == T_F64 else "i64"
            out_off[i] = out_tot[key]
            out_tot[key] += int(n[i])
    return dict(n=n, type=typ, op=op, in_off=in_off, out_off=out_off, in_tot=in_tot, out_tot=out_tot)


def batch_bytes(wl: dict) -> int:
    """Algorithmic HBM bytes of one step: every input element read once,
    every scan output element written once (8-byte accumulators)."""
    b = 0
    for n, t, o in zip(wl["n"], wl["type"], wl["op"]):
        b += int(n) * ELEM_BYTES[int(t)]
        if o in (OP_SCAN_EXCL, OP_SCAN_INCL, OP_AXPB):
            b += int(n) * 8
    return b
