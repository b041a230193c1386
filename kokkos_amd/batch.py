"""Batched uneven dispatch through the C-ABI (kkmx_batch_*): device pools for
a workload from kokkos_amd.workloads, the batch handle, and result download."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import DeviceBuffer, check, load
from .workloads import OP_SCAN_EXCL, POOL_SEEDS, T_F64, T_I32, T_I64

MODE_PLAIN, MODE_FUSED, MODE_FUSED_GRAPH, MODE_PLAIN_GRAPH = range(4)


class BatchRun:
    def __init__(self, wl: dict, shape: int | None = None):
        P = load("kkmx_paths")
        self.P, self.wl = P, wl
        tot = wl["in_tot"]
        self.pool = {
            T_I32: DeviceBuffer(max(tot[T_I32], 1) * 4),
            T_I64: DeviceBuffer(max(tot[T_I64], 1) * 8),
            T_F64: DeviceBuffer(max(tot[T_F64], 1) * 8),
        }
        check(P.kkmx_fill_uniform_i32(self.pool[T_I32].ptr, tot[T_I32], POOL_SEEDS[T_I32], 0, 15), P)
        check(P.kkmx_fill_uniform_i64(self.pool[T_I64].ptr, tot[T_I64], POOL_SEEDS[T_I64], 0, 15), P)
        check(P.kkmx_fill_uniform_f64(self.pool[T_F64].ptr, tot[T_F64], POOL_SEEDS[T_F64], 0.0, 1.0), P)
        self.out = {
            "i64": DeviceBuffer(max(wl["out_tot"]["i64"], 1) * 8),
            "f64": DeviceBuffer(max(wl["out_tot"]["f64"], 1) * 8),
        }
        nops = len(wl["n"])
        self.res = DeviceBuffer(max(nops, 1) * 8)
        h = C.c_void_p()
        check(P.kkmx_batch_create(C.byref(h)), P)
        self.h = h
        esz = {T_I32: 4, T_I64: 8, T_F64: 8}
        for i in range(nops):
            t, n, op = int(wl["type"][i]), int(wl["n"][i]), int(wl["op"][i])
            src = self.pool[t].ptr + int(wl["in_off"][i]) * esz[t]
            dst = None
            if op == OP_SCAN_EXCL:
                dst = self.out["f64" if t == T_F64 else "i64"].ptr + int(wl["out_off"][i]) * 8
            check(P.kkmx_batch_add(h, op, t, src, n, dst, self.res.ptr + 8 * i, 1.0, 0.0), P)
        if shape is not None:
            check(P.kkmx_batch_set_shape(h, shape), P)
        tiles = C.c_int64()
        check(P.kkmx_batch_plan(h, C.byref(tiles)), P)
        self.tiles = tiles.value
        sh = C.c_int()
        check(P.kkmx_batch_get_shape(h, C.byref(sh)), P)
        self.shape = sh.value
        check(P.kkmx_fence(), P)

    def run(self, mode: int, steps: int = 1) -> None:
        check(self.P.kkmx_batch_run(self.h, mode, steps), self.P)

    def results(self):
        """(int64 view, float64 view) of the per-op result slots."""
        nops = len(self.wl["n"])
        raw = self.res.download(np.int64, nops)
        return raw, raw.view(np.float64)

    def inputs(self):
        tot = self.wl["in_tot"]
        return {
            T_I32: self.pool[T_I32].download(np.int32, tot[T_I32]),
            T_I64: self.pool[T_I64].download(np.int64, tot[T_I64]),
            T_F64: self.pool[T_F64].download(np.float64, tot[T_F64]),
        }

    def outputs(self):
        o = self.wl["out_tot"]
        return self.out["i64"].download(np.int64, o["i64"]), self.out["f64"].download(np.float64, o["f64"])

    def clear_outputs(self):
        rt = load("kkmx_rt")
        s = self.P.kkmx_stream()
        for b in (self.res, self.out["i64"], self.out["f64"]):
            check(rt.mx_memset_async(b.ptr, 0xFF, b.nbytes, s))
        check(self.P.kkmx_fence(), self.P)

    def close(self):
        if self.h:
            check(self.P.kkmx_batch_destroy(self.h), self.P)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
