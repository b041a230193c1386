"""Python access to the MI355X Kokkos execution space through its C ABI.

The product is C++/HIP: the Kokkos-API headers under kokkos_amd/include, the
C-ABI runtime (libkkmx_rt.so, include/kkmx_runtime.h) and the hot-path entry
points (libkkmx_paths.so, include/kkmx_paths.h).  This module only loads those
libraries with ctypes for tests and benchmarks.  There is no CPU fallback:
loading fails loudly when a library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIBDIR = PKG / "lib"

_libs: dict[str, C.CDLL] = {}


class KkmxError(RuntimeError):
    pass


def lib_path(name: str) -> Path:
    return LIBDIR / f"lib{name}.so"


def load(name: str) -> C.CDLL:
    """Load lib<name>.so from kokkos_amd/lib (build it with kokkos_amd.build)."""
    if name in _libs:
        return _libs[name]
    p = lib_path(name)
    if not p.exists():
        raise KkmxError(
            f"{p} is missing: build the HIP libraries first "
            "(python -m kokkos_amd.build or __graft_entry__.build())"
        )
    if name != "kkmx_rt":
        load("kkmx_rt")
    lib = C.CDLL(str(p), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    _libs[name] = lib
    _declare(name, lib)
    return lib


i64, u64, i32, f64, vp = C.c_int64, C.c_uint64, C.c_int32, C.c_double, C.c_void_p


def _sig(lib, fname, args, res=C.c_int):
    f = getattr(lib, fname)
    f.argtypes = args
    f.restype = res


def _declare(name: str, lib: C.CDLL) -> None:
    if name == "kkmx_rt":
        _sig(lib, "mx_init", [C.c_int])
        _sig(lib, "mx_device_count", [C.POINTER(C.c_int)])
        _sig(lib, "mx_malloc", [C.POINTER(vp), C.c_size_t, vp])
        _sig(lib, "mx_free", [vp, vp])
        _sig(lib, "mx_memcpy", [vp, vp, C.c_size_t])
        _sig(lib, "mx_memcpy_async", [vp, vp, C.c_size_t, vp])
        _sig(lib, "mx_memset_async", [vp, C.c_int, C.c_size_t, vp])
        _sig(lib, "mx_stream_sync", [vp])
        _sig(lib, "mx_device_sync", [])
        _sig(lib, "mx_event_create", [C.POINTER(vp)])
        _sig(lib, "mx_event_destroy", [vp])
        _sig(lib, "mx_event_record", [vp, vp])
        _sig(lib, "mx_event_sync", [vp])
        _sig(lib, "mx_event_query", [vp])
        _sig(lib, "mx_stream_capturing", [vp, C.POINTER(C.c_int)])
        _sig(lib, "mx_event_elapsed_ms", [C.POINTER(C.c_float), vp, vp])
        _sig(lib, "mx_error_string", [C.c_int], C.c_char_p)
        _sig(lib, "mx_default_stream", [C.POINTER(vp)])
        _sig(lib, "mx_device_props_get", [vp])
    elif name == "kkmx_paths":
        _sig(lib, "kkmx_init", [C.c_int])
        _sig(lib, "kkmx_finalize", [])
        _sig(lib, "kkmx_last_error", [], C.c_char_p)
        _sig(lib, "kkmx_stream", [], vp)
        _sig(lib, "kkmx_fence", [])
        _sig(lib, "kkmx_fill_uniform_f64", [vp, i64, u64, f64, f64])
        _sig(lib, "kkmx_fill_uniform_i64", [vp, i64, u64, i64, i64])
        _sig(lib, "kkmx_fill_uniform_i32", [vp, i64, u64, i32, i32])
        _sig(lib, "kkmx_stream_set", [vp, i64, f64])
        _sig(lib, "kkmx_stream_copy", [vp, vp, i64])
        _sig(lib, "kkmx_deep_copy_f64", [vp, vp, i64, i64, i64, C.c_int, C.c_int])
        _sig(lib, "kkmx_deep_fill_f64", [vp, i64, f64])
        _sig(lib, "kkmx_scatter_view_f64", [vp, vp, i64, vp, i32, C.c_int])
        _sig(lib, "kkmx_scatter_view_histogram", [vp, i64, vp, i32, C.c_int])
        _sig(lib, "kkmx_sort", [vp, i64, C.c_int])
        _sig(lib, "kkmx_sort_by_key", [vp, vp, i64, C.c_int, C.c_int, C.c_int])
        _sig(lib, "kkmx_binsort_f64", [vp, vp, i64, C.c_int, f64, f64, C.c_int, vp])
        _sig(lib, "kkmx_std_inclusive_scan_i32", [vp, vp, i64])
        _sig(lib, "kkmx_std_exclusive_scan_i64", [vp, vp, i64, i64])
        _sig(lib, "kkmx_std_inclusive_scan_max_i32", [vp, vp, i64])
        _sig(lib, "kkmx_std_transform_exclusive_scan_sq_f64", [vp, vp, i64, f64])
        _sig(lib, "kkmx_std_reduce_i64", [vp, i64, i64, C.POINTER(i64)])
        _sig(lib, "kkmx_std_reduce_max_i64", [vp, i64, i64, C.POINTER(i64)])
        _sig(lib, "kkmx_std_transform_reduce_dot_f64", [vp, vp, i64, f64, C.POINTER(f64)])
        _sig(lib, "kkmx_stream_scale", [vp, vp, i64, f64])
        _sig(lib, "kkmx_stream_add", [vp, vp, vp, i64])
        _sig(lib, "kkmx_stream_triad", [vp, vp, vp, i64, f64])
        _sig(lib, "kkmx_reduce_sum_f64", [vp, i64, vp])
        _sig(lib, "kkmx_reduce_i64", [vp, i64, C.c_int, vp])
        _sig(lib, "kkmx_reduce_i32", [vp, i64, C.c_int, vp])
        _sig(lib, "kkmx_reduce_minmaxsum_i64", [vp, i64, vp])
        _sig(lib, "kkmx_scan_excl_i32_i64", [vp, vp, i64, vp])
        _sig(lib, "kkmx_scan_incl_i32", [vp, vp, i64])
        _sig(lib, "kkmx_scan_incl_i32_traits", [vp, vp, i64, C.c_int])
        _sig(lib, "kkmx_set_scan_body", [C.c_int])
        _sig(lib, "kkmx_scan_incl_i32_body", [C.c_int])
        _sig(lib, "kkmx_scan_incl_i32_noalloc", [vp, vp, i64])
        _sig(lib, "kkmx_scan_excl_i64", [vp, vp, i64, vp])
        _sig(lib, "kkmx_scan_excl_f64", [vp, vp, i64, vp])
        _sig(lib, "kkmx_set_range_config", [C.c_int, C.c_int, C.c_int, C.c_int])
        _sig(lib, "kkmx_range_candidate", [C.c_int, C.c_int, i64] + [C.POINTER(C.c_int)] * 3)
        _sig(lib, "kkmx_stencil7", [vp, vp, i64, f64, f64, C.c_int, C.c_int, C.c_int])
        _sig(lib, "kkmx_stencil7_sweeps", [vp, vp, i64, C.c_int, f64, f64, C.c_int, C.c_int, C.c_int])
        _sig(lib, "kkmx_csr_generate", [vp, vp, vp, i64, i64, i64, i64, u64, C.POINTER(i64)])
        _sig(lib, "kkmx_csr_fill", [vp, vp, i64, i64, u64])
        _sig(lib, "kkmx_spmv_csr", [vp, vp, vp, vp, vp, i64, i64, C.c_int, C.c_int, C.c_int])
        _sig(lib, "kkmx_xorshift64_stream", [vp, i64, u64])
        _sig(lib, "kkmx_gups", [vp, i64, vp, i64])
        _sig(lib, "kkmx_zipf_keys", [vp, i64, f64, i32, u64])
        _sig(lib, "kkmx_histogram", [vp, i64, vp, i32, C.c_int, C.c_int])
        _sig(lib, "kkmx_scatter_add_f64", [vp, vp, i64, vp, i32])
        _sig(lib, "kkmx_batch_create", [C.POINTER(vp)])
        _sig(lib, "kkmx_batch_destroy", [vp])
        _sig(lib, "kkmx_batch_add", [vp, C.c_int, C.c_int, vp, i64, vp, vp, f64, f64])
        _sig(lib, "kkmx_batch_plan", [vp, C.POINTER(i64)])
        _sig(lib, "kkmx_batch_set_shape", [vp, C.c_int])
        _sig(lib, "kkmx_batch_get_shape", [vp, C.POINTER(C.c_int)])
        _sig(lib, "kkmx_set_mdrange_depth", [C.c_int])
        _sig(lib, "kkmx_batch_run", [vp, C.c_int, C.c_int])
    elif name == "kkmx_peak":
        _sig(lib, "kkmx_peak_copy", [vp, vp, i64, C.c_int, C.c_int, vp])


def check(rc: int, lib: C.CDLL | None = None) -> None:
    if rc != 0:
        msg = ""
        if lib is not None and hasattr(lib, "kkmx_last_error"):
            msg = (lib.kkmx_last_error() or b"").decode()
        if not msg:
            msg = (load("kkmx_rt").mx_error_string(rc) or b"").decode()
        raise KkmxError(f"status {rc}: {msg}")


class DeviceBuffer:
    """Device allocation through the C-ABI runtime (mx_malloc / mx_free)."""

    def __init__(self, nbytes: int):
        rt = load("kkmx_rt")
        self.nbytes = int(nbytes)
        p = vp()
        check(rt.mx_malloc(C.byref(p), max(self.nbytes, 1), None))
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            load("kkmx_rt").mx_free(self.ptr, None)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(load("kkmx_rt").mx_memcpy(self.ptr, arr.ctypes.data, arr.nbytes))
        return self

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(load("kkmx_rt").mx_memcpy(out.ctypes.data, self.ptr, out.nbytes))
        return out

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        return cls(arr.nbytes).upload(arr)


class DeviceProps(C.Structure):
    """mx_device_props of include/kkmx_runtime.h."""

    _fields_ = [
        ("name", C.c_char * 256),
        ("arch", C.c_char * 64),
        ("device", C.c_int),
        ("cu_count", C.c_int),
        ("wave_size", C.c_int),
        ("max_threads_per_block", C.c_int),
        ("max_threads_per_cu", C.c_int),
        ("regs_per_block", C.c_int),
        ("clock_khz", C.c_int),
        ("mem_clock_khz", C.c_int),
        ("mem_bus_width", C.c_int),
        ("l2_bytes", C.c_int),
        ("xcc_count", C.c_int),
        ("lds_per_block", C.c_size_t),
        ("lds_per_cu", C.c_size_t),
        ("global_mem", C.c_size_t),
        ("max_grid_x", C.c_int),
        ("concurrent_kernels", C.c_int),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["name"] = d["name"].decode()
        d["arch"] = d["arch"].decode()
        return d


def device_props() -> dict:
    init()
    p = DeviceProps()
    check(load("kkmx_rt").mx_device_props_get(C.byref(p)))
    return p.as_dict()


def init(device: int = 0) -> C.CDLL:
    """Initialise Kokkos on the MI355X space and return the paths library."""
    p = load("kkmx_paths")
    check(p.kkmx_init(device), p)
    return p


def fence() -> None:
    p = load("kkmx_paths")
    check(p.kkmx_fence(), p)


class EventTimer:
    """HIP events on the stream the entry points are ordered on."""

    def __init__(self):
        self.rt = load("kkmx_rt")
        self.stream = load("kkmx_paths").kkmx_stream()
        self.a, self.b = vp(), vp()
        check(self.rt.mx_event_create(C.byref(self.a)))
        check(self.rt.mx_event_create(C.byref(self.b)))

    def start(self):
        check(self.rt.mx_event_record(self.a, self.stream))

    def stop(self) -> float:
        check(self.rt.mx_event_record(self.b, self.stream))
        check(self.rt.mx_event_sync(self.b))
        ms = C.c_float()
        check(self.rt.mx_event_elapsed_ms(C.byref(ms), self.a, self.b))
        return float(ms.value)

    def __del__(self):
        try:
            self.rt.mx_event_destroy(self.a)
            self.rt.mx_event_destroy(self.b)
        except Exception:
            pass
