"""Loader for the restated reference unit tests (tests/cpp/selftest.hip,
built into tests/cpp/lib/libkkmx_selftest.so by kokkos_amd.build).

The library is test code: it links the product's runtime library and runs
each restated test as ordinary Kokkos user code on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import kokkos_amd as K

LIB = Path(__file__).resolve().parent / "cpp" / "lib" / "libkkmx_selftest.so"
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise K.KkmxError(f"{LIB} is missing: build it with python -m kokkos_amd.build kkmx_selftest")
        K.load("kkmx_rt")
        lib = C.CDLL(str(LIB), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        lib.kkmx_selftest_count.argtypes = []
        lib.kkmx_selftest_count.restype = C.c_int
        lib.kkmx_selftest_name.argtypes = [C.c_int]
        lib.kkmx_selftest_name.restype = C.c_char_p
        lib.kkmx_selftest_run.argtypes = [C.c_int, C.c_char_p, C.c_int]
        lib.kkmx_selftest_run.restype = C.c_int
        _lib = lib
    return _lib


def selftests() -> list[str]:
    """Names of the restated reference unit tests."""
    L = load()
    return [L.kkmx_selftest_name(i).decode() for i in range(L.kkmx_selftest_count())]


def run_selftest(name: str) -> tuple[int, str]:
    """Run one restated unit test on the GPU: (status, message); 0 = pass."""
    L = load()
    idx = selftests().index(name)
    buf = C.create_string_buffer(4096)
    rc = L.kkmx_selftest_run(idx, buf, len(buf))
    return rc, buf.value.decode(errors="replace")
