"""CPU oracle: TEST INFRASTRUCTURE ONLY.

ctypes bindings for oracle/liboracle.so (a plain-C restatement of the
reference kokkos Serial/OpenMP semantics on the hot path, see kkmx_oracle.c).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it, and
only as the checker.  The product (kokkos_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

i64, u64, i32, f64, vp = C.c_int64, C.c_uint64, C.c_int32, C.c_double, C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sig = {
            "or_splitmix64": ([u64, u64], u64),
            "or_fill_uniform_f64": ([vp, i64, u64, f64, f64], None),
            "or_fill_uniform_i64": ([vp, i64, u64, i64, i64], None),
            "or_fill_uniform_i32": ([vp, i64, u64, i32, i32], None),
            "or_xorshift64star": ([vp, i64, u64], None),
            "or_checksum64": ([vp, i64], u64),
            "or_sum_f64_serial": ([vp, i64], f64),
            "or_sum_f64_openmp": ([vp, i64, C.c_int], f64),
            "or_sum_f64_neumaier": ([vp, i64], f64),
            "or_sum_f64_exact": ([vp, i64], f64),
            "or_reduce_i64": ([vp, i64, C.c_int, C.c_int], i64),
            "or_reduce_i32": ([vp, i64, C.c_int, C.c_int], i64),
            "or_scan_excl_i32_i64": ([vp, vp, i64, C.c_int, vp], None),
            "or_scan_incl_i32": ([vp, vp, i64, C.c_int, vp], None),
            "or_scan_excl_i64": ([vp, vp, i64, C.c_int, vp], None),
            "or_scan_excl_f64": ([vp, vp, i64, C.c_int, vp], None),
            "or_scan_excl_f64_compensated": ([vp, vp, i64, vp], None),
            "or_stream_copy": ([vp, vp, i64, C.c_int], None),
            "or_stream_scale": ([vp, vp, i64, f64, C.c_int], None),
            "or_stream_add": ([vp, vp, vp, i64, C.c_int], None),
            "or_stream_triad": ([vp, vp, vp, i64, f64, C.c_int, C.c_int], None),
            "or_max_threads": ([], C.c_int),
            "or_stencil7": ([vp, vp, i64, f64, f64, C.c_int], None),
            "or_csr_rowptr": ([vp, i64, i64, i64, u64], i64),
            "or_csr_fill": ([vp, vp, i64, i64, u64], None),
            "or_spmv_csr": ([vp, vp, vp, vp, vp, vp, i64, C.c_int], None),
            "or_gups": ([vp, i64, vp, i64, C.c_int], None),
            "or_zipf_cdf": ([vp, f64, i32], None),
            "or_zipf_keys": ([vp, i64, vp, i32, u64], None),
            "or_histogram": ([vp, i64, vp, i32], None),
            "or_scatter_f64": ([vp, vp, i64, vp, vp, i32], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data


# ---------------------------------------------------------------- numpy helpers
def fill_uniform_f64(n, seed, lo, hi):
    x = np.empty(n, np.float64)
    lib().or_fill_uniform_f64(_p(x), n, seed, lo, hi)
    return x


def fill_uniform_i64(n, seed, lo, hi):
    x = np.empty(n, np.int64)
    lib().or_fill_uniform_i64(_p(x), n, seed, lo, hi)
    return x


def fill_uniform_i32(n, seed, lo, hi):
    x = np.empty(n, np.int32)
    lib().or_fill_uniform_i32(_p(x), n, seed, lo, hi)
    return x


def xorshift64star(n, seed):
    x = np.empty(n, np.uint64)
    lib().or_xorshift64star(_p(x), n, seed)
    return x


def checksum(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a)
    return int(lib().or_checksum64(_p(a), a.nbytes))


def sum_serial(x):
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_sum_f64_serial(_p(x), x.size)


def sum_openmp(x, nthreads):
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_sum_f64_openmp(_p(x), x.size, nthreads)


def sum_neumaier(x):
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_sum_f64_neumaier(_p(x), x.size)


def sum_exact(x):
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_sum_f64_exact(_p(x), x.size)


def reduce_i64(x, op, nthreads=1):
    x = np.ascontiguousarray(x, np.int64)
    return int(lib().or_reduce_i64(_p(x), x.size, op, nthreads))


def reduce_i32(x, op, nthreads=1):
    x = np.ascontiguousarray(x, np.int32)
    return int(lib().or_reduce_i32(_p(x), x.size, op, nthreads))


def scan_excl_i32_i64(x, nthreads=1):
    x = np.ascontiguousarray(x, np.int32)
    out = np.empty(x.size, np.int64)
    tot = np.zeros(1, np.int64)
    lib().or_scan_excl_i32_i64(_p(x), _p(out), x.size, nthreads, _p(tot))
    return out, int(tot[0])


def scan_incl_i32(x, nthreads=1):
    x = np.ascontiguousarray(x, np.int32)
    out = np.empty(x.size, np.int32)
    tot = np.zeros(1, np.int32)
    lib().or_scan_incl_i32(_p(x), _p(out), x.size, nthreads, _p(tot))
    return out


def scan_excl_i64(x, nthreads=1):
    x = np.ascontiguousarray(x, np.int64)
    out = np.empty(x.size, np.int64)
    tot = np.zeros(1, np.int64)
    lib().or_scan_excl_i64(_p(x), _p(out), x.size, nthreads, _p(tot))
    return out, int(tot[0])


def scan_excl_f64(x, nthreads=1):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(x.size, np.float64)
    tot = np.zeros(1, np.float64)
    lib().or_scan_excl_f64(_p(x), _p(out), x.size, nthreads, _p(tot))
    return out, float(tot[0])


def scan_excl_f64_compensated(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(x.size, np.float64)
    tot = np.zeros(1, np.float64)
    lib().or_scan_excl_f64_compensated(_p(x), _p(out), x.size, _p(tot))
    return out, float(tot[0])


def stream_copy(a, nthreads=1):
    a = np.ascontiguousarray(a, np.float64)
    b = np.empty_like(a)
    lib().or_stream_copy(_p(a), _p(b), a.size, nthreads)
    return b


def stream_scale(c, s, nthreads=1):
    c = np.ascontiguousarray(c, np.float64)
    b = np.empty_like(c)
    lib().or_stream_scale(_p(b), _p(c), c.size, s, nthreads)
    return b


def stream_add(a, b, nthreads=1):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    c = np.empty_like(a)
    lib().or_stream_add(_p(a), _p(b), _p(c), a.size, nthreads)
    return c


def stream_triad(b, c, s, fused=True, nthreads=1):
    a = np.empty_like(b)
    lib().or_stream_triad(_p(a), _p(b), _p(c), b.size, s, 1 if fused else 0, nthreads)
    return a


def stencil7(u, n, c0, c1, nthreads=8):
    v = u.copy()
    lib().or_stencil7(_p(u), _p(v), n, c0, c1, nthreads)
    return v


def csr(nrows, ncols, lo, hi, seed):
    rp = np.empty(nrows + 1, np.int64)
    nnz = lib().or_csr_rowptr(_p(rp), nrows, lo, hi, seed)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib().or_csr_fill(_p(col), _p(val), nnz, ncols, seed)
    return rp, col, val


def spmv(rp, col, val, x, nthreads=8):
    n = rp.size - 1
    y = np.empty(n)
    a = np.empty(n)
    lib().or_spmv_csr(_p(rp), _p(col), _p(val), _p(x), _p(y), _p(a), n, nthreads)
    return y, a


def gups(table_size, stream, nupdates, nthreads=8):
    t = np.zeros(table_size, np.int64)
    lib().or_gups(_p(t), table_size, _p(stream), nupdates, nthreads)
    return t


def zipf_keys(n, s, nbins, seed):
    cdf = np.empty(nbins)
    lib().or_zipf_cdf(_p(cdf), s, nbins)
    k = np.empty(n, np.int32)
    lib().or_zipf_keys(_p(k), n, _p(cdf), nbins, seed)
    return k


def histogram(keys, nbins):
    keys = np.ascontiguousarray(keys, np.int32)
    b = np.empty(nbins, np.int32)
    lib().or_histogram(_p(keys), keys.size, _p(b), nbins)
    return b


def scatter_f64(vals, idx, nbins):
    vals = np.ascontiguousarray(vals, np.float64)
    idx = np.ascontiguousarray(idx, np.int32)
    b = np.empty(nbins)
    a = np.empty(nbins)
    lib().or_scatter_f64(_p(vals), _p(idx), vals.size, _p(b), _p(a), nbins)
    return b, a
