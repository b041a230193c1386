"""Benchmark of the MI355X Kokkos execution space on its bandwidth roofline.

Workload (one GPU): STREAM copy / scale / add / triad over three
View<double*> of N = 2^28 (U[-1,1), seed 2, scalar 3.0), parallel_reduce
max / min / sum over View<int64_t*> of 2^28 (U[-2^40, 2^40], seed 3) and
parallel_scan inclusive over View<int32_t*> of 2^28 (U[0,8), seed 4).  Inputs
are generated on the device before timing.  A step is one pass of all eight
dispatches; `value` is the algorithmic bytes of all steps divided by the
time, in GB/s.  Per-dispatch times come from HIP events on the stream the
dispatches run on.

Beside the headline line the same run reports, as extra fields, the other
workloads of the path on the same GPU: the small plain-method pair (reduce
2^24 doubles, exclusive scan 2^24 int32 -> int64), the MDRange stencil and
Team SpMV, the atomics (GUPS, Zipf histogram, FP64 scatter) and batched
uneven dispatch (4096 reduce/scan ops per step: per-call launches, the
persistent fused kernel replayed from a hipGraph, and per-call launches
with the learned tuner).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0  # MI355X published HBM3E bandwidth
N = 1 << 28

# algorithmic bytes per element of each dispatch
OPS = [
    ("copy", 16),
    ("scale", 16),
    ("add", 24),
    ("triad", 24),
    ("reduce_max_i64", 8),
    ("reduce_min_i64", 8),
    ("reduce_sum_i64", 8),
    ("scan_incl_i32", 8),
]


def measured_copy_peak(K, a, b, n_bytes, timer):
    """Best device-to-device copy among hand-tuned variants (GB/s, 2x bytes)."""
    peak = K.load("kkmx_peak")
    s = K.load("kkmx_paths").kkmx_stream()
    best = 0.0
    for variant, bpc in [(6, 0), (5, 0), (2, 0), (0, 0), (2, 8)]:
        for _ in range(2):
            K.check(peak.kkmx_peak_copy(a, b, n_bytes, variant, bpc, s))
        ts = []
        for _ in range(5):
            timer.start()
            K.check(peak.kkmx_peak_copy(a, b, n_bytes, variant, bpc, s))
            ts.append(timer.stop())
        best = max(best, 2 * n_bytes / (min(ts) * 1e-3) / 1e9)
    return best


def cpu_baseline(n=N, steps=20, warmup=3):
    """The oracle (OpenMP restatement of the reference's OpenMP backend) on the
    host cores: the same 8 dispatches on arrays of the same size (2^28), each
    step timed on its own; `warmup` untimed steps first, then the median (and
    minimum) of `steps` timed steps, GB/s of algorithmic bytes."""
    import numpy as np

    import oracle as O

    L = O.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or L.or_max_threads()
    threads = max(1, min(threads, L.or_max_threads()))
    a = O.fill_uniform_f64(n, 2, -1.0, 1.0)
    b = O.fill_uniform_f64(n, 3, -1.0, 1.0)
    c = O.fill_uniform_f64(n, 4, -1.0, 1.0)
    x = O.fill_uniform_i64(n, 3, -(1 << 40), 1 << 40)
    s = O.fill_uniform_i32(n, 4, 0, 7)
    so = np.empty(n, np.int32)
    tot = np.zeros(1, np.int32)
    p = lambda arr: arr.ctypes.data  # noqa: E731

    def step():
        L.or_stream_copy(p(a), p(c), n, threads)
        L.or_stream_scale(p(b), p(c), n, 3.0, threads)
        L.or_stream_add(p(a), p(b), p(c), n, threads)
        L.or_stream_triad(p(a), p(b), p(c), n, 3.0, 0, threads)
        L.or_reduce_i64(p(x), n, 2, threads)
        L.or_reduce_i64(p(x), n, 1, threads)
        L.or_reduce_i64(p(x), n, 0, threads)
        L.or_scan_incl_i32(p(s), p(so), n, threads, p(tot))

    for _ in range(warmup):
        step()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        step()
        ts.append(time.perf_counter() - t0)
    bytes_per_step = sum(bpe for _, bpe in OPS) * n
    med, mn = float(np.median(ts)), float(min(ts))
    return {
        "value": round(bytes_per_step / med / 1e9, 2),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "host_cpus": os.cpu_count(),
        "median_step_s": round(med, 4),
        "min_step_s": round(mn, 4),
        "best_gbps": round(bytes_per_step / mn / 1e9, 2),
        "sample": f"the same 8 dispatches on 2^{n.bit_length() - 1}-element arrays (oracle/kkmx_oracle.c, OpenMP, "
        f"{threads} threads): {warmup} untimed steps, then the median of {steps} timed steps "
        f"({sum(ts):.1f} s)",
    }


def triad_exact(b, c, s=3.0, chunk=1 << 24):
    """Correctly rounded fma(s, c, b) in float64 for |s| with at most 11
    significant bits (s = 3.0 here), from numpy alone: s*c is exact in
    x87 long double (64-bit significand), TwoSum gives b + s*c = S + e exactly,
    and S is rounded to double with the double-rounding case (S exactly half
    way between two doubles, e != 0) corrected toward e."""
    LD = np.longdouble
    if np.finfo(LD).nmant < 63:
        raise RuntimeError("long double is not the x87 extended format")
    out = np.empty_like(b)
    for i in range(0, b.size, chunk):
        bl = b[i:i + chunk].astype(LD)
        yl = c[i:i + chunk].astype(LD) * LD(s)
        S = bl + yl
        bp = S - bl
        e = (bl - (S - bp)) + (yl - bp)
        d = S.astype(np.float64)
        r = S - d.astype(LD)
        dn = np.nextafter(d, np.where(r > 0, np.inf, -np.inf))
        mid = (r != 0) & (r * 2 == dn.astype(LD) - d.astype(LD))
        fix = mid & (e != 0) & (np.sign(e) == np.sign(r))
        d[fix] = dn[fix]
        out[i:i + chunk] = d
    return out


def verify_step(K, P, calls, a, b, c, x, s, so, res):
    """Checks one more step of the timed dispatches, bit for bit, with numpy
    on the downloaded inputs: copy, scale and add exactly; the triad against
    the correctly rounded fused multiply-add the device performs; the three
    int64 reductions (sums wrap like the device's); the int32 inclusive scan
    against numpy's int32 cumulative sum (the total, 7 * 2^28 at most, fits)."""
    ok = {}
    ha = a.download(np.float64, N)
    K.check(calls[0](), P)
    hc = c.download(np.float64, N)
    ok["copy"] = bool(np.array_equal(hc, ha))
    K.check(calls[1](), P)
    hb = b.download(np.float64, N)
    ok["scale"] = bool(np.array_equal(hb, 3.0 * hc))
    K.check(calls[2](), P)
    hc = c.download(np.float64, N)
    ok["add"] = bool(np.array_equal(hc, ha + hb))
    del ha
    K.check(calls[3](), P)
    ok["triad"] = bool(np.array_equal(a.download(np.float64, N), triad_exact(hb, hc)))
    del hb, hc
    for k in (4, 5, 6):
        K.check(calls[k](), P)
    K.check(calls[7](), P)
    K.fence()
    hx = x.download(np.int64, N)
    got = res.download(np.int64, 3)
    ok["reduce_max_i64"] = int(got[0]) == int(hx.max())
    ok["reduce_min_i64"] = int(got[1]) == int(hx.min())
    ok["reduce_sum_i64"] = int(got[2]) == int(hx.sum())
    del hx
    ok["scan_incl_i32"] = bool(np.array_equal(so.download(np.int32, N), np.cumsum(s.download(np.int32, N),
                                                                                   dtype=np.int32)))
    return ok


def _timeit(K, timer, fn, reps=5, warm=2):
    for _ in range(warm):
        fn()
    K.fence()
    ts = []
    for _ in range(reps):
        timer.start()
        fn()
        ts.append(timer.stop())
    return min(ts), float(np.median(ts))


def scan_view_traits(K, P, s, so, n):
    """The 2^28 int32 inclusive scan with its input read through a plain,
    a RandomAccess and a Restrict View (kkmx_scan_incl_i32_traits), each after
    its functor type's tile-body trials: GB/s (8 bytes per element), the body
    the runtime kept, and exactness against the headline scan's output."""
    import ctypes as C
    timer = K.EventTimer()
    ref = so.download(np.int32, n)
    out = K.DeviceBuffer(n * 4)
    res = {}
    for traits, name in [(0, "plain"), (1, "random_access"), (2, "restrict")]:
        for _ in range(8):
            K.check(P.kkmx_scan_incl_i32_traits(s.ptr, out.ptr, n, traits), P)
            K.fence()
        mn, med = _timeit(K, timer, lambda: K.check(P.kkmx_scan_incl_i32_traits(s.ptr, out.ptr, n, traits), P),
                          reps=10, warm=1)
        res[name] = {"ms": round(mn, 4), "gbps": round(8 * n / mn / 1e6, 1), "median_ms": round(med, 4),
                     "tile_body": int(P.kkmx_scan_incl_i32_body(traits)),
                     "exact": bool(np.array_equal(out.download(np.int32, n), ref))}
    del out
    return res


def leg_small(K, P, timer):
    """Plain-method pair: reduce sum of 2^24 U[0,1) doubles (seed 0) and the
    exclusive scan of 2^24 U[0,15] int32 (seed 1) into int64."""
    n = 1 << 24
    chk = lambda rc: K.check(rc, P)  # noqa: E731
    x, r = K.DeviceBuffer(8 * n), K.DeviceBuffer(8)
    s, o, t = K.DeviceBuffer(4 * n), K.DeviceBuffer(8 * n), K.DeviceBuffer(8)
    chk(P.kkmx_fill_uniform_f64(x.ptr, n, 0, 0.0, 1.0))
    chk(P.kkmx_fill_uniform_i32(s.ptr, n, 1, 0, 15))
    K.fence()
    red = _timeit(K, timer, lambda: chk(P.kkmx_reduce_sum_f64(x.ptr, n, r.ptr)), reps=20)
    scn = _timeit(K, timer, lambda: chk(P.kkmx_scan_excl_i32_i64(s.ptr, o.ptr, n, t.ptr)), reps=20)
    import math

    got = float(r.download(np.float64, 1)[0])
    exact = math.fsum(x.download(np.float64, n))
    tot = int(t.download(np.int64, 1)[0])
    tot_ref = int(s.download(np.int32, n).astype(np.int64).sum())
    return {
        "reduce_sum_f64_2^24": {"ms": round(red[0], 4), "gbps": round(8 * n / red[0] / 1e6, 1),
                                "rel_err_vs_exact": abs(got - exact) / exact},
        "scan_excl_i32_i64_2^24": {"ms": round(scn[0], 4), "gbps": round(12 * n / scn[0] / 1e6, 1),
                                   "total_exact": tot == tot_ref},
    }


def leg_md_team(K, P, timer):
    """7-point stencil on 512^3 (+halo), 10 sweeps; CSR SpMV with 2^20 rows."""
    import ctypes as C

    chk = lambda rc: K.check(rc, P)  # noqa: E731
    n, m = 512, 514
    U, V = K.DeviceBuffer(m**3 * 8), K.DeviceBuffer(m**3 * 8)
    chk(P.kkmx_fill_uniform_f64(U.ptr, m**3, 5, 0.0, 1.0))
    chk(P.kkmx_fill_uniform_f64(V.ptr, m**3, 5, 0.0, 1.0))
    st = _timeit(K, timer, lambda: chk(P.kkmx_stencil7_sweeps(U.ptr, V.ptr, n, 10, 0.4, 0.1, 0, 0, 0)), reps=3, warm=1)
    del U, V
    nrows = 1 << 20
    rp = K.DeviceBuffer((nrows + 1) * 8)
    nnz = C.c_int64()
    chk(P.kkmx_csr_generate(rp.ptr, None, None, nrows, nrows, 4, 64, 6, C.byref(nnz)))
    col, val = K.DeviceBuffer(nnz.value * 4), K.DeviceBuffer(nnz.value * 8)
    chk(P.kkmx_csr_fill(col.ptr, val.ptr, nnz.value, nrows, 6))
    X, Y = K.DeviceBuffer(nrows * 8), K.DeviceBuffer(nrows * 8)
    chk(P.kkmx_fill_uniform_f64(X.ptr, nrows, 606, -1.0, 1.0))
    sp = _timeit(K, timer, lambda: chk(P.kkmx_spmv_csr(rp.ptr, col.ptr, val.ptr, X.ptr, Y.ptr, nrows, nnz.value,
                                                       1, 0, 0)), reps=20)
    spmv_bytes = nnz.value * 12 + nrows * 16 + nrows * 8
    return {
        "stencil7_512^3_x10": {"ms": round(st[0], 3), "gbps": round(16 * n**3 * 10 / st[0] / 1e6, 1),
                               "bytes_per_point_sweep": 16},
        "spmv_team_2^20_rows": {"ms": round(sp[0], 4), "gbps": round(spmv_bytes / sp[0] / 1e6, 1),
                                "nnz": int(nnz.value), "algorithmic_bytes": int(spmv_bytes)},
    }


def leg_deep_copy(K, P, timer):
    """deep_copy between device Views: a contiguous 2 GiB copy (against the
    runtime's own device-to-device memcpy), a 2 GiB fill, and a 512^3 copy
    from LayoutLeft to LayoutRight."""
    chk = lambda rc: K.check(rc, P)  # noqa: E731
    rt = K.load("kkmx_rt")
    n = 1 << 28
    A, B = K.DeviceBuffer(n * 8), K.DeviceBuffer(n * 8)
    chk(P.kkmx_fill_uniform_f64(A.ptr, n, 44, -1.0, 1.0))
    cp = _timeit(K, timer, lambda: chk(P.kkmx_deep_copy_f64(B.ptr, A.ptr, n, 1, 1, 0, 0)), reps=10)
    mc = _timeit(K, timer, lambda: K.check(rt.mx_memcpy_async(B.ptr, A.ptr, n * 8, P.kkmx_stream())), reps=10)
    fl = _timeit(K, timer, lambda: chk(P.kkmx_deep_fill_f64(B.ptr, n, 0.5)), reps=10)
    m = 512
    tr = _timeit(K, timer, lambda: chk(P.kkmx_deep_copy_f64(B.ptr, A.ptr, m, m, m, 1, 0)), reps=10)
    del A, B
    return {
        "copy_contiguous_2GiB": {"ms": round(cp[0], 4), "gbps": round(16 * n / cp[0] / 1e6, 1)},
        "runtime_memcpy_2GiB": {"ms": round(mc[0], 4), "gbps": round(16 * n / mc[0] / 1e6, 1)},
        "fill_2GiB": {"ms": round(fl[0], 4), "gbps": round(8 * n / fl[0] / 1e6, 1)},
        "copy_left_to_right_512^3": {"ms": round(tr[0], 4), "gbps": round(16 * m**3 / tr[0] / 1e6, 1)},
    }


def leg_std_algorithms(K, P, timer):
    """Kokkos::Experimental std algorithms at 2^28 elements: inclusive_scan
    int32, exclusive_scan int64 with an initial value, a running-maximum
    inclusive_scan (user operator, no neutral element), reduce int64 and the
    float64 dot product transform_reduce."""
    import ctypes as C

    chk = lambda rc: K.check(rc, P)  # noqa: E731
    n = 1 << 28
    A, B = K.DeviceBuffer(n * 8), K.DeviceBuffer(n * 8)
    O = K.DeviceBuffer(n * 8)
    out_i, out_f = C.c_int64(), C.c_double()
    chk(P.kkmx_fill_uniform_i32(A.ptr, n, 61, 0, 7))
    isc = _timeit(K, timer, lambda: chk(P.kkmx_std_inclusive_scan_i32(A.ptr, O.ptr, n)), reps=10)
    msc = _timeit(K, timer, lambda: chk(P.kkmx_std_inclusive_scan_max_i32(A.ptr, O.ptr, n)), reps=10)
    chk(P.kkmx_fill_uniform_i64(A.ptr, n, 62, -(1 << 40), 1 << 40))
    esc = _timeit(K, timer, lambda: chk(P.kkmx_std_exclusive_scan_i64(A.ptr, O.ptr, n, 42)), reps=10)
    red = _timeit(K, timer, lambda: chk(P.kkmx_std_reduce_i64(A.ptr, n, 0, C.byref(out_i))), reps=10)
    chk(P.kkmx_fill_uniform_f64(A.ptr, n, 63, 0.0, 1.0))
    chk(P.kkmx_fill_uniform_f64(B.ptr, n, 64, 0.0, 1.0))
    dot = _timeit(K, timer, lambda: chk(P.kkmx_std_transform_reduce_dot_f64(A.ptr, B.ptr, n, 0.0, C.byref(out_f))),
                  reps=10)
    del A, B, O
    g = lambda bpe, t: round(bpe * n / t[0] / 1e6, 1)  # noqa: E731
    return {
        "inclusive_scan_i32": {"ms": round(isc[0], 4), "gbps": g(8, isc)},
        "inclusive_scan_max_i32": {"ms": round(msc[0], 4), "gbps": g(8, msc)},
        "exclusive_scan_i64_init": {"ms": round(esc[0], 4), "gbps": g(16, esc)},
        "reduce_i64": {"ms": round(red[0], 4), "gbps": g(8, red)},
        "transform_reduce_dot_f64": {"ms": round(dot[0], 4), "gbps": g(16, dot)},
    }


def leg_sort(K, P, timer):
    """Kokkos::sort (radix) of 2^28 int32 keys (full range, and below 2^16,
    where the two constant high digits are skipped) and 2^27 float64 keys,
    sort_by_key of 2^28 (int32, uint32) pairs, and BinSort of 2^24 float64
    keys over 2^20 bins with the within-bin sort."""
    chk = lambda rc: K.check(rc, P)  # noqa: E731
    n32, n64 = 1 << 28, 1 << 27
    A = K.DeviceBuffer(n32 * 4)
    fill32 = lambda: chk(P.kkmx_fill_uniform_i32(A.ptr, n32, 91, -(1 << 30), 1 << 30))  # noqa: E731
    fill32()
    t32 = []
    for _ in range(3):
        fill32()
        K.fence()
        timer.start()
        chk(P.kkmx_sort(A.ptr, n32, 0))
        t32.append(timer.stop())
    tsr = []
    for _ in range(3):
        chk(P.kkmx_fill_uniform_i32(A.ptr, n32, 95, 0, (1 << 16) - 1))
        K.fence()
        timer.start()
        chk(P.kkmx_sort(A.ptr, n32, 0))
        tsr.append(timer.stop())
    V = K.DeviceBuffer(n32 * 4)
    tkv = []
    for _ in range(3):
        fill32()
        chk(P.kkmx_fill_uniform_i32(V.ptr, n32, 96, 0, 1 << 30))
        K.fence()
        timer.start()
        chk(P.kkmx_sort_by_key(A.ptr, V.ptr, n32, 0, 4, 0))
        tkv.append(timer.stop())
    del A, V
    B = K.DeviceBuffer(n64 * 8)
    t64 = []
    for _ in range(3):
        chk(P.kkmx_fill_uniform_f64(B.ptr, n64, 92, -1.0, 1.0))
        K.fence()
        timer.start()
        chk(P.kkmx_sort(B.ptr, n64, 3))
        t64.append(timer.stop())
    del B
    nb, bins = 1 << 24, 1 << 20
    Kd, Vd = K.DeviceBuffer(nb * 8), K.DeviceBuffer(nb * 8)
    chk(P.kkmx_fill_uniform_f64(Kd.ptr, nb, 93, 0.0, 1.0))
    chk(P.kkmx_fill_uniform_i64(Vd.ptr, nb, 94, 0, 1000))
    bs = _timeit(K, timer, lambda: chk(P.kkmx_binsort_f64(Kd.ptr, Vd.ptr, nb, bins, 0.0, 1.0, 1, None)), reps=3,
                 warm=1)
    del Kd, Vd
    return {
        "sort_int32_2^28": {"ms": round(min(t32), 3), "keys_per_s": round(n32 / min(t32) * 1e3, 1)},
        "sort_int32_2^28_keys_below_2^16": {"ms": round(min(tsr), 3), "keys_per_s": round(n32 / min(tsr) * 1e3, 1),
                                            "passes": 2},
        "sort_by_key_int32_u32_2^28": {"ms": round(min(tkv), 3), "pairs_per_s": round(n32 / min(tkv) * 1e3, 1)},
        "sort_f64_2^27": {"ms": round(min(t64), 3), "keys_per_s": round(n64 / min(t64) * 1e3, 1)},
        "binsort_f64_2^24_keys_2^20_bins": {"ms": round(bs[0], 3), "keys_per_s": round(nb / bs[0] * 1e3, 1)},
    }


def leg_atomics(K, P, timer):
    """GUPS (2^30 int64 table, 2^28 updates), Zipf(1.1) histogram of 2^28
    keys over 65536 bins (LDS-privatised, and through ScatterView), FP64
    scatter of 2^26 into 2^20 (atomic_add, and through ScatterView)."""
    chk = lambda rc: K.check(rc, P)  # noqa: E731
    rt = K.load("kkmx_rt")
    T, U = 1 << 30, 1 << 28
    S, Tb = K.DeviceBuffer(2 * U * 8), K.DeviceBuffer(T * 8)
    chk(P.kkmx_xorshift64_stream(S.ptr, 2 * U, 7))
    K.check(rt.mx_memset_async(Tb.ptr, 0, T * 8, P.kkmx_stream()))
    g = _timeit(K, timer, lambda: chk(P.kkmx_gups(Tb.ptr, T, S.ptr, U)), reps=3, warm=1)
    del S, Tb
    nk, nb = 1 << 28, 65536
    Kb, Bb = K.DeviceBuffer(nk * 4), K.DeviceBuffer(nb * 4)
    chk(P.kkmx_zipf_keys(Kb.ptr, nk, 1.1, nb, 8))
    h = _timeit(K, timer, lambda: chk(P.kkmx_histogram(Kb.ptr, nk, Bb.ptr, nb, 3, 32768)), reps=5)
    h2 = _timeit(K, timer, lambda: chk(P.kkmx_histogram(Kb.ptr, nk, Bb.ptr, nb, 2, 32768)), reps=5)
    h1 = _timeit(K, timer, lambda: chk(P.kkmx_histogram(Kb.ptr, nk, Bb.ptr, nb, 1, 32768)), reps=5)
    hsv = _timeit(K, timer, lambda: chk(P.kkmx_scatter_view_histogram(Kb.ptr, nk, Bb.ptr, nb, 0)), reps=2, warm=1)
    hdup = _timeit(K, timer, lambda: chk(P.kkmx_scatter_view_histogram(Kb.ptr, nk, Bb.ptr, nb, 1)), reps=5)
    del Kb, Bb
    ns, nbins = 1 << 26, 1 << 20
    Vb, Ib, Bb = K.DeviceBuffer(ns * 8), K.DeviceBuffer(ns * 4), K.DeviceBuffer(nbins * 8)
    chk(P.kkmx_fill_uniform_f64(Vb.ptr, ns, 9, 0.0, 1.0))
    chk(P.kkmx_fill_uniform_i32(Ib.ptr, ns, 9 ^ 0xABCDEF, 0, nbins - 1))
    sc = _timeit(K, timer, lambda: chk(P.kkmx_scatter_add_f64(Vb.ptr, Ib.ptr, ns, Bb.ptr, nbins)), reps=5)
    ssv = _timeit(K, timer, lambda: chk(P.kkmx_scatter_view_f64(Vb.ptr, Ib.ptr, ns, Bb.ptr, nbins, 0)), reps=5)
    del Vb, Ib, Bb
    return {
        "gups_2^30_table_2^28_updates": {"ms": round(g[0], 3), "gups": round(U / g[0] / 1e6, 3)},
        "histogram_zipf_2^28_keys": {"ms": round(h[0], 3), "gbps": round(4 * nk / h[0] / 1e6, 1),
                                     "keys_per_s": round(nk / h[0] * 1e3, 1),
                                     "kernel": "every bin in LDS: the 256 hottest as 16 replicated 32-bit counters, "
                                               "the rest as 16-bit halves with exact carry-out; 16-byte key loads; "
                                               "team rows summed by a second kernel (no device-scope atomics)"},
        "histogram_zipf_2^28_keys_global_atomic_flush": {"ms": round(h2[0], 3), "gbps": round(4 * nk / h2[0] / 1e6, 1)},
        "histogram_zipf_2^28_keys_one_key_per_load": {"ms": round(h1[0], 3), "gbps": round(4 * nk / h1[0] / 1e6, 1)},
        "scatter_f64_2^26_into_2^20": {"ms": round(sc[0], 3), "updates_per_s": round(ns / sc[0] * 1e3, 1),
                                       "gbps": round(12 * ns / sc[0] / 1e6, 1)},
        "scatter_view_f64_2^26_into_2^20": {"ms": round(ssv[0], 3), "updates_per_s": round(ns / ssv[0] * 1e3, 1)},
        "histogram_scatter_view_atomic_2^28_keys": {"ms": round(hsv[0], 3),
                                                    "keys_per_s": round(nk / hsv[0] * 1e3, 1)},
        "histogram_scatter_view_duplicated_2^28_keys": {"ms": round(hdup[0], 3),
                                                        "keys_per_s": round(nk / hdup[0] * 1e3, 1),
                                                        "gbps": round(4 * nk / hdup[0] / 1e6, 1)},
    }


def leg_batched(K, P, timer, replays=1000, plain_steps=5):
    """4096 uneven reduce/scan ops per step (seed 10): per-call launches,
    the persistent fused kernel (direct and replayed from a hipGraph) and,
    when a trained model is present, the learned tuner loaded as a Kokkos
    Tools library: per-call launches with tuned launch shapes, the fused
    kernel with the tuned tile shape, and a batch of small ops (2^4..2^12
    elements) with the default against the tuned tile shape."""
    from kokkos_amd import batch as B
    from kokkos_amd import workloads as W

    wl = W.batch_workload(4096, 10)
    nbytes = W.batch_bytes(wl)
    run = B.BatchRun(wl)
    out = {"ops_per_step": 4096, "elements_per_step": int(wl["n"].sum()), "bytes_per_step": int(nbytes),
           "tiles": int(run.tiles)}

    def per_step(mode, steps):
        run.run(mode, 1)
        K.fence()
        timer.start()
        run.run(mode, steps)
        return timer.stop() / steps

    ms = per_step(B.MODE_FUSED_GRAPH, replays)
    gi, gf = (a.copy() for a in run.results())
    out["fused_graph"] = {"ms_per_step": round(ms, 4), "ops_per_s": round(4096 / ms * 1e3, 1),
                          "gbps": round(nbytes / ms / 1e6, 1), "replays": replays, "tile_shape": run.shape}
    ms = per_step(B.MODE_FUSED, 50)
    out["fused"] = {"ms_per_step": round(ms, 4), "ops_per_s": round(4096 / ms * 1e3, 1)}
    ms = per_step(B.MODE_PLAIN, plain_steps)
    pi, pf = (a.copy() for a in run.results())
    out["plain"] = {"ms_per_step": round(ms, 3), "ops_per_s": round(4096 / ms * 1e3, 1)}
    isint = wl["type"] != W.T_F64
    out["int_results_identical"] = bool(np.array_equal(gi[isint], pi[isint]))
    ref = np.abs(pf[~isint])
    out["f64_max_rel_diff_fused_vs_plain"] = float(np.max(np.abs(gf[~isint] - pf[~isint]) / np.maximum(ref, 1e-300)))
    ms = per_step(B.MODE_PLAIN_GRAPH, plain_steps)
    out["plain_graph"] = {"ms_per_step": round(ms, 3), "ops_per_s": round(4096 / ms * 1e3, 1)}
    model = ROOT / "kokkos_amd" / "tuner" / "mlp_gfx950.bin"
    if model.exists():
        from kokkos_amd import tuner as T

        lib = str(K.lib_path("kkmx_tuner")).encode()
        K.check(P.kkmx_tools_load(lib), P)
        try:
            T.load_into_tool(model)
            ms = per_step(B.MODE_PLAIN, plain_steps)
            ti, tf = run.results()
            out["plain_tuned"] = {"ms_per_step": round(ms, 3), "ops_per_s": round(4096 / ms * 1e3, 1),
                                  "tool": T.tool_stats(),
                                  "int_results_identical": bool(np.array_equal(ti[isint], pi[isint]))}
            # the fused kernel with its tile shape chosen by the learned tuner
            # at plan time (the default shape is the heuristic)
            run.close()
            run = B.BatchRun(wl)
            ms = per_step(B.MODE_FUSED_GRAPH, replays)
            ui, uf = run.results()
            out["fused_graph_tuned"] = {
                "ms_per_step": round(ms, 4), "ops_per_s": round(4096 / ms * 1e3, 1), "tile_shape": run.shape,
                "default_tile_shape": out["fused_graph"]["tile_shape"],
                "int_results_identical": bool(np.array_equal(ui[isint], pi[isint])),
                "f64_max_rel_diff_vs_plain": float(np.max(np.abs(uf[~isint] - pf[~isint]) / np.maximum(ref, 1e-300)))}
            # a batch of small ops (4096 ops of 2^4..2^12 elements): default
            # tile shape against the tuner's choice
            run.close()
            small = W.batch_workload(4096, 11, 4.0, 12.0)
            sm = {}
            for name, shape in (("default", None), ("tuned", -1)):
                if shape is None:
                    K.check(P.kkmx_tools_unload(), P)
                else:
                    K.check(P.kkmx_tools_load(lib), P)
                    T.load_into_tool(model)
                run = B.BatchRun(small)
                ms = per_step(B.MODE_FUSED_GRAPH, replays)
                sm[name] = {"ms_per_step": round(ms, 4), "ops_per_s": round(4096 / ms * 1e3, 1), "tile_shape": run.shape}
                if name == "default":
                    si = run.results()[0].copy()
                else:
                    sm["int_results_identical"] = bool(np.array_equal(run.results()[0][small["type"] != W.T_F64],
                                                                      si[small["type"] != W.T_F64]))
                run.close()
            sm["speedup_tuned_vs_default"] = round(sm["tuned"]["ops_per_s"] / sm["default"]["ops_per_s"], 3)
            out["small_ops_fused_graph"] = sm
            run = None
        finally:
            K.check(P.kkmx_tools_unload(), P)
    out["speedup_fused_graph_vs_plain"] = round(out["fused_graph"]["ops_per_s"] / out["plain"]["ops_per_s"], 2)
    if run is not None:
        run.close()
    if model.exists():
        from kokkos_amd import tuner as T

        rep = json.loads(T.MODEL_META.read_text()).get("report", {}) if T.MODEL_META.exists() else {}
        out["tuner"] = {
            "batch_4096_ops": tuner_in_situ_batch(K, P, timer, wl),
            "batch_4096_small_ops": tuner_in_situ_batch(K, P, timer, W.batch_workload(4096, 11, 4.0, 12.0)),
            "offline_heldout_regret": rep.get("test"),
        }
    return out


def tuner_in_situ_range(K, P, calls, patterns, n, reps=5):
    """Regret of the learned tuner and of the heuristic against exhaustive
    search on the headline's own dispatches: each dispatch is timed under
    the 16 launch-shape candidates (fixed with kkmx_set_range_config), under
    the heuristic (no override) and with the tuning tool loaded; the tool's
    choice is read back from it."""
    import ctypes as C

    from kokkos_amd import tuner as T

    model = ROOT / "kokkos_amd" / "tuner" / "mlp_gfx950.bin"
    if not model.exists():
        return None
    timer = K.EventTimer()
    chk = lambda rc: K.check(rc, P)  # noqa: E731
    out = {}
    for (name, _), f, pat in zip(OPS, calls, patterns):
        t_cand = []
        for c in range(16):
            b, it, mb = C.c_int(), C.c_int(), C.c_int()
            chk(P.kkmx_range_candidate(pat, c, n, C.byref(b), C.byref(it), C.byref(mb)))
            chk(P.kkmx_set_range_config(pat, b.value, it.value, mb.value))
            t_cand.append(_timeit(K, timer, lambda: chk(f()), reps=reps, warm=1)[0])
        chk(P.kkmx_set_range_config(pat, -1, 0, 0))
        t_heur = _timeit(K, timer, lambda: chk(f()), reps=reps, warm=1)[0]
        chk(P.kkmx_tools_load(str(K.lib_path("kkmx_tuner")).encode()))
        try:
            T.load_into_tool(model)
            chk(f())
            K.fence()
            ans = T.last_answer()
            t_tool = _timeit(K, timer, lambda: chk(f()), reps=reps, warm=1)[0]
        finally:
            chk(P.kkmx_tools_unload())
        best = min(t_cand)
        chosen = ans[2] if ans and ans[0] == 0 else None
        heur_c = ans[1] if ans and ans[0] == 0 else None
        out[name] = {
            "best_candidate": int(np.argmin(t_cand)), "best_ms": round(best, 4),
            "heuristic_candidate": heur_c, "heuristic_ms": round(t_heur, 4),
            "tuner_candidate": chosen, "tuner_ms": round(t_tool, 4),
            "regret_tuner": round(t_tool / best - 1, 4), "regret_heuristic": round(t_heur / best - 1, 4),
        }
    rt = [v["regret_tuner"] for v in out.values()]
    rh = [v["regret_heuristic"] for v in out.values()]
    return {"dispatches": out, "mean_regret_tuner": round(float(np.mean(rt)), 4),
            "mean_regret_heuristic": round(float(np.mean(rh)), 4)}


def tuner_in_situ_batch(K, P, timer, wl, replays=100):
    """Exhaustive search over the fused batch kernel's 8 tile shapes for one
    workload (graph replays), against the heuristic's shape (5) and the shape
    the learned tuner picks at plan time."""
    from kokkos_amd import batch as B
    from kokkos_amd import tuner as T

    def step_ms(run):
        run.run(B.MODE_FUSED_GRAPH, 1)
        K.fence()
        timer.start()
        run.run(B.MODE_FUSED_GRAPH, replays)
        return timer.stop() / replays

    t = []
    for shape in range(8):
        run = B.BatchRun(wl, shape)
        t.append(step_ms(run))
        run.close()
    K.check(P.kkmx_tools_load(str(K.lib_path("kkmx_tuner")).encode()), P)
    try:
        T.load_into_tool(ROOT / "kokkos_amd" / "tuner" / "mlp_gfx950.bin")
        run = B.BatchRun(wl)
        chosen = run.shape
        t_tool = step_ms(run)
        run.close()
    finally:
        K.check(P.kkmx_tools_unload(), P)
    best = min(t)
    return {"ms_per_shape": [round(x, 4) for x in t], "best_shape": int(np.argmin(t)), "heuristic_shape": 5,
            "tuner_shape": chosen, "tuner_ms": round(t_tool, 4),
            "regret_tuner": round(t_tool / best - 1, 4), "regret_heuristic": round(t[5] / best - 1, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--only", choices=["small_reduce_scan", "mdrange_team", "atomics", "deep_copy", "std_algorithms",
                                       "sort", "batched"],
                    help="run one secondary workload alone and print its JSON (for profiling); no headline line")
    ap.add_argument("--scan-shape", type=int, nargs=2, metavar=("BLOCK", "ITEMS"),
                    help="time only the 2^28 int32 inclusive scan under this launch shape (for profiling); no headline line")
    ap.add_argument("--scan-body", type=int, default=0, choices=[0, 1],
                    help="with --scan-shape: the scan tile body (0 keeps wave scans, 1 keeps contributions)")
    ap.add_argument("--scan-traits", type=int, default=1, choices=[0, 1, 2],
                    help="with --scan-shape: the input View's traits (0 plain, 1 RandomAccess, 2 Restrict)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
            "the multi-GPU sharding is a design note, not a run",
            file=sys.stderr,
        )
        sys.exit(2)

    import numpy as np

    import kokkos_amd as K

    P = K.init(0)
    if args.scan_shape:
        S, O = K.DeviceBuffer(N * 4), K.DeviceBuffer(N * 4)
        K.check(P.kkmx_fill_uniform_i32(S.ptr, N, 4, 0, 7), P)
        K.check(P.kkmx_set_range_config(2, args.scan_shape[0], args.scan_shape[1], 0), P)
        K.check(P.kkmx_set_scan_body(args.scan_body), P)
        fn = lambda: K.check(P.kkmx_scan_incl_i32_traits(S.ptr, O.ptr, N, args.scan_traits), P)  # noqa: E731
        t = _timeit(K, K.EventTimer(), fn, reps=3, warm=1)
        print(json.dumps({"scan_incl_i32_2^28": {"block": args.scan_shape[0], "items": args.scan_shape[1],
                                                 "tile_body": args.scan_body, "input_traits": args.scan_traits,
                                                 "ms": round(t[0], 4), "gbps": round(8 * N / t[0] / 1e6, 1)}}))
        return
    if args.only:
        legs = {"small_reduce_scan": leg_small, "mdrange_team": leg_md_team, "atomics": leg_atomics,
                "deep_copy": leg_deep_copy, "std_algorithms": leg_std_algorithms, "sort": leg_sort,
                "batched": leg_batched}
        print(json.dumps({args.only: legs[args.only](K, P, K.EventTimer())}), flush=True)
        return
    props = K.device_props()
    stream = P.kkmx_stream()
    rt = K.load("kkmx_rt")

    a, b, c = K.DeviceBuffer(N * 8), K.DeviceBuffer(N * 8), K.DeviceBuffer(N * 8)
    x = K.DeviceBuffer(N * 8)
    s, so = K.DeviceBuffer(N * 4), K.DeviceBuffer(N * 4)
    res = K.DeviceBuffer(64)
    chk = lambda rc: K.check(rc, P)  # noqa: E731
    chk(P.kkmx_fill_uniform_f64(a.ptr, N, 2, -1.0, 1.0))
    chk(P.kkmx_fill_uniform_f64(b.ptr, N, 3, -1.0, 1.0))
    chk(P.kkmx_fill_uniform_f64(c.ptr, N, 4, -1.0, 1.0))
    chk(P.kkmx_fill_uniform_i64(x.ptr, N, 3, -(1 << 40), 1 << 40))
    chk(P.kkmx_fill_uniform_i32(s.ptr, N, 4, 0, 7))
    K.fence()

    calls = [
        lambda: P.kkmx_stream_copy(a.ptr, c.ptr, N),
        lambda: P.kkmx_stream_scale(b.ptr, c.ptr, N, 3.0),
        lambda: P.kkmx_stream_add(a.ptr, b.ptr, c.ptr, N),
        lambda: P.kkmx_stream_triad(a.ptr, b.ptr, c.ptr, N, 3.0),
        lambda: P.kkmx_reduce_i64(x.ptr, N, 2, res.ptr),
        lambda: P.kkmx_reduce_i64(x.ptr, N, 1, res.ptr + 8),
        lambda: P.kkmx_reduce_i64(x.ptr, N, 0, res.ptr + 16),
        lambda: P.kkmx_scan_incl_i32(s.ptr, so.ptr, N),
    ]
    import ctypes as C

    events = []
    for _ in range((len(calls) + 1) * args.steps):
        e = C.c_void_p()
        K.check(rt.mx_event_create(C.byref(e)))
        events.append(e)

    # The scan's tile body is chosen per functor type from the runtime's own
    # timings of its first six large launches (both bodies give identical
    # results); let that happen before the timed steps, as it does early in
    # any long-running program.
    for _ in range(8):
        chk(P.kkmx_scan_incl_i32(s.ptr, so.ptr, N))
        K.fence()
    for _ in range(args.warmup):
        for f in calls:
            chk(f())
    K.fence()

    t0 = time.perf_counter()
    ei = 0
    for _ in range(args.steps):
        K.check(rt.mx_event_record(events[ei], stream))
        ei += 1
        for f in calls:
            chk(f())
            K.check(rt.mx_event_record(events[ei], stream))
            ei += 1
    K.fence()
    wall = time.perf_counter() - t0

    per_op_ms = np.zeros((args.steps, len(calls)))
    ei = 0
    for st in range(args.steps):
        base = st * (len(calls) + 1)
        for k in range(len(calls)):
            ms = C.c_float()
            K.check(rt.mx_event_elapsed_ms(C.byref(ms), events[base + k], events[base + k + 1]))
            per_op_ms[st, k] = ms.value
    step_ms = per_op_ms.sum(axis=1)
    bytes_per_step = sum(bpe for _, bpe in OPS) * N
    total_s = step_ms.sum() * 1e-3
    value = bytes_per_step * args.steps / total_s / 1e9
    avg_ms = per_op_ms.mean(axis=0)
    op_gbps = {name: round(bpe * N / (avg_ms[k] * 1e-3) / 1e9, 1) for k, (name, bpe) in enumerate(OPS)}

    # correctness of the timed dispatches: one more step, every output checked
    checks = verify_step(K, P, calls, a, b, c, x, s, so, res)
    ok = all(checks.values())

    scan_views = scan_view_traits(K, P, s, so, N)
    tuner_range = None if args.no_extra else tuner_in_situ_range(K, P, calls, [0, 0, 0, 0, 1, 1, 1, 2], N)
    peak = measured_copy_peak(K, a.ptr, b.ptr, N * 8, K.EventTimer())
    del a, b, c, x, s, so

    extra = {}
    if not args.no_extra:
        timer = K.EventTimer()
        for name, fn in [("small_reduce_scan", lambda: leg_small(K, P, timer)),
                         ("mdrange_team", lambda: leg_md_team(K, P, timer)),
                         ("atomics", lambda: leg_atomics(K, P, timer)),
                         ("deep_copy", lambda: leg_deep_copy(K, P, timer)),
                         ("std_algorithms", lambda: leg_std_algorithms(K, P, timer)),
                         ("sort", lambda: leg_sort(K, P, timer)),
                         ("batched", lambda: leg_batched(K, P, timer))]:
            t_leg = time.perf_counter()
            extra[name] = fn()
            print(f"bench: {name} done in {time.perf_counter() - t_leg:.1f} s", file=sys.stderr, flush=True)

    # dominant kernel: the dispatch with the largest share of the step
    dom = int(np.argmax(avg_ms))
    dom_name, dom_bpe = OPS[dom]
    dom_bytes = dom_bpe * N
    achieved = dom_bytes / (avg_ms[dom] * 1e-3) / 1e9

    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(dom_name)
        except Exception:
            traffic = None

    line = {
        "metric": "HBM GB/s (% peak) reduce/scan/triad; batched ops/s learned-tuned vs plain",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(step_ms.mean()), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64,int64,int32",
        "data": "synthetic (device-generated splitmix64 streams, seeds 2/3/4)",
        "config": {
            "workload": "STREAM copy/scale/add/triad f64 + reduce max/min/sum int64 + inclusive scan int32, N=2^28 each",
            "n": N,
            "global_batch": 1,
            "seq_len": N,
            "parallelism": "single GPU",
        },
        "per_op_gbps": op_gbps,
        "pct_of_spec_peak": {k: round(100 * v / HBM_PEAK_GBPS, 1) for k, v in op_gbps.items()},
        "measured_copy_peak_gbps": round(peak, 1),
        "pct_of_measured_peak": {k: round(100 * v / peak, 1) for k, v in op_gbps.items()},
        "results_bit_exact": bool(ok),
        "results_checked": checks,
        "scan_incl_i32_by_input_view": scan_views,
        "wall_s": round(wall, 3),
        "device": {"arch": props["arch"], "cus": props["cu_count"]},
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_source": "profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel, "
            "scaled by the counters' readings for a plain copy of known size (not measured in this run)",
        },
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    if extra:
        line["workloads"] = extra
        bt = extra["batched"]
        if tuner_range is not None and "tuner" in bt:
            bt["tuner"]["headline_range_dispatches"] = tuner_range
        line["batched_ops_per_s"] = {
            "fused_graph": bt["fused_graph"]["ops_per_s"],
            "plain": bt["plain"]["ops_per_s"],
            "plain_tuned": bt.get("plain_tuned", {}).get("ops_per_s"),
            "fused_graph_tuned": bt.get("fused_graph_tuned", {}).get("ops_per_s"),
            "speedup_vs_plain": bt["speedup_fused_graph_vs_plain"],
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
