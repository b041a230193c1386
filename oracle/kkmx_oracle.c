/*
 * kkmx_oracle.c -- TEST INFRASTRUCTURE ONLY.  A CPU restatement of what the
 * reference kokkos Serial and OpenMP backends compute on the hot path, used
 * as the parity checker by tests/, by __graft_entry__.smoke() and as the
 * CPU baseline in bench.py.  Nothing in the product (kokkos_amd/) links,
 * loads or calls this code.
 *
 * Semantics restated (paths relative to the kokkos source tree):
 *   reduce  : init, left fold over the range, final
 *             (core/src/Serial/Kokkos_Serial_Parallel_Range.hpp:86-140);
 *             OpenMP: each thread folds a contiguous slice, the partials are
 *             joined in thread order
 *             (core/src/OpenMP/Kokkos_OpenMP_Parallel_Reduce.hpp:113-160,
 *              core/src/impl/Kokkos_HostThreadTeam.hpp:357-389).
 *   scan    : update = identity; for i: f(i, update, final=true)
 *             (core/src/Serial/Kokkos_Serial_Parallel_Range.hpp:178-226);
 *             OpenMP: per-thread totals, serial combine, second pass
 *             (core/src/OpenMP/Kokkos_OpenMP_Parallel_Scan.hpp:87-140).
 *   STREAM  : benchmarks/stream/stream-kokkos.cpp:36-77 element-wise bodies.
 *   xorshift64*: algorithms/src/Kokkos_Random.hpp:788-808 (Random_XorShift64,
 *             state 0 replaced by 1318319, urand64 = state * M - 1).
 *
 * Floating point: compiled with -ffp-contract=off; where the device kernels
 * contract a*b+c into one fused multiply-add (hipcc's default for device
 * code), the oracle calls fma() explicitly so the comparison is bit-exact,
 * and a second, unfused form is kept for the 1-ulp cross-check.
 *
 * Parity pins (tests/test_oracle.py): closed forms from the reference's unit
 * tests (core/unit_test/TestParallelScanRangePolicy.hpp:42-88,
 * core/unit_test/TestReduce.hpp:27-140), Python's math.fsum (Shewchuk's exact
 * summation, an independent implementation) for the exact sum, brute force on
 * tiny inputs, and published known-answer outputs where they exist.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------ generators */

/* SplitMix64 (Steele, Lea, Flood 2014): output i of the stream seeded with s. */
EXPORT uint64_t or_splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t mulhi64(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

EXPORT void or_fill_uniform_f64(double* x, int64_t n, uint64_t seed, double lo, double hi) {
  const double w = hi - lo;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double u = (double)(or_splitmix64(seed, (uint64_t)i) >> 11) * 0x1.0p-53;
    x[i] = fma(w, u, lo);
  }
}

EXPORT void or_fill_uniform_i64(int64_t* x, int64_t n, uint64_t seed, int64_t lo, int64_t hi) {
  const uint64_t range = (uint64_t)hi - (uint64_t)lo + 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t r = or_splitmix64(seed, (uint64_t)i);
    x[i] = (int64_t)((uint64_t)lo + (range ? mulhi64(r, range) : r));
  }
}

EXPORT void or_fill_uniform_i32(int32_t* x, int64_t n, uint64_t seed, int32_t lo, int32_t hi) {
  const uint64_t range = (uint64_t)((int64_t)hi - lo + 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t r = or_splitmix64(seed, (uint64_t)i);
    x[i] = (int32_t)(uint32_t)((uint64_t)(int64_t)lo + mulhi64(r, range));
  }
}

/* Random_XorShift64::urand64 sequence starting from `seed`. */
EXPORT void or_xorshift64star(uint64_t* out, int64_t n, uint64_t seed) {
  uint64_t s = seed == 0 ? 1318319ull : seed;
  for (int64_t i = 0; i < n; ++i) {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    out[i] = s * 2685821657736338717ull - 1;
  }
}

/* ------------------------------------------------------------ checksum */
/* Position-aware 64-bit checksum of a byte buffer (length multiple of 8):
 * sum over words of mix(word ^ (index * golden)).  Order-independent to
 * compute (so it parallelises) but sensitive to any changed, moved or
 * swapped word. */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
EXPORT uint64_t or_checksum64(const void* p, int64_t bytes) {
  const uint64_t* w = (const uint64_t*)p;
  const int64_t nw = bytes / 8;
  uint64_t h = 0;
#pragma omp parallel for schedule(static) reduction(+ : h)
  for (int64_t i = 0; i < nw; ++i) h += mix64(w[i] ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull));
  int64_t rem = bytes - nw * 8;
  if (rem) {
    uint64_t t = 0;
    memcpy(&t, (const char*)p + nw * 8, (size_t)rem);
    h += mix64(t ^ ((uint64_t)nw * 0x9E3779B97F4A7C15ull) ^ 0x5bd1e995ull);
  }
  return h;
}

/* ------------------------------------------------------------ FP64 sums */

/* Serial semantics: left fold from 0.0. */
EXPORT double or_sum_f64_serial(const double* x, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  return s;
}

/* OpenMP semantics: contiguous slice per thread, partials joined in order. */
EXPORT double or_sum_f64_openmp(const double* x, int64_t n, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  double* part = (double*)calloc((size_t)nthreads, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    const int t = omp_get_thread_num();
    const int T = omp_get_num_threads();
    const int64_t b = n * t / T, e = n * (t + 1) / T;
    double s = 0.0;
    for (int64_t i = b; i < e; ++i) s += x[i];
    part[t] = s;
  }
  double s = 0.0;
  for (int t = 0; t < nthreads; ++t) s += part[t];
  free(part);
  return s;
}

/* Neumaier's compensated sum (Neumaier 1974, improved Kahan-Babuska). */
EXPORT double or_sum_f64_neumaier(const double* x, int64_t n) {
  double s = 0.0, c = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double t = s + x[i];
    if (fabs(s) >= fabs(x[i]))
      c += (s - t) + x[i];
    else
      c += (x[i] - t) + s;
    s = t;
  }
  return s + c;
}

/* Correctly rounded sum: Shewchuk's non-overlapping partials ("Adaptive
 * Precision Floating-Point Arithmetic", 1997) with the half-way correction of
 * the final rounding, as published in Python's math.fsum. */
EXPORT double or_sum_f64_exact(const double* x, int64_t n) {
  int cap = 64, m = 0;
  double* p = (double*)malloc(sizeof(double) * (size_t)cap);
  for (int64_t k = 0; k < n; ++k) {
    double v = x[k];
    int i = 0;
    for (int j = 0; j < m; ++j) {
      double y = p[j];
      if (fabs(v) < fabs(y)) {
        double t = v;
        v = y;
        y = t;
      }
      const double hi = v + y;
      const double lo = y - (hi - v);
      if (lo != 0.0) p[i++] = lo;
      v = hi;
    }
    if (i >= cap) {
      cap *= 2;
      p = (double*)realloc(p, sizeof(double) * (size_t)cap);
    }
    p[i++] = v;
    m = i;
  }
  double hi = 0.0;
  if (m > 0) {
    int j = m - 1;
    hi = p[j];
    double lo = 0.0;
    while (j > 0) {
      const double v = hi;
      const double y = p[--j];
      hi = v + y;
      const double yr = hi - v;
      lo = y - yr;
      if (lo != 0.0) break;
    }
    /* Round half to even correctly: if the remainder and the next partial
     * have the same sign, the true value is beyond the half-way point. */
    if (j > 0 && ((lo < 0.0 && p[j - 1] < 0.0) || (lo > 0.0 && p[j - 1] > 0.0))) {
      const double y = lo * 2.0;
      const double v = hi + y;
      const double yr = v - hi;
      if (y == yr) hi = v;
    }
  }
  free(p);
  return hi;
}

/* ------------------------------------------------------------ integer reductions */
/* op 0 = Sum (two's-complement wrap), 1 = Min, 2 = Max. */
EXPORT int64_t or_reduce_i64(const int64_t* x, int64_t n, int op, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  uint64_t s = 0;
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  if (op == 0) {
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; ++i) s += (uint64_t)x[i];
    return (int64_t)s;
  } else if (op == 1) {
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min : mn)
    for (int64_t i = 0; i < n; ++i) mn = x[i] < mn ? x[i] : mn;
    return mn;
  }
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(max : mx)
  for (int64_t i = 0; i < n; ++i) mx = x[i] > mx ? x[i] : mx;
  return mx;
}

EXPORT int64_t or_reduce_i32(const int32_t* x, int64_t n, int op, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  uint64_t s = 0;
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  if (op == 0) {
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; ++i) s += (uint64_t)(int64_t)x[i];
    return (int64_t)s;
  } else if (op == 1) {
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min : mn)
    for (int64_t i = 0; i < n; ++i) mn = x[i] < mn ? x[i] : mn;
    return mn;
  }
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(max : mx)
  for (int64_t i = 0; i < n; ++i) mx = x[i] > mx ? x[i] : mx;
  return mx;
}

/* ------------------------------------------------------------ scans */
/* Two-pass OpenMP scan: per-thread totals of contiguous slices, serial
 * exclusive combine in thread order, second pass writes.  With one thread it
 * is the Serial scan. */
#define OR_SCAN_BODY(IN_T, ACC_T, OUT_T, INCLUSIVE)                          \
  if (nthreads < 1) nthreads = 1;                                            \
  ACC_T* tot = (ACC_T*)calloc((size_t)nthreads + 1, sizeof(ACC_T));         \
  _Pragma("omp parallel num_threads(nthreads)") {                            \
    const int t = omp_get_thread_num();                                      \
    const int T = omp_get_num_threads();                                     \
    const int64_t b = n * t / T, e = n * (t + 1) / T;                        \
    ACC_T s = 0;                                                             \
    for (int64_t i = b; i < e; ++i) s += (ACC_T)in[i];                       \
    tot[t + 1] = s;                                                          \
    _Pragma("omp barrier") _Pragma("omp single") {                           \
      for (int k = 1; k <= T; ++k) tot[k] += tot[k - 1];                     \
    }                                                                        \
    ACC_T u = tot[t];                                                        \
    for (int64_t i = b; i < e; ++i) {                                        \
      if (!(INCLUSIVE)) out[i] = (OUT_T)u;                                   \
      u += (ACC_T)in[i];                                                     \
      if (INCLUSIVE) out[i] = (OUT_T)u;                                      \
    }                                                                        \
  }                                                                          \
  if (total) *total = tot[nthreads];                                         \
  free(tot);

EXPORT void or_scan_excl_i32_i64(const int32_t* in, int64_t* out, int64_t n, int nthreads, int64_t* total) {
  OR_SCAN_BODY(int32_t, int64_t, int64_t, 0)
}
EXPORT void or_scan_incl_i32(const int32_t* in, int32_t* out, int64_t n, int nthreads, int32_t* total) {
  /* int32 accumulate: wraps like the device's int32 update. */
  OR_SCAN_BODY(int32_t, uint32_t, int32_t, 1)
}
EXPORT void or_scan_excl_i64(const int64_t* in, int64_t* out, int64_t n, int nthreads, int64_t* total) {
  OR_SCAN_BODY(int64_t, uint64_t, int64_t, 0)
}
EXPORT void or_scan_excl_f64(const double* in, double* out, int64_t n, int nthreads, double* total) {
  OR_SCAN_BODY(double, double, double, 0)
}

/* Exclusive prefix sums of doubles with a Neumaier-compensated running sum
 * (Neumaier 1974): out[i] is the compensated sum of in[0..i-1], accurate to
 * about one rounding, the reference value for a device scan's tolerance. */
EXPORT void or_scan_excl_f64_compensated(const double* in, double* out, int64_t n, double* total) {
  double s = 0.0, c = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    out[i] = s + c;
    const double x = in[i], t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  if (total) *total = s + c;
}

/* ------------------------------------------------------------ STREAM */
EXPORT void or_stream_copy(const double* a, double* b, int64_t n, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) b[i] = a[i];
}
EXPORT void or_stream_scale(double* b, const double* c, int64_t n, double s, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) b[i] = s * c[i];
}
EXPORT void or_stream_add(const double* a, const double* b, double* c, int64_t n, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) c[i] = a[i] + b[i];
}
/* fused: a = fma(s, c, b), the contraction the device code performs. */
EXPORT void or_stream_triad(double* a, const double* b, const double* c, int64_t n, double s, int fused,
                            int nthreads) {
  if (fused) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) a[i] = fma(s, c[i], b[i]);
  } else {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) a[i] = b[i] + s * c[i];
  }
}

EXPORT int or_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------ config 3 */
/* 7-point update over the interior of an (n+2)^3 grid, first index fastest
 * (LayoutLeft); neighbour sum in the order the product functor writes it:
 * (((((u[i-1]+u[i+1])+u[j-1])+u[j+1])+u[k-1])+u[k+1]).  out = c0*u + c1*s,
 * unfused (the device may fuse one product; the comparison is toleranced). */
EXPORT void or_stencil7(const double* u, double* v, int64_t n, double c0, double c1, int nthreads) {
  const int64_t m = n + 2, sj = m, sk = m * m;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t k = 1; k <= n; ++k)
    for (int64_t j = 1; j <= n; ++j)
      for (int64_t i = 1; i <= n; ++i) {
        const int64_t o = i + j * sj + k * sk;
        const double s = u[o - 1] + u[o + 1] + u[o - sj] + u[o + sj] + u[o - sk] + u[o + sk];
        v[o] = c0 * u[o] + c1 * s;
      }
}

/* Seeded CSR, the same formulas as the device generator. */
#define OR_COL_SALT 0x5bd1e9955bd1e995ull
#define OR_VAL_SALT 0xc2b2ae3d27d4eb4full
EXPORT int64_t or_csr_rowptr(int64_t* rp, int64_t nrows, int64_t lo, int64_t hi, uint64_t seed) {
  const uint64_t range = (uint64_t)(hi - lo + 1);
  int64_t acc = 0;
  for (int64_t r = 0; r < nrows; ++r) {
    rp[r] = acc;
    acc += lo + (int64_t)mulhi64(or_splitmix64(seed, (uint64_t)r), range);
  }
  rp[nrows] = acc;
  return acc;
}
EXPORT void or_csr_fill(int32_t* col, double* val, int64_t nnz, int64_t ncols, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nnz; ++k) {
    col[k] = (int32_t)mulhi64(or_splitmix64(seed ^ OR_COL_SALT, (uint64_t)k), (uint64_t)ncols);
    const double u = (double)(or_splitmix64(seed ^ OR_VAL_SALT, (uint64_t)k) >> 11) * 0x1.0p-53;
    val[k] = fma(2.0, u, -1.0);
  }
}
/* y = A x, each row summed left to right; also the per-row normwise scale
 * sum |a_ij x_j| used as the tolerance denominator. */
EXPORT void or_spmv_csr(const int64_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                        double* absrow, int64_t nrows, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t r = 0; r < nrows; ++r) {
    double s = 0.0, a = 0.0;
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
      s += val[k] * x[col[k]];
      a += fabs(val[k] * x[col[k]]);
    }
    y[r] = s;
    if (absrow) absrow[r] = a;
  }
}

/* ------------------------------------------------------------ config 4 */
/* GUPS: table[stream[2u] & mask] += stream[2u+1] (wrapping). */
EXPORT void or_gups(int64_t* table, int64_t table_size, const uint64_t* stream, int64_t nupdates, int nthreads) {
  const uint64_t mask = (uint64_t)table_size - 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t u = 0; u < nupdates; ++u)
    __atomic_fetch_add((uint64_t*)&table[stream[2 * u] & mask], stream[2 * u + 1], __ATOMIC_RELAXED);
}
/* Zipf(s) over [0, nbins): cdf accumulated left to right, normalised, last
 * entry forced to 1; key = first k with u < cdf[k]. */
EXPORT void or_zipf_cdf(double* cdf, double s, int32_t nbins) {
  double acc = 0.0;
  for (int32_t k = 0; k < nbins; ++k) {
    acc += pow((double)(k + 1), -s);
    cdf[k] = acc;
  }
  for (int32_t k = 0; k < nbins; ++k) cdf[k] /= acc;
  cdf[nbins - 1] = 1.0;
}
EXPORT void or_zipf_keys(int32_t* keys, int64_t n, const double* cdf, int32_t nbins, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double u = (double)(or_splitmix64(seed, (uint64_t)i) >> 11) * 0x1.0p-53;
    int32_t lo = 0, hi = nbins - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (u < cdf[mid]) hi = mid;
      else lo = mid + 1;
    }
    keys[i] = lo;
  }
}
EXPORT void or_histogram(const int32_t* keys, int64_t n, int32_t* bins, int32_t nbins) {
  memset(bins, 0, sizeof(int32_t) * (size_t)nbins);
  for (int64_t i = 0; i < n; ++i) bins[keys[i]]++;
}
/* Per-bin compensated (Neumaier) sums of the scattered values, in index
 * order, and per-bin sum of |v| for relative errors. */
EXPORT void or_scatter_f64(const double* vals, const int32_t* idx, int64_t n, double* bins, double* babs,
                           int32_t nbins) {
  double* c = (double*)calloc((size_t)nbins, sizeof(double));
  memset(bins, 0, sizeof(double) * (size_t)nbins);
  if (babs) memset(babs, 0, sizeof(double) * (size_t)nbins);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t b = idx[i];
    const double x = vals[i], s = bins[b], t = s + x;
    if (fabs(s) >= fabs(x)) c[b] += (s - t) + x;
    else c[b] += (x - t) + s;
    bins[b] = t;
    if (babs) babs[b] += fabs(x);
  }
  for (int32_t b = 0; b < nbins; ++b) bins[b] += c[b];
  free(c);
}
