// C-ABI entry points of the hot path (declared in include/kkmx_paths.h).
//
// Every function here is plain Kokkos user code: unmanaged Views over the
// caller's device pointers, functors with the reference benchmarks' bodies,
// and parallel_for / parallel_reduce / parallel_scan on the default MI355X
// instance.  Nothing here calls HIP directly.
#include <Kokkos_Core.hpp>

#include <string>

#include "kkmx_paths.h"

namespace {

thread_local std::string g_err;

template <class Fn>
int guarded(Fn&& fn) {
  try {
    if (!Kokkos::is_initialized()) Kokkos::initialize();
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int reject(const char* msg) {
  g_err = msg;
  return -1;
}

template <class T>
using DevView = Kokkos::View<T*, Kokkos::LayoutRight, Kokkos::MI355XSpace,
                             Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
template <class T>
using DevScalar = Kokkos::View<T, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;

// ---------------------------------------------------------------- seeded inputs
// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014): output i of the stream
// seeded with s is mix(s + (i + 1) * golden_gamma), so any index can be
// generated independently and the host oracle reproduces it exactly.
KOKKOS_INLINE_FUNCTION uint64_t splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
KOKKOS_INLINE_FUNCTION uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) >> 64);
#endif
}

struct FillF64 {
  DevView<double> x;
  uint64_t seed;
  double lo, width;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    const double u = static_cast<double>(splitmix64(seed, static_cast<uint64_t>(i)) >> 11) * 0x1.0p-53;
    x(i)           = fma(width, u, lo);
  }
};
template <class T>
struct FillInt {
  DevView<T> x;
  uint64_t seed;
  int64_t lo;
  uint64_t range;  // hi - lo + 1 (0 means the full 2^64 range)
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    const uint64_t r = splitmix64(seed, static_cast<uint64_t>(i));
    const uint64_t o = range ? mulhi64(r, range) : r;
    x(i)             = static_cast<T>(static_cast<uint64_t>(lo) + o);
  }
};

// ---------------------------------------------------------------- STREAM
// Bodies of benchmarks/stream/stream-kokkos.cpp:36-77, with the reference's
// RangePolicy<IndexType<int>>.
using StreamPolicy = Kokkos::RangePolicy<Kokkos::IndexType<int>>;

// ---------------------------------------------------------------- reductions
struct SumF64 {
  DevView<const double> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, double& s) const { s += x(i); }
};
template <class T>
struct SumInt {
  DevView<const T> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& s) const {
    // Two's-complement wrap is associative: accumulate unsigned to keep the
    // overflow defined and the result order-independent.
    s = static_cast<int64_t>(static_cast<uint64_t>(s) + static_cast<uint64_t>(static_cast<int64_t>(x(i))));
  }
};
template <class T>
struct MinInt {
  DevView<const T> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& m) const {
    const int64_t v = x(i);
    if (v < m) m = v;
  }
};
template <class T>
struct MaxInt {
  DevView<const T> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& m) const {
    const int64_t v = x(i);
    if (v > m) m = v;
  }
};
struct MinMaxSum {
  struct value_type {
    int64_t mn, mx, sum;
  };
  DevView<const int64_t> x;
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.mn  = Kokkos::reduction_identity<int64_t>::min();
    v.mx  = Kokkos::reduction_identity<int64_t>::max();
    v.sum = 0;
  }
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (s.mn < d.mn) d.mn = s.mn;
    if (s.mx > d.mx) d.mx = s.mx;
    d.sum = static_cast<int64_t>(static_cast<uint64_t>(d.sum) + static_cast<uint64_t>(s.sum));
  }
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, value_type& v) const {
    const int64_t e = x(i);
    if (e < v.mn) v.mn = e;
    if (e > v.mx) v.mx = e;
    v.sum = static_cast<int64_t>(static_cast<uint64_t>(v.sum) + static_cast<uint64_t>(e));
  }
};

// ---------------------------------------------------------------- scans
template <class In, class Acc>
struct ExclusiveScan {
  DevView<const In> in;
  DevView<Acc> out;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, Acc& update, const bool final) const {
    if (final) out(i) = update;
    update += static_cast<Acc>(in(i));
  }
};
template <class T>
struct InclusiveScan {
  DevView<const T> in;
  DevView<T> out;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, T& update, const bool final) const {
    update += in(i);
    if (final) out(i) = update;
  }
};

using Range64 = Kokkos::RangePolicy<Kokkos::IndexType<int64_t>>;

}  // namespace

namespace {
template <class T>
int reduce_int(const T* x, int64_t n, int op, int64_t* out) {
  if (n < 0 || (n > 0 && !x) || !out || op < 0 || op > 2) return reject("kkmx_reduce: bad arguments");
  return guarded([&] {
    DevScalar<int64_t> r(out);
    DevView<const T> X(x, n);
    if (op == 0) {
      Kokkos::parallel_reduce("kkmx::sum", Range64(0, n), SumInt<T>{X}, r);
    } else if (op == 1) {
      Kokkos::parallel_reduce("kkmx::min", Range64(0, n), MinInt<T>{X}, Kokkos::Min<int64_t, Kokkos::MI355XSpace>(r));
    } else {
      Kokkos::parallel_reduce("kkmx::max", Range64(0, n), MaxInt<T>{X}, Kokkos::Max<int64_t, Kokkos::MI355XSpace>(r));
    }
  });
}
}  // namespace

extern "C" {

int kkmx_init(int device) {
  try {
    if (!Kokkos::is_initialized()) Kokkos::initialize(Kokkos::InitializationSettings().set_device_id(device));
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
int kkmx_finalize(void) {
  try {
    if (Kokkos::is_initialized()) Kokkos::finalize();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
const char* kkmx_last_error(void) { return g_err.c_str(); }
void* kkmx_stream(void) {
  if (!Kokkos::is_initialized() && kkmx_init(0) != 0) return nullptr;
  return Kokkos::MI355X().stream();
}
int kkmx_fence(void) {
  return guarded([] { Kokkos::fence("kkmx_fence"); });
}

int kkmx_fill_uniform_f64(double* x, int64_t n, uint64_t seed, double lo, double hi) {
  if (n < 0 || (n > 0 && !x)) return reject("kkmx_fill_uniform_f64: bad arguments");
  return guarded([&] {
    Kokkos::parallel_for("kkmx::fill_f64", Range64(0, n), FillF64{DevView<double>(x, n), seed, lo, hi - lo});
  });
}
int kkmx_fill_uniform_i64(int64_t* x, int64_t n, uint64_t seed, int64_t lo, int64_t hi) {
  if (n < 0 || (n > 0 && !x) || hi < lo) return reject("kkmx_fill_uniform_i64: bad arguments");
  return guarded([&] {
    const uint64_t range = static_cast<uint64_t>(hi) - static_cast<uint64_t>(lo) + 1;
    Kokkos::parallel_for("kkmx::fill_i64", Range64(0, n), FillInt<int64_t>{DevView<int64_t>(x, n), seed, lo, range});
  });
}
int kkmx_fill_uniform_i32(int32_t* x, int64_t n, uint64_t seed, int32_t lo, int32_t hi) {
  if (n < 0 || (n > 0 && !x) || hi < lo) return reject("kkmx_fill_uniform_i32: bad arguments");
  return guarded([&] {
    const uint64_t range = static_cast<uint64_t>(static_cast<int64_t>(hi) - lo + 1);
    Kokkos::parallel_for("kkmx::fill_i32", Range64(0, n), FillInt<int32_t>{DevView<int32_t>(x, n), seed, lo, range});
  });
}

int kkmx_stream_set(double* a, int64_t n, double scalar) {
  if (n < 0 || n > 0x7fffffff) return reject("kkmx_stream_set: n must fit IndexType<int>");
  return guarded([&] {
    DevView<double> A(a, n);
    Kokkos::parallel_for("set", StreamPolicy(0, int(n)), KOKKOS_LAMBDA(const int i) { A[i] = scalar; });
  });
}
int kkmx_stream_copy(const double* a, double* b, int64_t n) {
  if (n < 0 || n > 0x7fffffff) return reject("kkmx_stream_copy: n must fit IndexType<int>");
  return guarded([&] {
    DevView<const double> A(a, n);
    DevView<double> B(b, n);
    Kokkos::parallel_for("copy", StreamPolicy(0, int(n)), KOKKOS_LAMBDA(const int i) { B[i] = A[i]; });
  });
}
int kkmx_stream_scale(double* b, const double* c, int64_t n, double scalar) {
  if (n < 0 || n > 0x7fffffff) return reject("kkmx_stream_scale: n must fit IndexType<int>");
  return guarded([&] {
    DevView<double> B(b, n);
    DevView<const double> C(c, n);
    Kokkos::parallel_for("scale", StreamPolicy(0, int(n)), KOKKOS_LAMBDA(const int i) { B[i] = scalar * C[i]; });
  });
}
int kkmx_stream_add(const double* a, const double* b, double* c, int64_t n) {
  if (n < 0 || n > 0x7fffffff) return reject("kkmx_stream_add: n must fit IndexType<int>");
  return guarded([&] {
    DevView<const double> A(a, n), B(b, n);
    DevView<double> C(c, n);
    Kokkos::parallel_for("add", StreamPolicy(0, int(n)), KOKKOS_LAMBDA(const int i) { C[i] = A[i] + B[i]; });
  });
}
int kkmx_stream_triad(double* a, const double* b, const double* c, int64_t n, double scalar) {
  if (n < 0 || n > 0x7fffffff) return reject("kkmx_stream_triad: n must fit IndexType<int>");
  return guarded([&] {
    DevView<double> A(a, n);
    DevView<const double> B(b, n), C(c, n);
    Kokkos::parallel_for("triad", StreamPolicy(0, int(n)),
                         KOKKOS_LAMBDA(const int i) { A[i] = B[i] + scalar * C[i]; });
  });
}

int kkmx_reduce_sum_f64(const double* x, int64_t n, double* out) {
  if (n < 0 || (n > 0 && !x) || !out) return reject("kkmx_reduce_sum_f64: bad arguments");
  return guarded([&] {
    DevScalar<double> r(out);
    Kokkos::parallel_reduce("kkmx::sum_f64", Range64(0, n), SumF64{DevView<const double>(x, n)}, r);
  });
}

int kkmx_reduce_i64(const int64_t* x, int64_t n, int op, int64_t* out) { return reduce_int(x, n, op, out); }
int kkmx_reduce_i32(const int32_t* x, int64_t n, int op, int64_t* out) { return reduce_int(x, n, op, out); }

int kkmx_reduce_minmaxsum_i64(const int64_t* x, int64_t n, int64_t* out) {
  if (n < 0 || (n > 0 && !x) || !out) return reject("kkmx_reduce_minmaxsum_i64: bad arguments");
  return guarded([&] {
    using VT = MinMaxSum::value_type;
    DevScalar<VT> r(reinterpret_cast<VT*>(out));
    Kokkos::parallel_reduce("kkmx::minmaxsum_i64", Range64(0, n), MinMaxSum{DevView<const int64_t>(x, n)}, r);
  });
}

int kkmx_scan_excl_i32_i64(const int32_t* in, int64_t* out, int64_t n, int64_t* total) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_excl_i32_i64: bad arguments");
  return guarded([&] {
    ExclusiveScan<int32_t, int64_t> f{DevView<const int32_t>(in, n), DevView<int64_t>(out, n)};
    if (total) {
      DevScalar<int64_t> t(total);
      Kokkos::parallel_scan("kkmx::scan_excl_i32_i64", Range64(0, n), f, t);
    } else {
      Kokkos::parallel_scan("kkmx::scan_excl_i32_i64", Range64(0, n), f);
    }
  });
}
int kkmx_scan_incl_i32(const int32_t* in, int32_t* out, int64_t n) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_incl_i32: bad arguments");
  return guarded([&] {
    Kokkos::parallel_scan("kkmx::scan_incl_i32", Range64(0, n),
                          InclusiveScan<int32_t>{DevView<const int32_t>(in, n), DevView<int32_t>(out, n)});
  });
}
int kkmx_scan_excl_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* total) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_excl_i64: bad arguments");
  return guarded([&] {
    ExclusiveScan<int64_t, int64_t> f{DevView<const int64_t>(in, n), DevView<int64_t>(out, n)};
    if (total) {
      DevScalar<int64_t> t(total);
      Kokkos::parallel_scan("kkmx::scan_excl_i64", Range64(0, n), f, t);
    } else {
      Kokkos::parallel_scan("kkmx::scan_excl_i64", Range64(0, n), f);
    }
  });
}
int kkmx_scan_excl_f64(const double* in, double* out, int64_t n, double* total) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_excl_f64: bad arguments");
  return guarded([&] {
    ExclusiveScan<double, double> f{DevView<const double>(in, n), DevView<double>(out, n)};
    if (total) {
      DevScalar<double> t(total);
      Kokkos::parallel_scan("kkmx::scan_excl_f64", Range64(0, n), f, t);
    } else {
      Kokkos::parallel_scan("kkmx::scan_excl_f64", Range64(0, n), f);
    }
  });
}

int kkmx_set_range_config(int pattern, int block, int items, int max_blocks) {
  if (pattern < 0 || pattern > 2) return reject("kkmx_set_range_config: pattern must be 0, 1 or 2");
  auto& o = Kokkos::Impl::range_overrides();
  if (block < 0) {
    o.set[pattern] = false;
    return 0;
  }
  o.cfg[pattern] = Kokkos::Impl::RangeConfig{block, items < 1 ? 1 : items, max_blocks < 0 ? 0 : max_blocks};
  o.set[pattern] = true;
  return 0;
}

}  // extern "C"
