// C-ABI entry points of the hot path (declared in include/kkmx_paths.h).
//
// Every function here is plain Kokkos user code: unmanaged Views over the
// caller's device pointers, functors with the reference benchmarks' bodies,
// and parallel_for / parallel_reduce / parallel_scan on the default MI355X
// instance.  Nothing here calls HIP directly.
#include <Kokkos_Core.hpp>
#include <Kokkos_Batch.hpp>
#include <Kokkos_StdAlgorithms.hpp>
#include <Kokkos_ScatterView.hpp>
#include <Kokkos_Sort.hpp>

#include <cmath>
#include <memory>
#include <string>
#include <vector>

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
// The scan input is a View type parameter.  The entry points read it through
// a const RandomAccess View when it shares no memory with the output: the
// kernel never changes it, and those loads go through the constant address
// space, which lets the scan's final pass reuse the first pass's values
// instead of reading the input again.  Overlapping (e.g. in-place) calls use
// a plain const View.
template <class T>
using ReadOnlyView = Kokkos::View<const T*, Kokkos::LayoutRight, Kokkos::MI355XSpace,
                                  Kokkos::MemoryTraits<Kokkos::Unmanaged | Kokkos::RandomAccess>>;
template <class T>
using RestrictView = Kokkos::View<const T*, Kokkos::LayoutRight, Kokkos::MI355XSpace,
                                  Kokkos::MemoryTraits<Kokkos::Unmanaged | Kokkos::Restrict>>;
template <class In, class Acc, class InView = DevView<const In>>
struct ExclusiveScan {
  InView in;
  DevView<Acc> out;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, Acc& update, const bool final) const {
    if (final) out(i) = update;
    update += static_cast<Acc>(in(i));
  }
};
template <class T, class InView = DevView<const T>>
struct InclusiveScan {
  InView in;
  DevView<T> out;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, T& update, const bool final) const {
    update += in(i);
    if (final) out(i) = update;
  }
};
inline bool buffers_overlap(const void* a, size_t abytes, const void* b, size_t bbytes) {
  const auto pa = reinterpret_cast<uintptr_t>(a), pb = reinterpret_cast<uintptr_t>(b);
  return pa < pb + bbytes && pb < pa + abytes;
}

// Exclusive scan of n values into out, with an optional device total.
template <class In, class Acc, class InView>
void exclusive_scan_run(const char* label, const In* in, Acc* out, int64_t n, Acc* total) {
  ExclusiveScan<In, Acc, InView> f{InView(in, n), DevView<Acc>(out, n)};
  if (total) {
    DevScalar<Acc> t(total);
    Kokkos::parallel_scan(label, Kokkos::RangePolicy<Kokkos::IndexType<int64_t>>(0, n), f, t);
  } else {
    Kokkos::parallel_scan(label, Kokkos::RangePolicy<Kokkos::IndexType<int64_t>>(0, n), f);
  }
}
template <class In, class Acc>
void exclusive_scan_any(const char* label, const In* in, Acc* out, int64_t n, Acc* total) {
  if (buffers_overlap(in, sizeof(In) * size_t(n), out, sizeof(Acc) * size_t(n)))
    exclusive_scan_run<In, Acc, DevView<const In>>(label, in, out, n, total);
  else
    exclusive_scan_run<In, Acc, ReadOnlyView<In>>(label, in, out, n, total);
}


// ---------------------------------------------------------------- stencil
// 7-point Laplacian-type update over the interior of an (n+2)^3 LayoutLeft
// grid (first index fastest), ping-ponged between two Views.
// The input grid is read-only during a sweep, which the functor states with
// a const RandomAccess View (its loads may then move across the stores to v).
using Grid3   = Kokkos::View<double***, Kokkos::LayoutLeft, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
using Grid3In = Kokkos::View<const double***, Kokkos::LayoutLeft, Kokkos::MI355XSpace,
                             Kokkos::MemoryTraits<Kokkos::Unmanaged | Kokkos::RandomAccess>>;
struct Stencil7 {
  Grid3In u;
  Grid3 v;
  double c0, c1;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j, const int64_t k) const {
    const double s = u(i - 1, j, k) + u(i + 1, j, k) + u(i, j - 1, k) + u(i, j + 1, k) + u(i, j, k - 1) + u(i, j, k + 1);
    v(i, j, k)     = c0 * u(i, j, k) + c1 * s;
  }
};
struct Stencil7Residual {  // max |u - v| over the interior, for checks
  Grid3 u, v;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j, const int64_t k, double& m) const {
    const double d = fabs(u(i, j, k) - v(i, j, k));
    if (d > m) m = d;
  }
};

// ---------------------------------------------------------------- CSR SpMV
// One team per block of rows; the team first gathers x[col[k]] for every
// nonzero of its rows into level-0 scratch (LDS) with all lanes of the team
// (coalesced reads of col, many independent gathers in flight), then each
// thread forms one row's dot product from LDS with its vector lanes.
struct SpMVTeam {
  DevView<const int64_t> rp;
  DevView<const int32_t> col;
  DevView<const double> val, x;
  DevView<double> y;
  int64_t nrows;
  int rows_per_team;
  int max_team_nnz;  // scratch capacity in values
  using member_type = Kokkos::TeamPolicy<Kokkos::MI355X>::member_type;
  KOKKOS_INLINE_FUNCTION void operator()(const member_type& m) const {
    const int64_t r0 = int64_t(m.league_rank()) * rows_per_team;
    int64_t r1       = r0 + rows_per_team;
    if (r1 > nrows) r1 = nrows;
    const int64_t k0 = rp(r0), k1 = rp(r1);
    Kokkos::View<double*, Kokkos::MI355X::scratch_memory_space, Kokkos::MemoryTraits<Kokkos::Unmanaged>> xs(
        m.team_scratch(0), max_team_nnz);
    const bool staged = (k1 - k0) <= max_team_nnz;
    if (staged) {
      Kokkos::parallel_for(Kokkos::TeamVectorRange(m, k0, k1), [&](const int64_t k) { xs(k - k0) = x(col(k)); });
      m.team_barrier();
    }
    for (int64_t row = r0 + m.team_rank(); row < r1; row += m.team_size()) {
      double sum = 0.0;
      Kokkos::parallel_reduce(
          Kokkos::ThreadVectorRange(m, rp(row), rp(row + 1)),
          [&](const int64_t k, double& t) { t += val(k) * (staged ? xs(k - k0) : x(col(k))); }, sum);
      Kokkos::single(Kokkos::PerThread(m), [&]() { y(row) = sum; });
    }
  }
};
// Plain row-per-thread SpMV (no scratch), for comparison.
struct SpMVRow {
  DevView<const int64_t> rp;
  DevView<const int32_t> col;
  DevView<const double> val, x;
  DevView<double> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t row) const {
    double s = 0.0;
    for (int64_t k = rp(row); k < rp(row + 1); ++k) s += val(k) * x(col(k));
    y(row) = s;
  }
};
// CSR generation: nnz of row r = lo + mulhi(splitmix(seed, r), hi-lo+1);
// column of nonzero k = mulhi(splitmix(seed ^ C1, k), ncols); value of k from
// splitmix(seed ^ C2, k) in [-1, 1).
constexpr uint64_t kColSalt = 0x5bd1e9955bd1e995ull;
constexpr uint64_t kValSalt = 0xc2b2ae3d27d4eb4full;
struct CsrRowNnz {
  DevView<int64_t> cnt;
  uint64_t seed;
  int64_t lo;
  uint64_t range;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t r) const {
    cnt(r) = lo + static_cast<int64_t>(mulhi64(splitmix64(seed, static_cast<uint64_t>(r)), range));
  }
};
struct CsrRowPtr {  // exclusive scan of counts into rp[0..n], rp[n] = total
  DevView<const int64_t> cnt;
  DevView<int64_t> rp;
  int64_t n;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t r, int64_t& upd, const bool final) const {
    if (final) rp(r) = upd;
    upd += cnt(r);
    if (final && r == n - 1) rp(n) = upd;
  }
};
struct CsrFill {
  DevView<int32_t> col;
  DevView<double> val;
  uint64_t seed;
  uint64_t ncols;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t k) const {
    col(k) = static_cast<int32_t>(mulhi64(splitmix64(seed ^ kColSalt, static_cast<uint64_t>(k)), ncols));
    const double u = static_cast<double>(splitmix64(seed ^ kValSalt, static_cast<uint64_t>(k)) >> 11) * 0x1.0p-53;
    val(k)         = fma(2.0, u, -1.0);
  }
};

// ---------------------------------------------------------------- atomics
// Random_XorShift64 (algorithms/src/Kokkos_Random.hpp:788-808) outputs,
// generated in parallel: chunk t starts from the state M^(t*K) s0, where M
// is the GF(2) matrix of one xorshift step (the host computes the starts).
struct XorShiftChunks {
  DevView<uint64_t> out;
  DevView<const uint64_t> starts;
  int64_t n, chunk;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t t) const {
    uint64_t s      = starts(t);
    const int64_t b = t * chunk;
    int64_t e       = b + chunk;
    if (e > n) e = n;
    for (int64_t i = b; i < e; ++i) {
      s ^= s >> 12;
      s ^= s << 25;
      s ^= s >> 27;
      out(i) = s * 2685821657736338717ull - 1;
    }
  }
};
struct Gups {
  DevView<int64_t> table;
  DevView<const uint64_t> stream;
  uint64_t mask;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t u) const {
    const uint64_t idx = stream(2 * u) & mask;
    Kokkos::atomic_add(&table(idx), static_cast<int64_t>(stream(2 * u + 1)));
  }
};
// Zipf(s) keys over [0, nbins) by inverse CDF: key = first k with u < cdf[k].
struct ZipfKeys {
  DevView<int32_t> keys;
  DevView<const double> cdf;
  uint64_t seed;
  int32_t nbins;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    const double u = static_cast<double>(splitmix64(seed, static_cast<uint64_t>(i)) >> 11) * 0x1.0p-53;
    int32_t lo = 0, hi = nbins - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (u < cdf(mid)) hi = mid;
      else lo = mid + 1;
    }
    keys(i) = lo;
  }
};
struct HistGlobal {
  DevView<const int32_t> keys;
  DevView<int32_t> bins;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { Kokkos::atomic_add(&bins(keys(i)), 1); }
};
// LDS-privatised histogram: every team keeps bins [0, lds_bins) in level-0
// scratch, counts its share of the keys there with LDS atomics, and keys at
// or above lds_bins go straight to global atomics; the team then adds its
// nonzero LDS bins to the global histogram.
struct HistPrivatised {
  DevView<const int32_t> keys;
  DevView<int32_t> bins;
  int64_t n;
  int32_t lds_bins;
  using member_type = Kokkos::TeamPolicy<Kokkos::MI355X>::member_type;
  KOKKOS_INLINE_FUNCTION void operator()(const member_type& m) const {
    Kokkos::View<int32_t*, Kokkos::MI355X::scratch_memory_space, Kokkos::MemoryTraits<Kokkos::Unmanaged>> h(
        m.team_scratch(0), lds_bins);
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, lds_bins), [&](const int b) { h(b) = 0; });
    m.team_barrier();
    const int64_t nteams = m.league_size();
    const int64_t chunk  = (n + nteams - 1) / nteams;
    const int64_t b0     = chunk * m.league_rank();
    int64_t b1           = b0 + chunk;
    if (b1 > n) b1 = n;
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, b0, b1), [&](const int64_t i) {
      const int32_t k = keys(i);
      if (k < lds_bins) Kokkos::atomic_add(&h(k), 1);
      else Kokkos::atomic_add(&bins(k), 1);
    });
    m.team_barrier();
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, lds_bins), [&](const int b) {
      const int32_t c = h(b);
      if (c) Kokkos::atomic_add(&bins(b), c);
    });
  }
};
// Mode 2 of the histogram: the same team-private LDS bins, fed by 16-byte
// loads (four keys per load, kWideLoads loads in flight per lane before the
// first LDS atomic), with the lowest `hot` bins replicated kHotCopies times
// so that lanes counting the same frequent key hit different LDS words (lane
// l uses copy l % kHotCopies; copies are hot+1 words apart, which also spreads
// one bin's copies over the banks).  The team's share is a contiguous run of
// 4-key groups; team 0 also counts the n % 4 trailing keys.  The keys must be
// 16-byte aligned (checked by the entry point).
struct HistWide {
  static constexpr int kHotCopies = 16;
  static constexpr int kWideLoads = 4;
  DevView<const int32_t> keys;
  DevView<int32_t> bins;
  int64_t n;
  int32_t lds_bins;
  int32_t hot;
  using member_type = Kokkos::TeamPolicy<Kokkos::MI355X>::member_type;
  using Lds = Kokkos::View<int32_t*, Kokkos::MI355X::scratch_memory_space, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
  static size_t scratch_bytes(int32_t lds_bins, int32_t hot) {
    return (size_t(lds_bins) + size_t(kHotCopies) * (hot + 1)) * 4;
  }
  KOKKOS_INLINE_FUNCTION void count(const Lds& h, int32_t copy_base, int32_t k) const {
    if (k < hot) Kokkos::atomic_add(&h(lds_bins + copy_base + k), 1);
    else if (k < lds_bins) Kokkos::atomic_add(&h(k), 1);
    else Kokkos::atomic_add(&bins(k), 1);
  }
  KOKKOS_INLINE_FUNCTION void operator()(const member_type& m) const {
    const int32_t stride_copies = hot + 1;
    const int32_t words         = lds_bins + kHotCopies * stride_copies;
    Lds h(m.team_scratch(0), words);
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, words), [&](const int b) { h(b) = 0; });
    m.team_barrier();
    const int lane          = m.team_rank();
    const int T             = m.team_size();
    const int32_t copy_base = (lane % kHotCopies) * stride_copies;
    const int64_t groups    = n / 4;
    const int64_t nteams    = m.league_size();
    const int64_t chunk     = (groups + nteams - 1) / nteams;
    const int64_t g0        = chunk * m.league_rank();
    const int64_t g1        = g0 + chunk < groups ? g0 + chunk : groups;
    const int4* kv          = reinterpret_cast<const int4*>(keys.data());
    for (int64_t g = g0 + lane; g < g1; g += int64_t(kWideLoads) * T) {
      int4 q[kWideLoads];
#pragma unroll
      for (int u = 0; u < kWideLoads; ++u) {
        const int64_t gi = g + int64_t(u) * T;
        q[u]             = gi < g1 ? kv[gi] : int4{0, 0, 0, 0};
      }
#pragma unroll
      for (int u = 0; u < kWideLoads; ++u) {
        if (g + int64_t(u) * T >= g1) continue;  // past the team's share
        count(h, copy_base, q[u].x);
        count(h, copy_base, q[u].y);
        count(h, copy_base, q[u].z);
        count(h, copy_base, q[u].w);
      }
    }
    if (m.league_rank() == 0 && lane < n - groups * 4) count(h, copy_base, keys(groups * 4 + lane));
    m.team_barrier();
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, lds_bins), [&](const int b) {
      int32_t c = h(b);
      if (b < hot)
        for (int r = 0; r < kHotCopies; ++r) c += h(lds_bins + r * stride_copies + b);
      if (c) Kokkos::atomic_add(&bins(b), c);
    });
  }
};
// Mode 3 of the histogram: no device-scope atomics.  Each team owns one row
// of a [teams x nbins] table in HBM.  It counts bins [0, lds_bins) in LDS as
// mode 2 does and the remaining bins straight into its own row with
// workgroup-scope atomics (desul::MemoryScopeCore: no other team touches the
// row, so the update resolves in the XCD's L2), then stores its LDS counts
// into the row.  HistRowSum adds the column of every bin into `bins`.
struct HistRows {
  // Replicas of the hot bins as in mode 2 (a variant with one replica of the
  // 128 hottest bins per lane measured slower: 389 against 371 us for 2^28
  // keys, as its 33 KiB of replicas push more keys out of LDS).
  static constexpr int kHotCopies = HistWide::kHotCopies;
  static constexpr int kHotBins   = 256;
  static size_t scratch_bytes(int32_t lds_bins, int32_t hot) {
    return (size_t(lds_bins) + size_t(kHotCopies) * (hot + 1)) * 4;
  }
  DevView<const int32_t> keys;
  Kokkos::View<int32_t**, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>> rows;
  int64_t n;
  int32_t nbins;
  int32_t lds_bins;
  int32_t hot;
  using member_type = Kokkos::TeamPolicy<Kokkos::MI355X>::member_type;
  using Lds         = HistWide::Lds;
  KOKKOS_INLINE_FUNCTION void count(const Lds& h, int32_t* row, int32_t copy_base, int32_t k) const {
    if (k < hot) Kokkos::atomic_add(&h(lds_bins + copy_base + k), 1);
    else if (k < lds_bins) Kokkos::atomic_add(&h(k), 1);
    else desul::atomic_add(row + k, 1, desul::MemoryOrderRelaxed(), desul::MemoryScopeCore());
  }
  KOKKOS_INLINE_FUNCTION void operator()(const member_type& m) const {
    constexpr int kCopies       = kHotCopies;
    constexpr int kLoads        = HistWide::kWideLoads;
    const int32_t stride_copies = hot + 1;
    const int32_t words         = lds_bins + kCopies * stride_copies;
    int32_t* row                = &rows(m.league_rank(), 0);
    Lds h(m.team_scratch(0), words);
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, words), [&](const int b) { h(b) = 0; });
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, lds_bins, nbins), [&](const int b) { row[b] = 0; });
    m.team_barrier();
    const int lane          = m.team_rank();
    const int T             = m.team_size();
    const int32_t copy_base = (lane % kCopies) * stride_copies;
    const int64_t groups    = n / 4;
    const int64_t nteams    = m.league_size();
    const int64_t chunk     = (groups + nteams - 1) / nteams;
    const int64_t g0        = chunk * m.league_rank();
    const int64_t g1        = g0 + chunk < groups ? g0 + chunk : groups;
    const int4* kv          = reinterpret_cast<const int4*>(keys.data());
    for (int64_t g = g0 + lane; g < g1; g += int64_t(kLoads) * T) {
      int4 q[kLoads];
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        const int64_t gi = g + int64_t(u) * T;
        q[u]             = gi < g1 ? kv[gi] : int4{0, 0, 0, 0};
      }
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        if (g + int64_t(u) * T >= g1) continue;
        count(h, row, copy_base, q[u].x);
        count(h, row, copy_base, q[u].y);
        count(h, row, copy_base, q[u].z);
        count(h, row, copy_base, q[u].w);
      }
    }
    if (m.league_rank() == 0 && lane < n - groups * 4) count(h, row, copy_base, keys(groups * 4 + lane));
    m.team_barrier();
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, lds_bins), [&](const int b) {
      int32_t c = h(b);
      if (b < hot)
        for (int r = 0; r < kCopies; ++r) c += h(lds_bins + r * stride_copies + b);
      row[b] = c;
    });
  }
};
// Mode 3 when every bin fits in LDS (up to about 70 K bins): no key leaves
// the CU until the team's row is written.
//  * Bins below 256 (the head of skewed key distributions) are 32-bit
//    counters replicated 16 times, lane l adding to copy l mod 16.
//  * Every other bin is a 16-bit half of a 32-bit LDS word, incremented with
//    a returning LDS add of 1 or 1 << 16.  An add whose returned half ends in
//    0x3fff carries the half across a multiple of 16384; the lane that sees
//    it subtracts 16384 from the half again and appends the bin to the team's
//    carry list, which becomes one 16384 row atomic per entry at the end.  A
//    half stays below 65536 while each carry's subtraction lands within 49151
//    further adds to it.  A wavefront issues its subtraction right after its
//    own 16 adds return, so at most two rounds of the team's adds come in
//    between (2 x 16 wavefronts x 16 adds x 64 lanes = 32768).  The team
//    handles fewer than 1024 x 16384 keys, so the list never fills (larger
//    extents take the team-row path below).
//  * The 16 adds of a lane's four 16-byte key loads are issued back to back;
//    their returns are checked together, with one branch per 16 keys that is
//    taken only for a carry.
// Keys past the end of the team's share count into a spare word per lane.
// The time goes to streaming the keys: 2^28 Zipf keys take 0.24 ms, against
// 0.21 ms for the loads alone, and 0.38 ms when 3 % of the keys (bins above
// the 32 K that fit as 32-bit counters) left the loop for the team row.
struct HistPacked {
  static constexpr int kCopies = 16;
  static constexpr int kHot    = 256;
  static constexpr int kList   = 1024;
  static constexpr uint32_t kCarry = 16384;
  static constexpr int64_t kMaxKeysPerTeam = int64_t(kList) * kCarry;
  DevView<const int32_t> keys;
  Kokkos::View<int32_t**, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>> rows;
  int64_t n;
  int32_t nbins;
  using member_type = Kokkos::TeamPolicy<Kokkos::MI355X>::member_type;
  using Lds = Kokkos::View<uint32_t*, Kokkos::MI355X::scratch_memory_space, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
  KOKKOS_INLINE_FUNCTION static int32_t packed_words(int32_t nbins) { return (nbins - kHot + 1) / 2; }
  static size_t scratch_words(int32_t nbins) {
    return size_t(kCopies) * (kHot + 1) + size_t(packed_words(nbins)) + 64 + 1 + kList;
  }
  KOKKOS_INLINE_FUNCTION void operator()(const member_type& m) const {
    constexpr int kLoads   = HistWide::kWideLoads;
    const int32_t hot_w    = kCopies * (kHot + 1);
    const int32_t pw       = packed_words(nbins);
    const int32_t dummy0   = hot_w + pw;
    const int32_t list_cnt = dummy0 + 64;
    const int32_t list0    = list_cnt + 1;
    const int32_t words    = list0 + kList;
    int32_t* row           = &rows(m.league_rank(), 0);
    Lds h(m.team_scratch(0), words);
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, words), [&](const int b) { h(b) = 0u; });
    m.team_barrier();
    const int lane        = m.team_rank();
    const int T           = m.team_size();
    const int wl          = lane & 63;
    const int32_t hot_off = (lane % kCopies) * (kHot + 1);
    const int64_t groups  = n / 4;
    const int64_t nteams  = m.league_size();
    const int64_t chunk   = (groups + nteams - 1) / nteams;
    const int64_t g0      = chunk * m.league_rank();
    const int64_t g1      = g0 + chunk < groups ? g0 + chunk : groups;
    const int4* kv        = reinterpret_cast<const int4*>(keys.data());
    // LDS word and increment of key k: a hot replica (1), a packed half
    // (1 or 1 << 16), or this lane's spare word for an empty slot (k < 0).
    auto slot = [&](int32_t k, int32_t& w, uint32_t& inc, uint32_t& sh) {
      const int32_t j = k - kHot;
      sh              = (k >= kHot) ? uint32_t(j & 1) * 16u : 0u;
      w               = k < 0 ? dummy0 + wl : (k < kHot ? hot_off + k : hot_w + (j >> 1));
      inc             = 1u << sh;
    };
    auto carry = [&](int32_t k, int32_t w, uint32_t sh) {
      Kokkos::atomic_sub(&h(w), kCarry << sh);
      const uint32_t s2 = Kokkos::atomic_fetch_add(&h(list_cnt), 1u);
      h(list0 + int32_t(s2)) = uint32_t(k);
    };
    for (int64_t gb = g0; gb < g1; gb += int64_t(kLoads) * T) {
      int32_t k[4 * kLoads];
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        const int64_t gi = gb + lane + int64_t(u) * T;
        const int4 q     = gi < g1 ? kv[gi] : int4{-1, -1, -1, -1};
        k[4 * u] = q.x, k[4 * u + 1] = q.y, k[4 * u + 2] = q.z, k[4 * u + 3] = q.w;
      }
      uint32_t old[4 * kLoads];
#pragma unroll
      for (int r = 0; r < 4 * kLoads; ++r) {
        int32_t w;
        uint32_t inc, sh;
        slot(k[r], w, inc, sh);
        old[r] = Kokkos::atomic_fetch_add(&h(w), inc);
      }
      bool any = false;
#pragma unroll
      for (int r = 0; r < 4 * kLoads; ++r) {
        int32_t w;
        uint32_t inc, sh;
        slot(k[r], w, inc, sh);
        any = any || (k[r] >= kHot && ((old[r] >> sh) & (kCarry - 1)) == kCarry - 1);
      }
      if (__ballot(any)) {
#pragma unroll
        for (int r = 0; r < 4 * kLoads; ++r) {
          int32_t w;
          uint32_t inc, sh;
          slot(k[r], w, inc, sh);
          if (k[r] >= kHot && ((old[r] >> sh) & (kCarry - 1)) == kCarry - 1) carry(k[r], w, sh);
        }
      }
    }
    if (m.league_rank() == 0 && lane < n - groups * 4) {
      int32_t w;
      uint32_t inc, sh;
      const int32_t kk = keys(groups * 4 + lane);
      slot(kk, w, inc, sh);
      const uint32_t o = Kokkos::atomic_fetch_add(&h(w), inc);
      if (kk >= kHot && ((o >> sh) & (kCarry - 1)) == kCarry - 1) carry(kk, w, sh);
    }
    m.team_barrier();
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, nbins), [&](const int b) {
      uint32_t c;
      if (b < kHot) {
        c = 0;
        for (int r = 0; r < kCopies; ++r) c += h(r * (kHot + 1) + b);
      } else {
        const int32_t j = b - kHot;
        c               = (h(hot_w + (j >> 1)) >> ((j & 1) * 16)) & 0xffffu;
      }
      row[b] = int32_t(c);
    });
    m.team_barrier();
    const int32_t nlist = int32_t(h(list_cnt));
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, nlist), [&](const int s2) {
      desul::atomic_add(row + int32_t(h(list0 + s2)), int32_t(kCarry), desul::MemoryOrderRelaxed(), desul::MemoryScopeCore());
    });
  }
};
// One team of 8 wavefronts per 64 consecutive bins: wavefront w sums rows
// w, w+8, ... of the team's bins (64 lanes = 64 consecutive bins, so each
// load is one contiguous 256-byte run), and the 8 partial sums meet in LDS.
struct HistRowSum {
  static constexpr int kBins = 64, kWaves = 8;
  Kokkos::View<const int32_t**, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>> rows;
  DevView<int32_t> bins;
  using member_type = Kokkos::TeamPolicy<Kokkos::MI355X>::member_type;
  KOKKOS_INLINE_FUNCTION void operator()(const member_type& m) const {
    Kokkos::View<int32_t**, Kokkos::MI355X::scratch_memory_space, Kokkos::MemoryTraits<Kokkos::Unmanaged>> part(
        m.team_scratch(0), kWaves, kBins);
    const int w = m.team_rank() / kBins, l = m.team_rank() % kBins;
    const int64_t b = int64_t(m.league_rank()) * kBins + l;
    const int64_t nrows = int64_t(rows.extent(0));
    int32_t c = 0;
    if (b < int64_t(bins.extent(0))) {
#pragma unroll 8
      for (int64_t r = w; r < nrows; r += kWaves) c += rows(r, b);
    }
    part(w, l) = c;
    m.team_barrier();
    if (w == 0 && b < int64_t(bins.extent(0))) {
      int32_t t = 0;
      for (int k = 0; k < kWaves; ++k) t += part(k, l);
      bins(b) += t;
    }
  }
};
// Grow-only table of team rows for mode 3, released by Kokkos::finalize.
int32_t* hist_rows_space(size_t elems) {
  static Kokkos::View<int32_t*, Kokkos::MI355XSpace>* rows = nullptr;
  if (!rows) {
    rows = new Kokkos::View<int32_t*, Kokkos::MI355XSpace>();
    Kokkos::push_finalize_hook([] {
      delete rows;
      rows = nullptr;
    });
  }
  if (rows->extent(0) < elems)
    *rows = Kokkos::View<int32_t*, Kokkos::MI355XSpace>(
        Kokkos::view_alloc(Kokkos::WithoutInitializing, std::string("kkmx_hist_rows")), elems);
  return rows->data();
}
struct ScatterAdd {
  DevView<const double> vals;
  DevView<const int32_t> idx;
  DevView<double> bins;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { Kokkos::atomic_add(&bins(idx(i)), vals(i)); }
};

using Range64 = Kokkos::RangePolicy<Kokkos::IndexType<int64_t>>;

}  // namespace


// GF(2) 64x64 matrices as 64 column words: M * s = xor of columns where s has 1s.
struct Gf2Mat {
  uint64_t col[64];
};
static uint64_t gf2_apply(const Gf2Mat& m, uint64_t s) {
  uint64_t r = 0;
  for (int b = 0; b < 64; ++b)
    if ((s >> b) & 1) r ^= m.col[b];
  return r;
}
static Gf2Mat gf2_mul(const Gf2Mat& a, const Gf2Mat& b) {  // a * b
  Gf2Mat r;
  for (int c = 0; c < 64; ++c) r.col[c] = gf2_apply(a, b.col[c]);
  return r;
}
static uint64_t xorshift_step(uint64_t s) {
  s ^= s >> 12;
  s ^= s << 25;
  s ^= s >> 27;
  return s;
}
static Gf2Mat xorshift_matrix_pow(uint64_t k) {
  Gf2Mat base, acc;
  for (int c = 0; c < 64; ++c) {
    base.col[c] = xorshift_step(uint64_t(1) << c);
    acc.col[c]  = uint64_t(1) << c;
  }
  while (k) {
    if (k & 1) acc = gf2_mul(base, acc);
    base = gf2_mul(base, base);
    k >>= 1;
  }
  return acc;
}

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

namespace {
template <class SV>
struct ScatterViewF64 {
  DevView<const double> vals;
  DevView<const int32_t> idx;
  SV sv;
  int op;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    auto a = sv.access();
    if constexpr (std::is_same<typename SV::original_value_type, double>::value) a(idx(i)).update(vals(i));
  }
};
template <class SV>
struct ScatterViewHist {
  DevView<const int32_t> keys;
  SV sv;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { sv.access()(keys(i)) += 1; }
};
template <class Op>
void scatter_view_f64(const double* vals, const int32_t* idx, int64_t n, double* bins, int32_t nbins) {
  namespace KE = Kokkos::Experimental;
  DevView<double> B(bins, nbins);
  auto sv = KE::create_scatter_view(Op{}, B);
  Kokkos::parallel_for("kkmx::scatter_view_f64", Range64(0, n),
                       ScatterViewF64<decltype(sv)>{DevView<const double>(vals, n), DevView<const int32_t>(idx, n), sv, 0});
  KE::contribute(B, sv);
}
}  // namespace

namespace {
// Kokkos::Experimental::sort_by_key over raw device words (C ABI below).
template <class T>
struct GreaterOp {
  KOKKOS_FUNCTION bool operator()(const T& a, const T& b) const { return a > b; }
};
template <class K, class V>
void sort_by_key_values(K* keys, V* values, int64_t n, int descending) {
  namespace KE = Kokkos::Experimental;
  Kokkos::MI355X ex;
  if (descending) KE::sort_by_key(ex, DevView<K>(keys, n), DevView<V>(values, n), GreaterOp<K>{});
  else KE::sort_by_key(ex, DevView<K>(keys, n), DevView<V>(values, n));
  ex.fence();
}
template <class K>
void sort_by_key_typed(void* keys, void* values, int64_t n, int value_bytes, int descending) {
  if (value_bytes == 4) sort_by_key_values(static_cast<K*>(keys), static_cast<uint32_t*>(values), n, descending);
  else sort_by_key_values(static_cast<K*>(keys), static_cast<uint64_t*>(values), n, descending);
}
}  // namespace

// ---------------------------------------------------------------- no-allocation scan
// kkmx_scan_incl_i32_noalloc is kkmx_scan_incl_i32 for callers that cannot
// allocate (graph capture, tight serving loops).  The scan's scratch -- a
// 256-byte control block holding the tile-claim word, then one 8-byte status
// word per tile -- lives in the fixed device array `scratch` below instead of
// the instance's grow-only buffers.  Each GPU holds its own copy of the array
// until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (kkmx_init looks it up,
// or the first call on a device).  Every call clears the control block and
// the status words it will use with a small kernel (no memset, so the call
// may be captured into a graph) and then runs the same kernel as the
// instance path for extents of 2^22 and more (1024 lanes x 32 values per
// tile).  Calls that overlap on one GPU share the array and race.
namespace {
constexpr int64_t kNoAllocTileElems   = 1024 * 32;
constexpr int64_t kNoAllocMaxTiles    = int64_t(1) << 16;  // 2^31 elements
constexpr size_t kNoAllocControlBytes = 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocControlBytes + size_t(kNoAllocMaxTiles) * 8;
constexpr int kNoAllocMaxDevices      = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mtx;
unsigned char* g_scratch_dev[kNoAllocMaxDevices] = {};

// This device's address of `scratch`; call with the device current.
int noalloc_scratch_setup(int dev) {
  if (dev < 0 || dev >= kNoAllocMaxDevices) return reject("kkmx_scan_incl_i32_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  if (g_scratch_dev[dev]) return 0;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != dev)
    return reject("kkmx_scan_incl_i32_noalloc: the Kokkos device is not current on this thread");
  void* p              = nullptr;
  const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (err != hipSuccess || p == nullptr) {
    g_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(err);
    return 1;
  }
  g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  return 0;
}

__global__ __launch_bounds__(256) void k_noalloc_clear(uint64_t* words, int64_t nwords) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nwords) words[i] = 0;
}
}  // namespace

extern "C" {

int kkmx_init(int device) {
  try {
    if (!Kokkos::is_initialized()) Kokkos::initialize(Kokkos::InitializationSettings().set_device_id(device));
    return noalloc_scratch_setup(Kokkos::Impl::mi355x_globals().props.device);
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
int kkmx_deep_copy_f64(double* dst, const double* src, int64_t n0, int64_t n1, int64_t n2, int dst_layout,
                       int src_layout) {
  if (n0 < 0 || n1 < 0 || n2 < 0 || dst_layout < 0 || dst_layout > 1 || src_layout < 0 || src_layout > 1 ||
      (n0 * n1 * n2 > 0 && (!dst || !src)))
    return reject("kkmx_deep_copy_f64: bad arguments");
  return guarded([&] {
    using L = Kokkos::View<double***, Kokkos::LayoutLeft, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
    using R = Kokkos::View<double***, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
    using CL = Kokkos::View<const double***, Kokkos::LayoutLeft, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
    using CR = Kokkos::View<const double***, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
    const size_t a = size_t(n0), b = size_t(n1), c = size_t(n2);
    Kokkos::MI355X ex;
    if (dst_layout == 0 && src_layout == 0) Kokkos::deep_copy(ex, L(dst, a, b, c), CL(src, a, b, c));
    else if (dst_layout == 0) Kokkos::deep_copy(ex, L(dst, a, b, c), CR(src, a, b, c));
    else if (src_layout == 0) Kokkos::deep_copy(ex, R(dst, a, b, c), CL(src, a, b, c));
    else Kokkos::deep_copy(ex, R(dst, a, b, c), CR(src, a, b, c));
  });
}
int kkmx_deep_fill_f64(double* dst, int64_t n, double value) {
  if (n < 0 || (n > 0 && !dst)) return reject("kkmx_deep_fill_f64: bad arguments");
  return guarded([&] { Kokkos::deep_copy(Kokkos::MI355X(), DevView<double>(dst, n), value); });
}
namespace {
struct MaxI32Op {
  KOKKOS_FUNCTION int32_t operator()(const int32_t a, const int32_t b) const { return a > b ? a : b; }
};
struct MaxI64Op {
  KOKKOS_FUNCTION int64_t operator()(const int64_t a, const int64_t b) const { return a > b ? a : b; }
};
struct PlusF64Op {
  KOKKOS_FUNCTION double operator()(const double a, const double b) const { return a + b; }
};
struct SquareF64Op {
  KOKKOS_FUNCTION double operator()(const double a) const { return a * a; }
};
}  // namespace
int kkmx_sort(void* keys, int64_t n, int type) {
  if (n < 0 || (n > 0 && !keys) || type < 0 || type > 5) return reject("kkmx_sort: bad arguments");
  return guarded([&] {
    Kokkos::MI355X ex;
    switch (type) {
      case 0: Kokkos::sort(ex, DevView<int32_t>(static_cast<int32_t*>(keys), n)); break;
      case 1: Kokkos::sort(ex, DevView<int64_t>(static_cast<int64_t*>(keys), n)); break;
      case 2: Kokkos::sort(ex, DevView<float>(static_cast<float*>(keys), n)); break;
      case 3: Kokkos::sort(ex, DevView<double>(static_cast<double*>(keys), n)); break;
      case 4: Kokkos::sort(ex, DevView<uint32_t>(static_cast<uint32_t*>(keys), n)); break;
      default: Kokkos::sort(ex, DevView<uint64_t>(static_cast<uint64_t*>(keys), n)); break;
    }
  });
}
int kkmx_sort_by_key(void* keys, void* values, int64_t n, int key_type, int value_bytes, int descending) {
  if (n < 0 || (n > 0 && (!keys || !values)) || key_type < 0 || key_type > 5 ||
      (value_bytes != 4 && value_bytes != 8) || descending < 0 || descending > 1)
    return reject("kkmx_sort_by_key: bad arguments");
  return guarded([&] {
    switch (key_type) {
      case 0: sort_by_key_typed<int32_t>(keys, values, n, value_bytes, descending); break;
      case 1: sort_by_key_typed<int64_t>(keys, values, n, value_bytes, descending); break;
      case 2: sort_by_key_typed<float>(keys, values, n, value_bytes, descending); break;
      case 3: sort_by_key_typed<double>(keys, values, n, value_bytes, descending); break;
      case 4: sort_by_key_typed<uint32_t>(keys, values, n, value_bytes, descending); break;
      default: sort_by_key_typed<uint64_t>(keys, values, n, value_bytes, descending); break;
    }
  });
}
int kkmx_binsort_f64(const double* keys, int64_t* values, int64_t n, int nbins, double kmin, double kmax,
                     int sort_within_bins, int32_t* perm_out) {
  if (n < 0 || n > 0x7fffffff || nbins < 1 || (n > 0 && (!keys || !values)) || !(kmax > kmin))
    return reject("kkmx_binsort_f64: bad arguments");
  return guarded([&] {
    using KV = DevView<const double>;
    KV K(keys, n);
    Kokkos::MI355X ex;
    Kokkos::BinSort<KV, Kokkos::BinOp1D<KV>> bs(ex, K, Kokkos::BinOp1D<KV>(nbins, kmin, kmax), sort_within_bins != 0);
    bs.create_permute_vector(ex);
    bs.sort(ex, DevView<int64_t>(values, n));
    if (perm_out && n > 0)
      KKMX_SAFE_CALL(mx_memcpy_async(perm_out, bs.get_permute_vector().data(), size_t(n) * 4, ex.stream()));
    ex.fence();
  });
}
int kkmx_std_inclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_std_inclusive_scan_i32: bad arguments");
  return guarded([&] {
    namespace KE = Kokkos::Experimental;
    KE::inclusive_scan(Kokkos::MI355X(), DevView<const int32_t>(in, n), DevView<int32_t>(out, n));
  });
}
int kkmx_std_exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, int64_t init) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_std_exclusive_scan_i64: bad arguments");
  return guarded([&] {
    namespace KE = Kokkos::Experimental;
    KE::exclusive_scan(Kokkos::MI355X(), DevView<const int64_t>(in, n), DevView<int64_t>(out, n), init);
  });
}
int kkmx_std_inclusive_scan_max_i32(const int32_t* in, int32_t* out, int64_t n) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_std_inclusive_scan_max_i32: bad arguments");
  return guarded([&] {
    namespace KE = Kokkos::Experimental;
    KE::inclusive_scan(Kokkos::MI355X(), DevView<const int32_t>(in, n), DevView<int32_t>(out, n), MaxI32Op{});
  });
}
int kkmx_std_transform_exclusive_scan_sq_f64(const double* in, double* out, int64_t n, double init) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_std_transform_exclusive_scan_sq_f64: bad arguments");
  return guarded([&] {
    namespace KE = Kokkos::Experimental;
    KE::transform_exclusive_scan(Kokkos::MI355X(), DevView<const double>(in, n), DevView<double>(out, n), init,
                                 PlusF64Op{}, SquareF64Op{});
  });
}
int kkmx_std_reduce_i64(const int64_t* in, int64_t n, int64_t init, int64_t* out) {
  if (n < 0 || (n > 0 && !in) || !out) return reject("kkmx_std_reduce_i64: bad arguments");
  return guarded([&] { *out = Kokkos::Experimental::reduce(Kokkos::MI355X(), DevView<const int64_t>(in, n), init); });
}
int kkmx_std_reduce_max_i64(const int64_t* in, int64_t n, int64_t init, int64_t* out) {
  if (n < 0 || (n > 0 && !in) || !out) return reject("kkmx_std_reduce_max_i64: bad arguments");
  return guarded([&] {
    *out = Kokkos::Experimental::reduce(Kokkos::MI355X(), DevView<const int64_t>(in, n), init, MaxI64Op{});
  });
}
int kkmx_std_transform_reduce_dot_f64(const double* a, const double* b, int64_t n, double init, double* out) {
  if (n < 0 || (n > 0 && (!a || !b)) || !out) return reject("kkmx_std_transform_reduce_dot_f64: bad arguments");
  return guarded([&] {
    *out = Kokkos::Experimental::transform_reduce(Kokkos::MI355X(), DevView<const double>(a, n),
                                                  DevView<const double>(b, n), init);
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
  return guarded([&] { exclusive_scan_any("kkmx::scan_excl_i32_i64", in, out, n, total); });
}
int kkmx_scan_incl_i32_traits(const int32_t* in, int32_t* out, int64_t n, int in_traits) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_incl_i32_traits: bad arguments");
  if (in_traits < 0 || in_traits > 2) return reject("kkmx_scan_incl_i32_traits: in_traits must be 0, 1 or 2");
  if (in_traits != 0 && n > 0 && buffers_overlap(in, 4 * size_t(n), out, 4 * size_t(n)))
    return reject("kkmx_scan_incl_i32_traits: RandomAccess / Restrict input must not overlap the output");
  return guarded([&] {
    if (in_traits == 1)
      Kokkos::parallel_scan("kkmx::scan_incl_i32", Range64(0, n),
                            InclusiveScan<int32_t, ReadOnlyView<int32_t>>{ReadOnlyView<int32_t>(in, n),
                                                                          DevView<int32_t>(out, n)});
    else if (in_traits == 2)
      Kokkos::parallel_scan("kkmx::scan_incl_i32", Range64(0, n),
                            InclusiveScan<int32_t, RestrictView<int32_t>>{RestrictView<int32_t>(in, n),
                                                                          DevView<int32_t>(out, n)});
    else
      Kokkos::parallel_scan("kkmx::scan_incl_i32", Range64(0, n),
                            InclusiveScan<int32_t>{DevView<const int32_t>(in, n), DevView<int32_t>(out, n)});
  });
}
int kkmx_scan_incl_i32(const int32_t* in, int32_t* out, int64_t n) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_incl_i32: bad arguments");
  return kkmx_scan_incl_i32_traits(in, out, n, buffers_overlap(in, 4 * size_t(n), out, 4 * size_t(n)) ? 0 : 1);
}
int kkmx_scan_incl_i32_body(int in_traits) {
  using namespace Kokkos::Impl;
  using R = Kokkos::RangePolicy<Kokkos::IndexType<int64_t>>;
  auto chosen = [](auto f) {
    using CR = typename ScanTypes<R, decltype(f), void>::CR;
    std::lock_guard<std::mutex> lk(scan_body_choice<CR>().m);
    const int c = scan_body_choice<CR>().chosen;
    return c < 0 ? -2 : c;
  };
  switch (in_traits) {
    case 0: return chosen(InclusiveScan<int32_t>{});
    case 1: return chosen(InclusiveScan<int32_t, ReadOnlyView<int32_t>>{});
    case 2: return chosen(InclusiveScan<int32_t, RestrictView<int32_t>>{});
    default: return reject("kkmx_scan_incl_i32_body: in_traits must be 0, 1 or 2");
  }
}
int kkmx_set_scan_body(int body) {
  if (body < -1 || body > 1) return reject("kkmx_set_scan_body: body must be -1, 0 or 1");
  Kokkos::Impl::range_overrides().scan_body = body;
  return 0;
}
int kkmx_scan_incl_i32_noalloc(const int32_t* in, int32_t* out, int64_t n) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_incl_i32_noalloc: bad arguments");
  if (n > kNoAllocMaxTiles * kNoAllocTileElems)
    return reject("kkmx_scan_incl_i32_noalloc: n exceeds 2^31 elements, the most its fixed scratch "
                  "(65536 tile-status words) covers");
  if (!Kokkos::is_initialized()) return reject("kkmx_scan_incl_i32_noalloc: call kkmx_init first");
  if (n == 0) return 0;
  try {
    const int dev = Kokkos::Impl::mi355x_globals().props.device;
    if (int rc = noalloc_scratch_setup(dev)) return rc;
    unsigned char* base = g_scratch_dev[dev];
    using F             = InclusiveScan<int32_t>;
    using CR            = Kokkos::Impl::ScanTypes<Range64, F, void>::CR;
    CR cr(F{DevView<const int32_t>(in, n), DevView<int32_t>(out, n)}, typename CR::reducer_store{}, 1);
    int64_t begin = 0, end = n, ntiles = (n + kNoAllocTileElems - 1) / kNoAllocTileElems;
    int64_t nwords = int64_t(kNoAllocControlBytes / 8) + ntiles;
    uint64_t* words = reinterpret_cast<uint64_t*>(base);
    mx_stream_t stream = Kokkos::MI355X().stream();
    Kokkos::Impl::mx_launch_kernel(reinterpret_cast<const void*>(&k_noalloc_clear),
                                   static_cast<unsigned>((nwords + 255) / 256), 256, 0, stream, words, nwords);
    Kokkos::Impl::ScanTileState<int32_t> st;
    st.bind(base + kNoAllocControlBytes, nullptr);
    unsigned* control = reinterpret_cast<unsigned*>(base);
    int32_t* total    = nullptr;
    const void* k     = Kokkos::Impl::range_scan_kernel<CR, int64_t, false>(32);
    Kokkos::Impl::mx_launch_kernel(k, static_cast<unsigned>(ntiles), 1024, 0, stream, cr, begin, end, ntiles, st,
                                   control, total);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
int kkmx_scan_excl_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* total) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_excl_i64: bad arguments");
  return guarded([&] { exclusive_scan_any("kkmx::scan_excl_i64", in, out, n, total); });
}
int kkmx_scan_excl_f64(const double* in, double* out, int64_t n, double* total) {
  if (n < 0 || (n > 0 && (!in || !out))) return reject("kkmx_scan_excl_f64: bad arguments");
  return guarded([&] { exclusive_scan_any("kkmx::scan_excl_f64", in, out, n, total); });
}


// ---------------------------------------------------------------- config 3
int kkmx_stencil7(const double* u, double* v, int64_t n, double c0, double c1, int t0, int t1, int t2) {
  if (n < 1 || !u || !v) return reject("kkmx_stencil7: bad arguments");
  return guarded([&] {
    Grid3In U(u, n + 2, n + 2, n + 2);
    Grid3 V(v, n + 2, n + 2, n + 2);
    using P = Kokkos::MDRangePolicy<Kokkos::Rank<3>, Kokkos::IndexType<int64_t>>;
    if (t0 > 0)
      Kokkos::parallel_for("kkmx::stencil7", P({1, 1, 1}, {n + 1, n + 1, n + 1}, {t0, t1, t2}), Stencil7{U, V, c0, c1});
    else
      Kokkos::parallel_for("kkmx::stencil7", P({1, 1, 1}, {n + 1, n + 1, n + 1}), Stencil7{U, V, c0, c1});
  });
}
int kkmx_stencil7_sweeps(double* u, double* v, int64_t n, int sweeps, double c0, double c1, int t0, int t1, int t2) {
  for (int s = 0; s < sweeps; ++s) {
    int rc = (s % 2 == 0) ? kkmx_stencil7(u, v, n, c0, c1, t0, t1, t2) : kkmx_stencil7(v, u, n, c0, c1, t0, t1, t2);
    if (rc) return rc;
  }
  return 0;
}
int kkmx_csr_generate(int64_t* rowptr, int32_t* col, double* val, int64_t nrows, int64_t ncols, int64_t nnz_lo,
                      int64_t nnz_hi, uint64_t seed, int64_t* nnz_out) {
  if (nrows < 1 || !rowptr || nnz_hi < nnz_lo) return reject("kkmx_csr_generate: bad arguments");
  return guarded([&] {
    Kokkos::View<int64_t*, Kokkos::MI355XSpace> cnt("csr_counts", nrows);
    DevView<int64_t> RP(rowptr, nrows + 1);
    Kokkos::parallel_for("kkmx::csr_counts", Range64(0, nrows),
                         CsrRowNnz{DevView<int64_t>(cnt.data(), nrows), seed, nnz_lo, uint64_t(nnz_hi - nnz_lo + 1)});
    Kokkos::parallel_scan("kkmx::csr_rowptr", Range64(0, nrows), CsrRowPtr{DevView<const int64_t>(cnt.data(), nrows), RP, nrows});
    int64_t nnz = 0;
    KKMX_SAFE_CALL(mx_stream_sync(Kokkos::MI355X().stream()));
    KKMX_SAFE_CALL(mx_memcpy(&nnz, rowptr + nrows, sizeof(int64_t)));
    *nnz_out = nnz;
    if (col && val)
      Kokkos::parallel_for("kkmx::csr_fill", Range64(0, nnz),
                           CsrFill{DevView<int32_t>(col, nnz), DevView<double>(val, nnz), seed, uint64_t(ncols)});
  });
}
int kkmx_csr_fill(int32_t* col, double* val, int64_t nnz, int64_t ncols, uint64_t seed) {
  if (nnz < 0 || (nnz && (!col || !val))) return reject("kkmx_csr_fill: bad arguments");
  return guarded([&] {
    Kokkos::parallel_for("kkmx::csr_fill", Range64(0, nnz),
                         CsrFill{DevView<int32_t>(col, nnz), DevView<double>(val, nnz), seed, uint64_t(ncols)});
  });
}
int kkmx_spmv_csr(const int64_t* rowptr, const int32_t* col, const double* val, const double* x, double* y,
                  int64_t nrows, int64_t nnz, int variant, int team_size, int vector_length) {
  if (nrows < 1 || !rowptr || !x || !y) return reject("kkmx_spmv_csr: bad arguments");
  return guarded([&] {
    DevView<const int64_t> RP(rowptr, nrows + 1);
    DevView<const int32_t> C(col, nnz);
    DevView<const double> V(val, nnz), X(x, nrows);
    DevView<double> Y(y, nrows);
    if (variant == 0) {
      Kokkos::parallel_for("kkmx::spmv_row", Range64(0, nrows), SpMVRow{RP, C, V, X, Y});
      return;
    }
    const int ts = team_size > 0 ? team_size : 64;
    const int vl = vector_length > 0 ? vector_length : 8;
    const int rows_per_team = ts;
    const int cap  = rows_per_team * 72;  // scratch values per team (mean row length 34)
    const int64_t league = (nrows + rows_per_team - 1) / rows_per_team;
    auto pol = Kokkos::TeamPolicy<Kokkos::MI355X>(static_cast<int>(league), ts, vl)
                   .set_scratch_size(0, Kokkos::PerTeam(size_t(cap) * sizeof(double)));
    Kokkos::parallel_for("kkmx::spmv_team", pol, SpMVTeam{RP, C, V, X, Y, nrows, rows_per_team, cap});
  });
}

// ---------------------------------------------------------------- config 4
int kkmx_xorshift64_stream(uint64_t* out, int64_t n, uint64_t seed) {
  if (n < 0 || (n && !out)) return reject("kkmx_xorshift64_stream: bad arguments");
  return guarded([&] {
    const int64_t chunk = 4096;
    const int64_t nchunks = (n + chunk - 1) / chunk;
    std::vector<uint64_t> starts(static_cast<size_t>(nchunks));
    const Gf2Mat jump = xorshift_matrix_pow(static_cast<uint64_t>(chunk));
    uint64_t s = seed == 0 ? 1318319ull : seed;
    for (int64_t t = 0; t < nchunks; ++t) {
      starts[t] = s;
      s         = gf2_apply(jump, s);
    }
    Kokkos::View<uint64_t*, Kokkos::MI355XSpace> d_starts(Kokkos::view_alloc(Kokkos::WithoutInitializing, std::string("xs_starts")), nchunks);
    KKMX_SAFE_CALL(mx_memcpy(d_starts.data(), starts.data(), starts.size() * sizeof(uint64_t)));
    Kokkos::parallel_for("kkmx::xorshift64", Range64(0, nchunks),
                         XorShiftChunks{DevView<uint64_t>(out, n), DevView<const uint64_t>(d_starts.data(), nchunks), n, chunk});
    Kokkos::fence();
  });
}
int kkmx_gups(int64_t* table, int64_t table_size, const uint64_t* stream, int64_t nupdates) {
  if (table_size < 1 || (table_size & (table_size - 1)) || !table || nupdates < 0 || (nupdates && !stream))
    return reject("kkmx_gups: table_size must be a power of two");
  return guarded([&] {
    Kokkos::parallel_for("kkmx::gups", Range64(0, nupdates),
                         Gups{DevView<int64_t>(table, table_size), DevView<const uint64_t>(stream, 2 * nupdates),
                              uint64_t(table_size - 1)});
  });
}
int kkmx_zipf_keys(int32_t* keys, int64_t n, double s, int32_t nbins, uint64_t seed) {
  if (n < 0 || nbins < 1 || (n && !keys)) return reject("kkmx_zipf_keys: bad arguments");
  return guarded([&] {
    // cdf[k] = sum_{j<=k} (j+1)^-s / H, accumulated left to right in double
    std::vector<double> cdf(static_cast<size_t>(nbins));
    double acc = 0.0;
    for (int32_t k = 0; k < nbins; ++k) {
      acc += std::pow(double(k + 1), -s);
      cdf[k] = acc;
    }
    for (int32_t k = 0; k < nbins; ++k) cdf[k] /= acc;
    cdf[nbins - 1] = 1.0;
    Kokkos::View<double*, Kokkos::MI355XSpace> d_cdf(Kokkos::view_alloc(Kokkos::WithoutInitializing, std::string("zipf_cdf")), nbins);
    KKMX_SAFE_CALL(mx_memcpy(d_cdf.data(), cdf.data(), cdf.size() * sizeof(double)));
    Kokkos::parallel_for("kkmx::zipf_keys", Range64(0, n),
                         ZipfKeys{DevView<int32_t>(keys, n), DevView<const double>(d_cdf.data(), nbins), seed, nbins});
    Kokkos::fence();
  });
}
int kkmx_histogram(const int32_t* keys, int64_t n, int32_t* bins, int32_t nbins, int mode, int lds_bins) {
  if (n < 0 || nbins < 1 || !bins || (n && !keys) || mode < 0 || mode > 3)
    return reject("kkmx_histogram: bad arguments");
  return guarded([&] {
    DevView<const int32_t> K(keys, n);
    DevView<int32_t> B(bins, nbins);
    if (mode == 0) {
      Kokkos::parallel_for("kkmx::hist_global", Range64(0, n), HistGlobal{K, B});
      return;
    }
    int lb = lds_bins > 0 ? lds_bins : 32768;
    if (lb > nbins) lb = nbins;
    const int cus = Kokkos::MI355X::impl_cu_count();
    if (mode >= 2) {
      // Modes 2 and 3 keep as many bins in LDS as fit beside the hot copies.
      const int32_t hot0  = lb < 256 ? lb : 256;
      const size_t budget = Kokkos::TeamPolicy<Kokkos::MI355X>(cus, 1024).scratch_size_max(0);
      const int64_t fit   = int64_t(budget / 4) - int64_t(HistWide::kHotCopies) * (hot0 + 1);
      if (lb > fit) lb = int32_t(fit);
    }
    if (mode == 3 && reinterpret_cast<uintptr_t>(keys) % 16 == 0 && nbins > HistPacked::kHot) {
      const int cus_ = Kokkos::MI355X::impl_cu_count();
      const size_t budget = Kokkos::TeamPolicy<Kokkos::MI355X>(cus_, 1024).scratch_size_max(0);
      if (HistPacked::scratch_words(nbins) * 4 <= budget && n / cus_ < HistPacked::kMaxKeysPerTeam - 4096) {
        int32_t* rows = hist_rows_space(size_t(cus_) * nbins);
        using Rows = Kokkos::View<int32_t**, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
        auto pol = Kokkos::TeamPolicy<Kokkos::MI355X>(cus_, 1024)
                       .set_scratch_size(0, Kokkos::PerTeam(HistPacked::scratch_words(nbins) * 4));
        Kokkos::parallel_for("kkmx::hist_packed", pol, HistPacked{K, Rows(rows, cus_, nbins), n, nbins});
        const int sum_teams = (nbins + HistRowSum::kBins - 1) / HistRowSum::kBins;
        auto sum_pol = Kokkos::TeamPolicy<Kokkos::MI355X>(sum_teams, HistRowSum::kBins * HistRowSum::kWaves)
                           .set_scratch_size(0, Kokkos::PerTeam(size_t(HistRowSum::kBins) * HistRowSum::kWaves * 4));
        Kokkos::parallel_for("kkmx::hist_row_sum", sum_pol, HistRowSum{decltype(HistRowSum::rows)(rows, cus_, nbins), B});
        return;
      }
    }
    if (mode == 3 && reinterpret_cast<uintptr_t>(keys) % 16 == 0) {
      int32_t lb3       = lds_bins > 0 ? lds_bins : 32768;
      if (lb3 > nbins) lb3 = nbins;
      const int32_t hot = lb3 < HistRows::kHotBins ? lb3 : HistRows::kHotBins;
      const size_t budget = Kokkos::TeamPolicy<Kokkos::MI355X>(cus, 1024).scratch_size_max(0);
      const int64_t fit   = int64_t(budget / 4) - int64_t(HistRows::kHotCopies) * (hot + 1);
      if (lb3 > fit) lb3 = int32_t(fit);
      int32_t* rows = hist_rows_space(size_t(cus) * nbins);
      using Rows = Kokkos::View<int32_t**, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
      auto pol = Kokkos::TeamPolicy<Kokkos::MI355X>(cus, 1024)
                     .set_scratch_size(0, Kokkos::PerTeam(HistRows::scratch_bytes(lb3, hot)));
      Kokkos::parallel_for("kkmx::hist_rows", pol, HistRows{K, Rows(rows, cus, nbins), n, nbins, lb3, hot});
      const int sum_teams = (nbins + HistRowSum::kBins - 1) / HistRowSum::kBins;
      auto sum_pol = Kokkos::TeamPolicy<Kokkos::MI355X>(sum_teams, HistRowSum::kBins * HistRowSum::kWaves)
                         .set_scratch_size(0, Kokkos::PerTeam(size_t(HistRowSum::kBins) * HistRowSum::kWaves * 4));
      Kokkos::parallel_for("kkmx::hist_row_sum", sum_pol, HistRowSum{decltype(HistRowSum::rows)(rows, cus, nbins), B});
      return;
    }
    if (mode >= 2 && reinterpret_cast<uintptr_t>(keys) % 16 == 0) {
      const int32_t hot = lb < 256 ? lb : 256;
      auto pol = Kokkos::TeamPolicy<Kokkos::MI355X>(cus, 1024)
                     .set_scratch_size(0, Kokkos::PerTeam(HistWide::scratch_bytes(lb, hot)));
      Kokkos::parallel_for("kkmx::hist_lds_wide", pol, HistWide{K, B, n, lb, hot});
      return;
    }
    auto pol = Kokkos::TeamPolicy<Kokkos::MI355X>(cus, 1024).set_scratch_size(0, Kokkos::PerTeam(size_t(lb) * 4));
    Kokkos::parallel_for("kkmx::hist_lds", pol, HistPrivatised{K, B, n, lb});
  });
}
int kkmx_scatter_view_f64(const double* vals, const int32_t* idx, int64_t n, double* bins, int32_t nbins, int op) {
  if (n < 0 || nbins < 1 || !bins || (n && (!vals || !idx)) || op < 0 || op > 2)
    return reject("kkmx_scatter_view_f64: bad arguments");
  return guarded([&] {
    namespace KE = Kokkos::Experimental;
    if (op == 0) scatter_view_f64<KE::ScatterSum>(vals, idx, n, bins, nbins);
    else if (op == 1) scatter_view_f64<KE::ScatterMin>(vals, idx, n, bins, nbins);
    else scatter_view_f64<KE::ScatterMax>(vals, idx, n, bins, nbins);
  });
}
int kkmx_scatter_view_histogram(const int32_t* keys, int64_t n, int32_t* bins, int32_t nbins, int duplicated) {
  if (n < 0 || nbins < 1 || !bins || (n && !keys)) return reject("kkmx_scatter_view_histogram: bad arguments");
  return guarded([&] {
    namespace KE = Kokkos::Experimental;
    DevView<int32_t> B(bins, nbins);
    auto run = [&](auto sv) {
      Kokkos::parallel_for("kkmx::scatter_view_histogram", Range64(0, n),
                           ScatterViewHist<decltype(sv)>{DevView<const int32_t>(keys, n), sv});
      KE::contribute(B, sv);
    };
    if (duplicated) run(KE::create_scatter_view(KE::ScatterSum{}, KE::ScatterDuplicated{}, KE::ScatterAtomic{}, B));
    else run(KE::create_scatter_view(B));
  });
}
int kkmx_scatter_add_f64(const double* vals, const int32_t* idx, int64_t n, double* bins, int32_t nbins) {
  if (n < 0 || nbins < 1 || !bins || (n && (!vals || !idx))) return reject("kkmx_scatter_add_f64: bad arguments");
  return guarded([&] {
    Kokkos::parallel_for("kkmx::scatter_f64", Range64(0, n),
                         ScatterAdd{DevView<const double>(vals, n), DevView<const int32_t>(idx, n), DevView<double>(bins, nbins)});
  });
}

int kkmx_set_mdrange_depth(int depth) {
  Kokkos::Impl::md_depth_override() = depth > 0 ? depth : 0;
  return 0;
}

int kkmx_range_candidate(int pattern, int candidate, int64_t n, int* block, int* items, int* max_blocks) {
  if (pattern < 0 || pattern > 2 || candidate < 0 || candidate >= Kokkos::Impl::kRangeCandidates || n < 0 ||
      !block || !items || !max_blocks)
    return reject("kkmx_range_candidate: bad arguments");
  return guarded([&] {
    const auto c = Kokkos::Impl::range_candidate(pattern, candidate, n);
    *block       = c.block;
    *items       = c.items;
    *max_blocks  = c.max_blocks;
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

// ---------------------------------------------------------------- batched dispatch
struct kkmx_batch {
  Kokkos::Experimental::BatchDispatch batch;
  std::unique_ptr<Kokkos::Experimental::Graph<Kokkos::MI355X>> plain_graph;
  mx_graph_t fused_graph     = nullptr;
  mx_graph_exec_t fused_exec = nullptr;
  ~kkmx_batch() {
    (void)mx_stream_sync(Kokkos::MI355X().stream());
    if (fused_exec) (void)mx_graph_exec_destroy(fused_exec);
    if (fused_graph) (void)mx_graph_destroy(fused_graph);
  }
};

int kkmx_batch_create(kkmx_batch** out) {
  if (!out) return reject("kkmx_batch_create: null handle");
  return guarded([&] { *out = new kkmx_batch(); });
}
int kkmx_batch_destroy(kkmx_batch* b) {
  return guarded([&] { delete b; });
}
int kkmx_batch_add(kkmx_batch* b, int op, int type, const void* in, int64_t n, void* out, void* result, double alpha,
                   double beta) {
  if (!b) return reject("kkmx_batch_add: null handle");
  if (op < 0 || op > 5) return reject("kkmx_batch_add: op must be 0..5");
  if (type < 0 || type > 2) return reject("kkmx_batch_add: type must be 0, 1 or 2");
  return guarded([&] {
    b->batch.add(Kokkos::Experimental::BatchOp(op), Kokkos::Experimental::BatchType(type), in, n, out, result, alpha,
                 beta);
  });
}
int kkmx_batch_set_shape(kkmx_batch* b, int shape) {
  if (!b) return reject("kkmx_batch_set_shape: null handle");
  return guarded([&] { b->batch.set_shape(shape); });
}
int kkmx_batch_get_shape(kkmx_batch* b, int* shape) {
  if (!b || !shape) return reject("kkmx_batch_get_shape: null argument");
  *shape = b->batch.shape();
  return 0;
}
int kkmx_batch_plan(kkmx_batch* b, int64_t* tiles_out) {
  if (!b) return reject("kkmx_batch_plan: null handle");
  return guarded([&] {
    b->batch.plan();
    if (tiles_out) *tiles_out = b->batch.tiles();
  });
}
int kkmx_batch_run(kkmx_batch* b, int mode, int steps) {
  if (!b) return reject("kkmx_batch_run: null handle");
  if (mode < 0 || mode > 3) return reject("kkmx_batch_run: mode must be 0..3");
  if (steps < 0) return reject("kkmx_batch_run: negative step count");
  return guarded([&] {
    Kokkos::MI355X ex;
    b->batch.plan();
    switch (mode) {
      case 0:
        for (int s = 0; s < steps; ++s) b->batch.submit_plain(ex);
        break;
      case 1:
        for (int s = 0; s < steps; ++s) b->batch.submit(ex);
        break;
      case 2:
        if (!b->fused_exec) {
          KKMX_SAFE_CALL(mx_stream_sync(ex.stream()));
          KKMX_SAFE_CALL(mx_graph_capture_begin(ex.stream()));
          b->batch.submit(ex);
          KKMX_SAFE_CALL(mx_graph_capture_end(ex.stream(), &b->fused_graph));
          KKMX_SAFE_CALL(mx_graph_instantiate(&b->fused_exec, b->fused_graph));
        }
        for (int s = 0; s < steps; ++s) {
          b->batch.impl_note_replays(ex, 1);
          KKMX_SAFE_CALL(mx_graph_launch(b->fused_exec, ex.stream()));
        }
        break;
      case 3:
        if (!b->plain_graph) {
          b->plain_graph.reset(new Kokkos::Experimental::Graph<Kokkos::MI355X>(b->batch.plain_graph(ex)));
          b->plain_graph->instantiate();
        }
        for (int s = 0; s < steps; ++s) b->plain_graph->submit(ex);
        break;
    }
  });
}

// ---------------------------------------------------------------- tools
// Loads Kokkos Tools libraries (';'-separated paths) into the running
// process, as KOKKOS_TOOLS_LIBS does at initialisation; unload stops them.
int kkmx_tools_load(const char* libs) {
  if (!libs) return reject("kkmx_tools_load: null path");
  return guarded([&] { Kokkos::Tools::Impl::tools_start(libs, ""); });
}
int kkmx_tools_unload(void) {
  return guarded([&] { Kokkos::Tools::Impl::tools_stop(); });
}

}  // extern "C"
