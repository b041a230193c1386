// Launch-shape sweep: the timings the learned tuner is trained on, collected
// by the runtime itself.
//
// A family is one Range dispatch (parallel_for, parallel_reduce or
// parallel_scan) with a fixed functor whose bytes and flops per index are
// declared as cost hints.  kkmx_sweep_time runs one family at one extent
// under each of the 16 launch-shape candidates the tuning hooks choose from
// and returns the best-of-reps kernel time of each, measured with HIP events
// on the instance stream.  Plain Kokkos user code: the candidates are applied
// through the same launch-shape override the tuning hooks use.
#include <Kokkos_Core.hpp>
#include <Kokkos_Batch.hpp>

#include <algorithm>
#include <cmath>
#include <string>

namespace {

thread_local std::string g_err;

template <class T>
using Vec = Kokkos::View<T*, Kokkos::LayoutRight, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
template <class T>
using Scalar = Kokkos::View<T, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;

struct Buffers {
  int64_t n   = 0;
  double* a   = nullptr;
  double* b   = nullptr;
  double* c   = nullptr;
  double* res = nullptr;  // 64 bytes of reduction results
} g_buf;

// ---------------------------------------------------------------- families
struct CopyF64 {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 0;
  Vec<const double> x;
  Vec<double> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { y(i) = x(i); }
};
struct TriadF64 {
  static constexpr double kkmx_bytes_per_index = 24, kkmx_flops_per_index = 2;
  Vec<double> a;
  Vec<const double> b, c;
  double s;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { a(i) = b(i) + s * c(i); }
};
struct ScaleI32 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 1;
  Vec<const int32_t> x;
  Vec<int32_t> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { y(i) = 3 * x(i); }
};
template <int K>
struct PolyF64 {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 2 * K;
  Vec<const double> x;
  Vec<double> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    const double v = x(i);
    double p       = 1.0;
#pragma unroll 8
    for (int k = 0; k < K; ++k) p = p * v + 0.5;
    y(i) = p;
  }
};
struct FillI64 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 1;
  Vec<int64_t> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { y(i) = i * 7; }
};
struct Stencil1D {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 2;
  Vec<const double> x;
  Vec<double> y;
  int64_t n;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    const int64_t l = i > 0 ? i - 1 : 0, r = i + 1 < n ? i + 1 : n - 1;
    y(i)            = x(l) + x(i) + x(r);
  }
};
struct SumF64 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 1;
  Vec<const double> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, double& s) const { s += x(i); }
};
struct SumI32 {
  static constexpr double kkmx_bytes_per_index = 4, kkmx_flops_per_index = 1;
  Vec<const int32_t> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& s) const { s += x(i); }
};
struct SumI64 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 1;
  Vec<const int64_t> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& s) const { s += x(i); }
};
struct SumF32 {
  static constexpr double kkmx_bytes_per_index = 4, kkmx_flops_per_index = 1;
  Vec<const float> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, float& s) const { s += x(i); }
};
struct DotF64 {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 2;
  Vec<const double> x, y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, double& s) const { s += x(i) * y(i); }
};
struct MaxI64 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 1;
  Vec<const int64_t> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& m) const {
    if (x(i) > m) m = x(i);
  }
};
struct MinMaxSum {
  double mn, mx, sum;
};
struct MinMaxSumF64 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 3;
  using value_type = MinMaxSum;
  Vec<const double> x;
  KOKKOS_INLINE_FUNCTION void init(MinMaxSum& v) const {
    v.mn  = 1e300;
    v.mx  = -1e300;
    v.sum = 0;
  }
  KOKKOS_INLINE_FUNCTION void join(MinMaxSum& a, const MinMaxSum& b) const {
    a.mn = b.mn < a.mn ? b.mn : a.mn;
    a.mx = b.mx > a.mx ? b.mx : a.mx;
    a.sum += b.sum;
  }
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, MinMaxSum& v) const {
    const double t = x(i);
    v.mn           = t < v.mn ? t : v.mn;
    v.mx           = t > v.mx ? t : v.mx;
    v.sum += t;
  }
};
template <int K>
struct PolySumF64 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 2 * K;
  Vec<const double> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, double& s) const {
    const double v = x(i);
    double p       = 1.0;
#pragma unroll 8
    for (int k = 0; k < K; ++k) p = p * v + 0.5;
    s += p;
  }
};
struct ScanExclI32I64 {
  static constexpr double kkmx_bytes_per_index = 12, kkmx_flops_per_index = 1;
  Vec<const int32_t> x;
  Vec<int64_t> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& u, const bool final) const {
    if (final) y(i) = u;
    u += x(i);
  }
};
struct ScanInclI32 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 1;
  Vec<const int32_t> x;
  Vec<int32_t> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int32_t& u, const bool final) const {
    u += x(i);
    if (final) y(i) = u;
  }
};
struct ScanExclI64 {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 1;
  Vec<const int64_t> x;
  Vec<int64_t> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& u, const bool final) const {
    if (final) y(i) = u;
    u += x(i);
  }
};
struct ScanExclF64 {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 1;
  Vec<const double> x;
  Vec<double> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, double& u, const bool final) const {
    if (final) y(i) = u;
    u += x(i);
  }
};
struct CountSum {
  int64_t count, sum;
  KOKKOS_INLINE_FUNCTION CountSum& operator+=(const CountSum& o) {
    count += o.count;
    sum += o.sum;
    return *this;
  }
};
struct ScanPairI64 {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 2;
  Vec<const int64_t> x;
  Vec<int64_t> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, CountSum& u, const bool final) const {
    u.count += x(i) & 1;
    u.sum += x(i);
    if (final) y(i) = u.sum - u.count;
  }
};

// ---- TeamPolicy families (extent = league size; one row of `width` per team)
using TeamP  = Kokkos::TeamPolicy<Kokkos::MI355X>;
using Member = TeamP::member_type;
struct TeamRowSum {  // y(r) = sum of row r
  static constexpr double kkmx_bytes_per_index = 256 * 8 + 8, kkmx_flops_per_index = 256;
  Vec<const double> a;
  Vec<double> y;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m) const {
    const int64_t r = m.league_rank();
    double s        = 0;
    Kokkos::parallel_reduce(
        Kokkos::TeamThreadRange(m, 256), [&](const int k, double& u) { u += a(r * 256 + k); }, s);
    Kokkos::single(Kokkos::PerTeam(m), [&]() { y(r) = s; });
  }
};
struct TeamGather {  // y(r) = sum_k a(r*32+k) * x(col(r*32+k)), 32 random columns per row
  static constexpr double kkmx_bytes_per_index = 32 * 20 + 8, kkmx_flops_per_index = 64;
  Vec<const double> a, x;
  Vec<const int32_t> col;
  Vec<double> y;
  int64_t ncols;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m) const {
    const int64_t r = m.league_rank();
    double s        = 0;
    Kokkos::parallel_reduce(
        Kokkos::TeamThreadRange(m, 32),
        [&](const int k, double& u) {
          const int64_t j = r * 32 + k;
          u += a(j) * x(int64_t(uint32_t(col(j))) % ncols);
        },
        s);
    Kokkos::single(Kokkos::PerTeam(m), [&]() { y(r) = s; });
  }
};
struct TeamCopy {  // 512 doubles per team, thread x vector lanes
  static constexpr double kkmx_bytes_per_index = 512 * 16, kkmx_flops_per_index = 0;
  Vec<const double> a;
  Vec<double> b;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m) const {
    const int64_t r = m.league_rank();
    Kokkos::parallel_for(Kokkos::TeamVectorRange(m, 512), [&](const int k) { b(r * 512 + k) = a(r * 512 + k); });
  }
};
struct TeamTotal {  // total of all rows: nested reduce inside a team reduction
  static constexpr double kkmx_bytes_per_index = 128 * 8, kkmx_flops_per_index = 128;
  Vec<const double> a;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m, double& u) const {
    const int64_t r = m.league_rank();
    double s        = 0;
    Kokkos::parallel_reduce(
        Kokkos::TeamThreadRange(m, 128), [&](const int k, double& v) { v += a(r * 128 + k); }, s);
    Kokkos::single(Kokkos::PerTeam(m), [&]() { u += s; });
  }
};
// ---- MDRangePolicy families (extent = points)
using Grid3 = Kokkos::View<double***, Kokkos::LayoutLeft, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
using Grid2 = Kokkos::View<double**, Kokkos::LayoutLeft, Kokkos::MI355XSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
struct Md7Point {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 8;
  Grid3 u, v;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j, const int64_t k) const {
    v(i, j, k) = 0.4 * u(i, j, k) + 0.1 * (u(i - 1, j, k) + u(i + 1, j, k) + u(i, j - 1, k) + u(i, j + 1, k) +
                                           u(i, j, k - 1) + u(i, j, k + 1));
  }
};
struct Md5Point2D {
  static constexpr double kkmx_bytes_per_index = 16, kkmx_flops_per_index = 6;
  Grid2 u, v;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j) const {
    v(i, j) = 0.5 * u(i, j) + 0.125 * (u(i - 1, j) + u(i + 1, j) + u(i, j - 1) + u(i, j + 1));
  }
};
struct MdAxpy3 {
  static constexpr double kkmx_bytes_per_index = 24, kkmx_flops_per_index = 2;
  Grid3 x, y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j, const int64_t k) const {
    y(i, j, k) += 2.5 * x(i, j, k);
  }
};
struct MdSumSq3 {
  static constexpr double kkmx_bytes_per_index = 8, kkmx_flops_per_index = 2;
  Grid3 x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j, const int64_t k, double& s) const {
    s += x(i, j, k) * x(i, j, k);
  }
};

struct InitBuffers {
  Vec<double> a, b, c;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    a(i) = double(i % 1000) * 1e-3;
    b(i) = double(i % 77) * 1e-2;
    c(i) = 0.25;
  }
};

constexpr int kRangeFamilies = 20;
constexpr int kFirstBatch    = 28;
constexpr int kFamilies      = 32;

// pattern, value bytes, floating, functor bytes, bytes/index, flops/index
struct FamilyInfo {
  int pattern;
  int value_bytes;
  int floating;
  int functor_bytes;
  double bytes, flops;
  int kind;  // 0 RangePolicy, 1 TeamPolicy, 2 MDRangePolicy, 3 batched dispatch
};

template <class F>
FamilyInfo info(int pattern, int value_bytes, int floating, int kind = 0) {
  return FamilyInfo{pattern, value_bytes, floating, int(sizeof(F)), F::kkmx_bytes_per_index, F::kkmx_flops_per_index,
                    kind};
}

FamilyInfo family_info(int f) {
  switch (f) {
    case 0: return info<CopyF64>(0, 0, 0);
    case 1: return info<TriadF64>(0, 0, 0);
    case 2: return info<ScaleI32>(0, 0, 0);
    case 3: return info<PolyF64<8>>(0, 0, 0);
    case 4: return info<PolyF64<64>>(0, 0, 0);
    case 5: return info<FillI64>(0, 0, 0);
    case 6: return info<Stencil1D>(0, 0, 0);
    case 7: return info<SumF64>(1, 8, 1);
    case 8: return info<SumI32>(1, 8, 0);
    case 9: return info<SumI64>(1, 8, 0);
    case 10: return info<SumF32>(1, 4, 1);
    case 11: return info<DotF64>(1, 8, 1);
    case 12: return info<MaxI64>(1, 8, 0);
    case 13: return info<MinMaxSumF64>(1, 24, 1);
    case 14: return info<PolySumF64<32>>(1, 8, 1);
    case 15: return info<ScanExclI32I64>(2, 8, 0);
    case 16: return info<ScanInclI32>(2, 4, 0);
    case 17: return info<ScanExclI64>(2, 8, 0);
    case 18: return info<ScanExclF64>(2, 8, 1);
    case 19: return info<ScanPairI64>(2, 16, 0);
    case 20: return info<TeamRowSum>(0, 0, 0, 1);
    case 21: return info<TeamGather>(0, 0, 0, 1);
    case 22: return info<TeamCopy>(0, 0, 0, 1);
    case 23: return info<TeamTotal>(1, 8, 1, 1);
    case 24: return info<Md7Point>(0, 0, 0, 2);
    case 25: return info<Md5Point2D>(0, 0, 0, 2);
    case 26: return info<MdAxpy3>(0, 0, 0, 2);
    case 27: return info<MdSumSq3>(1, 8, 1, 2);
    // batched dispatch: the features depend on the generated op list and
    // come from kkmx_sweep_last_features
    default: return FamilyInfo{3, 0, 0, 0, 0.0, 1.0, 3};
  }
}

// ---- batched dispatch families (extent = element budget of the batch):
// ops with log-uniform extents in [2^lo, 2^hi], half sums and half exclusive
// scans, element types int32 / int64 / float64 in turn (or float64 only).
struct BatchFamily {
  double lo, hi;
  bool f64_only;
};
BatchFamily batch_family(int f) {
  switch (f - kFirstBatch) {
    case 0: return {8, 20, false};
    case 1: return {4, 12, false};
    case 2: return {12, 22, false};
    default: return {8, 16, true};
  }
}
uint64_t sweep_mix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
// Adds the ops of family f with element budget `budget` (<= the sweep
// buffers): inputs are consecutive slices of buffer a, scan outputs of b,
// results slots of c.  Returns the elements added.
int64_t batch_ops(int f, int64_t budget, Kokkos::Experimental::BatchDispatch& bd) {
  using namespace Kokkos::Experimental;
  const BatchFamily bf = batch_family(f);
  uint64_t st          = sweep_mix(uint64_t(f) * 0x100000001ull + uint64_t(budget));
  int64_t used = 0, out = 0, nres = 0;
  while (nres < 16384 && nres < g_buf.n) {
    st               = sweep_mix(st);
    const double u   = double(st >> 11) * 0x1.0p-53;
    int64_t size     = int64_t(std::exp2(bf.lo + (bf.hi - bf.lo) * u));
    if (used + size > budget) {
      if (nres > 0) break;
      size = budget;
    }
    const BatchType t = bf.f64_only ? BatchType::Float64 : BatchType(nres % 3);
    const bool scan   = (st >> 7) & 1;
    double* in        = g_buf.a + used;
    if (scan) {
      bd.add(BatchOp::ScanExclusive, t, in, size, g_buf.b + out, g_buf.c + nres);
      out += size;
    } else {
      bd.add(BatchOp::ReduceSum, t, in, size, nullptr, g_buf.c + nres);
    }
    used += size;
    ++nres;
  }
  return used;
}
double g_last_features[7] = {0, 0, 0, 0, 0, 0, 0};

using P = Kokkos::RangePolicy<Kokkos::MI355X, Kokkos::IndexType<int64_t>>;

void run_family(int f, int64_t n) {
  const P p(0, n);
  double *a = g_buf.a, *b = g_buf.b, *c = g_buf.c;
  Scalar<double> rd(g_buf.res);
  Scalar<int64_t> ri(reinterpret_cast<int64_t*>(g_buf.res));
  Scalar<float> rf(reinterpret_cast<float*>(g_buf.res));
  Scalar<MinMaxSum> rm(reinterpret_cast<MinMaxSum*>(g_buf.res));
  auto i32 = [](double* q, int64_t m) { return Vec<int32_t>(reinterpret_cast<int32_t*>(q), m); };
  auto i64 = [](double* q, int64_t m) { return Vec<int64_t>(reinterpret_cast<int64_t*>(q), m); };
  switch (f) {
    case 0: Kokkos::parallel_for("sweep.copy", p, CopyF64{Vec<const double>(a, n), Vec<double>(b, n)}); break;
    case 1:
      Kokkos::parallel_for("sweep.triad", p, TriadF64{Vec<double>(a, n), Vec<const double>(b, n), Vec<const double>(c, n), 3.0});
      break;
    case 2: Kokkos::parallel_for("sweep.scale_i32", p, ScaleI32{i32(a, n), i32(b, n)}); break;
    case 3: Kokkos::parallel_for("sweep.poly8", p, PolyF64<8>{Vec<const double>(a, n), Vec<double>(b, n)}); break;
    case 4: Kokkos::parallel_for("sweep.poly64", p, PolyF64<64>{Vec<const double>(a, n), Vec<double>(b, n)}); break;
    case 5: Kokkos::parallel_for("sweep.fill_i64", p, FillI64{i64(a, n)}); break;
    case 6: Kokkos::parallel_for("sweep.stencil1d", p, Stencil1D{Vec<const double>(a, n), Vec<double>(b, n), n}); break;
    case 7: Kokkos::parallel_reduce("sweep.sum_f64", p, SumF64{Vec<const double>(a, n)}, rd); break;
    case 8: Kokkos::parallel_reduce("sweep.sum_i32", p, SumI32{i32(a, n)}, ri); break;
    case 9: Kokkos::parallel_reduce("sweep.sum_i64", p, SumI64{i64(a, n)}, ri); break;
    case 10:
      Kokkos::parallel_reduce("sweep.sum_f32", p, SumF32{Vec<const float>(reinterpret_cast<float*>(a), n)}, rf);
      break;
    case 11: Kokkos::parallel_reduce("sweep.dot", p, DotF64{Vec<const double>(a, n), Vec<const double>(b, n)}, rd); break;
    case 12:
      Kokkos::parallel_reduce("sweep.max_i64", p, MaxI64{i64(a, n)}, Kokkos::Max<int64_t, Kokkos::MI355XSpace>(ri));
      break;
    case 13: Kokkos::parallel_reduce("sweep.minmaxsum", p, MinMaxSumF64{Vec<const double>(a, n)}, rm); break;
    case 14: Kokkos::parallel_reduce("sweep.polysum32", p, PolySumF64<32>{Vec<const double>(a, n)}, rd); break;
    case 15: Kokkos::parallel_scan("sweep.scan_i32_i64", p, ScanExclI32I64{i32(a, n), i64(b, n)}); break;
    case 16: Kokkos::parallel_scan("sweep.scan_incl_i32", p, ScanInclI32{i32(a, n), i32(b, n)}); break;
    case 17: Kokkos::parallel_scan("sweep.scan_i64", p, ScanExclI64{i64(a, n), i64(b, n)}); break;
    case 18: Kokkos::parallel_scan("sweep.scan_f64", p, ScanExclF64{Vec<const double>(a, n), Vec<double>(b, n)}); break;
    default: Kokkos::parallel_scan("sweep.scan_pair", p, ScanPairI64{i64(a, n), i64(b, n)}); break;
  }
}

// Effective extent a family runs at for a requested extent (team families
// bound the league by the buffer size; MDRange families use a cube or
// square of the nearest side).
// Columns of the gathered vector of the team gather family: 2^20, or a
// quarter of the buffers when they are smaller.
int64_t gather_cols() { return std::min<int64_t>(int64_t(1) << 20, g_buf.n / 4); }

int64_t effective_extent(int f, int64_t n) {
  const int64_t cap = g_buf.n;
  switch (f) {
    case 20: return std::min<int64_t>(n, cap / 256);
    case 21: return std::min<int64_t>(n, (cap - gather_cols()) / 32);
    case 22: return std::min<int64_t>(n, cap / 512);
    case 23: return std::min<int64_t>(n, cap / 128);
    case 24:
    case 26:
    case 27: {
      int64_t sd = std::max<int64_t>(2, int64_t(std::cbrt(double(n))));
      while ((sd + 2) * (sd + 2) * (sd + 2) > cap) --sd;
      return sd * sd * sd;
    }
    case 25: {
      int64_t sd = std::max<int64_t>(2, int64_t(std::sqrt(double(n))));
      while ((sd + 2) * (sd + 2) > cap) --sd;
      return sd * sd;
    }
    default: return std::min<int64_t>(n, cap);
  }
}

// Team shape candidate c (or the policy's AUTO defaults for c < 0).
TeamP team_policy(int64_t league, int c) {
  if (c < 0) return TeamP(int(league), Kokkos::AUTO, Kokkos::AUTO);
  return TeamP(int(league), Kokkos::Impl::kTeamCandidates[c][0], Kokkos::Impl::kTeamCandidates[c][1]);
}

void run_team(int f, int64_t L, int c) {
  double *a = g_buf.a, *b = g_buf.b, *cc = g_buf.c;
  const TeamP tp = team_policy(L, c);
  switch (f) {
    case 20: Kokkos::parallel_for("sweep.team_rowsum", tp, TeamRowSum{Vec<const double>(a, L * 256), Vec<double>(b, L)}); break;
    case 21: {
      const int64_t ncols = gather_cols();
      Kokkos::parallel_for("sweep.team_gather", tp,
                           TeamGather{Vec<const double>(a, L * 32), Vec<const double>(cc, ncols),
                                      Vec<const int32_t>(reinterpret_cast<int32_t*>(b), L * 32),
                                      Vec<double>(cc + ncols, L), ncols});
      break;
    }
    case 22: Kokkos::parallel_for("sweep.team_copy", tp, TeamCopy{Vec<const double>(a, L * 512), Vec<double>(b, L * 512)}); break;
    default:
      Kokkos::parallel_reduce("sweep.team_total", tp, TeamTotal{Vec<const double>(a, L * 128)},
                              Scalar<double>(g_buf.res));
      break;
  }
}

void run_md(int f, int64_t n, int c) {
  double *a = g_buf.a, *b = g_buf.b;
  if (f == 25) {
    const int64_t sd = int64_t(std::llround(std::sqrt(double(n))));
    using MP         = Kokkos::MDRangePolicy<Kokkos::MI355X, Kokkos::Rank<2>, Kokkos::IndexType<int64_t>>;
    MP mp            = c < 0 ? MP({1, 1}, {sd + 1, sd + 1}) : MP({1, 1}, {sd + 1, sd + 1}, {16 << (c % 4), 1 << (c / 4)});
    Kokkos::parallel_for("sweep.md_5pt", mp, Md5Point2D{Grid2(a, sd + 2, sd + 2), Grid2(b, sd + 2, sd + 2)});
    return;
  }
  const int64_t sd = int64_t(std::llround(std::cbrt(double(n))));
  using MP         = Kokkos::MDRangePolicy<Kokkos::MI355X, Kokkos::Rank<3>, Kokkos::IndexType<int64_t>>;
  MP mp = c < 0 ? MP({1, 1, 1}, {sd + 1, sd + 1, sd + 1})
                : MP({1, 1, 1}, {sd + 1, sd + 1, sd + 1}, {16 << (c % 4), 1 << (c / 4), 1});
  const int64_t m = sd + 2;
  switch (f) {
    case 24: Kokkos::parallel_for("sweep.md_7pt", mp, Md7Point{Grid3(a, m, m, m), Grid3(b, m, m, m)}); break;
    case 26: Kokkos::parallel_for("sweep.md_axpy", mp, MdAxpy3{Grid3(a, m, m, m), Grid3(b, m, m, m)}); break;
    default:
      Kokkos::parallel_reduce("sweep.md_sumsq", mp, MdSumSq3{Grid3(a, m, m, m)}, Scalar<double>(g_buf.res));
      break;
  }
}

}  // namespace

extern "C" {

const char* kkmx_sweep_last_error(void) { return g_err.c_str(); }
int kkmx_sweep_family_count(void) { return kFamilies; }

// Features of the last timed batched-dispatch group: pattern, extent (total
// elements), value bytes, functor bytes (mean op extent), floating, bytes
// per element, flops per element -- what BatchDispatch::plan() sends the
// tuning tool for that batch.
int kkmx_sweep_last_features(double* out) {
  if (!out) return -1;
  std::copy(g_last_features, g_last_features + 7, out);
  return 0;
}

// pattern, value bytes, floating, functor bytes, bytes/index, flops/index of a family.
int kkmx_sweep_family_info(int f, int* pattern, int* value_bytes, int* floating, int* functor_bytes, double* bytes,
                           double* flops, int* kind) {
  if (f < 0 || f >= kFamilies) return -1;
  const FamilyInfo i = family_info(f);
  *kind              = i.kind;
  *pattern           = i.pattern;
  *value_bytes       = i.value_bytes;
  *floating          = i.floating;
  *functor_bytes     = i.functor_bytes;
  *bytes             = i.bytes;
  *flops             = i.flops;
  return 0;
}

// Allocates three arrays of max_n doubles (inputs filled with small values
// so every family's arithmetic stays finite).
int kkmx_sweep_setup(int64_t max_n) {
  if (max_n < 1) return -1;
  try {
    if (!Kokkos::is_initialized()) Kokkos::initialize();
    double** bufs[3] = {&g_buf.a, &g_buf.b, &g_buf.c};
    for (auto* q : bufs) KKMX_SAFE_CALL(mx_malloc(reinterpret_cast<void**>(q), size_t(max_n) * 8, nullptr));
    KKMX_SAFE_CALL(mx_malloc(reinterpret_cast<void**>(&g_buf.res), 64, nullptr));
    g_buf.n = max_n;
    Kokkos::parallel_for("sweep.init", P(0, max_n), InitBuffers{Vec<double>(g_buf.a, max_n), Vec<double>(g_buf.b, max_n),
                                                                Vec<double>(g_buf.c, max_n)});
    Kokkos::fence();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int kkmx_sweep_teardown(void) {
  double** bufs[4] = {&g_buf.a, &g_buf.b, &g_buf.c, &g_buf.res};
  for (auto* q : bufs) {
    if (*q) (void)mx_free(*q, nullptr);
    *q = nullptr;
  }
  g_buf.n = 0;
  return 0;
}

// Best-of-`reps` kernel time (ms) of family f at extent n under each of the
// 16 launch-shape candidates (ms_out[16]), after one untimed warm-up each.
// candidate < 0 times all 16; otherwise only that one (ms_out[0]).  heuristic
// = 1 adds the time of the untuned heuristic shape in ms_out[16].
int kkmx_sweep_time(int f, int64_t n, int reps, int candidate, int heuristic, float* ms_out, int64_t* n_eff) {
  if (f < 0 || f >= kFamilies || n < 1 || reps < 1 || !ms_out || g_buf.n < 4096 ||
      (f < 20 && n > g_buf.n)) {
    g_err = "kkmx_sweep_time: bad arguments";
    return -1;
  }
  n = effective_extent(f, n);
  if (n_eff) *n_eff = n;
  try {
    Kokkos::MI355X ex;
    mx_event_t e0, e1;
    KKMX_SAFE_CALL(mx_event_create(&e0));
    KKMX_SAFE_CALL(mx_event_create(&e1));
    const FamilyInfo fi = family_info(f);
    if (fi.kind == 3) {
      // One plan per tile shape; candidates past the shape count are not
      // valid answers and get a 1000 ms sentinel.
      auto time_shape = [&](int shape) {
        Kokkos::Experimental::BatchDispatch bd(ex);
        const int64_t used = batch_ops(f, n, bd);
        if (n_eff) *n_eff = used;
        bd.set_shape(shape);
        bd.plan();
        const auto x = bd.tuning_features();
        const double feat[7] = {double(x.pattern), double(x.extent), double(x.value_bytes), double(x.functor_bytes),
                                x.floating ? 1.0 : 0.0, x.bytes, x.flops};
        std::copy(feat, feat + 7, g_last_features);
        float best = 1e30f;
        bd.submit();
        for (int r = 0; r < reps; ++r) {
          KKMX_SAFE_CALL(mx_event_record(e0, ex.stream()));
          bd.submit();
          KKMX_SAFE_CALL(mx_event_record(e1, ex.stream()));
          KKMX_SAFE_CALL(mx_event_sync(e1));
          float ms = 0;
          KKMX_SAFE_CALL(mx_event_elapsed_ms(&ms, e0, e1));
          best = ms < best ? ms : best;
        }
        return best;
      };
      const int c0 = candidate < 0 ? 0 : candidate, c1 = candidate < 0 ? Kokkos::Impl::kRangeCandidates : candidate + 1;
      for (int c = c0; c < c1; ++c) ms_out[c - c0] = c < Kokkos::Impl::kBatchShapes ? time_shape(c) : 1000.0f;
      if (heuristic) ms_out[c1 - c0] = time_shape(Kokkos::Impl::kBatchDefaultShape);
      KKMX_SAFE_CALL(mx_event_destroy(e0));
      KKMX_SAFE_CALL(mx_event_destroy(e1));
      return 0;
    }
    const int pattern   = fi.pattern;
    auto& ov            = Kokkos::Impl::range_overrides();
    // cand < 0: the untuned default shape.  Range families take the
    // candidate through the launch-shape override, Team and MDRange
    // families through their policy.
    auto run = [&](int cand) {
      if (fi.kind == 1) run_team(f, n, cand);
      else if (fi.kind == 2) run_md(f, n, cand);
      else run_family(f, n);
    };
    auto time_one = [&](int cand) {
      float best = 1e30f;
      run(cand);
      for (int r = 0; r < reps; ++r) {
        KKMX_SAFE_CALL(mx_event_record(e0, ex.stream()));
        run(cand);
        KKMX_SAFE_CALL(mx_event_record(e1, ex.stream()));
        KKMX_SAFE_CALL(mx_event_sync(e1));
        float ms = 0;
        KKMX_SAFE_CALL(mx_event_elapsed_ms(&ms, e0, e1));
        best = ms < best ? ms : best;
      }
      return best;
    };
    const int c0 = candidate < 0 ? 0 : candidate, c1 = candidate < 0 ? Kokkos::Impl::kRangeCandidates : candidate + 1;
    for (int c = c0; c < c1; ++c) {
      if (fi.kind == 0) {
        ov.cfg[pattern] = Kokkos::Impl::range_candidate(pattern, c, n);
        ov.set[pattern] = true;
      }
      ms_out[c - c0] = time_one(c);
    }
    ov.set[pattern] = false;
    if (heuristic) ms_out[c1 - c0] = time_one(-1);
    KKMX_SAFE_CALL(mx_event_destroy(e0));
    KKMX_SAFE_CALL(mx_event_destroy(e1));
    return 0;
  } catch (const std::exception& e) {
    Kokkos::Impl::range_overrides().set[0] = Kokkos::Impl::range_overrides().set[1] =
        Kokkos::Impl::range_overrides().set[2] = false;
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"
