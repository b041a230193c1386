// Restated unit tests of the reference (core/unit_test/*.hpp), compiled
// against the MI355X execution space as ordinary Kokkos user code.
//
// Each test is a host function that dispatches kernels and checks closed-form
// answers; a failed check throws with the file:line and the values.  The
// tests are exposed through a small C-ABI (kkmx_selftest_*) so pytest can
// run them one by one on the GPU.  Closed forms and invariants are the ones
// the reference's tests assert; each test names the reference test it follows.
#include <Kokkos_Core.hpp>
#include <Kokkos_StdAlgorithms.hpp>
#include <Kokkos_ScatterView.hpp>
#include <Kokkos_Sort.hpp>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <cmath>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Entry {
  const char* name;
  void (*fn)();
};
std::vector<Entry>& registry() {
  static std::vector<Entry> r;
  return r;
}
struct Registrar {
  Registrar(const char* n, void (*f)()) { registry().push_back({n, f}); }
};

#define KKMX_TEST(NAME)                                 \
  static void NAME();                                   \
  static Registrar KOKKOS_IMPL_CONCAT(reg_, NAME)(#NAME, NAME); \
  static void NAME()

template <class A, class B>
void check_eq(const A& a, const B& b, const char* ea, const char* eb, const char* file, int line) {
  if (!(a == b)) {
    std::ostringstream os;
    os << file << ":" << line << ": " << ea << " == " << eb << " failed: " << a << " vs " << b;
    throw Failure(os.str());
  }
}
#define CHECK_EQ(a, b) check_eq((a), (b), #a, #b, __FILE__, __LINE__)
#define CHECK(c)                                                                                  \
  do {                                                                                            \
    if (!(c)) throw Failure(std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " #c); \
  } while (0)

using Exec = Kokkos::MI355X;

template <class V>
typename V::non_const_value_type read0(const Exec& ex, const V& v) {
  typename V::non_const_value_type x;
  Kokkos::deep_copy(ex, x, v);
  ex.fence();
  return x;
}

// =========================================================================
// Graph (core/unit_test/TestGraph.hpp:27-420)
// =========================================================================
template <class T>
using AtomicScalar = Kokkos::View<T, Exec, Kokkos::MemoryTraits<Kokkos::Atomic>>;

struct CountTestFunctor {
  using value_type = int;
  AtomicScalar<int> count;
  AtomicScalar<int> bugs;
  int expected_count_min;
  int expected_count_max;
  template <class... Ts>
  KOKKOS_FUNCTION void operator()(Ts&&...) const noexcept {
    bugs() += int(count() > expected_count_max || count() < expected_count_min);
    count()++;
  }
};
struct SetViewToValueFunctor {
  using value_type = int;
  AtomicScalar<int> v;
  int value;
  template <class... Ts>
  KOKKOS_FUNCTION void operator()(Ts&&...) const noexcept {
    v() = value;
  }
};
struct SetResultToViewFunctor {
  using value_type = int;
  AtomicScalar<int> v;
  template <class U>
  KOKKOS_FUNCTION void operator()(U&&, value_type& val) const noexcept {
    val += v();
  }
};
// The reference aborts if a zero-work reduction calls the functor; here a
// call is counted instead so the check fails without trapping the GPU.
struct NoOpReduceFunctor {
  AtomicScalar<int> bugs;
  KOKKOS_FUNCTION void operator()(int, int&) const { bugs() += 1; }
  KOKKOS_FUNCTION void operator()(int, int, int&) const { bugs() += 1; }
  KOKKOS_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type&, int&) const { bugs() += 1; }
};

struct GraphFixture {
  Exec ex{};
  AtomicScalar<int> count{Kokkos::view_alloc("count", ex)};
  AtomicScalar<int> bugs{Kokkos::view_alloc("bugs", ex)};
  GraphFixture() {
    Kokkos::deep_copy(ex, count, 0);
    Kokkos::deep_copy(ex, bugs, 0);
    ex.fence();
  }
};

KKMX_TEST(graph_submit_once) {
  GraphFixture t;
  auto graph = Kokkos::Experimental::create_graph<Exec>([&](auto root) {
    root.then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 0});
  });
  graph.submit();
  CHECK_EQ(read0(graph.get_execution_space(), t.count), 1);
  CHECK_EQ(read0(graph.get_execution_space(), t.bugs), 0);
}

KKMX_TEST(graph_submit_once_rvalue) {
  GraphFixture t;
  Kokkos::Experimental::create_graph(t.ex, [&](auto root) {
    root.then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 0});
  }).submit();
  CHECK_EQ(read0(t.ex, t.count), 1);
  CHECK_EQ(read0(t.ex, t.bugs), 0);
}

KKMX_TEST(graph_instantiate_and_submit_once) {
  GraphFixture t;
  auto graph = Kokkos::Experimental::create_graph(t.ex, [&](auto root) {
    root.then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 0});
  });
  graph.instantiate();
  graph.submit();
  CHECK_EQ(read0(t.ex, t.count), 1);
  CHECK_EQ(read0(t.ex, t.bugs), 0);
}

// Submitting on another instance than the one the graph was built with
// (TestGraph.hpp:203-230).
KKMX_TEST(graph_submit_onto_another_instance) {
  GraphFixture t;
  mx_stream_t s;
  KKMX_SAFE_CALL(mx_stream_create(&s));
  {
    Exec other(s, true);
    auto graph = Kokkos::Experimental::create_graph(t.ex, [&](auto root) {
      root.then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 0});
    });
    graph.instantiate();
    t.ex.fence();
    graph.submit(other);
    CHECK_EQ(read0(other, t.count), 1);
    CHECK_EQ(read0(other, t.bugs), 0);
  }
}

KKMX_TEST(graph_create_graph_no_closure_with_exec) {
  GraphFixture t;
  Kokkos::Experimental::Graph<Exec> graph{t.ex};
  auto root = graph.root_node();
  root.then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 0});
  graph.submit(t.ex);
  CHECK_EQ(read0(t.ex, t.count), 1);
  CHECK_EQ(read0(t.ex, t.bugs), 0);
}

KKMX_TEST(graph_submit_six) {
  GraphFixture t;
  auto graph = Kokkos::Experimental::create_graph(t.ex, [&](auto root) {
    auto f_setup_count = root.then_parallel_for(1, SetViewToValueFunctor{t.count, 0});
    auto f_setup_bugs  = root.then_parallel_for(1, SetViewToValueFunctor{t.bugs, 0});
    auto ready         = Kokkos::Experimental::when_all(f_setup_count, f_setup_bugs);
    ready.then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 6});
    ready.then_parallel_for(Kokkos::RangePolicy<Exec>{0, 1}, CountTestFunctor{t.count, t.bugs, 0, 6});
    ready.then_parallel_for(Kokkos::MDRangePolicy<Exec, Kokkos::Rank<2>>{{0, 0}, {1, 1}},
                            CountTestFunctor{t.count, t.bugs, 0, 6});
    ready.then_parallel_for(Kokkos::TeamPolicy<Exec>{1, 1}, CountTestFunctor{t.count, t.bugs, 0, 6});
    ready.then_parallel_for(2, CountTestFunctor{t.count, t.bugs, 0, 6});
  });
  graph.submit();
  CHECK_EQ(read0(t.ex, t.count), 6);
  CHECK_EQ(read0(t.ex, t.bugs), 0);
}

KKMX_TEST(graph_when_all_cycle) {
  GraphFixture t;
  Kokkos::View<int, Exec> reduction_out{"reduction_out"};
  Kokkos::Experimental::create_graph(t.ex, [&](auto root) {
    auto f1 = root.then_parallel_for(1, SetViewToValueFunctor{t.count, 0});
    auto f2 = f1.then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 0});
    auto f3 = f2.then_parallel_for(5, CountTestFunctor{t.count, t.bugs, 1, 5});
    auto f4 = Kokkos::Experimental::when_all(f2, f3).then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 6, 6});
    Kokkos::Experimental::when_all(f1, f4, f3).then_parallel_reduce(6, SetResultToViewFunctor{t.count}, reduction_out);
  }).submit();
  CHECK_EQ(read0(t.ex, t.bugs), 0);
  CHECK_EQ(read0(t.ex, t.count), 7);
  CHECK_EQ(read0(t.ex, reduction_out), 42);
}

// repeat_chain (TestGraph.hpp:333-372) with device-resident results: the
// MI355X cannot write HostSpace, so the results are copied back after each
// replay instead.
KKMX_TEST(graph_repeat_chain) {
  GraphFixture t;
  Kokkos::View<int, Exec> count_out{"count_out"}, bugs_out{"bugs_out"};
  auto graph = Kokkos::Experimental::create_graph(t.ex, [&](auto root) {
    root.then_parallel_for(1, SetViewToValueFunctor{t.count, 0})
        .then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 0, 0})
        .then_parallel_for(1, CountTestFunctor{t.count, t.bugs, 1, 1})
        .then_parallel_reduce(1, SetResultToViewFunctor{t.count}, count_out)
        .then_parallel_reduce(1, SetResultToViewFunctor{t.bugs}, Kokkos::Sum<int, Kokkos::MI355XSpace>{bugs_out});
  });
  for (int i = 0; i < 10; ++i) {
    graph.submit();
    t.ex.fence();
    CHECK_EQ(read0(t.ex, count_out), 2);
    CHECK_EQ(read0(t.ex, bugs_out), 0);
  }
}

KKMX_TEST(graph_zero_work_reduce) {
  GraphFixture t;
  auto graph = Kokkos::Experimental::create_graph(t.ex, [&](Kokkos::Experimental::GraphNodeRef<Exec> root) {
    NoOpReduceFunctor no_op{t.bugs};
    root.then_parallel_reduce(Kokkos::RangePolicy<Exec>(0, 0), no_op, t.count)
        .then_parallel_reduce(Kokkos::MDRangePolicy<Exec, Kokkos::Rank<2>>{{0, 0}, {0, 0}}, no_op, t.count)
        .then_parallel_reduce(Kokkos::TeamPolicy<Exec>{0, Kokkos::AUTO}, no_op, t.count);
  });
  graph.submit();
  Kokkos::deep_copy(t.ex, t.count, 1);
  graph.submit();
  CHECK_EQ(read0(t.ex, t.count), 0);
  CHECK_EQ(read0(t.ex, t.bugs), 0);
}

KKMX_TEST(graph_empty_graph) {
  Exec ex;
  auto graph = Kokkos::Experimental::create_graph(ex, [](auto) {});
  graph.instantiate();
  graph.submit(ex);
  ex.fence();
}

// A DAG of a fill, a scan, a reduce of the scan and a dependent for, replayed
// many times: every replay must equal the eager result (scan state and
// reduction counters are self-cleaning, so replays need no resets).
struct FillIota {
  Kokkos::View<int64_t*, Exec> x;
  int64_t mult;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { x(i) = (i % 13) * mult - 5; }
};
struct ExclScan {
  Kokkos::View<int64_t*, Exec> x, y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& upd, const bool final) const {
    if (final) y(i) = upd;
    upd += x(i);
  }
};
struct SumOf {
  Kokkos::View<int64_t*, Exec> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& s) const { s += y(i) ^ i; }
};
struct MaxOf {
  Kokkos::View<int64_t*, Exec> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int64_t& m) const {
    if (x(i) * 3 + i > m) m = x(i) * 3 + i;
  }
};
struct CombineInto {
  Kokkos::View<int64_t, Exec> s, m;
  Kokkos::View<int64_t[2], Exec> out;
  KOKKOS_INLINE_FUNCTION void operator()() const {
    out(0) = s();
    out(1) = m() - s();
  }
};

KKMX_TEST(graph_dag_replay_matches_eager) {
  Exec ex;
  const int64_t n = 3000017;
  Kokkos::View<int64_t*, Exec> x("x", n), y("y", n);
  Kokkos::View<int64_t, Exec> s("s"), m("m");
  Kokkos::View<int64_t[2], Exec> out("out");
  // eager
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(ex, 0, n), FillIota{x, 7});
  Kokkos::parallel_scan(Kokkos::RangePolicy<Exec>(ex, 0, n), ExclScan{x, y});
  Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec>(ex, 0, n), SumOf{y}, s);
  Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec>(ex, 0, n), MaxOf{x}, Kokkos::Max<int64_t, Kokkos::MI355XSpace>(m));
  int64_t s_eager = read0(ex, s), m_eager = read0(ex, m);
  auto y_eager = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), y);
  Kokkos::deep_copy(ex, y, 0);
  Kokkos::deep_copy(ex, s, 0);
  ex.fence();
  auto graph = Kokkos::Experimental::create_graph(ex, [&](auto root) {
    auto f = root.then_parallel_for("fill", Kokkos::RangePolicy<Exec>(0, n), FillIota{x, 7});
    auto sc = f.then_parallel_scan("scan", Kokkos::RangePolicy<Exec>(0, n), ExclScan{x, y});
    auto r1 = sc.then_parallel_reduce("sum", Kokkos::RangePolicy<Exec>(0, n), SumOf{y}, s);
    auto r2 = f.then_parallel_reduce("max", Kokkos::RangePolicy<Exec>(0, n), MaxOf{x},
                                     Kokkos::Max<int64_t, Kokkos::MI355XSpace>(m));
    Kokkos::Experimental::when_all(r1, r2).then("combine", CombineInto{s, m, out});
  });
  for (int rep = 0; rep < 50; ++rep) {
    graph.submit();
    ex.fence();
    CHECK_EQ(read0(ex, s), s_eager);
    CHECK_EQ(read0(ex, m), m_eager);
  }
  auto y_graph = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), y);
  for (int64_t i = 0; i < n; ++i)
    if (y_graph(i) != y_eager(i)) CHECK_EQ(y_graph(i), y_eager(i));
  auto o = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
  CHECK_EQ(o(0), s_eager);
  CHECK_EQ(o(1), m_eager - s_eager);
  size_t nodes = 0;
  KKMX_SAFE_CALL(mx_graph_node_count(graph.native_graph(), &nodes));
  CHECK(nodes >= 5);
}

// The same DAG on the Serial space runs nodes in order at submit.
KKMX_TEST(graph_serial_backend) {
  Kokkos::Serial sx;
  Kokkos::View<int, Kokkos::HostSpace> v("v"), r("r");
  auto graph = Kokkos::Experimental::create_graph(sx, [&](auto root) {
    auto a = root.then_parallel_for(Kokkos::RangePolicy<Kokkos::Serial>(0, 10), [=](int i) { v() += i; });
    a.then_parallel_reduce(Kokkos::RangePolicy<Kokkos::Serial>(0, 4), [=](int i, int& s) { s += v() + i; }, r);
  });
  graph.submit();
  graph.submit();
  CHECK_EQ(v(), 90);
  CHECK_EQ(r(), 4 * 90 + 6);
}

// =========================================================================
// Reduce (core/unit_test/TestReduce.hpp:27-260,520-700)
// =========================================================================
template <class ScalarType>
struct ReduceFunctor {
  using execution_space = Exec;
  using size_type       = Exec::size_type;
  struct value_type {
    ScalarType value[3];
  };
  const size_type nwork;
  ReduceFunctor(size_type n) : nwork(n) {}
  KOKKOS_INLINE_FUNCTION void join(value_type& dst, const value_type& src) const {
    dst.value[0] += src.value[0];
    dst.value[1] += src.value[1];
    dst.value[2] += src.value[2];
  }
  KOKKOS_INLINE_FUNCTION void operator()(size_type iwork, value_type& dst) const {
    dst.value[0] += 1;
    dst.value[1] += iwork + 1;
    dst.value[2] += nwork - iwork;
  }
};
struct ReduceFunctorFinal : ReduceFunctor<int64_t> {
  using value_type = ReduceFunctor<int64_t>::value_type;
  ReduceFunctorFinal(size_t n) : ReduceFunctor<int64_t>(n) {}
  KOKKOS_INLINE_FUNCTION void final(value_type& dst) const {
    dst.value[0] = -dst.value[0];
    dst.value[1] = -dst.value[1];
    dst.value[2] = -dst.value[2];
  }
};
struct ReducerTag {};
struct ReduceFunctorFinalTag {
  using execution_space = Exec;
  using size_type       = Exec::size_type;
  struct value_type {
    int64_t value[3];
  };
  const size_type nwork;
  ReduceFunctorFinalTag(size_type n) : nwork(n) {}
  KOKKOS_INLINE_FUNCTION void join(ReducerTag, value_type& dst, const value_type& src) const {
    dst.value[0] += src.value[0];
    dst.value[1] += src.value[1];
    dst.value[2] += src.value[2];
  }
  KOKKOS_INLINE_FUNCTION void operator()(ReducerTag, size_type iwork, value_type& dst) const {
    dst.value[0] -= 1;
    dst.value[1] -= iwork + 1;
    dst.value[2] -= nwork - iwork;
  }
  KOKKOS_INLINE_FUNCTION void final(ReducerTag, value_type& dst) const {
    ++dst.value[0];
    ++dst.value[1];
    ++dst.value[2];
  }
};

template <class ScalarType>
void reduce_struct_value(uint64_t nw) {
  const uint64_t nsum = nw % 2 ? nw * ((nw + 1) / 2) : (nw / 2) * (nw + 1);
  for (int rep = 0; rep < 4; ++rep) {
    typename ReduceFunctor<ScalarType>::value_type r;
    Kokkos::parallel_reduce(nw, ReduceFunctor<ScalarType>(nw), r);
    for (int j = 0; j < 3; ++j) CHECK_EQ(r.value[j], (ScalarType)(j == 0 ? nw : nsum));
    ReduceFunctorFinal::value_type rf;
    if (rep % 2) Kokkos::parallel_reduce(nw, ReduceFunctorFinal(nw), rf);
    else Kokkos::parallel_reduce("Reduce", nw, ReduceFunctorFinal(nw), rf);
    for (int j = 0; j < 3; ++j) CHECK_EQ(-rf.value[j], (int64_t)(j == 0 ? nw : nsum));
    ReduceFunctorFinalTag::value_type rt;
    Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec, ReducerTag>(0, nw), ReduceFunctorFinalTag(nw), rt);
    for (int j = 0; j < 3; ++j) CHECK_EQ(rt.value[j], 1 - (int64_t)(j == 0 ? nw : nsum));
  }
}
KKMX_TEST(reduce_int64_struct_value) {
  reduce_struct_value<int64_t>(0);
  reduce_struct_value<int64_t>(1000000);
}
KKMX_TEST(reduce_double_struct_value) {
  reduce_struct_value<double>(0);
  reduce_struct_value<double>(1000000);
}

template <class ScalarType>
struct RuntimeReduceFunctor {
  using execution_space = Exec;
  using value_type      = ScalarType[];
  using size_type       = Exec::size_type;
  const unsigned value_count;
  const size_type nwork;
  RuntimeReduceFunctor(size_type n, unsigned c) : value_count(c), nwork(n) {}
  KOKKOS_INLINE_FUNCTION void init(ScalarType dst[]) const {
    for (unsigned i = 0; i < value_count; ++i) dst[i] = 0;
  }
  KOKKOS_INLINE_FUNCTION void join(ScalarType dst[], const ScalarType src[]) const {
    for (unsigned i = 0; i < value_count; ++i) dst[i] += src[i];
  }
  KOKKOS_INLINE_FUNCTION void operator()(size_type iwork, ScalarType dst[]) const {
    const size_type tmp[3] = {1, iwork + 1, nwork - iwork};
    for (size_type i = 0; i < static_cast<size_type>(value_count); ++i) dst[i] += tmp[i % 3];
  }
};
template <class ScalarType>
struct RuntimeReduceMinMax {
  using execution_space = Exec;
  using value_type      = ScalarType[];
  using size_type       = Exec::size_type;
  const unsigned value_count;
  const size_type nwork;
  const ScalarType amin, amax;
  RuntimeReduceMinMax(size_type n, unsigned c)
      : value_count(c), nwork(n), amin(std::numeric_limits<ScalarType>::min()),
        amax(std::numeric_limits<ScalarType>::max()) {}
  KOKKOS_INLINE_FUNCTION void init(ScalarType dst[]) const {
    for (unsigned i = 0; i < value_count; ++i) dst[i] = i % 2 ? amax : amin;
  }
  KOKKOS_INLINE_FUNCTION void join(ScalarType dst[], const ScalarType src[]) const {
    for (unsigned i = 0; i < value_count; ++i)
      dst[i] = i % 2 ? (dst[i] < src[i] ? dst[i] : src[i]) : (dst[i] > src[i] ? dst[i] : src[i]);
  }
  KOKKOS_INLINE_FUNCTION void operator()(size_type iwork, ScalarType dst[]) const {
    const ScalarType tmp[2] = {ScalarType(iwork + 1), ScalarType(nwork - iwork)};
    for (size_type i = 0; i < static_cast<size_type>(value_count); ++i)
      dst[i] = i % 2 ? (dst[i] < tmp[i % 2] ? dst[i] : tmp[i % 2]) : (dst[i] > tmp[i % 2] ? dst[i] : tmp[i % 2]);
  }
};
struct RuntimeReduceFunctorFinal : RuntimeReduceFunctor<int64_t> {
  using value_type = int64_t[];
  RuntimeReduceFunctorFinal(size_t n, unsigned c) : RuntimeReduceFunctor<int64_t>(n, c) {}
  KOKKOS_INLINE_FUNCTION void final(int64_t dst[]) const {
    for (unsigned i = 0; i < value_count; ++i) dst[i] = -dst[i];
  }
};

// TestReduceDynamic (TestReduce.hpp:381-470).
template <class ScalarType>
void reduce_dynamic(uint64_t nw) {
  const uint64_t nsum = nw % 2 ? nw * ((nw + 1) / 2) : (nw / 2) * (nw + 1);
  const ScalarType amin = nw ? ScalarType(1) : std::numeric_limits<ScalarType>::max();
  const ScalarType amax = nw ? ScalarType(nw) : std::numeric_limits<ScalarType>::min();
  for (unsigned count = 1; count <= 9; count += 2) {
    std::vector<ScalarType> r(count), rmm(count);
    std::vector<int64_t> rf(count);
    Kokkos::parallel_reduce(nw, RuntimeReduceFunctor<ScalarType>(nw, count), r.data());
    Kokkos::parallel_reduce(nw, RuntimeReduceMinMax<ScalarType>(nw, count), rmm.data());
    Kokkos::parallel_reduce(nw, RuntimeReduceFunctorFinal(nw, count), rf.data());
    for (unsigned j = 0; j < count; ++j) {
      const uint64_t correct = 0 == j % 3 ? nw : nsum;
      CHECK_EQ(r[j], (ScalarType)correct);
      CHECK_EQ(rf[j], -(int64_t)correct);
      CHECK_EQ(rmm[j], (j % 2) ? amin : amax);
    }
  }
}
KKMX_TEST(reduce_dynamic_int64) {
  reduce_dynamic<int64_t>(0);
  reduce_dynamic<int64_t>(1000000);
}
KKMX_TEST(reduce_dynamic_double) {
  reduce_dynamic<double>(0);
  reduce_dynamic<double>(1000000);
}
// TestReduceDynamicView (TestReduce.hpp:472-518): the result is a View.
KKMX_TEST(reduce_dynamic_view) {
  for (uint64_t nw : {uint64_t(0), uint64_t(1000000)}) {
    const uint64_t nsum = nw % 2 ? nw * ((nw + 1) / 2) : (nw / 2) * (nw + 1);
    const unsigned count = 7;
    Kokkos::View<int64_t*, Kokkos::HostSpace> host("host", count);
    Kokkos::View<int64_t*, Exec> dev("dev", count);
    Kokkos::parallel_reduce(nw, RuntimeReduceFunctor<int64_t>(nw, count), host);
    Kokkos::parallel_reduce(nw, RuntimeReduceFunctor<int64_t>(nw, count), dev);
    Kokkos::fence();
    auto hd = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), dev);
    for (unsigned j = 0; j < count; ++j) {
      const uint64_t correct = 0 == j % 3 ? nw : nsum;
      CHECK_EQ(host(j), (int64_t)correct);
      CHECK_EQ(hd(j), (int64_t)correct);
    }
  }
}

struct CombinedReduceFunctorSameType {
  using execution_space = Exec;
  using size_type       = Exec::size_type;
  const size_type nwork;
  explicit CombinedReduceFunctorSameType(size_type n) : nwork(n) {}
  KOKKOS_INLINE_FUNCTION void operator()(size_type iwork, int64_t& d1, int64_t& d2, int64_t& d3) const {
    d1 += 1;
    d2 += iwork + 1;
    d3 += nwork - iwork;
  }
  KOKKOS_INLINE_FUNCTION void operator()(size_type iwork, size_type z1, size_type z2, int64_t& d1, int64_t& d2,
                                         int64_t& d3) const {
    d1 += 1 + z1;
    d2 += iwork + 1 + z2;
    d3 += nwork - iwork;
  }
};
KKMX_TEST(reduce_int_combined) {
  constexpr uint64_t nw = 1000, nsum = (nw / 2) * (nw + 1);
  int64_t r1 = 0, r2 = 0, r3 = 0;
  Kokkos::parallel_reduce("int_combined_reduce", Kokkos::RangePolicy<Exec>(0, nw),
                          CombinedReduceFunctorSameType(nw), r1, r2, r3);
  CHECK_EQ(uint64_t(r1), nw);
  CHECK_EQ(uint64_t(r2), nsum);
  CHECK_EQ(uint64_t(r3), nsum);
}
KKMX_TEST(reduce_mdrange_combined) {
  constexpr uint64_t nw = 1000, nsum = (nw / 2) * (nw + 1);
  int64_t r1 = 0, r2 = 0, r3 = 0;
  Kokkos::parallel_reduce("int_combined_reduce_mdrange",
                          Kokkos::MDRangePolicy<Exec, Kokkos::Rank<3>>({{0, 0, 0}}, {{int64_t(nw), 1, 1}}),
                          CombinedReduceFunctorSameType(nw), r1, r2, r3);
  CHECK_EQ(uint64_t(r1), nw);
  CHECK_EQ(uint64_t(r2), nsum);
  CHECK_EQ(uint64_t(r3), nsum);
  r1 = r2 = r3 = 0;
  Kokkos::parallel_reduce(
      "int_combined_reduce_mdrange",
      Kokkos::MDRangePolicy<Exec, Kokkos::Rank<3>, Kokkos::Schedule<Kokkos::Dynamic>>({{0, 0, 0}}, {{int64_t(nw), 1, 1}}),
      CombinedReduceFunctorSameType(nw), r1, r2, r3);
  CHECK_EQ(uint64_t(r1), nw);
  CHECK_EQ(uint64_t(r2), nsum);
  CHECK_EQ(uint64_t(r3), nsum);
}
KKMX_TEST(reduce_int_combined_mixed) {
  constexpr uint64_t nw = 1000, nsum = (nw / 2) * (nw + 1);
  {
    Kokkos::View<int64_t, Kokkos::HostSpace> r1("r1"), r3("r3");
    int64_t r2 = 0;
    Kokkos::parallel_reduce("mixed", Kokkos::RangePolicy<Exec>(0, nw), CombinedReduceFunctorSameType(nw), r1, r2,
                            Kokkos::Sum<int64_t, Kokkos::HostSpace>{r3});
    CHECK_EQ(r1(), int64_t(nw));
    CHECK_EQ(r2, int64_t(nsum));
    CHECK_EQ(r3(), int64_t(nsum));
  }
  {
    Kokkos::View<int64_t, Kokkos::MI355XSpace> r1("r1"), r3("r3");
    int64_t r2 = 0;
    Kokkos::parallel_reduce("mixed", Kokkos::RangePolicy<Exec>(0, nw), CombinedReduceFunctorSameType(nw), r1, r2,
                            Kokkos::Sum<int64_t, Kokkos::MI355XSpace>{r3});
    int64_t h1, h3;
    Kokkos::deep_copy(h1, r1);
    Kokkos::deep_copy(h3, r3);
    CHECK_EQ(h1, int64_t(nw));
    CHECK_EQ(r2, int64_t(nsum));
    CHECK_EQ(h3, int64_t(nsum));
  }
}

// reduction_with_large_iteration_count (TestReduce.hpp:656-680): 2^39-2^8+1
// ones summed in double are exact.
struct OnesF64 {
  KOKKOS_FUNCTION void operator()(const int64_t, double& u) const { u += 1.0; }
};
KKMX_TEST(reduce_large_iteration_count) {
  const int64_t N = (int64_t(1) << 39) - (int64_t(1) << 8) + 1;
  double nu = 0;
  Kokkos::parallel_reduce("sample reduction", Kokkos::RangePolicy<Exec, Kokkos::IndexType<int64_t>>(0, N), OnesF64(),
                          nu);
  CHECK_EQ(nu, double(N));
}

// =========================================================================
// Reducers (core/unit_test/TestReducers.hpp): each reducer over a known array
// against a host fold; Loc reducers keep the first index among ties.
// =========================================================================
template <class T>
struct ValuesOf {
  Kokkos::View<T*, Exec> v;
};
template <class T, class R>
struct ReduceWith {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, typename R::value_type& u) const;
};
template <class T>
Kokkos::View<T*, Exec> upload(const std::vector<T>& h) {
  Kokkos::View<T*, Exec> d("d", h.size());
  Kokkos::View<const T*, Kokkos::HostSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>> hv(h.data(), h.size());
  Kokkos::deep_copy(d, hv);
  return d;
}
template <class T>
struct SumF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const { u += v(i); }
};
template <class T>
struct ProdF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const { u *= v(i); }
};
template <class T>
struct MinF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const {
    if (v(i) < u) u = v(i);
  }
};
template <class T>
struct MaxF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const {
    if (v(i) > u) u = v(i);
  }
};
template <class T>
struct LAndF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const { u = u && v(i); }
};
template <class T>
struct LOrF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const { u = u || v(i); }
};
template <class T>
struct BAndF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const { u = u & v(i); }
};
template <class T>
struct BOrF {
  Kokkos::View<T*, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, T& u) const { u = u | v(i); }
};
template <class T>
struct MinLocF {
  Kokkos::View<T*, Exec> v;
  using VT = Kokkos::ValLocScalar<T, int>;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, VT& u) const {
    if (v(i) < u.val) {
      u.val = v(i);
      u.loc = i;
    }
  }
};
template <class T>
struct MaxLocF {
  Kokkos::View<T*, Exec> v;
  using VT = Kokkos::ValLocScalar<T, int>;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, VT& u) const {
    if (v(i) > u.val) {
      u.val = v(i);
      u.loc = i;
    }
  }
};
template <class T>
struct MinMaxF {
  Kokkos::View<T*, Exec> v;
  using VT = Kokkos::MinMaxScalar<T>;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, VT& u) const {
    if (v(i) < u.min_val) u.min_val = v(i);
    if (v(i) > u.max_val) u.max_val = v(i);
  }
};
template <class T>
struct MinMaxLocF {
  Kokkos::View<T*, Exec> v;
  using VT = Kokkos::MinMaxLocScalar<T, int>;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, VT& u) const {
    if (v(i) < u.min_val) {
      u.min_val = v(i);
      u.min_loc = i;
    }
    if (v(i) > u.max_val) {
      u.max_val = v(i);
      u.max_loc = i;
    }
  }
};

template <class T>
void reducers_for(const std::vector<T>& h) {
  const int n = static_cast<int>(h.size());
  auto d      = upload(h);
  using P     = Kokkos::RangePolicy<Exec>;
  T sum = 0, prod = 1, mn = std::numeric_limits<T>::max(), mx = std::numeric_limits<T>::lowest();
  int mnl = -1, mxl = -1;
  for (int i = 0; i < n; ++i) {
    sum += h[i];
    prod *= (i < 20 ? h[i] : T(1));
    if (h[i] < mn) { mn = h[i]; mnl = i; }
    if (h[i] > mx) { mx = h[i]; mxl = i; }
  }
  {
    T r = 0;
    Kokkos::parallel_reduce(P(0, n), SumF<T>{d}, Kokkos::Sum<T>(r));
    if (std::is_integral<T>::value) CHECK_EQ(r, sum);
    else CHECK(std::abs(double(r - sum)) <= 1e-12 * std::abs(double(sum)) + 1e-12);
  }
  {
    T r = 0;
    Kokkos::parallel_reduce(P(0, std::min(n, 20)), ProdF<T>{d}, Kokkos::Prod<T>(r));
    if (std::is_integral<T>::value) CHECK_EQ(r, prod);
    else CHECK(std::abs(double(r - prod)) <= 1e-12 * std::abs(double(prod)));
  }
  {
    T r = 0;
    Kokkos::parallel_reduce(P(0, n), MinF<T>{d}, Kokkos::Min<T>(r));
    CHECK_EQ(r, mn);
    Kokkos::parallel_reduce(P(0, n), MaxF<T>{d}, Kokkos::Max<T>(r));
    CHECK_EQ(r, mx);
  }
  {
    typename MinLocF<T>::VT r;
    Kokkos::parallel_reduce(P(0, n), MinLocF<T>{d}, Kokkos::MinLoc<T, int>(r));
    CHECK_EQ(r.val, mn);
    CHECK_EQ(r.loc, mnl);
    Kokkos::parallel_reduce(P(0, n), MaxLocF<T>{d}, Kokkos::MaxLoc<T, int>(r));
    CHECK_EQ(r.val, mx);
    CHECK_EQ(r.loc, mxl);
  }
  {
    typename MinMaxF<T>::VT r;
    Kokkos::parallel_reduce(P(0, n), MinMaxF<T>{d}, Kokkos::MinMax<T>(r));
    CHECK_EQ(r.min_val, mn);
    CHECK_EQ(r.max_val, mx);
    typename MinMaxLocF<T>::VT rl;
    Kokkos::parallel_reduce(P(0, n), MinMaxLocF<T>{d}, Kokkos::MinMaxLoc<T, int>(rl));
    CHECK_EQ(rl.min_val, mn);
    CHECK_EQ(rl.max_val, mx);
    CHECK_EQ(rl.min_loc, mnl);
    CHECK_EQ(rl.max_loc, mxl);
  }
  // Empty range: every reducer returns its identity.
  {
    T r = 7;
    Kokkos::parallel_reduce(P(0, 0), MinF<T>{d}, Kokkos::Min<T>(r));
    CHECK_EQ(r, Kokkos::reduction_identity<T>::min());
    Kokkos::parallel_reduce(P(0, 0), MaxF<T>{d}, Kokkos::Max<T>(r));
    CHECK_EQ(r, Kokkos::reduction_identity<T>::max());
    Kokkos::parallel_reduce(P(0, 0), ProdF<T>{d}, Kokkos::Prod<T>(r));
    CHECK_EQ(r, T(1));
  }
}

KKMX_TEST(reducers_int) {
  std::vector<int> h(100003);
  uint64_t s = 12345;
  for (auto& x : h) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    x = int((s >> 33) % 100001) - 50000;
  }
  // ties for min and max at several places: the first index wins
  h[5] = -60000; h[777] = -60000; h[90000] = -60000;
  h[11] = 60000; h[12] = 60000; h[99999] = 60000;
  for (int i = 0; i < 20; ++i) h[i] = (i % 3) + 1;
  h[5] = -60000; h[11] = 60000; h[12] = 60000;
  reducers_for<int>(h);
}
KKMX_TEST(reducers_int64) {
  std::vector<int64_t> h(300007);
  uint64_t s = 777;
  for (auto& x : h) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    x = int64_t(s >> 20) - (int64_t(1) << 43);
  }
  for (int i = 0; i < 20; ++i) h[i] = (i % 2) ? 3 : -2;
  h[1000] = std::numeric_limits<int64_t>::min() / 4;
  h[200000] = std::numeric_limits<int64_t>::min() / 4;
  h[3000] = std::numeric_limits<int64_t>::max() / 4;
  reducers_for<int64_t>(h);
}
KKMX_TEST(reducers_double) {
  std::vector<double> h(65537);
  uint64_t s = 99;
  for (auto& x : h) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    x = double(s >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  }
  for (int i = 0; i < 20; ++i) h[i] = 0.5 + 0.25 * (i % 4);
  h[4096] = -3.0; h[60000] = -3.0;
  h[7] = 3.0; h[65536] = 3.0;
  reducers_for<double>(h);
}
KKMX_TEST(reducers_logical_bitwise) {
  const int n = 50000;
  std::vector<int> h(n, 1);
  auto d = upload(h);
  using P = Kokkos::RangePolicy<Exec>;
  int r = 0;
  Kokkos::parallel_reduce(P(0, n), LAndF<int>{d}, Kokkos::LAnd<int>(r));
  CHECK_EQ(r, 1);
  h[31337] = 0;
  d = upload(h);
  Kokkos::parallel_reduce(P(0, n), LAndF<int>{d}, Kokkos::LAnd<int>(r));
  CHECK_EQ(r, 0);
  Kokkos::parallel_reduce(P(0, n), LOrF<int>{d}, Kokkos::LOr<int>(r));
  CHECK_EQ(r, 1);
  std::vector<int> z(n, 0);
  z[49999] = 1;
  auto dz = upload(z);
  Kokkos::parallel_reduce(P(0, n - 1), LOrF<int>{dz}, Kokkos::LOr<int>(r));
  CHECK_EQ(r, 0);
  Kokkos::parallel_reduce(P(0, n), LOrF<int>{dz}, Kokkos::LOr<int>(r));
  CHECK_EQ(r, 1);
  std::vector<unsigned> b(n);
  unsigned band = ~0u, bor = 0;
  for (int i = 0; i < n; ++i) {
    b[i] = ~(1u << (i % 29)) & (0xF0F0F0F0u | (1u << (i % 31)));
    band &= b[i];
    bor |= b[i];
  }
  auto db = upload(b);
  unsigned rb = 0;
  Kokkos::parallel_reduce(P(0, n), BAndF<unsigned>{db}, Kokkos::BAnd<unsigned>(rb));
  CHECK_EQ(rb, band);
  Kokkos::parallel_reduce(P(0, n), BOrF<unsigned>{db}, Kokkos::BOr<unsigned>(rb));
  CHECK_EQ(rb, bor);
}

// =========================================================================
// Scan (core/unit_test/TestParallelScanRangePolicy.hpp:24-270)
// =========================================================================
template <class ValueType>
struct ScanRangeFunctor {
  using execution_space = Exec;
  using ViewType        = Kokkos::View<ValueType*, Exec>;
  ViewType prefix_results, postfix_results;
  KOKKOS_INLINE_FUNCTION void operator()(const size_t i, ValueType& update, bool final_pass) const {
    if (final_pass) prefix_results(i) = update;
    update += i;
    if (final_pass) postfix_results(i) = update;
  }
  KOKKOS_INLINE_FUNCTION void init(ValueType& update) const { update = 0; }
  KOKKOS_INLINE_FUNCTION void join(ValueType& update, const ValueType& input) const { update += input; }

  void check(size_t work_size) const {
    auto pre  = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), prefix_results);
    auto post = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), postfix_results);
    for (size_t i = 0; i < work_size; ++i) {
      CHECK_EQ(pre(i), static_cast<ValueType>((static_cast<ValueType>(i) * (i - 1)) / 2));
      CHECK_EQ(post(i), static_cast<ValueType>((static_cast<ValueType>(i) * (i + 1)) / 2));
    }
    Kokkos::deep_copy(prefix_results, 0);
    Kokkos::deep_copy(postfix_results, 0);
  }
  ValueType total(size_t work_size) const {
    return static_cast<ValueType>(static_cast<ValueType>(work_size) * (work_size - 1) / 2);
  }

  template <class... Args>
  void test_scan(const size_t work_size) {
    prefix_results  = ViewType("prefix_results", work_size);
    postfix_results = ViewType("postfix_results", work_size);
    if constexpr (sizeof...(Args) == 0) {
      Kokkos::parallel_scan("TestWithStrArg1", work_size, *this);
      check(work_size);
      Kokkos::parallel_scan(work_size, *this);
      check(work_size);
      {
        ValueType rv = 0;
        Kokkos::parallel_scan("TestWithStrArg2", work_size, *this, rv);
        check(work_size);
        CHECK_EQ(rv, total(work_size));
      }
      {
        ValueType rv = 0;
        Kokkos::parallel_scan(work_size, *this, rv);
        check(work_size);
        CHECK_EQ(rv, total(work_size));
      }
      {
        Kokkos::View<ValueType, Kokkos::HostSpace> rview("return_view");
        Kokkos::parallel_scan(work_size, *this, rview);
        check(work_size);
        CHECK_EQ(rview(), total(work_size));
      }
    } else {
      Kokkos::RangePolicy<Exec, Args...> policy(0, work_size);
      Kokkos::parallel_scan("TestWithStrArg3", policy, *this);
      check(work_size);
      Kokkos::parallel_scan(policy, *this);
      check(work_size);
      {
        ValueType rv = 0;
        Kokkos::parallel_scan("TestWithStrArg4", policy, *this, rv);
        check(work_size);
        CHECK_EQ(rv, total(work_size));
      }
      {
        Kokkos::View<ValueType, Exec> rview("return_view");
        Kokkos::parallel_scan(policy, *this, rview);
        check(work_size);
        ValueType t;
        Kokkos::deep_copy(t, rview);
        CHECK_EQ(t, total(work_size));
      }
      {
        const auto pr = Kokkos::Experimental::require(policy, Kokkos::Experimental::WorkItemProperty::HintLightWeight);
        ValueType rv = 0;
        Kokkos::parallel_scan(pr, *this, rv);
        check(work_size);
        CHECK_EQ(rv, total(work_size));
      }
    }
  }
  template <class... Args>
  void test_scan(const std::vector<size_t>& sizes) {
    for (size_t s : sizes) test_scan<Args...>(s);
  }
  void all(const std::vector<size_t>& sizes) {
    test_scan<>(sizes);
    test_scan<Kokkos::Schedule<Kokkos::Static>>(sizes);
    test_scan<Kokkos::Schedule<Kokkos::Dynamic>>(sizes);
  }
};
KKMX_TEST(scan_range_policy_char) { ScanRangeFunctor<char>{}.all({5, 10}); }
KKMX_TEST(scan_range_policy_short) { ScanRangeFunctor<short>{}.all({50, 100}); }
KKMX_TEST(scan_range_policy_int) { ScanRangeFunctor<int>{}.all({0, 1, 2, 1000, 1001}); }
KKMX_TEST(scan_range_policy_long) { ScanRangeFunctor<long>{}.all({1000, 10000}); }
KKMX_TEST(scan_range_policy_float) { ScanRangeFunctor<float>{}.all({13, 34}); }
KKMX_TEST(scan_range_policy_double) { ScanRangeFunctor<double>{}.all({17, 59}); }
KKMX_TEST(scan_range_policy_large) { ScanRangeFunctor<int64_t>{}.all({(size_t(1) << 22) + 12345, 10000019}); }

// Tile status words keep 30 bits of the launch epoch, and the instance zeroes
// them once every 2^29 scan launches.  The next launch's epoch is set to
// alias the previous launch's (previous + 2^30) and the launch count to the
// period: the scan over new data must still be exact, and only its own tiles
// may hold status words afterwards.
struct ScanEpochProbe {
  Kokkos::View<int*, Exec> a, out;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int& u, const bool final) const {
    if (final) out(i) = u;
    u += a(i);
  }
};
KKMX_TEST(scan_epoch_wrap_clears_status) {
  Exec ex;
  Kokkos::Impl::MI355XInstance* inst = ex.impl_internal_space_instance();
  const int64_t nbig = int64_t(1) << 22, nsmall = 4096;
  Kokkos::View<int*, Exec> a("a", nbig), out("out", nbig);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(ex, 0, nbig), KOKKOS_LAMBDA(const int64_t i) { a(i) = int(i % 7) - 3; });
  Kokkos::parallel_scan(Kokkos::RangePolicy<Exec>(ex, 0, nbig), ScanEpochProbe{a, out});
  ex.fence();
  auto nonzero_words = [&]() {
    std::vector<uint64_t> w(inst->scan_status_bytes / 8);
    KKMX_SAFE_CALL(mx_memcpy(w.data(), inst->scan_status, w.size() * 8));
    int64_t c = 0;
    for (uint64_t x : w) c += (x != 0);
    return c;
  };
  const int64_t before = nonzero_words();
  CHECK(before > 64);
  unsigned next_epoch = 0;
  KKMX_SAFE_CALL(mx_memcpy(&next_epoch, inst->control + 3, sizeof(unsigned)));
  CHECK(next_epoch >= 2u);
  const unsigned alias = (next_epoch - 1u) + (1u << 30);
  KKMX_SAFE_CALL(mx_memcpy(inst->control + 3, &alias, sizeof(unsigned)));
  inst->scan_since_clear = Kokkos::Impl::kScanClearPeriod;
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(ex, 0, nsmall), KOKKOS_LAMBDA(const int64_t i) { a(i) = int(i % 5) + 1; });
  Kokkos::parallel_scan(Kokkos::RangePolicy<Exec>(ex, 0, nsmall), ScanEpochProbe{a, out});
  ex.fence();
  CHECK_EQ(inst->scan_since_clear, uint64_t(1));
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
  int64_t bad = 0, acc = 0;
  for (int64_t i = 0; i < nsmall; ++i) {
    bad += (h(i) != acc);
    acc += i % 5 + 1;
  }
  CHECK_EQ(bad, int64_t(0));
  const int64_t after = nonzero_words();
  CHECK(after >= 1);
  CHECK(after <= nsmall / 512);  // tiles hold at least 64 lanes x 8 items
  unsigned epoch_now = 0;
  KKMX_SAFE_CALL(mx_memcpy(&epoch_now, inst->control + 3, sizeof(unsigned)));
  CHECK_EQ(epoch_now, alias + 1u);
}

// The contribution-keeping tile body at 64 values per lane runs as a loop of
// one resident workgroup per CU over the tiles: totals, a ragged last tile, a
// non-zero begin, more tiles than workgroups, back-to-back launches of both
// bodies, and the claim word left at {0, epoch + 1} by every launch.
struct ScanIntTotal {
  Kokkos::View<const int*, Exec, Kokkos::MemoryTraits<Kokkos::RandomAccess>> a;
  Kokkos::View<int*, Exec> out;
  int64_t begin;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, int& u, const bool final) const {
    u += a(i - begin);
    if (final) out(i - begin) = u;
  }
};
KKMX_TEST(scan_keep_loop_totals_and_claims) {
  Exec ex;
  Kokkos::Impl::MI355XInstance* inst = ex.impl_internal_space_instance();
  const int saved = Kokkos::Impl::range_overrides().scan_body;
  for (int body : {1, 0, 1}) {
    Kokkos::Impl::range_overrides().scan_body = body;
    for (int64_t n : {(int64_t(1) << 22) + 12345, (int64_t(1) << 25) + 12345, 3 * (int64_t(1) << 24) + 1}) {
      const int64_t b = 77;
      Kokkos::View<int*, Exec> a("a", n), out("out", n);
      Kokkos::parallel_for(Kokkos::RangePolicy<Exec, Kokkos::IndexType<int64_t>>(ex, 0, n),
                           KOKKOS_LAMBDA(const int64_t i) { a(i) = int((uint64_t(i) * 2654435761u) % 201) - 100; });
      unsigned before[2], after[2];
      KKMX_SAFE_CALL(mx_memcpy(before, inst->control + 2, sizeof(before)));
      int total = 0;
      Kokkos::parallel_scan(Kokkos::RangePolicy<Exec, Kokkos::IndexType<int64_t>>(ex, b, b + n),
                            ScanIntTotal{a, out, b}, total);
      ex.fence();
      KKMX_SAFE_CALL(mx_memcpy(after, inst->control + 2, sizeof(after)));
      auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
      int64_t bad = 0, acc = 0;
      for (int64_t k = 0; k < n; ++k) {
        acc += int64_t((uint64_t(k) * 2654435761u) % 201) - 100;
        bad += (h(k) != acc);
      }
      CHECK_EQ(bad, int64_t(0));
      CHECK_EQ(int64_t(total), acc);
      CHECK_EQ(after[0], 0u);
      CHECK_EQ(after[1], before[1] + 1u);
    }
  }
  Kokkos::Impl::range_overrides().scan_body = saved;
}

// A graph holding the looping contribution-keeping scan (2^24 + 5 int32
// indices, body pinned): every replay equals the host scan, and so does an
// eager scan between replays, which leaves its instance's claim word at
// {0, epoch + 1} (graph nodes run on instances of their own).
KKMX_TEST(graph_scan_keep_loop_replays) {
  Exec ex;
  Kokkos::Impl::MI355XInstance* inst = ex.impl_internal_space_instance();
  const int saved = Kokkos::Impl::range_overrides().scan_body;
  Kokkos::Impl::range_overrides().scan_body = 1;
  const int64_t n = (int64_t(1) << 24) + 5;
  Kokkos::View<int*, Exec> a("a", n), out("out", n);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec, Kokkos::IndexType<int64_t>>(ex, 0, n),
                       KOKKOS_LAMBDA(const int64_t i) { a(i) = int(i % 13) - 6; });
  auto graph = Kokkos::Experimental::create_graph(ex, [&](auto root) {
    root.then_parallel_scan("scan", Kokkos::RangePolicy<Exec, Kokkos::IndexType<int64_t>>(0, n),
                            ScanIntTotal{a, out, 0});
  });
  std::vector<int> want(n);
  int acc = 0;
  for (int64_t k = 0; k < n; ++k) want[k] = (acc += int(k % 13) - 6);
  Kokkos::View<int*, Exec> out2("out2", n);
  const auto mismatches = [&](const Kokkos::View<int*, Exec>& v) {
    auto h      = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), v);
    int64_t bad = 0;
    for (int64_t k = 0; k < n; ++k) bad += (h(k) != want[k]);
    return bad;
  };
  for (int rep = 0; rep < 4; ++rep) {
    Kokkos::deep_copy(ex, out, -1);
    Kokkos::deep_copy(ex, out2, -1);
    ex.fence();
    graph.submit();
    ex.fence();
    CHECK_EQ(mismatches(out), int64_t(0));
    unsigned before[2], after[2];
    KKMX_SAFE_CALL(mx_memcpy(before, inst->control + 2, sizeof(before)));
    Kokkos::parallel_scan(Kokkos::RangePolicy<Exec, Kokkos::IndexType<int64_t>>(ex, 0, n), ScanIntTotal{a, out2, 0});
    ex.fence();
    KKMX_SAFE_CALL(mx_memcpy(after, inst->control + 2, sizeof(after)));
    CHECK_EQ(mismatches(out2), int64_t(0));
    CHECK_EQ(after[0], 0u);
    CHECK_EQ(after[1], before[1] + 1u);
  }
  Kokkos::Impl::range_overrides().scan_body = saved;
}

// A scan whose value is a struct (generic tile state) and a scan with a tag
// and a non-zero begin, with both tile bodies (at 2^21 + 9 indices the
// contribution-keeping body gives each lane 4 consecutive indices a round).
struct Pair64 {
  int64_t a, b;
  KOKKOS_INLINE_FUNCTION Pair64& operator+=(const Pair64& o) {
    a += o.a;
    b += o.b;
    return *this;
  }
};
struct ScanTag {};
struct StructScan {
  Kokkos::View<int64_t*, Exec> outa, outb;
  int64_t begin;
  KOKKOS_INLINE_FUNCTION void operator()(ScanTag, const int64_t i, Pair64& u, const bool final) const {
    u.a += 1;
    u.b += i;
    if (final) {
      outa(i - begin) = u.a;
      outb(i - begin) = u.b;
    }
  }
};
KKMX_TEST(scan_struct_value_with_tag_and_offset) {
  const int saved = Kokkos::Impl::range_overrides().scan_body;
  for (int body : {0, 1})
  for (int64_t n : {int64_t(1), int64_t(777), int64_t(1000003), (int64_t(1) << 21) + 9}) {
    Kokkos::Impl::range_overrides().scan_body = body;
    const int64_t b = 1000;
    Kokkos::View<int64_t*, Exec> a("a", n), c("c", n);
    Pair64 total{0, 0};
    Kokkos::parallel_scan(Kokkos::RangePolicy<Exec, ScanTag>(b, b + n), StructScan{a, c, b}, total);
    auto ha = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), a);
    auto hc = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), c);
    for (int64_t k = 0; k < n; ++k) {
      const int64_t i = b + k;
      if (ha(k) != k + 1) CHECK_EQ(ha(k), k + 1);
      const int64_t want = (i * (i + 1)) / 2 - (b * (b - 1)) / 2;
      if (hc(k) != want) CHECK_EQ(hc(k), want);
    }
    CHECK_EQ(total.a, n);
    CHECK_EQ(total.b, ((b + n - 1) * (b + n)) / 2 - (b * (b - 1)) / 2);
  }
  Kokkos::Impl::range_overrides().scan_body = saved;
}

// =========================================================================
// MDRange (core/unit_test/TestMDRange.hpp): every point visited once, for and
// reduce, ranks 2..6, both iteration orders, explicit tiles, negative begins.
// =========================================================================
template <int R>
struct MDLinear {
  int64_t lo[6], ext[6];
  KOKKOS_INLINE_FUNCTION int64_t lin(const int64_t* idx) const {
    int64_t l = 0;
    for (int r = 0; r < R; ++r) l = l * ext[r] + (idx[r] - lo[r]);
    return l;
  }
};
template <int R>
struct MDCount : MDLinear<R> {
  Kokkos::View<int*, Exec> hits;
  template <class... I>
  KOKKOS_INLINE_FUNCTION void operator()(const I... i) const {
    const int64_t idx[] = {int64_t(i)...};
    Kokkos::atomic_add(&hits(this->lin(idx)), 1);
  }
};
// Sum over the box of (linear index + 1) = P(P+1)/2 for P points.
template <int R>
struct MDLinSum : MDLinear<R> {
  using value_type = int64_t;
  KOKKOS_INLINE_FUNCTION void add(const int64_t* x, int64_t& u) const { u += this->lin(x) + 1; }
  KOKKOS_INLINE_FUNCTION void operator()(int64_t a, int64_t b, int64_t& u) const {
    const int64_t x[] = {a, b};
    add(x, u);
  }
  KOKKOS_INLINE_FUNCTION void operator()(int64_t a, int64_t b, int64_t c, int64_t& u) const {
    const int64_t x[] = {a, b, c};
    add(x, u);
  }
  KOKKOS_INLINE_FUNCTION void operator()(int64_t a, int64_t b, int64_t c, int64_t d, int64_t& u) const {
    const int64_t x[] = {a, b, c, d};
    add(x, u);
  }
  KOKKOS_INLINE_FUNCTION void operator()(int64_t a, int64_t b, int64_t c, int64_t d, int64_t e, int64_t& u) const {
    const int64_t x[] = {a, b, c, d, e};
    add(x, u);
  }
  KOKKOS_INLINE_FUNCTION void operator()(int64_t a, int64_t b, int64_t c, int64_t d, int64_t e, int64_t f,
                                         int64_t& u) const {
    const int64_t x[] = {a, b, c, d, e, f};
    add(x, u);
  }
};

template <int R, Kokkos::Iterate It>
void md_case(const std::vector<int64_t>& lo, const std::vector<int64_t>& hi, const std::vector<int64_t>& tile) {
  using Pol = Kokkos::MDRangePolicy<Exec, Kokkos::Rank<R, It, It>, Kokkos::IndexType<int64_t>>;
  typename Pol::point_type b, e;
  typename Pol::tile_type t;
  MDCount<R> f;
  MDLinSum<R> fs;
  int64_t total = 1;
  for (int r = 0; r < R; ++r) {
    b[r]      = lo[r];
    e[r]      = hi[r];
    t[r]      = tile.empty() ? 0 : tile[r];
    f.lo[r]   = fs.lo[r] = lo[r];
    f.ext[r]  = fs.ext[r] = hi[r] - lo[r];
    total *= f.ext[r];
  }
  f.hits = Kokkos::View<int*, Exec>("hits", total > 0 ? total : 1);
  Kokkos::parallel_for(Pol(b, e, t), f);
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), f.hits);
  for (int64_t k = 0; k < total; ++k)
    if (h(k) != 1) CHECK_EQ(h(k), 1);
  int64_t sum = -1;
  Kokkos::parallel_reduce(Pol(b, e, t), fs, sum);
  CHECK_EQ(sum, total * (total + 1) / 2);
  Kokkos::View<int64_t, Exec> dsum("dsum");
  Kokkos::parallel_reduce(Pol(b, e, t), fs, dsum);
  int64_t hs;
  Kokkos::deep_copy(hs, dsum);
  CHECK_EQ(hs, total * (total + 1) / 2);
}
template <int R>
void md_both(const std::vector<int64_t>& lo, const std::vector<int64_t>& hi, const std::vector<int64_t>& tile) {
  md_case<R, Kokkos::Iterate::Left>(lo, hi, tile);
  md_case<R, Kokkos::Iterate::Right>(lo, hi, tile);
}
KKMX_TEST(mdrange_2d) {
  md_both<2>({0, 0}, {100, 100}, {});
  md_both<2>({0, 0}, {1000, 3}, {});
  md_both<2>({3, 7}, {517, 130}, {32, 8});
  md_both<2>({0, 0}, {33, 1025}, {1, 256});
  md_both<2>({0, 0}, {0, 10}, {});
}
KKMX_TEST(mdrange_3d) {
  md_both<3>({0, 0, 0}, {64, 64, 64}, {});
  md_both<3>({1, 2, 3}, {101, 52, 33}, {4, 4, 4});
  md_both<3>({0, 0, 0}, {130, 7, 9}, {64, 2, 2});
}
KKMX_TEST(mdrange_neg_idx) {
  md_both<2>({-1, -1}, {12, 7}, {});
  md_both<3>({-5, -2, -3}, {20, 9, 4}, {8, 2, 2});
}
KKMX_TEST(mdrange_4d) {
  md_both<4>({0, 0, 0, 0}, {20, 10, 9, 7}, {});
  md_both<4>({2, 1, 0, 3}, {19, 14, 6, 9}, {4, 4, 2, 2});
}
KKMX_TEST(mdrange_5d) {
  md_both<5>({0, 0, 0, 0, 0}, {8, 7, 6, 5, 9}, {});
  md_both<5>({1, 0, 2, 0, 1}, {9, 6, 7, 4, 5}, {2, 2, 2, 2, 2});
}
KKMX_TEST(mdrange_6d) {
  md_both<6>({0, 0, 0, 0, 0, 0}, {6, 5, 4, 7, 3, 5}, {});
  md_both<6>({1, 1, 0, 0, 2, 0}, {5, 6, 4, 3, 5, 4}, {2, 2, 2, 1, 2, 2});
}

// A fixed-size struct value over an MDRange (the compile-time counterpart of
// the runtime-length arrays below).
struct MDScalarArray {
  struct value_type {
    int64_t v[4];
  };
  KOKKOS_INLINE_FUNCTION void init(value_type& u) const {
    for (int k = 0; k < 4; ++k) u.v[k] = 0;
  }
  KOKKOS_INLINE_FUNCTION void join(value_type& a, const value_type& b) const {
    for (int k = 0; k < 4; ++k) a.v[k] += b.v[k];
  }
  KOKKOS_INLINE_FUNCTION void operator()(const int i, const int j, value_type& u) const {
    u.v[0] += 1;
    u.v[1] += i;
    u.v[2] += j;
    u.v[3] += int64_t(i) * j;
  }
};
KKMX_TEST(mdrange_struct_reduce) {
  const int N0 = 513, N1 = 77;
  MDScalarArray::value_type r;
  Kokkos::parallel_reduce(Kokkos::MDRangePolicy<Exec, Kokkos::Rank<2>>({0, 0}, {N0, N1}), MDScalarArray(), r);
  const int64_t s0 = int64_t(N0) * (N0 - 1) / 2, s1 = int64_t(N1) * (N1 - 1) / 2;
  CHECK_EQ(r.v[0], int64_t(N0) * N1);
  CHECK_EQ(r.v[1], s0 * N1);
  CHECK_EQ(r.v[2], s1 * N0);
  CHECK_EQ(r.v[3], s0 * s1);
}

// mdrange_array_reduce (TestMDRange_c.hpp:30, TestMDRange.hpp:31-215):
// value_type = double[] with a runtime value_count, once with the default
// init/join (zero, +) and once with the functor's own; the InitTag
// parallel_for sets every element to 3, so the sums are 6*N and 3*N.
struct MDArrayReduce {
  using scalar_type = double;
  using value_type  = scalar_type[];
  Kokkos::View<int***, Exec> in;
  unsigned value_count;
  struct InitTag {};
  KOKKOS_INLINE_FUNCTION void operator()(const InitTag&, const int i, const int j, const int k) const {
    in(i, j, k) = 3;
  }
  KOKKOS_INLINE_FUNCTION void operator()(const int i, const int j, const int k, value_type lsum) const {
    lsum[0] += in(i, j, k) * 2;
    lsum[1] += in(i, j, k);
    for (unsigned c = 2; c < value_count; ++c) lsum[c] += double(i) * (c == 2) + double(j) * (c == 3) + k * (c > 3);
  }
};
// Same functor with its own init and join (TestMDRange.hpp:114-126).
struct MDArrayReduceOwnJoin : MDArrayReduce {
  KOKKOS_INLINE_FUNCTION void init(scalar_type dst[]) const {
    for (unsigned c = 0; c < value_count; ++c) dst[c] = 0.0;
  }
  KOKKOS_INLINE_FUNCTION void join(scalar_type dst[], const scalar_type src[]) const {
    for (unsigned c = 0; c < value_count; ++c) dst[c] += src[c];
  }
};
template <class F>
void md_array_reduce_case(int N0, int N1, int N2, unsigned count, bool device_result) {
  F f;
  f.in          = Kokkos::View<int***, Exec>("in", N0, N1, N2);
  f.value_count = count;
  using Init    = Kokkos::MDRangePolicy<Exec, Kokkos::Rank<3>, Kokkos::IndexType<int>, typename F::InitTag>;
  using Pol     = Kokkos::MDRangePolicy<Exec, Kokkos::Rank<3>, Kokkos::IndexType<int>>;
  Kokkos::parallel_for(Init({0, 0, 0}, {N0, N1, N2}, {3, 3, 3}), f);
  std::vector<double> sums(count, -1.0);
  if (device_result) {
    Kokkos::View<double*, Exec> r("r", count);
    Kokkos::parallel_reduce(Pol({0, 0, 0}, {N0, N1, N2}), f, r);
    auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), r);
    for (unsigned c = 0; c < count; ++c) sums[c] = h(c);
  } else {
    Kokkos::parallel_reduce(Pol({0, 0, 0}, {N0, N1, N2}, {3, 3, 3}), f, sums.data());
  }
  const double n = double(N0) * N1 * N2;
  CHECK_EQ(sums[0], 6 * n);
  CHECK_EQ(sums[1], 3 * n);
  const double s0 = double(N0) * (N0 - 1) / 2 * N1 * N2, s1 = double(N1) * (N1 - 1) / 2 * N0 * N2,
               s2 = double(N2) * (N2 - 1) / 2 * N0 * N1;
  for (unsigned c = 2; c < count; ++c) CHECK_EQ(sums[c], c == 2 ? s0 : c == 3 ? s1 : s2);
}
KKMX_TEST(mdrange_array_reduce) {
  md_array_reduce_case<MDArrayReduce>(4, 5, 1, 2, false);
  md_array_reduce_case<MDArrayReduceOwnJoin>(4, 5, 10, 2, false);
  md_array_reduce_case<MDArrayReduce>(130, 67, 33, 7, false);
  md_array_reduce_case<MDArrayReduceOwnJoin>(257, 31, 64, 5, true);
  md_array_reduce_case<MDArrayReduce>(1, 1, 1, 40, true);
}

// team_array_reduce (TestTeam.hpp:297-333,399-433): TeamPolicy with
// value_type = T[] and value_count = 3; each of the league*team threads takes
// a chunk of [0, nwork) and adds 1, i+1 and nwork-i, so the sums are nwork,
// nwork(nwork+1)/2 and the same.  Results into host arrays and device Views.
template <class T>
struct TeamArrayReduce {
  using value_type = T[];
  unsigned value_count = 3;
  int64_t nwork;
  KOKKOS_INLINE_FUNCTION void init(value_type dst) const {
    for (unsigned i = 0; i < value_count; ++i) dst[i] = 0;
  }
  KOKKOS_INLINE_FUNCTION void join(value_type dst, const value_type src) const {
    for (unsigned i = 0; i < value_count; ++i) dst[i] += src[i];
  }
  KOKKOS_INLINE_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type& team, value_type dst) const {
    const int64_t rank  = team.team_rank() + int64_t(team.team_size()) * team.league_rank();
    const int64_t size  = int64_t(team.team_size()) * team.league_size();
    const int64_t chunk = (nwork + size - 1) / size;
    int64_t i           = chunk * rank;
    const int64_t e     = i + chunk < nwork ? i + chunk : nwork;
    for (; i < e; ++i) {
      dst[0] += 1;
      dst[1] += i + 1;
      dst[2] += nwork - i;
    }
  }
};
template <class T>
void team_array_reduce_case(int64_t nwork, bool device_result) {
  using Pol = Kokkos::TeamPolicy<Exec>;
  TeamArrayReduce<T> f;
  f.nwork                = nwork;
  const unsigned ts      = Pol(1, 1).team_size_recommended(f, Kokkos::ParallelReduceTag());
  const int64_t league   = (nwork + ts - 1) / ts;
  const uint64_t nw      = uint64_t(nwork);
  const uint64_t nsum    = nw % 2 ? nw * ((nw + 1) / 2) : (nw / 2) * (nw + 1);
  for (int rep = 0; rep < 3; ++rep) {
    T r[3] = {T(-1), T(-1), T(-1)};
    if (device_result) {
      Kokkos::View<T*, Exec> d("d", 3);
      Kokkos::parallel_reduce(Pol(league > 0 ? league : 1, ts), f, d);
      Kokkos::deep_copy(Kokkos::View<T*, Kokkos::HostSpace, Kokkos::MemoryUnmanaged>(r, 3), d);
    } else {
      Kokkos::View<T*, Kokkos::HostSpace, Kokkos::MemoryUnmanaged> h(r, 3);
      Kokkos::parallel_reduce(Pol(league > 0 ? league : 1, ts), f, h);
    }
    CHECK_EQ(uint64_t(r[0]), nw);
    CHECK_EQ(uint64_t(r[1]), nsum);
    CHECK_EQ(uint64_t(r[2]), nsum);
  }
}
KKMX_TEST(team_array_reduce) {
  for (int64_t n : {int64_t(0), int64_t(1), int64_t(1000), int64_t(10007), int64_t(1) << 20}) {
    team_array_reduce_case<int64_t>(n, false);
    team_array_reduce_case<double>(n, true);
    team_array_reduce_case<long>(n, n % 2 == 0);
  }
}

// =========================================================================
// Team (core/unit_test/TestTeamBasic.hpp, TestTeamScan.hpp, TestTeamVector.hpp)
// =========================================================================
using TeamPol = Kokkos::TeamPolicy<Exec>;
using Member  = TeamPol::member_type;

// team_for: every (league_rank, team_rank) runs once.
struct TeamForF {
  Kokkos::View<int**, Exec> hits;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m) const {
    if (m.league_size() != int(hits.extent(0)) || m.team_size() != int(hits.extent(1))) {
      Kokkos::atomic_add(&hits(0, 0), 1000);
      return;
    }
    Kokkos::atomic_add(&hits(m.league_rank(), m.team_rank()), 1);
  }
};
KKMX_TEST(team_for) {
  for (int league : {1, 7, 1000, 10007})
    for (int ts : {1, 3, 64, 127, 256, 1024}) {
      Kokkos::View<int**, Exec> hits("hits", league, ts);
      Kokkos::parallel_for(TeamPol(league, ts), TeamForF{hits});
      auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), hits);
      for (int l = 0; l < league; ++l)
        for (int t = 0; t < ts; ++t)
          if (h(l, t) != 1) CHECK_EQ(h(l, t), 1);
    }
}
// team_reduce: sum of (league_rank * team_size + team_rank + 1) over all.
struct TeamReduceF {
  using value_type = int64_t;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m, int64_t& u) const {
    u += int64_t(m.league_rank()) * m.team_size() + m.team_rank() + 1;
  }
};
KKMX_TEST(team_reduce) {
  for (int league : {0, 1, 13, 4096})
    for (int ts : {1, 64, 100, 512}) {
      int64_t r = -5;
      Kokkos::parallel_reduce(TeamPol(league, ts), TeamReduceF(), r);
      const int64_t p = int64_t(league) * ts;
      CHECK_EQ(r, p * (p + 1) / 2);
    }
}
// Nested TeamThreadRange reduce + team_broadcast + single.
struct TeamNestedF {
  Kokkos::View<int64_t*, Exec> out;
  int n;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m) const {
    const int lr = m.league_rank();
    int64_t s    = 0;
    Kokkos::parallel_reduce(
        Kokkos::TeamThreadRange(m, n), [&](const int i, int64_t& u) { u += int64_t(lr) * n + i; }, s);
    int64_t mx = 0;
    Kokkos::parallel_reduce(
        Kokkos::TeamThreadRange(m, n), [&](const int i, int64_t& u) { if (i * 7 % n > u) u = i * 7 % n; },
        Kokkos::Max<int64_t>(mx));
    int64_t b = (m.team_rank() == m.team_size() - 1) ? s + lr : -1;
    m.team_broadcast(b, m.team_size() - 1);
    Kokkos::single(Kokkos::PerTeam(m), [&]() { out(lr) = s * 3 + (b - s - lr) + mx * 0; });
    if (b != s + lr || s != int64_t(lr) * n * n + int64_t(n) * (n - 1) / 2) Kokkos::atomic_add(&out(lr), int64_t(1) << 50);
  }
};
KKMX_TEST(team_nested_reduce_broadcast_single) {
  for (int ts : {32, 64, 96, 256})
    for (int n : {1, 5, 64, 1000}) {
      const int league = 301;
      Kokkos::View<int64_t*, Exec> out("out", league);
      Kokkos::parallel_for(TeamPol(league, ts), TeamNestedF{out, n});
      auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
      for (int l = 0; l < league; ++l) {
        const int64_t s = int64_t(l) * n * n + int64_t(n) * (n - 1) / 2;
        if (h(l) != 3 * s) CHECK_EQ(h(l), 3 * s);
      }
    }
}
// team_scan (TestTeamScan.hpp:24-114).
template <class DataType>
struct TeamScanF {
  Kokkos::View<DataType**, Exec> a_d, a_r;
  int32_t N;
  KOKKOS_FUNCTION void operator()(const Member& team) const {
    const int lr = team.league_rank();
    Kokkos::parallel_for(Kokkos::TeamThreadRange(team, 0, N), [&](const int i) { a_d(lr, i) = lr * N + i; });
    Kokkos::parallel_scan(Kokkos::TeamThreadRange(team, 0, N), [&](int i, DataType& val, const bool final) {
      val += a_d(lr, i);
      if (final) a_r(lr, i) = val;
    });
  }
};
template <class DataType>
void team_scan_case(int M, int N) {
  TeamScanF<DataType> f{Kokkos::View<DataType**, Exec>("a_d", M, N), Kokkos::View<DataType**, Exec>("a_r", M, N), N};
  Kokkos::parallel_for(TeamPol(M, 127), f);
  auto ai = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), f.a_d);
  auto ao = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), f.a_r);
  for (int i = 0; i < M; ++i) {
    DataType ref = 0;
    for (int j = 0; j < N; ++j) {
      ref += ai(i, j);
      if (ao(i, j) != ref) CHECK_EQ(ao(i, j), ref);
    }
  }
}
KKMX_TEST(team_scan) {
  for (int M : {0, 1, 10, 100})
    for (int N : {0, 1, 2, 127, 128, 1000}) {
      team_scan_case<int32_t>(M, N);
      team_scan_case<int64_t>(M, N);
    }
}
// ThreadVectorRange reduce and scan inside TeamThreadRange.
struct TeamVectorF {
  Kokkos::View<int64_t**, Exec> out;
  int n;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m) const {
    const int lr = m.league_rank();
    Kokkos::parallel_for(Kokkos::TeamThreadRange(m, n), [&](const int i) {
      int64_t s = 0;
      Kokkos::parallel_reduce(
          Kokkos::ThreadVectorRange(m, i + 1), [&](const int k, int64_t& u) { u += k + lr; }, s);
      // The scan total comes back to every vector lane; a variable written
      // by the lane that owns k == i would only be visible in that lane.
      int64_t last = 0;
      Kokkos::parallel_scan(
          Kokkos::ThreadVectorRange(m, i + 1), [&](const int k, int64_t& u, const bool) { u += k; }, last);
      Kokkos::single(Kokkos::PerThread(m), [&]() { out(lr, i) = s * 1000003 + last; });
    });
  }
};
KKMX_TEST(team_vector_reduce_scan) {
  for (int vl : {1, 4, 16, 64})
    for (int ts : {1, 4, 16}) {
      if (ts * vl > 1024) continue;
      const int league = 37, n = 70;
      Kokkos::View<int64_t**, Exec> out("out", league, n);
      Kokkos::parallel_for(TeamPol(league, ts, vl), TeamVectorF{out, n});
      auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
      for (int l = 0; l < league; ++l)
        for (int i = 0; i < n; ++i) {
          const int64_t tri = int64_t(i) * (i + 1) / 2;
          const int64_t want = (tri + int64_t(i + 1) * l) * 1000003 + tri;
          if (h(l, i) != want) CHECK_EQ(h(l, i), want);
        }
    }
}
// Team scratch: PerTeam and PerThread level-0 buffers (TestTeamBasic.hpp:182-215).
struct TeamScratchF {
  using shmem = Kokkos::View<int64_t*, Exec::scratch_memory_space, Kokkos::MemoryTraits<Kokkos::Unmanaged>>;
  Kokkos::View<int64_t*, Exec> out;
  int team_n, thread_n;
  KOKKOS_INLINE_FUNCTION void operator()(const Member& m) const {
    shmem team_buf(m.team_scratch(0), team_n);
    shmem mine(m.thread_scratch(0), thread_n);
    Kokkos::parallel_for(Kokkos::TeamThreadRange(m, team_n), [&](const int i) { team_buf(i) = i + m.league_rank(); });
    for (int k = 0; k < thread_n; ++k) mine(k) = m.team_rank() * 10 + k;
    m.team_barrier();
    int64_t s = 0;
    Kokkos::parallel_reduce(
        Kokkos::TeamThreadRange(m, team_n), [&](const int i, int64_t& u) { u += team_buf(team_n - 1 - i); }, s);
    int64_t t = 0;
    for (int k = 0; k < thread_n; ++k) t += mine(k) - m.team_rank() * 10;
    if (m.team_rank() == 0) out(m.league_rank()) = s;
    if (t != int64_t(thread_n) * (thread_n - 1) / 2) Kokkos::atomic_add(&out(m.league_rank()), int64_t(1) << 40);
  }
};
KKMX_TEST(team_scratch_per_team_per_thread) {
  for (int team_n : {1, 100, 4000})
    for (int thread_n : {1, 8, 32}) {
      const int league = 513, ts = 64;
      Kokkos::View<int64_t*, Exec> out("out", league);
      const size_t bytes_team   = TeamScratchF::shmem::shmem_size(team_n);
      const size_t bytes_thread = TeamScratchF::shmem::shmem_size(thread_n);
      Kokkos::parallel_for(
          TeamPol(league, ts).set_scratch_size(0, Kokkos::PerTeam(bytes_team), Kokkos::PerThread(bytes_thread)),
          TeamScratchF{out, team_n, thread_n});
      auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
      for (int l = 0; l < league; ++l) {
        const int64_t want = int64_t(team_n) * (team_n - 1) / 2 + int64_t(team_n) * l;
        if (h(l) != want) CHECK_EQ(h(l), want);
      }
    }
}

// =========================================================================
// Atomics (core/unit_test/TestAtomics.hpp:26-455): add, CAS and exchange loops
// against the serial answer; min/max; sub-word and wide types.
// =========================================================================
template <class T>
struct AddLoop {
  Kokkos::View<T, Exec> data;
  KOKKOS_INLINE_FUNCTION void operator()(const int) const { Kokkos::atomic_fetch_add(&data(), (T)1); }
};
template <class T>
struct CASLoop {
  Kokkos::View<T, Exec> data;
  KOKKOS_INLINE_FUNCTION void operator()(const int) const {
    T* p = &data();
    T old = *p, assumed;
    do {
      assumed = old;
      old     = Kokkos::atomic_compare_exchange(p, assumed, T(assumed + (T)1));
    } while (old != assumed);
  }
};
template <class T>
struct ExchLoop {
  Kokkos::View<T, Exec> data, data2;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    T old = Kokkos::atomic_exchange(&data(), (T)i);
    Kokkos::atomic_fetch_add(&data2(), old);
  }
};
template <class T>
void atomic_loops(int loop) {
  Kokkos::View<T, Exec> d("d"), d2("d2");
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, loop), AddLoop<T>{d});
  T r;
  Kokkos::deep_copy(r, d);
  CHECK_EQ(r, (T)loop);
  Kokkos::deep_copy(d, T(0));
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, loop), CASLoop<T>{d});
  Kokkos::deep_copy(r, d);
  CHECK_EQ(r, (T)loop);
  // exchange: the values 0..loop-1 each leave once, either through an
  // exchange's old value (summed into d2) or as the final content of d.
  Kokkos::deep_copy(d, T(0));
  Kokkos::deep_copy(d2, T(0));
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, loop), ExchLoop<T>{d, d2});
  T r2;
  Kokkos::deep_copy(r, d);
  Kokkos::deep_copy(r2, d2);
  CHECK_EQ(T(r + r2), T((T)loop * (T)(loop - 1) / 2));
}
KKMX_TEST(atomics_loops) {
  const int loop = 10000;
  atomic_loops<int>(loop);
  atomic_loops<unsigned int>(loop);
  atomic_loops<long>(loop);
  atomic_loops<unsigned long long>(loop);
  atomic_loops<float>(1000);
  atomic_loops<double>(loop);
}

template <class T>
struct MinMaxOps {
  Kokkos::View<T*, Exec> v;  // 0: min, 1: max, 2: fetch_add sum, 3: sub
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    const T x = T((i * 7919) % 10007) - T(5000);
    Kokkos::atomic_min(&v(0), x);
    Kokkos::atomic_max(&v(1), x);
    Kokkos::atomic_add(&v(2), x);
    Kokkos::atomic_sub(&v(3), x);
  }
};
template <class T>
void atomic_minmax_case(int n) {
  Kokkos::View<T*, Exec> v("v", 4);
  T init[4] = {std::numeric_limits<T>::max(), std::numeric_limits<T>::lowest(), T(0), T(0)};
  Kokkos::View<T*, Kokkos::HostSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>> hi(init, 4);
  Kokkos::deep_copy(v, hi);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), MinMaxOps<T>{v});
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), v);
  T mn = std::numeric_limits<T>::max(), mx = std::numeric_limits<T>::lowest(), sm = 0;
  for (int i = 0; i < n; ++i) {
    const T x = T((i * 7919) % 10007) - T(5000);
    mn = x < mn ? x : mn;
    mx = x > mx ? x : mx;
    sm += x;
  }
  CHECK_EQ(h(0), mn);
  CHECK_EQ(h(1), mx);
  CHECK_EQ(h(2), sm);
  CHECK_EQ(h(3), T(-sm));
}
KKMX_TEST(atomics_min_max_add_sub) {
  atomic_minmax_case<int>(100000);
  atomic_minmax_case<long long>(100000);
  atomic_minmax_case<double>(100000);  // integer-valued: exact in any order
  atomic_minmax_case<float>(1000);
}

// Sub-word (char, short) and wide (16-byte) types.  These go through
// compare-exchange loops on the containing 32-bit word and through the lock
// array: every success makes all concurrent attempts on the same word or lock
// retry, so the retry count grows with the square of the threads per word.
// 4096 threads over 256 slots keep it to a few hundred thousand attempts.
struct Wide {
  double a, b;
  KOKKOS_INLINE_FUNCTION Wide operator+(const Wide& o) const { return Wide{a + o.a, b + o.b}; }
  KOKKOS_INLINE_FUNCTION bool operator==(const Wide& o) const { return a == o.a && b == o.b; }
};
constexpr int kSlots = 256;
struct SubWordWideF {
  Kokkos::View<int8_t*, Exec> c;
  Kokkos::View<int16_t*, Exec> s;
  Kokkos::View<Wide*, Exec> w;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    Kokkos::atomic_add(&c(i % kSlots), int8_t(1));
    Kokkos::atomic_add(&s(i % kSlots), int16_t(2));
    Kokkos::atomic_add(&w(i % kSlots), Wide{1.0, double(i)});
  }
};
KKMX_TEST(atomics_subword_and_wide) {
  const int n = 4096 + 77;
  Kokkos::View<int8_t*, Exec> c("c", kSlots);
  Kokkos::View<int16_t*, Exec> s("s", kSlots);
  Kokkos::View<Wide*, Exec> w("w", kSlots);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), SubWordWideF{c, s, w});
  auto hc = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), c);
  auto hs = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), s);
  auto hw = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), w);
  for (int k = 0; k < kSlots; ++k) {
    int cnt = 0;
    double sum = 0;
    for (int i = k; i < n; i += kSlots) {
      ++cnt;
      sum += i;
    }
    CHECK_EQ(int(hc(k)), int(int8_t(cnt)));
    CHECK_EQ(int(hs(k)), int(int16_t(2 * cnt)));
    CHECK_EQ(hw(k).a, double(cnt));
    CHECK_EQ(hw(k).b, sum);
  }
}

// MemoryTraits<Atomic> Views (core/unit_test/TestAtomicViews.hpp).
struct AtomicViewF {
  Kokkos::View<int64_t*, Exec, Kokkos::MemoryTraits<Kokkos::Atomic>> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    v(i % 17) += i;
    v(17 + i % 3)++;
  }
};
KKMX_TEST(atomic_views) {
  const int n = 200000;
  Kokkos::View<int64_t*, Exec, Kokkos::MemoryTraits<Kokkos::Atomic>> v("v", 20);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), AtomicViewF{v});
  Kokkos::View<int64_t*, Exec> plain = v;
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), plain);
  for (int k = 0; k < 17; ++k) {
    int64_t want = 0;
    for (int i = k; i < n; i += 17) want += i;
    CHECK_EQ(h(k), want);
  }
  for (int k = 0; k < 3; ++k) {
    int64_t want = 0;
    for (int i = k; i < n; i += 3) ++want;
    CHECK_EQ(h(17 + k), want);
  }
}

// deep_copy: ViewFill / ViewCopy (core/src/Kokkos_CopyViews.hpp:58-420)
KKMX_TEST(deep_copy_fill_contiguous_and_strided) {
  Kokkos::View<double*, Exec> a("a", 1000003);
  Kokkos::deep_copy(a, 2.5);
  auto ha = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), a);
  for (size_t i = 0; i < ha.extent(0); ++i) CHECK_EQ(ha(i), 2.5);
  // a strided 30 x 20 window of a larger allocation: only its elements change
  Kokkos::View<double*, Exec> base("base", 64 * 40);
  Kokkos::deep_copy(base, -1.0);
  Kokkos::View<double**, Kokkos::LayoutStride, Exec> sv(base.data(), Kokkos::LayoutStride(30, 2, 20, 128));
  Kokkos::deep_copy(sv, 7.0);
  auto hb = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), base);
  for (int k = 0; k < 64 * 40; ++k) {
    const bool in = (k % 128) % 2 == 0 && (k % 128) / 2 < 30 && k / 128 < 20;
    CHECK_EQ(hb(k), in ? 7.0 : -1.0);
  }
  Kokkos::MI355X ex;
  Kokkos::deep_copy(ex, sv, 3.0);
  ex.fence();
  Kokkos::deep_copy(hb, base);
  for (int k = 0; k < 64 * 40; ++k) {
    const bool in = (k % 128) % 2 == 0 && (k % 128) / 2 < 30 && k / 128 < 20;
    CHECK_EQ(hb(k), in ? 3.0 : -1.0);
  }
}
struct Set2 {
  Kokkos::View<int64_t**, Kokkos::LayoutLeft, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j) const { v(i, j) = i * 1000 + j; }
};
struct Set3 {
  Kokkos::View<int64_t***, Kokkos::LayoutRight, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, const int64_t j, const int64_t k) const {
    v(i, j, k) = (i * 100 + j) * 100 + k;
  }
};
struct Set5 {
  Kokkos::View<int32_t*****, Kokkos::LayoutLeft, Exec> v;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t a, const int64_t b, const int64_t c, const int64_t d,
                                         const int64_t e) const {
    v(a, b, c, d, e) = int32_t((((a * 10 + b) * 10 + c) * 10 + d) * 10 + e);
  }
};
KKMX_TEST(deep_copy_between_layouts) {
  using MD2 = Kokkos::MDRangePolicy<Exec, Kokkos::Rank<2>, Kokkos::IndexType<int64_t>>;
  using MD3 = Kokkos::MDRangePolicy<Exec, Kokkos::Rank<3>, Kokkos::IndexType<int64_t>>;
  using MD5 = Kokkos::MDRangePolicy<Exec, Kokkos::Rank<5>, Kokkos::IndexType<int64_t>>;
  Kokkos::View<int64_t**, Kokkos::LayoutLeft, Exec> L("L", 37, 53);
  Kokkos::parallel_for(MD2({0, 0}, {37, 53}), Set2{L});
  Kokkos::View<int64_t**, Kokkos::LayoutRight, Exec> R("R", 37, 53);
  Kokkos::deep_copy(R, L);
  auto hr = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), R);
  for (int i = 0; i < 37; ++i)
    for (int j = 0; j < 53; ++j) CHECK_EQ(hr(i, j), int64_t(i) * 1000 + j);
  Kokkos::View<int64_t***, Kokkos::LayoutRight, Exec> R3("R3", 13, 7, 29);
  Kokkos::parallel_for(MD3({0, 0, 0}, {13, 7, 29}), Set3{R3});
  Kokkos::View<int64_t***, Kokkos::LayoutLeft, Exec> L3("L3", 13, 7, 29);
  Kokkos::MI355X ex;
  Kokkos::deep_copy(ex, L3, R3);  // ordered on ex, no fence
  ex.fence();
  auto hl3 = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), L3);
  for (int i = 0; i < 13; ++i)
    for (int j = 0; j < 7; ++j)
      for (int k = 0; k < 29; ++k) CHECK_EQ(hl3(i, j, k), (int64_t(i) * 100 + j) * 100 + k);
  Kokkos::View<int32_t*****, Kokkos::LayoutLeft, Exec> L5("L5", 3, 4, 5, 6, 7);
  Kokkos::parallel_for(MD5({0, 0, 0, 0, 0}, {3, 4, 5, 6, 7}), Set5{L5});
  Kokkos::View<int32_t*****, Kokkos::LayoutRight, Exec> R5("R5", 3, 4, 5, 6, 7);
  Kokkos::deep_copy(R5, L5);
  auto h5 = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), R5);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 5; ++c)
        for (int d = 0; d < 6; ++d)
          for (int e = 0; e < 7; ++e) CHECK_EQ(h5(a, b, c, d, e), int32_t((((a * 10 + b) * 10 + c) * 10 + d) * 10 + e));
}
KKMX_TEST(deep_copy_contiguous_and_strided_sources) {
  // contiguous, 16-byte path and element path (odd byte count)
  Kokkos::View<double*, Exec> x("x", 4099), y("y", 4099);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, 4099), KOKKOS_LAMBDA(const int i) { x(i) = 0.5 * i; });
  Kokkos::deep_copy(y, x);
  auto hy = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), y);
  for (int i = 0; i < 4099; ++i) CHECK_EQ(hy(i), 0.5 * i);
  Kokkos::View<char*, Exec> c1("c1", 1001), c2("c2", 1001);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, 1001), KOKKOS_LAMBDA(const int i) { c1(i) = char(i * 7); });
  Kokkos::deep_copy(c2, c1);
  auto hc = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), c2);
  for (int i = 0; i < 1001; ++i) CHECK_EQ(int(hc(i)), int(char(i * 7)));
  // strided source window -> contiguous LayoutLeft destination
  Kokkos::View<double*, Exec> base("base", 64 * 40);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, 64 * 40), KOKKOS_LAMBDA(const int k) { base(k) = k; });
  Kokkos::View<double**, Kokkos::LayoutStride, Exec> sv(base.data(), Kokkos::LayoutStride(30, 2, 20, 128));
  Kokkos::View<double**, Kokkos::LayoutLeft, Exec> d("d", 30, 20);
  Kokkos::deep_copy(d, sv);
  auto hd = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), d);
  for (int i = 0; i < 30; ++i)
    for (int j = 0; j < 20; ++j) CHECK_EQ(hd(i, j), double(i * 2 + j * 128));
}

// std algorithms: iterator forms (algorithms/src/std_algorithms)
struct MinI64Op {
  KOKKOS_FUNCTION int64_t operator()(const int64_t a, const int64_t b) const { return a < b ? a : b; }
};
struct TimesTwo {
  KOKKOS_FUNCTION int64_t operator()(const int64_t a) const { return 2 * a; }
};
KKMX_TEST(std_iterators_and_subranges) {
  namespace KE = Kokkos::Experimental;
  const int n = 1000;
  Kokkos::View<int64_t*, Exec> x("x", n), y("y", n);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), KOKKOS_LAMBDA(const int i) { x(i) = (i * 37) % 101 - 50; });
  Kokkos::deep_copy(y, -1);
  CHECK_EQ(KE::distance(KE::begin(x), KE::end(x)), ptrdiff_t(n));
  Kokkos::MI355X ex;
  // inclusive scan of x[10, 995) into y[10, ...)
  auto last = KE::inclusive_scan(ex, KE::cbegin(x) + 10, KE::cend(x) - 5, KE::begin(y) + 10);
  CHECK_EQ(KE::distance(KE::begin(y), last), ptrdiff_t(n - 5));
  auto hx = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), x);
  auto hy = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), y);
  int64_t s = 0;
  for (int i = 0; i < n; ++i) {
    if (i >= 10 && i < n - 5) {
      s += hx(i);
      CHECK_EQ(hy(i), s);
    } else {
      CHECK_EQ(hy(i), int64_t(-1));
    }
  }
  // running minimum with an initial value, and a transform scan
  KE::inclusive_scan(ex, x, y, MinI64Op{}, int64_t(-7));
  Kokkos::deep_copy(hy, y);
  int64_t m = -7;
  for (int i = 0; i < n; ++i) {
    m = hx(i) < m ? hx(i) : m;
    CHECK_EQ(hy(i), m);
  }
  KE::transform_inclusive_scan(ex, x, y, KE::Impl::PlusOp{}, TimesTwo{});
  Kokkos::deep_copy(hy, y);
  s = 0;
  for (int i = 0; i < n; ++i) {
    s += 2 * hx(i);
    CHECK_EQ(hy(i), s);
  }
  CHECK_EQ(KE::reduce(ex, x, int64_t(3), MinI64Op{}), int64_t(-50));
  CHECK_EQ(KE::transform_reduce(ex, x, int64_t(0), KE::Impl::PlusOp{}, TimesTwo{}), 2 * s / 2);
}
KKMX_TEST(std_strided_views) {
  namespace KE = Kokkos::Experimental;
  // every third element of a buffer, scanned into every second of another
  Kokkos::View<int32_t*, Exec> a("a", 300), b("b", 200);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, 300), KOKKOS_LAMBDA(const int i) { a(i) = i; });
  Kokkos::deep_copy(b, 0);
  Kokkos::View<int32_t*, Kokkos::LayoutStride, Exec> as(a.data(), Kokkos::LayoutStride(100, 3));
  Kokkos::View<int32_t*, Kokkos::LayoutStride, Exec> bs(b.data(), Kokkos::LayoutStride(100, 2));
  Kokkos::MI355X ex;
  auto last = KE::exclusive_scan(ex, as, bs, int32_t(5));
  CHECK_EQ(KE::distance(KE::begin(bs), last), ptrdiff_t(100));
  auto hb = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), b);
  int32_t s = 5;
  for (int k = 0; k < 100; ++k) {
    CHECK_EQ(hb(2 * k), s);
    CHECK_EQ(hb(2 * k + 1), 0);
    s += 3 * k;
  }
  CHECK_EQ(KE::reduce(ex, as), int32_t(3 * 99 * 100 / 2));
}

// ScatterView (containers/src/Kokkos_ScatterView.hpp; restates the sum /
// prod / min / max checks of containers/unit_tests/TestScatterView.hpp)
template <class SV>
struct ScatterSumF {
  SV sv;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    auto a = sv.access();
    a(i % 37) += int64_t(i);
    a(i % 37 + 37) -= 1;
    a[74] += 1;
  }
};
KKMX_TEST(scatter_view_sum_into_original_and_contribute) {
  namespace KE = Kokkos::Experimental;
  const int n = 100000;
  Kokkos::View<int64_t*, Exec> bins("bins", 75);
  Kokkos::deep_copy(bins, 0);
  auto sv = KE::create_scatter_view(bins);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), ScatterSumF<decltype(sv)>{sv});
  KE::contribute(bins, sv);  // same data: nothing to do
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), bins);
  for (int b = 0; b < 37; ++b) {
    int64_t want = 0, cnt = 0;
    for (int i = b; i < n; i += 37) {
      want += i;
      ++cnt;
    }
    CHECK_EQ(h(b), want);
    CHECK_EQ(h(b + 37), -cnt);
  }
  CHECK_EQ(h(74), int64_t(n));
  // a separate ScatterView, reset, then contributed into an existing View
  KE::ScatterView<int64_t*, Kokkos::LayoutRight, Exec> sv2("sv2", 75);
  sv2.reset();
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), ScatterSumF<decltype(sv2)>{sv2});
  KE::contribute(bins, sv2);
  Kokkos::deep_copy(h, bins);
  CHECK_EQ(h(74), int64_t(2 * n));
  CHECK_EQ(h(0), 2 * ((n - 1) / 37) * ((n - 1) / 37 + 1) / 2 * 37);
}
template <class SMax, class SMin, class SProd>
struct ScatterMinMaxProdF {
  SMax mx;
  SMin mn;
  SProd pr;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    const double v = double((i * 7919) % 10007) - 5000.0;
    mx.access()(i % 16).update(v);
    mn.access()(i % 16).update(v);
    if (i < 64) pr.access()(i % 4) *= int64_t(2);
  }
};
KKMX_TEST(scatter_view_min_max_prod) {
  namespace KE = Kokkos::Experimental;
  const int n = 50000;
  Kokkos::View<double*, Exec> vmax("vmax", 16), vmin("vmin", 16);
  Kokkos::View<int64_t*, Exec> vprod("vprod", 4);
  auto smax  = KE::create_scatter_view(KE::ScatterMax{}, vmax);
  auto smin  = KE::create_scatter_view(KE::ScatterMin{}, vmin);
  auto sprod = KE::create_scatter_view(KE::ScatterProd{}, vprod);
  smax.reset();
  smin.reset();
  sprod.reset();
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n),
                       ScatterMinMaxProdF<decltype(smax), decltype(smin), decltype(sprod)>{smax, smin, sprod});
  auto hmax = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), vmax);
  auto hmin = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), vmin);
  auto hpr  = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), vprod);
  for (int b = 0; b < 16; ++b) {
    double wmax = -1e300, wmin = 1e300;
    for (int i = b; i < n; i += 16) {
      const double v = double((i * 7919) % 10007) - 5000.0;
      wmax = v > wmax ? v : wmax;
      wmin = v < wmin ? v : wmin;
    }
    CHECK_EQ(hmax(b), wmax);
    CHECK_EQ(hmin(b), wmin);
  }
  for (int b = 0; b < 4; ++b) CHECK_EQ(hpr(b), int64_t(1) << 16);
}

KKMX_TEST(scatter_view_duplicated_replicas) {
  namespace KE = Kokkos::Experimental;
  const int n = 300000;
  Kokkos::View<int64_t*, Exec> bins("bins", 75);
  Kokkos::deep_copy(bins, 5);  // contributions add to what is there
  auto sv = KE::create_scatter_view(KE::ScatterSum{}, KE::ScatterDuplicated{}, KE::ScatterAtomic{}, bins);
  CHECK(sv.replicas() > 1);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), ScatterSumF<decltype(sv)>{sv});
  KE::contribute(bins, sv);
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), bins);
  for (int b = 0; b < 37; ++b) {
    int64_t want = 5, cnt = 0;
    for (int i = b; i < n; i += 37) {
      want += i;
      ++cnt;
    }
    CHECK_EQ(h(b), want);
    CHECK_EQ(h(b + 37), 5 - cnt);
  }
  CHECK_EQ(h(74), int64_t(n) + 5);
  // reset_except, a second pass: the replicas start from zero again
  sv.reset_except(bins);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), ScatterSumF<decltype(sv)>{sv});
  KE::contribute(bins, sv);
  Kokkos::deep_copy(h, bins);
  CHECK_EQ(h(74), int64_t(2 * n) + 5);
}

// =========================================================================
// subview (core/unit_test/TestViewSubview.hpp: layouts of the results,
// values read and written through them, deep_copy of non-contiguous
// windows, the allocation kept alive by a subview).
// =========================================================================
template <class V2>
struct ReadSub2 {
  V2 s;
  Kokkos::View<double*, Exec> out;
  int n1;
  KOKKOS_FUNCTION void operator()(const int i) const { out(i) = s(i / n1, i % n1); }
};
template <class V1>
struct ReadSub1 {
  V1 s;
  Kokkos::View<double*, Exec> out;
  KOKKOS_FUNCTION void operator()(const int i) const { out(i) = s(i); }
};
template <class V2>
struct NegateSub2 {
  V2 s;
  int n1;
  KOKKOS_FUNCTION void operator()(const int i) const { s(i / n1, i % n1) = -s(i / n1, i % n1); }
};

template <class V2>
std::vector<double> read2(const V2& s) {
  const int n0 = int(s.extent(0)), n1 = int(s.extent(1));
  Kokkos::View<double*, Exec> out("out", size_t(n0 * n1));
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n0 * n1), ReadSub2<V2>{s, out, n1});
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
  return std::vector<double>(h.data(), h.data() + n0 * n1);
}
template <class V1>
std::vector<double> read1(const V1& s) {
  const int n = int(s.extent(0));
  Kokkos::View<double*, Exec> out("out", size_t(n));
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), ReadSub1<V1>{s, out});
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
  return std::vector<double>(h.data(), h.data() + n);
}

KKMX_TEST(subview_layouts_values_and_copies) {
  using R3 = Kokkos::View<double***, Kokkos::LayoutRight, Exec>;
  using L3 = Kokkos::View<double***, Kokkos::LayoutLeft, Exec>;
  const int N0 = 4, N1 = 5, N2 = 6;
  auto f = [](int i, int j, int k) { return 100.0 * i + 10.0 * j + k; };
  R3 v("v", N0, N1, N2);
  L3 w("w", N0, N1, N2);
  {
    auto hv = Kokkos::create_mirror_view(v);
    auto hw = Kokkos::create_mirror_view(w);
    for (int i = 0; i < N0; ++i)
      for (int j = 0; j < N1; ++j)
        for (int k = 0; k < N2; ++k) hv(i, j, k) = hw(i, j, k) = f(i, j, k);
    Kokkos::deep_copy(v, hv);
    Kokkos::deep_copy(w, hw);
  }
  // LayoutRight source
  auto s1 = Kokkos::subview(v, 2, Kokkos::make_pair(1, 4), Kokkos::ALL);
  static_assert(std::is_same<decltype(s1)::array_layout, Kokkos::LayoutRight>::value, "");
  CHECK_EQ(s1.extent(0), size_t(3));
  CHECK_EQ(s1.extent(1), size_t(N2));
  CHECK_EQ(s1.stride(0), size_t(N2));
  auto r1 = read2(s1);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < N2; ++b) CHECK_EQ(r1[a * N2 + b], f(2, 1 + a, b));
  // rank 2 ending at the last dimension: LayoutRight with a padded stride
  auto s2 = Kokkos::subview(v, Kokkos::ALL, 3, std::make_pair(2, 5));
  static_assert(std::is_same<decltype(s2)::array_layout, Kokkos::LayoutRight>::value, "");
  CHECK_EQ(s2.stride(0), size_t(N1 * N2));
  auto r2 = read2(s2);
  for (int a = 0; a < N0; ++a)
    for (int b = 0; b < 3; ++b) CHECK_EQ(r2[a * 3 + b], f(a, 3, 2 + b));
  auto s2b = Kokkos::subview(v, Kokkos::ALL, Kokkos::make_pair(1, 3), 4);
  static_assert(std::is_same<decltype(s2b)::array_layout, Kokkos::LayoutStride>::value, "");
  auto r2b = read2(s2b);
  for (int a = 0; a < N0; ++a)
    for (int b = 0; b < 2; ++b) CHECK_EQ(r2b[a * 2 + b], f(a, 1 + b, 4));
  Kokkos::View<double*, Kokkos::LayoutRight, Exec> row = Kokkos::subview(v, 1, 2, Kokkos::ALL);
  CHECK(row.span_is_contiguous());
  auto r3 = read1(row);
  for (int k = 0; k < N2; ++k) CHECK_EQ(r3[k], f(1, 2, k));
  auto s4 = Kokkos::subview(v, 1, 2, 3);
  static_assert(decltype(s4)::rank == 0, "");
  double x4 = 0;
  Kokkos::deep_copy(x4, s4);
  CHECK_EQ(x4, f(1, 2, 3));
  auto s5 = Kokkos::subview(s1, 1, Kokkos::ALL);  // a subview of a subview
  auto r5 = read1(s5);
  for (int k = 0; k < N2; ++k) CHECK_EQ(r5[k], f(2, 2, k));
  // LayoutLeft source: a padded rank-2 window keeps LayoutLeft
  auto t1 = Kokkos::subview(w, Kokkos::make_pair(1, 3), Kokkos::make_pair(0, 2), 4);
  static_assert(std::is_same<decltype(t1)::array_layout, Kokkos::LayoutLeft>::value, "");
  CHECK(!t1.span_is_contiguous());
  auto q1 = read2(t1);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) CHECK_EQ(q1[a * 2 + b], f(1 + a, b, 4));
  auto t2 = Kokkos::subview(w, Kokkos::ALL, 2, Kokkos::ALL);
  static_assert(std::is_same<decltype(t2)::array_layout, Kokkos::LayoutLeft>::value, "");
  auto q2 = read2(t2);
  for (int a = 0; a < N0; ++a)
    for (int b = 0; b < N2; ++b) CHECK_EQ(q2[a * N2 + b], f(a, 2, b));
  // write through a strided window; nothing else changes
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, N0 * 3), NegateSub2<decltype(s2)>{s2, 3});
  {
    auto hv = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), v);
    for (int i = 0; i < N0; ++i)
      for (int j = 0; j < N1; ++j)
        for (int k = 0; k < N2; ++k)
          CHECK_EQ(hv(i, j, k), (j == 3 && k >= 2 && k < 5) ? -f(i, j, k) : f(i, j, k));
  }
  // deep_copy into and out of non-contiguous windows
  Kokkos::View<double**, Kokkos::LayoutLeft, Exec> dense("dense", 2, 2);
  Kokkos::deep_copy(dense, t1);
  auto hd = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), dense);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) CHECK_EQ(hd(a, b), f(1 + a, b, 4));
  Kokkos::deep_copy(t1, 7.0);
  auto hw = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), w);
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < N1; ++j)
      for (int k = 0; k < N2; ++k)
        CHECK_EQ(hw(i, j, k), (i >= 1 && i < 3 && j < 2 && k == 4) ? 7.0 : f(i, j, k));
  // across spaces: a padded device window to a host mirror and back
  auto ht1 = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), t1);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) CHECK_EQ(ht1(a, b), 7.0);
  auto hs2 = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), s2b);
  for (int a = 0; a < N0; ++a)
    for (int b = 0; b < 2; ++b) CHECK_EQ(hs2(a, b), f(a, 1 + b, 4));
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) ht1(a, b) = 1000.0 + 10 * a + b;
  Kokkos::deep_copy(t1, ht1);
  auto hw2 = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), w);
  for (int i = 0; i < N0; ++i)
    for (int j = 0; j < N1; ++j)
      for (int k = 0; k < N2; ++k)
        CHECK_EQ(hw2(i, j, k), (i >= 1 && i < 3 && j < 2 && k == 4) ? 1000.0 + 10 * (i - 1) + j : f(i, j, k));
  // a subview keeps the allocation alive
  Kokkos::View<double*, Kokkos::LayoutRight, Exec> keep;
  {
    Kokkos::View<double**, Kokkos::LayoutRight, Exec> tmp("tmp", 3, 8);
    Kokkos::deep_copy(tmp, 2.5);
    keep = Kokkos::subview(tmp, 2, Kokkos::ALL);
    CHECK_EQ(tmp.use_count(), 2);
  }
  CHECK_EQ(keep.use_count(), 1);
  auto rk = read1(keep);
  for (int k = 0; k < 8; ++k) CHECK_EQ(rk[k], 2.5);
  // LayoutRight -> LayoutStride assignment keeps the strides
  Kokkos::View<double**, Kokkos::LayoutStride, Exec> st = s1;
  CHECK_EQ(st.stride(0), size_t(N2));
  CHECK_EQ(read2(st)[N2 + 1], f(2, 2, 1));
}

// resize keeps elements by index (core/unit_test/TestResize.hpp): rank 2 in
// both layouts, one extent growing and the other shrinking.
template <class L>
void resize_case() {
  Kokkos::View<int**, L, Exec> v("v", 3, 4);
  auto h = Kokkos::create_mirror_view(v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) h(i, j) = 10 * i + j;
  Kokkos::deep_copy(v, h);
  Kokkos::resize(v, 5, 3);
  CHECK_EQ(v.extent(0), size_t(5));
  CHECK_EQ(v.extent(1), size_t(3));
  auto g = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), v);
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 3; ++j) CHECK_EQ(g(i, j), i < 3 ? 10 * i + j : 0);
}
KKMX_TEST(resize_keeps_elements_by_index) {
  resize_case<Kokkos::LayoutLeft>();
  resize_case<Kokkos::LayoutRight>();
}

// =========================================================================
// Nested MD ranges (core/unit_test/TestTeamMDRange.hpp): every point of a
// per-team box visited exactly once at each level (threads, threads x
// vector, one thread's vector lanes), both iteration directions, and the
// sum reduction received by every lane.
// =========================================================================
struct TeamMDProbe {
  Kokkos::View<int****, Kokkos::LayoutRight, Exec> hits;  // (team, i, j, k)
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> sums;     // (team, level)
  KOKKOS_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type& t) const {
    using M = Kokkos::TeamPolicy<Exec>::member_type;
    const int r = t.league_rank();
    const int a = hits.extent_int(1), b = hits.extent_int(2), c = hits.extent_int(3);
    // a thread-level body runs on each of the thread's vector lanes: count once
    Kokkos::parallel_for(Kokkos::TeamThreadMDRange<Kokkos::Rank<3>, M>(t, a, b, c),
                         [&](const int i, const int j, const int k) {
                           Kokkos::single(Kokkos::PerThread(t), [&]() { Kokkos::atomic_add(&hits(r, i, j, k), 1); });
                         });
    t.team_barrier();
    Kokkos::parallel_for(
        Kokkos::TeamVectorMDRange<Kokkos::Rank<3, Kokkos::Iterate::Right>, M>(t, a, b, c),
        [&](const int i, const int j, const int k) { Kokkos::atomic_add(&hits(r, i, j, k), 10); });
    t.team_barrier();
    if (t.team_rank() == 0)
      Kokkos::parallel_for(Kokkos::ThreadVectorMDRange<Kokkos::Rank<2>, M>(t, a, b), [&](const int i, const int j) {
        for (int k = 0; k < c; ++k) Kokkos::atomic_add(&hits(r, i, j, k), 100);
      });
    t.team_barrier();
    int s0 = 0, s1 = 0, s2 = 0;
    Kokkos::parallel_reduce(
        Kokkos::TeamThreadMDRange<Kokkos::Rank<3>, M>(t, a, b, c),
        [&](const int i, const int j, const int k, int& u) { u += i * 100 + j * 10 + k; }, s0);
    Kokkos::parallel_reduce(
        Kokkos::TeamVectorMDRange<Kokkos::Rank<2, Kokkos::Iterate::Right>, M>(t, a, b),
        [&](const int i, const int j, int& u) { u += i * 10 + j; }, s1);
    Kokkos::parallel_reduce(Kokkos::ThreadVectorMDRange<Kokkos::Rank<2>, M>(t, b, c),
                            [&](const int j, const int k, int& u) { u += j + k; }, s2);
    // every lane holds the sums; lane 0 of each thread records them
    Kokkos::single(Kokkos::PerThread(t), [&]() {
      Kokkos::atomic_max(&sums(r, 0), s0);
      Kokkos::atomic_max(&sums(r, 1), s1);
      Kokkos::atomic_max(&sums(r, 2), s2);
    });
  }
};

KKMX_TEST(team_md_ranges) {
  const int L = 6, A = 5, B = 7, C = 3;
  Kokkos::View<int****, Kokkos::LayoutRight, Exec> hits("hits", L, A, B, C);
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> sums("sums", L, 3);
  Kokkos::deep_copy(sums, 0);
  Kokkos::parallel_for(Kokkos::TeamPolicy<Exec>(L, 8, 8), TeamMDProbe{hits, sums});
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), hits);
  auto s = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), sums);
  int w0 = 0, w1 = 0, w2 = 0;
  for (int i = 0; i < A; ++i)
    for (int j = 0; j < B; ++j)
      for (int k = 0; k < C; ++k) w0 += i * 100 + j * 10 + k;
  for (int i = 0; i < A; ++i)
    for (int j = 0; j < B; ++j) w1 += i * 10 + j;
  for (int j = 0; j < B; ++j)
    for (int k = 0; k < C; ++k) w2 += j + k;
  for (int r = 0; r < L; ++r) {
    for (int i = 0; i < A; ++i)
      for (int j = 0; j < B; ++j)
        for (int k = 0; k < C; ++k) CHECK_EQ(h(r, i, j, k), 111);
    CHECK_EQ(s(r, 0), w0);
    CHECK_EQ(s(r, 1), w1);
    CHECK_EQ(s(r, 2), w2);  // each thread reduced the whole box on its own vector lanes
  }
}

// =========================================================================
// Kokkos::complex (core/unit_test/TestComplex.hpp): arithmetic and the
// functions in a kernel against std::complex on the host (within a few ulp
// of the larger part), a Sum reduction of complex values and complex
// atomic_add (integer-valued parts, so the sums are exact).
// =========================================================================
using cplx = Kokkos::complex<double>;
struct ComplexProbe {
  Kokkos::View<const cplx*, Exec> z;
  Kokkos::View<cplx**, Kokkos::LayoutRight, Exec> out;
  KOKKOS_FUNCTION void operator()(const int i) const {
    const cplx a = z(i), b = z((i + 3) % z.extent_int(0));
    int c         = 0;
    out(i, c++)   = a + b;
    out(i, c++)   = a - 2.0 * b;
    out(i, c++)   = a * b;
    out(i, c++)   = a / b;
    out(i, c++)   = Kokkos::conj(a) + cplx(Kokkos::abs(a), Kokkos::arg(a));
    out(i, c++)   = Kokkos::exp(a);
    out(i, c++)   = Kokkos::log(a);
    out(i, c++)   = Kokkos::sqrt(a);
    out(i, c++)   = Kokkos::pow(a, b);
    out(i, c++)   = Kokkos::sin(a) + Kokkos::cosh(b);
  }
};
struct ComplexSum {
  Kokkos::View<const cplx*, Exec> z;
  KOKKOS_FUNCTION void operator()(const int i, cplx& s) const { s += z(i); }
};
struct ComplexAtomic {
  Kokkos::View<cplx*, Exec> bins;
  KOKKOS_FUNCTION void operator()(const int i) const { Kokkos::atomic_add(&bins(i % 5), cplx(1.0, double(i % 3))); }
};

KKMX_TEST(complex_arithmetic_reduce_atomic) {
  const int n = 512;
  Kokkos::View<cplx*, Exec> z("z", n);
  auto hz = Kokkos::create_mirror_view(z);
  for (int i = 0; i < n; ++i) hz(i) = cplx(0.25 + (i % 37) / 8.0, ((i * 7) % 29) / 9.0 - 1.5);
  Kokkos::deep_copy(z, hz);
  Kokkos::View<cplx**, Kokkos::LayoutRight, Exec> out("out", n, 10);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), ComplexProbe{z, out});
  auto g = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
  for (int i = 0; i < n; ++i) {
    const std::complex<double> a(hz(i).real(), hz(i).imag());
    const std::complex<double> b(hz((i + 3) % n).real(), hz((i + 3) % n).imag());
    const std::complex<double> want[] = {a + b, a - 2.0 * b, a * b, a / b,
                                         std::conj(a) + std::complex<double>(std::abs(a), std::arg(a)),
                                         std::exp(a), std::log(a), std::sqrt(a), std::pow(a, b),
                                         std::sin(a) + std::cosh(b)};
    for (int c = 0; c < 10; ++c) {
      const double scale = std::max(std::abs(want[c].real()), std::abs(want[c].imag())) + 1e-300;
      CHECK(std::abs(g(i, c).real() - want[c].real()) <= 1e-14 * scale);
      CHECK(std::abs(g(i, c).imag() - want[c].imag()) <= 1e-14 * scale);
    }
  }
  // exact sums on integer-valued parts
  Kokkos::View<cplx*, Exec> w("w", n);
  auto hw = Kokkos::create_mirror_view(w);
  cplx want(0.0, 0.0);
  for (int i = 0; i < n; ++i) {
    hw(i) = cplx(double(i), double(-2 * i));
    want += hw(i);
  }
  Kokkos::deep_copy(w, hw);
  cplx s;
  Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec>(0, n), ComplexSum{w}, s);
  CHECK(s == want);
  Kokkos::View<cplx*, Exec> bins("bins", 5);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, 3000), ComplexAtomic{bins});
  auto hb = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), bins);
  for (int k = 0; k < 5; ++k) {
    double re = 0, im = 0;
    for (int i = k; i < 3000; i += 5) {
      re += 1.0;
      im += double(i % 3);
    }
    CHECK(hb(k) == cplx(re, im));
  }
}

// =========================================================================
// kokkos_malloc / kokkos_realloc / kokkos_free (core/unit_test/
// TestMalloc.hpp) and a runtime-length reduction into a C array
// (TestReduce.hpp:123-140, the array result decays to a pointer).
// =========================================================================
struct ArraySums {
  using value_type = double[];
  const unsigned value_count = 3;
  Kokkos::View<double*, Exec> x;
  KOKKOS_FUNCTION void operator()(const int i, double* s) const {
    s[0] += x(i);
    s[1] += 1;
    s[2] += x(i) * x(i);
  }
  KOKKOS_FUNCTION void init(double* s) const { s[0] = s[1] = s[2] = 0; }
  KOKKOS_FUNCTION void join(double* a, const double* b) const {
    for (int k = 0; k < 3; ++k) a[k] += b[k];
  }
};
struct FillRaw {
  int* p;
  KOKKOS_FUNCTION void operator()(const int i) const { p[i] = 3 * i; }
};

KKMX_TEST(raw_allocations_and_array_results) {
  int* p = static_cast<int*>(Kokkos::kokkos_malloc<>("raw", 1000 * sizeof(int)));
  CHECK(p != nullptr);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, 1000), FillRaw{p});
  Kokkos::fence();
  p = static_cast<int*>(Kokkos::kokkos_realloc<>(p, 2000 * sizeof(int)));  // keeps the first 1000
  std::vector<int> h(1000);
  KKMX_SAFE_CALL(mx_memcpy(h.data(), p, 1000 * sizeof(int)));
  for (int i = 0; i < 1000; ++i) CHECK_EQ(h[i], 3 * i);
  Kokkos::kokkos_free<>(p);
  void* hp = Kokkos::kokkos_malloc<Kokkos::HostSpace>(64);
  hp       = Kokkos::kokkos_realloc<Kokkos::HostSpace>(hp, 128);
  Kokkos::kokkos_free<Kokkos::HostSpace>(hp);

  const int n = 4096;
  Kokkos::View<double*, Exec> x("x", n);
  auto hx = Kokkos::create_mirror_view(x);
  for (int i = 0; i < n; ++i) hx(i) = i % 100;
  Kokkos::deep_copy(x, hx);
  double sums[3];
  Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec>(0, n), ArraySums{3, x}, sums);
  CHECK_EQ(sums[0], 202560.0);    // 40 * 4950 + sum 0..95
  CHECK_EQ(sums[1], double(n));
  CHECK_EQ(sums[2], 13424320.0);  // 40 * 328350 + sum of squares 0..95
}

// =========================================================================
// A typical application written with lambdas against the default spaces:
// Range/MDRange/Team dispatch, reducers, a scan with total, team scratch,
// atomics, subviews, mirrors, Kokkos::Array, sort, ScatterView, a std
// algorithm and resize; every result is checked against its closed form.
// =========================================================================
KKMX_TEST(typical_application_with_lambdas) {
  using ExecSpace = Kokkos::DefaultExecutionSpace;
  const int N = 1000;
  Kokkos::View<double*> a("a", N), b("b", N);
  Kokkos::View<double* [3]> pos("pos", N);
  Kokkos::View<double**, Kokkos::LayoutRight> m("m", N, 8);
  Kokkos::parallel_for("init", N, KOKKOS_LAMBDA(const int i) {
    a(i)      = i;
    b(i)      = 2.0 * i;
    pos(i, 0) = i;
    pos(i, 2) = -i;
  });
  Kokkos::parallel_for(Kokkos::RangePolicy<ExecSpace, Kokkos::Schedule<Kokkos::Dynamic>>(0, N),
                       KOKKOS_LAMBDA(const int i) { Kokkos::atomic_add(&a(i % 10), 1.0); });
  double sum = 0;
  Kokkos::parallel_reduce("dot", N, KOKKOS_LAMBDA(const int i, double& s) { s += a(i) * b(i); }, sum);
  CHECK_EQ(sum, 665676000.0);  // 2 * sum i^2 + 200 * sum_{i<10} i
  double mx = 0;
  Kokkos::parallel_reduce(
      N, KOKKOS_LAMBDA(const int i, double& v) { if (b(i) > v) v = b(i); }, Kokkos::Max<double>(mx));
  CHECK_EQ(mx, 1998.0);
  Kokkos::MinLoc<double, int>::value_type ml;
  Kokkos::parallel_reduce(
      N,
      KOKKOS_LAMBDA(const int i, Kokkos::MinLoc<double, int>::value_type& v) {
        if (b(i) < v.val) {
          v.val = b(i);
          v.loc = i;
        }
      },
      Kokkos::MinLoc<double, int>(ml));
  CHECK_EQ(ml.loc, 0);
  Kokkos::View<int*> cnt("cnt", N), off("off", N);
  Kokkos::deep_copy(cnt, 1);
  int total = 0;
  Kokkos::parallel_scan(
      "offsets", N,
      KOKKOS_LAMBDA(const int i, int& u, const bool final) {
        if (final) off(i) = u;
        u += cnt(i);
      },
      total);
  CHECK_EQ(total, N);
  Kokkos::parallel_for(Kokkos::MDRangePolicy<Kokkos::Rank<2>>({0, 0}, {N, 8}, {64, 4}),
                       KOKKOS_LAMBDA(const int i, const int j) { m(i, j) = i + j; });
  using team_t      = Kokkos::TeamPolicy<>::member_type;
  using scratch_t   = Kokkos::View<double*, ExecSpace::scratch_memory_space>;
  const int scratch = int(scratch_t::shmem_size(64));
  Kokkos::parallel_for(Kokkos::TeamPolicy<>(N / 10, Kokkos::AUTO).set_scratch_size(0, Kokkos::PerTeam(scratch)),
                       KOKKOS_LAMBDA(const team_t& t) {
                         scratch_t s(t.team_scratch(0), 64);
                         Kokkos::parallel_for(Kokkos::TeamThreadRange(t, 10), [&](const int k) { s(k) = k; });
                         t.team_barrier();
                         double r = 0;
                         Kokkos::parallel_reduce(
                             Kokkos::TeamThreadRange(t, 10), [&](const int k, double& u) { u += s(k); }, r);
                         Kokkos::single(Kokkos::PerTeam(t), [&]() { a(t.league_rank()) += r; });
                       });
  auto row = Kokkos::subview(m, 3, Kokkos::ALL);
  auto hr  = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), row);
  for (int j = 0; j < 8; ++j) CHECK_EQ(hr(j), 3.0 + j);
  auto hoff = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), off);
  CHECK_EQ(hoff(N - 1), N - 1);
  Kokkos::Array<double, 3> arr = {1, 2, 3};
  CHECK_EQ(arr[1], 2.0);
  Kokkos::sort(a);
  auto ha = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), a);
  for (int i = 1; i < N; ++i) CHECK(ha(i - 1) <= ha(i));
  Kokkos::Experimental::ScatterView<double*> sv(b);
  Kokkos::parallel_for(N, KOKKOS_LAMBDA(const int i) {
    auto acc = sv.access();
    acc(i % 7) += 1.0;
  });
  Kokkos::Experimental::contribute(b, sv);
  auto hb = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), b);
  CHECK_EQ(hb(0), 0.0 + 143);  // 1000 = 7 * 142 + 6: indices 0..5 get 143 hits
  CHECK_EQ(hb(6), 12.0 + 142);
  namespace KE = Kokkos::Experimental;
  const double r2 = KE::reduce(ExecSpace(), a);
  CHECK_EQ(r2, 505000.0);  // 499500 + 10 * 100 atomics + 100 teams * 45
  Kokkos::resize(a, 2 * N);
  auto hz = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), a);
  CHECK_EQ(hz(N - 1), ha(N - 1));
  CHECK_EQ(hz(2 * N - 1), 0.0);
}

// =========================================================================
// Math functions in kernels (core/unit_test/TestMathematicalFunctions.hpp):
// the correctly rounded ones (sqrt, fma, rounding, fabs, fmin/fmax,
// copysign, integer abs, min/max/clamp) equal the host's bit for bit; the
// others agree with the host C library within 4 ulp.  And partition_space
// (core/unit_test/TestExecSpacePartitioning.hpp): independent instances.
// =========================================================================
constexpr int kMathCols = 24;
struct MathProbe {
  Kokkos::View<const double*, Exec> x;
  Kokkos::View<double**, Kokkos::LayoutRight, Exec> out;
  KOKKOS_FUNCTION void operator()(const int i) const {
    const double v = x(i), w = x((i + 7) % x.extent_int(0));
    int c = 0;
    out(i, c++) = Kokkos::sqrt(Kokkos::fabs(v));
    out(i, c++) = Kokkos::fma(v, w, 0.25);
    out(i, c++) = Kokkos::floor(v);
    out(i, c++) = Kokkos::ceil(v);
    out(i, c++) = Kokkos::trunc(v);
    out(i, c++) = Kokkos::round(v);
    out(i, c++) = Kokkos::fmin(v, w);
    out(i, c++) = Kokkos::fmax(v, w);
    out(i, c++) = Kokkos::copysign(v, w);
    out(i, c++) = double(Kokkos::abs(int(v * 1000)));
    out(i, c++) = Kokkos::clamp(v, -1.0, 1.5) + Kokkos::min(v, w) - Kokkos::max(v, w);
    out(i, c++) = double(Kokkos::sqrt(float(Kokkos::abs(v))));  // float overload
    // within a few ulp of the host
    out(i, c++) = Kokkos::exp(v);
    out(i, c++) = Kokkos::log(Kokkos::fabs(v) + 1e-3);
    out(i, c++) = Kokkos::sin(v);
    out(i, c++) = Kokkos::cos(v);
    out(i, c++) = Kokkos::tanh(v);
    out(i, c++) = Kokkos::pow(Kokkos::fabs(v) + 0.5, w);
    out(i, c++) = Kokkos::atan2(v, w);
    out(i, c++) = Kokkos::cbrt(v);
    out(i, c++) = Kokkos::hypot(v, w);
    out(i, c++) = Kokkos::erf(v);
    out(i, c++) = Kokkos::expm1(v);
    out(i, c++) = Kokkos::log1p(Kokkos::fabs(v));
  }
};

struct PartSum {
  Kokkos::View<int*, Exec> a;
  KOKKOS_FUNCTION void operator()(const int i, int64_t& s) const { s += a(i); }
};

KKMX_TEST(math_functions_and_partition_space) {
  const int n = 4096;
  Kokkos::View<double*, Exec> x("x", n);
  auto hx = Kokkos::create_mirror_view(x);
  for (int i = 0; i < n; ++i) hx(i) = std::ldexp(double((int64_t(i) * 2654435761u) % 100003) - 50001.0, -13);
  Kokkos::deep_copy(x, hx);
  Kokkos::View<double**, Kokkos::LayoutRight, Exec> out("out", n, kMathCols);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), MathProbe{x, out});
  auto g = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
  for (int i = 0; i < n; ++i) {
    const double v = hx(i), w = hx((i + 7) % n);
    const double exact[] = {std::sqrt(std::fabs(v)), std::fma(v, w, 0.25), std::floor(v), std::ceil(v),
                            std::trunc(v), std::round(v), std::fmin(v, w), std::fmax(v, w), std::copysign(v, w),
                            double(std::abs(int(v * 1000))),
                            std::min(std::max(v, -1.0), 1.5) + std::min(v, w) - std::max(v, w),
                            double(std::sqrt(float(std::fabs(v))))};
    for (int c = 0; c < 12; ++c) CHECK_EQ(g(i, c), exact[c]);
    const double near[] = {std::exp(v), std::log(std::fabs(v) + 1e-3), std::sin(v), std::cos(v), std::tanh(v),
                           std::pow(std::fabs(v) + 0.5, w), std::atan2(v, w), std::cbrt(v), std::hypot(v, w),
                           std::erf(v), std::expm1(v), std::log1p(std::fabs(v))};
    for (int c = 0; c < 12; ++c) {
      const double a = g(i, 12 + c), b = near[c];
      CHECK(std::fabs(a - b) <= 4 * std::numeric_limits<double>::epsilon() * std::fabs(b) + 1e-300);
    }
  }
  CHECK_EQ(Kokkos::numbers::pi, M_PI);
  CHECK_EQ(Kokkos::numbers::e, M_E);

  // two partitions: own streams, independent work, each fenced on its own
  auto parts = Kokkos::Experimental::partition_space(Exec(), 1, 1);
  CHECK(parts[0].stream() != parts[1].stream());
  CHECK(parts[0].stream() != Exec().stream());
  Kokkos::View<int*, Exec> a("a", 1 << 20), b("b", 1 << 20);
  Kokkos::deep_copy(parts[0], a, 3);
  Kokkos::deep_copy(parts[1], b, 5);
  parts[0].fence();
  parts[1].fence();
  int64_t sa = 0, sb = 0;
  Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec>(parts[0], 0, 1 << 20), PartSum{a}, sa);
  Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec>(parts[1], 0, 1 << 20), PartSum{b}, sb);
  CHECK_EQ(sa, int64_t(3) << 20);
  CHECK_EQ(sb, int64_t(5) << 20);
  auto many = Kokkos::Experimental::partition_space(Exec(), std::vector<int>{1, 2, 3});
  CHECK_EQ(int(many.size()), 3);
}

// =========================================================================
// Team- and thread-level sorts (algorithms/unit_tests/TestNestedSort.hpp):
// every row sorted by its team (TeamVectorRange) or by one thread's vector
// lanes (ThreadVectorRange), keys with ties and a non-power-of-two length;
// keys and values checked against a host run of the same network, so the
// order of equal keys is checked too.
// =========================================================================
struct TeamSortRows {
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> keys;
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> vals;
  KOKKOS_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type& m) const {
    const int r = m.league_rank();
    Kokkos::Experimental::sort_by_key_team(m, Kokkos::subview(keys, r, Kokkos::ALL),
                                           Kokkos::subview(vals, r, Kokkos::ALL));
  }
};
struct ThreadSortRows {
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> keys;
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> vals;
  KOKKOS_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type& m) const {
    const int r = m.league_rank() * m.team_size() + m.team_rank();
    if (r < int(keys.extent(0)))
      Kokkos::Experimental::sort_by_key_thread(m, Kokkos::subview(keys, r, Kokkos::ALL),
                                               Kokkos::subview(vals, r, Kokkos::ALL));
  }
};
struct TeamSortDescending {
  Kokkos::View<double**, Kokkos::LayoutRight, Exec> keys;
  KOKKOS_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type& m) const {
    Kokkos::Experimental::sort_team(m, Kokkos::subview(keys, m.league_rank(), Kokkos::ALL),
                                    [](const double a, const double b) { return a > b; });
  }
};

// The network on the host, step by step (the device spreads each step's
// pairs over lanes; the pairs of a step touch disjoint elements).
void host_bitonic(std::vector<int>& k, std::vector<int>& v) {
  const size_t n  = k.size();
  unsigned levels = 0;
  while ((size_t(1) << levels) < n) ++levels;
  for (unsigned i = 0; i < levels; ++i)
    for (unsigned j = 0; j <= i; ++j) {
      const unsigned sh = i - j;
      const size_t half = size_t(1) << sh;
      for (size_t p = 0; p < (size_t(1) << levels) / 2; ++p) {
        const size_t lo = ((p >> sh) << (sh + 1)) | (p & (half - 1));
        const size_t hi = j == 0 ? (lo ^ (2 * half - 1)) : (lo | half);
        if (hi < n && k[hi] < k[lo]) {
          std::swap(k[lo], k[hi]);
          std::swap(v[lo], v[hi]);
        }
      }
    }
}

KKMX_TEST(nested_sort_team_and_thread) {
  const int rows = 37, n = 100;
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> keys("keys", rows, n), vals("vals", rows, n);
  auto hk = Kokkos::create_mirror_view(keys);
  auto hv = Kokkos::create_mirror_view(vals);
  auto fill = [&]() {
    for (int r = 0; r < rows; ++r)
      for (int i = 0; i < n; ++i) {
        hk(r, i) = int((int64_t(r + 3) * 7919 * (i + 1)) % 23);  // many ties
        hv(r, i) = i;
      }
    Kokkos::deep_copy(keys, hk);
    Kokkos::deep_copy(vals, hv);
  };
  auto check = [&]() {
    auto gk = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), keys);
    auto gv = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), vals);
    for (int r = 0; r < rows; ++r) {
      std::vector<int> k(n), v(n);
      for (int i = 0; i < n; ++i) {
        k[i] = hk(r, i);
        v[i] = hv(r, i);
      }
      host_bitonic(k, v);
      for (int i = 0; i < n; ++i) {
        CHECK_EQ(gk(r, i), k[i]);
        CHECK_EQ(gv(r, i), v[i]);
        if (i) CHECK(gk(r, i - 1) <= gk(r, i));
      }
    }
  };
  fill();
  Kokkos::parallel_for(Kokkos::TeamPolicy<Exec>(rows, 64), TeamSortRows{keys, vals});
  check();
  fill();
  Kokkos::parallel_for(Kokkos::TeamPolicy<Exec>((rows + 3) / 4, 4, 16), ThreadSortRows{keys, vals});
  check();
  // a user comparator: descending doubles
  Kokkos::View<double**, Kokkos::LayoutRight, Exec> d("d", rows, n);
  auto hd = Kokkos::create_mirror_view(d);
  for (int r = 0; r < rows; ++r)
    for (int i = 0; i < n; ++i) hd(r, i) = double((r * 31 + i * 17) % 97) / 8.0;
  Kokkos::deep_copy(d, hd);
  Kokkos::parallel_for(Kokkos::TeamPolicy<Exec>(rows, 32), TeamSortDescending{d});
  auto gd = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), d);
  for (int r = 0; r < rows; ++r) {
    std::vector<double> want(n);
    for (int i = 0; i < n; ++i) want[i] = hd(r, i);
    std::sort(want.begin(), want.end(), [](double a, double b) { return a > b; });
    for (int i = 0; i < n; ++i) CHECK_EQ(gd(r, i), want[i]);
  }
}

// =========================================================================
// BinSort over a sub-range of the keys (algorithms/unit_tests/
// TestBinSortB.hpp:75-185): distinct unsigned keys binned one per bin,
// values in a LayoutStride View; and double keys with many per bin sorted
// within bins, whose permutation must be the stable order of the sub-range.
// =========================================================================
KKMX_TEST(binsort_subrange_and_strided_values) {
  const int n = 6000, b = 1000, e = 4500;
  std::vector<unsigned> hk(n);
  for (int i = 0; i < n; ++i) hk[i] = unsigned((int64_t(i) * 2654435761u) % 1000003u);  // distinct
  Kokkos::View<unsigned*, Exec> keys("keys", n);
  auto hkv = Kokkos::create_mirror_view(keys);
  for (int i = 0; i < n; ++i) hkv(i) = hk[i];
  Kokkos::deep_copy(keys, hkv);
  unsigned lo = hk[b], hi = hk[b];
  for (int i = b; i < e; ++i) {
    lo = hk[i] < lo ? hk[i] : lo;
    hi = hk[i] > hi ? hk[i] : hi;
  }
  using KV = Kokkos::View<unsigned*, Exec>;
  Kokkos::BinOp1D<KV> binner(e - b + 5, lo, hi);
  Kokkos::BinSort<KV, Kokkos::BinOp1D<KV>> sorter(keys, b, e, binner, true);
  sorter.create_permute_vector(Exec());
  // values: column 0 of an n x 2 row-major table, LayoutStride with stride 2
  Kokkos::View<double*, Exec> store("store", 2 * n);
  Kokkos::View<double*, Kokkos::LayoutStride, Exec> vals(store.data(), Kokkos::LayoutStride(n, 2));
  auto hs = Kokkos::create_mirror_view(store);
  for (int i = 0; i < 2 * n; ++i) hs(i) = 0.5 * i;
  Kokkos::deep_copy(store, hs);
  sorter.sort(Exec(), keys, b, e);
  sorter.sort(Exec(), vals, b, e);
  Exec().fence();
  auto gk = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), keys);
  auto gs = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), store);
  std::vector<unsigned> want(hk.begin() + b, hk.begin() + e);
  std::sort(want.begin(), want.end());
  std::vector<int> row_of(1000003, -1);
  for (int i = 0; i < n; ++i) row_of[hk[i]] = i;
  for (int i = 0; i < n; ++i) {
    if (i < b || i >= e) {
      CHECK_EQ(gk(i), hk[i]);
      CHECK_EQ(gs(2 * i), 0.5 * (2 * i));
    } else {
      CHECK_EQ(gk(i), want[i - b]);
      CHECK_EQ(gs(2 * i), 0.5 * (2 * row_of[gk(i)]));  // the value moved with its key
    }
    CHECK_EQ(gs(2 * i + 1), 0.5 * (2 * i + 1));  // the other column stays put
  }

  // within-bin sort of a sub-range: 16 bins for 3500 keys on a coarse grid
  std::vector<double> hd(n);
  for (int i = 0; i < n; ++i) hd[i] = double((int64_t(i) * 7919) % 257) / 256.0;
  Kokkos::View<double*, Exec> dkeys("dkeys", n);
  auto hdv = Kokkos::create_mirror_view(dkeys);
  for (int i = 0; i < n; ++i) hdv(i) = hd[i];
  Kokkos::deep_copy(dkeys, hdv);
  using DV = Kokkos::View<double*, Exec>;
  Kokkos::BinSort<DV, Kokkos::BinOp1D<DV>> dsorter(dkeys, b, e, Kokkos::BinOp1D<DV>(16, 0.0, 1.0), true);
  dsorter.create_permute_vector(Exec());
  auto perm = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), dsorter.get_permute_vector());
  std::vector<int> order(e - b);
  for (int i = 0; i < e - b; ++i) order[i] = b + i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return hd[x] < hd[y]; });
  for (int i = 0; i < e - b; ++i) CHECK_EQ(perm(i), order[i]);
  auto cnt = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), dsorter.get_bin_count());
  auto off = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), dsorter.get_bin_offsets());
  int total = 0;
  for (int k = 0; k < 17; ++k) {
    CHECK_EQ(off(k), total);
    total += cnt(k);
  }
  CHECK_EQ(total, e - b);

  // without the within-bin sort (atomic placement): bins non-decreasing
  // over the sub-range, the same keys, the rest untouched
  Kokkos::deep_copy(keys, hkv);
  Kokkos::BinSort<KV, Kokkos::BinOp1D<KV>> usorter(keys, b, e, binner, false);
  usorter.create_permute_vector(Exec());
  usorter.sort(Exec(), keys, b, e);
  Exec().fence();
  Kokkos::deep_copy(gk, keys);
  std::vector<unsigned> got;
  for (int i = 0; i < n; ++i) {
    if (i < b || i >= e) {
      CHECK_EQ(gk(i), hk[i]);
      continue;
    }
    got.push_back(gk(i));
    if (i > b) CHECK(binner.bin(gk, i - 1) <= binner.bin(gk, i));
  }
  std::sort(got.begin(), got.end());
  CHECK(got == want);
}

// =========================================================================
// Kokkos::sort of strided keys (TestSort.hpp sorts subviews): every third
// element of a buffer, whole and over [begin, end), on the device; the
// elements in between stay put.
// =========================================================================
KKMX_TEST(sort_strided_keys) {
  const int n = 30001;
  Kokkos::View<int*, Exec> store("store", 3 * n);
  auto hs = Kokkos::create_mirror_view(store);
  for (int i = 0; i < 3 * n; ++i) hs(i) = int((int64_t(i) * 7919) % 100003) - 50000;
  Kokkos::deep_copy(store, hs);
  Kokkos::View<int*, Kokkos::LayoutStride, Exec> keys(store.data(), Kokkos::LayoutStride(n, 3));
  Kokkos::sort(Exec(), keys);
  auto g = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), store);
  std::vector<int> want;
  for (int i = 0; i < n; ++i) want.push_back(hs(3 * i));
  std::sort(want.begin(), want.end());
  for (int i = 0; i < n; ++i) {
    CHECK_EQ(g(3 * i), want[i]);
    CHECK_EQ(g(3 * i + 1), hs(3 * i + 1));
    CHECK_EQ(g(3 * i + 2), hs(3 * i + 2));
  }
  // a window [100, 20000) of the strided keys, after reversing them
  for (int i = 0; i < n; ++i) hs(3 * i) = -g(3 * i);
  Kokkos::deep_copy(store, hs);
  Kokkos::sort(Exec(), keys, 100, 20000);
  Kokkos::deep_copy(g, store);
  std::vector<int> win;
  for (int i = 100; i < 20000; ++i) win.push_back(hs(3 * i));
  std::sort(win.begin(), win.end());
  for (int i = 0; i < n; ++i) CHECK_EQ(g(3 * i), (i >= 100 && i < 20000) ? win[i - 100] : hs(3 * i));
}

// =========================================================================
// sort_by_key through the permutation path (values that are not contiguous
// 4- or 8-byte words): a LayoutStride column of a rank-2 View and 2-byte
// values (algorithms/unit_tests/TestSortByKey.hpp checks keys and values
// against a host sort of the pairs).
// =========================================================================
// A comparator that is not trivially copyable (TestSortByKey.hpp:33-60).
struct GreaterHoldingView {
  Kokkos::View<int*, Exec> dummy;
  template <class T>
  KOKKOS_INLINE_FUNCTION bool operator()(const T& a, const T& b) const {
    return a > b;
  }
};

// Empty Views on the device and on the host, and the comparator overload
// (TestSortByKey.hpp:105-200).
KKMX_TEST(sort_by_key_empty_and_comparator) {
  Kokkos::View<int*, Exec> k0("k0", 0);
  Kokkos::View<float*, Exec> v0("v0", 0);
  Kokkos::Experimental::sort_by_key(Exec(), k0, v0);
  Kokkos::View<int*, Kokkos::HostSpace> hk0("hk0", 0);
  Kokkos::View<float*, Kokkos::HostSpace> hv0("hv0", 0);
  Kokkos::Experimental::sort_by_key(Kokkos::Serial(), hk0, hv0);

  const int n = 4099;
  Kokkos::View<double*, Exec> keys("keys", n);
  Kokkos::View<int*, Exec> vals("vals", n);
  auto hk = Kokkos::create_mirror_view(keys);
  auto hv = Kokkos::create_mirror_view(vals);
  for (int i = 0; i < n; ++i) {
    hk(i) = double((i * 37) % 101);
    hv(i) = i;
  }
  Kokkos::deep_copy(keys, hk);
  Kokkos::deep_copy(vals, hv);
  Kokkos::Experimental::sort_by_key(Exec(), keys, vals, GreaterHoldingView{Kokkos::View<int*, Exec>("dummy", 1)});
  auto gk = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), keys);
  auto gv = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), vals);
  for (int i = 0; i < n; ++i) {
    CHECK_EQ(gk(i), hk(gv(i)));                     // values moved with their keys
    if (i + 1 < n) CHECK(!(gk(i + 1) > gk(i)));     // descending
    if (i + 1 < n && gk(i + 1) == gk(i)) CHECK(gv(i) < gv(i + 1));  // ties in input order
  }
}

KKMX_TEST(sort_by_key_strided_and_short_values) {
  const int n = 100003;
  std::vector<int> hk(n);
  for (int i = 0; i < n; ++i) hk[i] = int((int64_t(i) * 7919) % 1000) - 500;  // many ties
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hk[a] < hk[b]; });

  Kokkos::View<int*, Exec> keys("keys", n);
  Kokkos::View<double**, Kokkos::LayoutRight, Exec> table("table", n, 3);
  auto hkv = Kokkos::create_mirror_view(keys);
  auto ht  = Kokkos::create_mirror_view(table);
  for (int i = 0; i < n; ++i) {
    hkv(i) = hk[i];
    for (int c = 0; c < 3; ++c) ht(i, c) = 10.0 * i + c;
  }
  Kokkos::deep_copy(keys, hkv);
  Kokkos::deep_copy(table, ht);
  // column 1 of the table as a LayoutStride View (stride 3)
  Kokkos::View<double*, Kokkos::LayoutStride, Exec> column(table.data() + 1, Kokkos::LayoutStride(n, 3));
  Kokkos::Experimental::sort_by_key(Exec(), keys, column);
  Kokkos::deep_copy(hkv, keys);
  Kokkos::deep_copy(ht, table);
  for (int i = 0; i < n; ++i) {
    CHECK_EQ(hkv(i), hk[order[i]]);
    CHECK_EQ(ht(i, 1), 10.0 * order[i] + 1);
    CHECK_EQ(ht(i, 0), 10.0 * i);  // the other columns stay put
  }

  Kokkos::View<int*, Exec> keys2("keys2", n);
  Kokkos::View<int16_t*, Exec> small("small", n);
  auto hs = Kokkos::create_mirror_view(small);
  for (int i = 0; i < n; ++i) {
    hkv(i) = hk[i];
    hs(i)  = int16_t(i % 30000);
  }
  Kokkos::deep_copy(keys2, hkv);
  Kokkos::deep_copy(small, hs);
  Kokkos::Experimental::sort_by_key(Exec(), keys2, small);
  Kokkos::deep_copy(hkv, keys2);
  Kokkos::deep_copy(hs, small);
  for (int i = 0; i < n; ++i) {
    CHECK_EQ(hkv(i), hk[order[i]]);
    CHECK_EQ(int(hs(i)), order[i] % 30000);
  }
}

// =========================================================================
// Dispatch from two host threads onto one instance
// (core/unit_test/TestExecSpaceThreadSafety.hpp:56-380).  Kernels submitted
// to the same instance run one after another in submission order, so a
// kernel that counts a shared device counter from 0 to N always reads N
// back; if two ran at once, one would see too many increments.
// =========================================================================
constexpr int kTsN = 100000;  // increments per kernel
constexpr int kTsM = 10;      // kernels per host thread

struct TsCounter {
  Kokkos::View<int, Exec> view;
  KOKKOS_FUNCTION bool clash() const {
    Kokkos::atomic_store(view.data(), 0);
    for (int i = 0; i < kTsN; ++i) Kokkos::atomic_inc(view.data());
    return Kokkos::atomic_load(view.data()) != kTsN;
  }
};
struct TsRangeFor {
  TsCounter c;
  Kokkos::View<int, Exec> error;
  KOKKOS_FUNCTION void operator()(int) const {
    if (c.clash()) Kokkos::atomic_store(error.data(), 1);
  }
};
struct TsMDFor {
  TsCounter c;
  Kokkos::View<int, Exec> error;
  KOKKOS_FUNCTION void operator()(int, int) const {
    if (c.clash()) Kokkos::atomic_store(error.data(), 1);
  }
};
struct TsTeamFor {
  TsCounter c;
  Kokkos::View<int, Exec> error;
  KOKKOS_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type& m) const {
    Kokkos::single(Kokkos::PerTeam(m), [&]() {
      if (c.clash()) Kokkos::atomic_store(error.data(), 1);
    });
  }
};
struct TsRangeReduce {
  using value_type = int;
  TsCounter c;
  KOKKOS_FUNCTION void operator()(int, int& update) const {
    if (c.clash()) ++update;
  }
};
struct TsMDReduce {
  using value_type = int;
  TsCounter c;
  KOKKOS_FUNCTION void operator()(int, int, int& update) const {
    if (c.clash()) ++update;
  }
};
struct TsTeamReduce {
  using value_type = int;
  TsCounter c;
  KOKKOS_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type& m, int& update) const {
    Kokkos::single(Kokkos::PerTeam(m), [&]() {
      if (c.clash()) ++update;
    });
  }
};

// Runs body(exec) kTsM times on each of two host threads, both on the
// default instance, then checks the error flag.
template <class Body>
void two_threads_one_instance(const Body& body, const Kokkos::View<int, Exec>& error) {
  auto work = [&]() {
    Exec exec;
    for (int j = 0; j < kTsM; ++j) body(exec);
    exec.fence();
  };
  std::thread t1(work), t2(work);
  t1.join();
  t2.join();
  CHECK_EQ(read0(Exec(), error), 0);
}

KKMX_TEST(thread_safety_range_for) {
  Kokkos::View<int, Exec> view("view"), error("error");
  two_threads_one_instance(
      [&](const Exec& ex) { Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(ex, 0, 1), TsRangeFor{{view}, error}); },
      error);
}

KKMX_TEST(thread_safety_mdrange_for) {
  Kokkos::View<int, Exec> view("view"), error("error");
  two_threads_one_instance(
      [&](const Exec& ex) {
        Kokkos::parallel_for(Kokkos::MDRangePolicy<Exec, Kokkos::Rank<2>>(ex, {0, 0}, {1, 1}), TsMDFor{{view}, error});
      },
      error);
}

KKMX_TEST(thread_safety_team_for) {
  Kokkos::View<int, Exec> view("view"), error("error");
  two_threads_one_instance(
      [&](const Exec& ex) { Kokkos::parallel_for(Kokkos::TeamPolicy<Exec>(ex, 1, 1, 1), TsTeamFor{{view}, error}); },
      error);
}

KKMX_TEST(thread_safety_range_reduce) {
  Kokkos::View<int, Exec> view("view"), error("error");
  two_threads_one_instance(
      [&](const Exec& ex) { Kokkos::parallel_reduce(Kokkos::RangePolicy<Exec>(ex, 0, 1), TsRangeReduce{{view}}, error); },
      error);
}

KKMX_TEST(thread_safety_mdrange_reduce) {
  Kokkos::View<int, Exec> view("view"), error("error");
  two_threads_one_instance(
      [&](const Exec& ex) {
        Kokkos::parallel_reduce(Kokkos::MDRangePolicy<Exec, Kokkos::Rank<2>>(ex, {0, 0}, {1, 1}), TsMDReduce{{view}},
                                error);
      },
      error);
}

KKMX_TEST(thread_safety_team_reduce) {
  Kokkos::View<int, Exec> view("view"), error("error");
  two_threads_one_instance(
      [&](const Exec& ex) { Kokkos::parallel_reduce(Kokkos::TeamPolicy<Exec>(ex, 1, 1, 1), TsTeamReduce{{view}}, error); },
      error);
}

// =========================================================================
// View layout queries, LayoutStride allocations, rank-0 copies, local
// (in-kernel) deep copies and a vector-lane scan with reducers
// (core/unit_test/TestViewAPI.hpp:25-40 required_allocation_size,
// TestTeam.hpp:1310-1322 order_dimensions, TestLocalDeepCopy.hpp,
// TestTeamVector.hpp:880-960 ThreadVectorRange scans with reducers).
// =========================================================================
KKMX_TEST(view_layouts_and_local_deep_copy) {
  using V1  = Kokkos::View<int*, Exec>;
  using V1R = Kokkos::View<int*, Kokkos::LayoutRight, Exec>;
  CHECK_EQ(V1::required_allocation_size(100), size_t(400));
  CHECK_EQ(V1R::required_allocation_size(Kokkos::LayoutRight(100)), size_t(400));
  const int order[3] = {2, 0, 1}, dims[3] = {100, 10, 3};
  const Kokkos::LayoutStride ls = Kokkos::LayoutStride::order_dimensions(3, order, dims);
  CHECK_EQ(ls.stride[2], size_t(1));
  CHECK_EQ(ls.stride[0], size_t(3));
  CHECK_EQ(ls.stride[1], size_t(300));
  using SV3 = Kokkos::View<double***, Kokkos::LayoutStride, Exec>;
  using RV3 = Kokkos::View<double***, Kokkos::LayoutRight, Exec>;
  CHECK_EQ(SV3::shmem_size(ls), RV3::shmem_size(100, 10, 3));
  const int bad[3] = {0, 0, 1};
  CHECK_EQ(Kokkos::LayoutStride::order_dimensions(3, bad, dims).stride[0], size_t(0));

  // a managed LayoutStride allocation keeps the requested strides
  Kokkos::View<int***, Kokkos::LayoutStride, Exec> sv("sv", ls);
  CHECK_EQ(sv.span(), size_t(3000));
  CHECK_EQ(sv.stride(1), size_t(300));
  Kokkos::parallel_for(
      Kokkos::MDRangePolicy<Exec, Kokkos::Rank<3>>({0, 0, 0}, {100, 10, 3}),
      KOKKOS_LAMBDA(int i, int j, int k) { sv(i, j, k) = i * 10000 + j * 10 + k; });
  std::vector<int> raw(3000);
  Kokkos::View<int*, Kokkos::HostSpace, Kokkos::MemoryTraits<Kokkos::Unmanaged>> hraw(raw.data(), 3000);
  Kokkos::View<int*, Exec, Kokkos::MemoryTraits<Kokkos::Unmanaged>> draw(sv.data(), 3000);
  Kokkos::deep_copy(hraw, draw);
  for (int i = 0; i < 100; ++i)
    for (int j = 0; j < 10; ++j)
      for (int k = 0; k < 3; ++k) CHECK_EQ(raw[3 * i + 300 * j + k], i * 10000 + j * 10 + k);

  // rank-0 copies between layouts and spaces
  Kokkos::View<int, Kokkos::LayoutLeft, Exec> s0("s0");
  Kokkos::View<int, Kokkos::LayoutRight, Exec> s1("s1");
  Kokkos::deep_copy(s0, 41);
  Kokkos::deep_copy(s1, s0);
  auto h1 = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), s1);
  CHECK_EQ(h1(), 41);

  // team and sequential local_deep_copy of rank-2 slices: contiguous
  // (LayoutRight) and strided (LayoutLeft) slices, copies and fills
  const int L = 9, A = 17, B = 5;
  Kokkos::View<double***, Kokkos::LayoutRight, Exec> srcR("srcR", L, A, B), dstR("dstR", L, A, B);
  Kokkos::View<double***, Kokkos::LayoutLeft, Exec> srcL("srcL", L, A, B), dstL("dstL", L, A, B),
      seqL("seqL", L, A, B), fillL("fillL", L, A, B);
  Kokkos::View<double**, Kokkos::LayoutRight, Exec> seq1("seq1", L, A);
  Kokkos::parallel_for(
      Kokkos::MDRangePolicy<Exec, Kokkos::Rank<3>>({0, 0, 0}, {L, A, B}), KOKKOS_LAMBDA(int t, int a, int b) {
        srcR(t, a, b) = t * 1000.0 + a * 10.0 + b;
        srcL(t, a, b) = -(t * 1000.0 + a * 10.0 + b);
      });
  using TP = Kokkos::TeamPolicy<Exec>;
  Kokkos::parallel_for(
      TP(L, 4, 16), KOKKOS_LAMBDA(const TP::member_type& m) {
        const int t = m.league_rank();
        Kokkos::Experimental::local_deep_copy(m, Kokkos::subview(dstR, t, Kokkos::ALL, Kokkos::ALL),
                                              Kokkos::subview(srcR, t, Kokkos::ALL, Kokkos::ALL));
        Kokkos::Experimental::local_deep_copy(m, Kokkos::subview(dstL, t, Kokkos::ALL, Kokkos::ALL),
                                              Kokkos::subview(srcL, t, Kokkos::ALL, Kokkos::ALL));
        Kokkos::Experimental::local_deep_copy(m, Kokkos::subview(fillL, t, Kokkos::ALL, Kokkos::ALL), 0.5 * t);
      });
  Kokkos::parallel_for(
      Kokkos::RangePolicy<Exec>(0, L), KOKKOS_LAMBDA(int t) {
        Kokkos::Experimental::local_deep_copy(Kokkos::subview(seqL, t, Kokkos::ALL, Kokkos::ALL),
                                              Kokkos::subview(srcL, t, Kokkos::ALL, Kokkos::ALL));
        Kokkos::Experimental::local_deep_copy(Kokkos::subview(seq1, t, Kokkos::ALL), double(t) + 0.25);
      });
  auto hR = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), dstR);
  auto hL = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), dstL);
  auto hS = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), seqL);
  auto hF = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), fillL);
  auto h1d = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), seq1);
  for (int t = 0; t < L; ++t)
    for (int a = 0; a < A; ++a) {
      CHECK_EQ(h1d(t, a), double(t) + 0.25);
      for (int b = 0; b < B; ++b) {
        const double v = t * 1000.0 + a * 10.0 + b;
        CHECK_EQ(hR(t, a, b), v);
        CHECK_EQ(hL(t, a, b), -v);
        CHECK_EQ(hS(t, a, b), -v);
        CHECK_EQ(hF(t, a, b), 0.5 * t);
      }
    }

  // ThreadVectorRange scans with Max and Prod reducers (exact: integers and
  // powers of two), inclusive outputs and the reducer's total per thread
  const int teams = 6, ts = 4, n = 37;
  Kokkos::View<int*, Exec> in_i("in_i", teams * ts * n), out_i("out_i", teams * ts * n), tot_i("tot_i", teams * ts);
  Kokkos::View<double*, Exec> in_d("in_d", teams * ts * n), out_d("out_d", teams * ts * n),
      tot_d("tot_d", teams * ts);
  Kokkos::parallel_for(
      Kokkos::RangePolicy<Exec>(0, teams * ts * n), KOKKOS_LAMBDA(int i) {
        in_i(i) = int((i * 7919) % 1009) - 500;
        in_d(i) = (i % 3 == 0) ? 2.0 : (i % 3 == 1 ? 0.5 : (i % 7 == 0 ? 2.0 : 1.0));
      });
  Kokkos::parallel_for(
      TP(teams, ts, 16), KOKKOS_LAMBDA(const TP::member_type& m) {
        Kokkos::parallel_for(Kokkos::TeamThreadRange(m, ts), [&](int r) {
          const int base = (m.league_rank() * ts + r) * n;
          int mx         = 0;
          Kokkos::parallel_scan(
              Kokkos::ThreadVectorRange(m, n),
              [&](int j, int& upd, bool final) {
                if (in_i(base + j) > upd) upd = in_i(base + j);
                if (final) out_i(base + j) = upd;
              },
              Kokkos::Max<int, Exec>(mx));
          double pr = 0;
          Kokkos::parallel_scan(
              Kokkos::ThreadVectorRange(m, n),
              [&](int j, double& upd, bool final) {
                upd *= in_d(base + j);
                if (final) out_d(base + j) = upd;
              },
              Kokkos::Prod<double, Exec>(pr));
          Kokkos::single(Kokkos::PerThread(m), [&]() {
            tot_i(base / n) = mx;
            tot_d(base / n) = pr;
          });
        });
      });
  auto hin_i = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), in_i);
  auto hin_d = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), in_d);
  auto hout_i = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out_i);
  auto hout_d = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out_d);
  auto htot_i = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), tot_i);
  auto htot_d = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), tot_d);
  for (int r = 0; r < teams * ts; ++r) {
    int mx    = std::numeric_limits<int>::lowest();
    double pr = 1.0;
    for (int j = 0; j < n; ++j) {
      mx = std::max(mx, hin_i(r * n + j));
      pr *= hin_d(r * n + j);
      CHECK_EQ(hout_i(r * n + j), mx);
      CHECK_EQ(hout_d(r * n + j), pr);
    }
    CHECK_EQ(htot_i(r), mx);
    CHECK_EQ(htot_d(r), pr);
  }

  // MinMaxFirstLastLoc: ties take the first minimum and the last maximum
  const int nv = 100003;
  Kokkos::View<int*, Exec> vals("vals", nv);
  Kokkos::parallel_for(
      Kokkos::RangePolicy<Exec>(0, nv), KOKKOS_LAMBDA(int i) { vals(i) = (i * 37) % 11; });
  using MMFL = Kokkos::MinMaxFirstLastLoc<int, int, Kokkos::HostSpace>;
  MMFL::value_type mm;
  Kokkos::parallel_reduce(
      Kokkos::RangePolicy<Exec>(0, nv),
      KOKKOS_LAMBDA(int i, MMFL::value_type& u) {
        if (vals(i) < u.min_val || (vals(i) == u.min_val && i < u.min_loc)) {
          u.min_val = vals(i);
          u.min_loc = i;
        }
        if (vals(i) > u.max_val || (vals(i) == u.max_val && i > u.max_loc)) {
          u.max_val = vals(i);
          u.max_loc = i;
        }
      },
      MMFL(mm));
  int first_min = -1, last_max = -1;
  for (int i = 0; i < nv; ++i) {
    if (first_min < 0 && (i * 37) % 11 == 0) first_min = i;
    if ((i * 37) % 11 == 10) last_max = i;
  }
  CHECK_EQ(mm.min_val, 0);
  CHECK_EQ(mm.max_val, 10);
  CHECK_EQ(mm.min_loc, first_min);
  CHECK_EQ(mm.max_loc, last_max);
}

// =========================================================================
// Combined reductions inside a team: four results (plain sums, Sum and Max
// reducers) from one TeamThreadRange, ThreadVectorRange and TeamVectorRange
// pass each (core/unit_test/TestTeamCombinedReducers.hpp:100-260, closed
// forms n(n+1)/2, n and n*n).
// =========================================================================
KKMX_TEST(team_combined_reducers) {
  const int league = 10, n = 1000;
  using TP = Kokkos::TeamPolicy<Exec>;
  Kokkos::View<int**, Kokkos::LayoutRight, Exec> out("out", league, 12);
  Kokkos::parallel_for(
      TP(league, 8, 8), KOKKOS_LAMBDA(const TP::member_type& team) {
        int a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        Kokkos::parallel_reduce(
            Kokkos::TeamThreadRange(team, n),
            [=](int i, int& v0, int& v1, int& v2, int& v3) {
              v0 += i + 1;
              v1 += i + 1;
              v2 = (v2 < i + 1) ? i + 1 : v2;
              v3 += n;
            },
            a0, Kokkos::Sum<int>(a1), Kokkos::Max<int>(a2), a3);
        int b0 = 0, b1 = 0, b2 = 0, b3 = 0;
        Kokkos::parallel_reduce(
            Kokkos::ThreadVectorRange(team, n),
            [=](int i, int& v0, int& v1, int& v2, int& v3) {
              v0 += i + 1;
              v1 += i + 1;
              v2 = (v2 < i + 1) ? i + 1 : v2;
              v3 += n;
            },
            b0, Kokkos::Sum<int>(b1), Kokkos::Max<int>(b2), b3);
        int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        Kokkos::parallel_reduce(
            Kokkos::TeamVectorRange(team, n),
            [=](int i, int& v0, int& v1, int& v2, int& v3) {
              v0 += i + 1;
              v1 += i + 1;
              v2 = (v2 < i + 1) ? i + 1 : v2;
              v3 += n;
            },
            c0, Kokkos::Sum<int>(c1), Kokkos::Max<int>(c2), c3);
        Kokkos::single(Kokkos::PerTeam(team), [=]() {
          const int r = team.league_rank();
          const int v[12] = {a0, a1, a2, a3, b0, b1, b2, b3, c0, c1, c2, c3};
          for (int k = 0; k < 12; ++k) out(r, k) = v[k];
        });
      });
  auto h = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), out);
  for (int r = 0; r < league; ++r)
    for (int g = 0; g < 3; ++g) {
      CHECK_EQ(h(r, 4 * g + 0), n * (n + 1) / 2);
      CHECK_EQ(h(r, 4 * g + 1), n * (n + 1) / 2);
      CHECK_EQ(h(r, 4 * g + 2), n);
      CHECK_EQ(h(r, 4 * g + 3), n * n);
    }
}

}  // namespace

// ---------------------------------------------------------------- C-ABI
// =========================================================================
// Built-in tuners (core/unit_test/tools/TestBuiltinTuners.cpp:29-118,
// TestCategoricalTuner.cpp): the sparse problem maps coordinates in [0, 1)
// to points; with no tuning tool loaded begin() returns the first point and
// tune() leaves the policy as it was.
// =========================================================================
struct TunerTeamF {
  KOKKOS_INLINE_FUNCTION void operator()(const Kokkos::TeamPolicy<Exec>::member_type&) const {}
};
struct TunerMDF {
  KOKKOS_INLINE_FUNCTION void operator()(const int, const int) const {}
};
KKMX_TEST(builtin_tuners_points_and_defaults) {
  namespace KTE = Kokkos::Tools::Experimental;
  Kokkos::TeamPolicy<Exec> teamp(1, Kokkos::AUTO, Kokkos::AUTO);
  Kokkos::MDRangePolicy<Exec, Kokkos::Rank<2>> mdp({0, 0}, {1, 1});
  KTE::TeamSizeTuner team_tuner("team_tuner", teamp, TunerTeamF{}, Kokkos::ParallelForTag{},
                                KTE::Impl::Impl::SimpleTeamSizeCalculator{});
  KTE::MDRangeTuner<2> md_tuner("md_tuner", mdp, TunerMDF{}, Kokkos::ParallelForTag{},
                                KTE::Impl::Impl::SimpleTeamSizeCalculator{});
  std::vector<int> options{1, 2, 3, 4, 5};
  auto team_opts = team_tuner.combine("options", options);
  auto md_opts   = md_tuner.combine("options", options);
  CHECK_EQ(std::get<0>(team_opts.get_point(0.0, 0.0, 0.0)), 1);
  CHECK_EQ(std::get<0>(team_opts.get_point(0.9, 0.0, 0.0)), 5);
  CHECK_EQ(std::get<0>(team_opts.get_point(0.45, 0.0, 0.0)), 3);
  // Vector lengths are listed in increasing key order: 0.0 picks the
  // smallest, 0.999 the largest (64); team sizes at that length count down
  // from the maximum.
  const auto low  = team_tuner.get_point(0.0, 0.0);
  const auto high = team_tuner.get_point(0.999, 0.0);
  Kokkos::TeamPolicy<Exec> p1(teamp), p64(teamp);
  p1.impl_set_vector_length(1);
  p64.impl_set_vector_length(64);
  CHECK_EQ(std::get<0>(low), int64_t(1));
  CHECK_EQ(std::get<1>(low), int64_t(p1.team_size_max(TunerTeamF{}, Kokkos::ParallelForTag{})));
  CHECK_EQ(std::get<0>(high), int64_t(64));
  CHECK_EQ(std::get<1>(high), int64_t(p64.team_size_max(TunerTeamF{}, Kokkos::ParallelForTag{})));
  // MD tiles: powers of two whose product stays below max_tile_size_product.
  const auto t0 = md_tuner.get_point(0.0, 0.0);
  CHECK_EQ(std::get<0>(t0), 1);
  CHECK_EQ(std::get<1>(t0), 1);
  const auto t1 = md_tuner.get_point(0.999, 0.999);
  CHECK(int64_t(std::get<0>(t1)) * std::get<1>(t1) < 1024);
  // No tool: begin() answers the first point, tune() keeps the policy.
  auto cfg = md_opts.begin();
  CHECK_EQ(std::get<0>(cfg), 1);
  md_opts.end();
  auto tuned = md_tuner.tune(mdp);
  md_tuner.end();
  Kokkos::parallel_for(tuned, TunerMDF{});
  auto cat = KTE::make_categorical_tuner("cat", std::vector<std::string>{"a", "b", "c"});
  CHECK(cat.begin() == "a");
  cat.end();
  KTE::SingleDimensionalRangeTuner<int64_t> r("rng", KTE::StatisticalCategory::kokkos_value_ratio, int64_t(42),
                                              int64_t(5), int64_t(100), int64_t(5));
  CHECK_EQ(r.begin(), int64_t(42));
  r.end();
}

// impl_hardware_thread_id: every lane's slot is below
// impl_max_hardware_threads, the lane field is the lane, and a launch that
// fills the GPU touches all 8 XCDs and many wave slots.
KKMX_TEST(hardware_thread_ids) {
  const int64_t n = int64_t(1) << 22;
  Kokkos::View<unsigned*, Exec> ids("ids", n);
  Kokkos::View<unsigned*, Exec> lanes("lanes", n);
  Kokkos::parallel_for(Kokkos::RangePolicy<Exec>(0, n), KOKKOS_LAMBDA(const int64_t i) {
    ids(i)   = Exec::impl_hardware_thread_id();
    lanes(i) = static_cast<unsigned>(__lane_id());
  });
  auto h  = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), ids);
  auto hl = Kokkos::create_mirror_view_and_copy(Kokkos::HostSpace(), lanes);
  const unsigned mx = Exec::impl_max_hardware_threads();
  std::vector<char> xcd(16, 0);
  std::vector<char> seen(mx / 64, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (h(i) >= mx) CHECK(h(i) < mx);
    if (h(i) % 64 != hl(i)) CHECK_EQ(h(i) % 64, hl(i));
    seen[h(i) / 64] = 1;
    xcd[(h(i) / 64) >> 14] = 1;  // slot = xcc * 2^14 + (se, sh, cu, simd, wave)
  }
  int nx = 0, nslots = 0;
  for (char c : xcd) nx += c;
  for (char c : seen) nslots += c;
  CHECK_EQ(nx, 8);
  CHECK(nslots >= 1024);
}

// =========================================================================
// Backend registration (core/src/impl/Kokkos_ExecSpaceManager.hpp:157-163):
// a host execution space registered from a namespace-scope initialiser is
// initialised by Kokkos::initialize (when the registration came first),
// fenced by Kokkos::fence and printed by Kokkos::print_configuration.
// =========================================================================
struct ProbeCounters {
  static inline std::atomic<int> inits{0}, finals{0}, fences{0};
};
class SelftestProbeSpace {
 public:
  using execution_space      = SelftestProbeSpace;
  using memory_space         = Kokkos::HostSpace;
  using device_type          = Kokkos::Device<execution_space, memory_space>;
  using array_layout         = Kokkos::LayoutRight;
  using size_type            = size_t;
  using scratch_memory_space = Kokkos::ScratchMemorySpace<Kokkos::Serial>;
  static constexpr const char* name() { return "SelftestProbe"; }
  int concurrency() const { return 1; }
  void fence(const std::string& = "") const {}
  static void impl_static_fence(const std::string&) { ++ProbeCounters::fences; }
  void print_configuration(std::ostream& os, bool = false) const { os << "SelftestProbe execution space\n"; }
  static void impl_initialize(const Kokkos::InitializationSettings&) { ++ProbeCounters::inits; }
  static void impl_finalize() { ++ProbeCounters::finals; }
  bool operator==(const SelftestProbeSpace&) const { return true; }
  bool operator!=(const SelftestProbeSpace&) const { return false; }
};
const bool g_probe_registered_before_init = !Kokkos::is_initialized();
const int g_probe_registered = Kokkos::Impl::initialize_space_factory<SelftestProbeSpace>("200_SelftestProbe");

KKMX_TEST(exec_space_registration) {
  CHECK_EQ(g_probe_registered, 1);
  auto& m = Kokkos::Impl::ExecSpaceManager::get_instance();
  CHECK(m.is_registered("100_Serial"));
  CHECK(m.is_registered("150_MI355X"));
  CHECK(m.is_registered("200_SelftestProbe"));
  CHECK_EQ(ProbeCounters::inits.load(), g_probe_registered_before_init ? 1 : 0);
  CHECK_EQ(ProbeCounters::finals.load(), 0);
  const int f0 = ProbeCounters::fences.load();
  Kokkos::fence("selftest: global fence reaches every registered space");
  CHECK_EQ(ProbeCounters::fences.load(), f0 + 1);
  std::ostringstream os;
  Kokkos::print_configuration(os, false);
  const std::string txt = os.str();
  const size_t ps = txt.find("Host Serial Execution Space"), pm = txt.find("KOKKOS_ENABLE_MI355X"),
               pp = txt.find("SelftestProbe execution space");
  CHECK(ps != std::string::npos && pm != std::string::npos && pp != std::string::npos);
  CHECK(ps < pm && pm < pp);  // in the order of the keys
}

extern "C" {

int kkmx_selftest_count(void) { return static_cast<int>(registry().size()); }

const char* kkmx_selftest_name(int i) {
  if (i < 0 || i >= static_cast<int>(registry().size())) return nullptr;
  return registry()[i].name;
}

// Runs test i.  Returns 0 on pass, 1 on a failed check, 2 on an exception
// from the runtime, -1 for a bad index; msg receives the failure text.
int kkmx_selftest_run(int i, char* msg, int msg_len) {
  if (i < 0 || i >= static_cast<int>(registry().size())) return -1;
  auto put = [&](const std::string& s) {
    if (msg && msg_len > 0) {
      std::strncpy(msg, s.c_str(), msg_len - 1);
      msg[msg_len - 1] = 0;
    }
  };
  put("");
  try {
    if (!Kokkos::is_initialized()) Kokkos::initialize();
    registry()[i].fn();
    Kokkos::fence();
    return 0;
  } catch (const Failure& f) {
    put(f.what());
    return 1;
  } catch (const std::exception& e) {
    put(std::string("exception: ") + e.what());
    return 2;
  }
}

}  // extern "C"
