// Kokkos::Serial dispatch closures: the sequential host execution the
// reference's Serial backend defines (core/src/Serial/Kokkos_Serial_Parallel_Range.hpp:25-320,
// Kokkos_Serial_Parallel_MDRange.hpp:27-166, Kokkos_Serial_Parallel_Team.hpp:212-382):
// reduce = init, left fold over the range, final; scan = f(i, update, true)
// once per index in order.  Used for host mirrors and host-side loops.
#ifndef KKMX_HOST_EXEC_HPP
#define KKMX_HOST_EXEC_HPP

#include <kkmx/FunctorAnalysis.hpp>
#include <kkmx/Policy.hpp>

namespace Kokkos {
namespace Impl {

template <class F, class... Traits>
class ParallelFor<F, RangePolicy<Traits...>, Serial> {
 public:
  using Policy = RangePolicy<Traits...>;
  ParallelFor(const F& f, const Policy& p) : m_f(f), m_p(p) {}
  void execute() const {
    using Tag = typename Policy::work_tag;
    for (auto i = m_p.begin(); i < m_p.end(); ++i) call_functor<Tag>(m_f, i);
  }

 private:
  F m_f;
  Policy m_p;
};

template <class CR, class... Traits>
class ParallelReduce<CR, RangePolicy<Traits...>, Serial> {
 public:
  using Policy = RangePolicy<Traits...>;
  using V      = typename CR::value_type;
  template <class Mem>
  ParallelReduce(const CR& cr, const Policy& p, V* result, Mem) : m_cr(cr), m_p(p), m_result(result) {}
  void execute() const {
    using Tag = typename CR::work_tag;
    if constexpr (CR::is_array) {
      const unsigned n = m_cr.value_count();
      std::vector<V> tmp(n);
      V* v = m_result ? m_result : tmp.data();
      m_cr.init_array(v);
      for (auto i = m_p.begin(); i < m_p.end(); ++i) call_functor<Tag>(m_cr.functor(), i, v);
      m_cr.final_array(v);
    } else {
      V v;
      m_cr.init(v);
      for (auto i = m_p.begin(); i < m_p.end(); ++i) call_functor<Tag>(m_cr.functor(), i, v);
      m_cr.final(v);
      if (m_result) *m_result = v;
    }
  }

 private:
  CR m_cr;
  Policy m_p;
  V* m_result;
};

template <class CR, class... Traits>
class ParallelScan<CR, RangePolicy<Traits...>, Serial> {
 public:
  using Policy = RangePolicy<Traits...>;
  ParallelScan(const CR& cr, const Policy& p) : m_cr(cr), m_p(p) {}
  void execute() const {
    using Tag = typename CR::work_tag;
    typename CR::value_type u;
    m_cr.init(u);
    for (auto i = m_p.begin(); i < m_p.end(); ++i) call_functor<Tag>(m_cr.functor(), i, u, true);
  }

 private:
  CR m_cr;
  Policy m_p;
};

template <class CR, class Ret, class... Traits>
class ParallelScanWithTotal<CR, RangePolicy<Traits...>, Ret, Serial> {
 public:
  using Policy = RangePolicy<Traits...>;
  using V      = typename CR::value_type;
  ParallelScanWithTotal(const CR& cr, const Policy& p, V* total, bool) : m_cr(cr), m_p(p), m_total(total) {}
  void execute() const {
    using Tag = typename CR::work_tag;
    V u;
    m_cr.init(u);
    for (auto i = m_p.begin(); i < m_p.end(); ++i) call_functor<Tag>(m_cr.functor(), i, u, true);
    if (m_total) *m_total = u;
  }

 private:
  CR m_cr;
  Policy m_p;
  V* m_total;
};

// MDRange on the host: nested loops, the first (Iterate::Left) or last
// (Iterate::Right) index fastest.
template <unsigned R, bool Left, class Fn>
void md_host_loop(const int64_t* lo, const int64_t* hi, const Fn& fn) {
  int64_t idx[8];
  for (unsigned r = 0; r < R; ++r) {
    if (hi[r] <= lo[r]) return;
    idx[r] = lo[r];
  }
  while (true) {
    fn(idx);
    // increment, fastest dimension first
    int r = Left ? 0 : int(R) - 1;
    while (true) {
      if (++idx[r] < hi[r]) break;
      idx[r] = lo[r];
      r += Left ? 1 : -1;
      if (r < 0 || r >= int(R)) return;
    }
  }
}

template <class Tag, class F, unsigned R, class... Extra>
KOKKOS_FORCEINLINE_FUNCTION void md_call(const F& f, const int64_t* i, Extra&&... extra) {
  if constexpr (R == 2) call_functor<Tag>(f, i[0], i[1], extra...);
  else if constexpr (R == 3) call_functor<Tag>(f, i[0], i[1], i[2], extra...);
  else if constexpr (R == 4) call_functor<Tag>(f, i[0], i[1], i[2], i[3], extra...);
  else if constexpr (R == 5) call_functor<Tag>(f, i[0], i[1], i[2], i[3], i[4], extra...);
  else call_functor<Tag>(f, i[0], i[1], i[2], i[3], i[4], i[5], extra...);
}

template <class F, class... Traits>
class ParallelFor<F, MDRangePolicy<Traits...>, Serial> {
 public:
  using Policy = MDRangePolicy<Traits...>;
  ParallelFor(const F& f, const Policy& p) : m_f(f), m_p(p) {}
  void execute() const {
    constexpr unsigned R = Policy::rank;
    int64_t lo[R], hi[R];
    for (unsigned r = 0; r < R; ++r) { lo[r] = m_p.m_lower[r]; hi[r] = m_p.m_upper[r]; }
    md_host_loop<R, Policy::inner_direction == Iterate::Left>(
        lo, hi, [&](const int64_t* i) { md_call<typename Policy::work_tag, F, R>(m_f, i); });
  }

 private:
  F m_f;
  Policy m_p;
};

template <class CR, class... Traits>
class ParallelReduce<CR, MDRangePolicy<Traits...>, Serial> {
 public:
  using Policy = MDRangePolicy<Traits...>;
  using V      = typename CR::value_type;
  template <class Mem>
  ParallelReduce(const CR& cr, const Policy& p, V* result, Mem) : m_cr(cr), m_p(p), m_result(result) {}
  void execute() const {
    constexpr unsigned R = Policy::rank;
    int64_t lo[R], hi[R];
    for (unsigned r = 0; r < R; ++r) { lo[r] = m_p.m_lower[r]; hi[r] = m_p.m_upper[r]; }
    if constexpr (CR::is_array) {
      std::vector<V> tmp(m_cr.value_count());
      V* v = m_result ? m_result : tmp.data();
      m_cr.init_array(v);
      md_host_loop<R, Policy::inner_direction == Iterate::Left>(lo, hi, [&](const int64_t* i) {
        md_call<typename Policy::work_tag, typename CR::functor_type, R>(m_cr.functor(), i, v);
      });
      m_cr.final_array(v);
    } else {
      V v;
      m_cr.init(v);
      md_host_loop<R, Policy::inner_direction == Iterate::Left>(lo, hi, [&](const int64_t* i) {
        md_call<typename Policy::work_tag, typename CR::functor_type, R>(m_cr.functor(), i, v);
      });
      m_cr.final(v);
      if (m_result) *m_result = v;
    }
  }

 private:
  CR m_cr;
  Policy m_p;
  V* m_result;
};

}  // namespace Impl
}  // namespace Kokkos

#endif
