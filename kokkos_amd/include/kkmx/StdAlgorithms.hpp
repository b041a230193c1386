// Kokkos::Experimental std-algorithm scans and reductions over rank-1 Views
// and their random-access iterators, built on this space's parallel_scan /
// parallel_reduce kernels:
//   inclusive_scan, exclusive_scan, transform_inclusive_scan,
//   transform_exclusive_scan, reduce, transform_reduce,
//   begin / end / cbegin / cend / distance.
//
// Reference: algorithms/src/std_algorithms/Kokkos_InclusiveScan.hpp,
// Kokkos_ExclusiveScan.hpp, Kokkos_TransformInclusiveScan.hpp,
// Kokkos_TransformExclusiveScan.hpp, Kokkos_Reduce.hpp,
// Kokkos_TransformReduce.hpp, Kokkos_BeginEnd.hpp, Kokkos_Distance.hpp and
// impl/Kokkos_RandomAccessIterator.hpp (overload sets: iterator or View
// arguments, optional label first, optional custom operators and initial
// values; scans return the iterator past the last element written).
//
// A user binary operator has no known neutral element, so scans and
// reductions with one carry {value, empty} pairs whose join skips the
// empty side (reference impl/Kokkos_ValueWrapperForNoNeutralElement.hpp).
// The default operator is +, with 0 as the neutral element.  Every call
// reads in[i] before writing out[i], so in-place scans (first == dest) work.
#ifndef KKMX_STD_ALGORITHMS_HPP
#define KKMX_STD_ALGORITHMS_HPP

#include <Kokkos_Core.hpp>

#include <iterator>

namespace Kokkos {
namespace Experimental {

template <class ViewType>
class RandomAccessIterator;

template <class D, class... P>
class RandomAccessIterator<View<D, P...>> {
 public:
  using view_type         = View<D, P...>;
  using iterator_category = std::random_access_iterator_tag;
  using value_type        = typename view_type::non_const_value_type;
  using difference_type   = ptrdiff_t;
  using pointer           = typename view_type::pointer_type;
  using reference         = typename view_type::value_type&;
  static_assert(view_type::rank == 1, "std algorithms take rank-1 Views");

  RandomAccessIterator() = default;
  KOKKOS_FUNCTION explicit RandomAccessIterator(const view_type& v, ptrdiff_t i = 0)
      : m_ptr(v.data() + i * ptrdiff_t(v.stride_0())), m_stride(ptrdiff_t(v.stride_0())) {}
  // non-const -> const iterator
  template <class OD, class... OP,
            class = std::enable_if_t<std::is_convertible<View<OD, OP...>, view_type>::value>>
  KOKKOS_FUNCTION RandomAccessIterator(const RandomAccessIterator<View<OD, OP...>>& o)
      : m_ptr(o.data()), m_stride(o.stride()) {}

  KOKKOS_FUNCTION reference operator*() const { return *m_ptr; }
  KOKKOS_FUNCTION reference operator[](difference_type n) const { return m_ptr[n * m_stride]; }
  KOKKOS_FUNCTION RandomAccessIterator& operator++() {
    m_ptr += m_stride;
    return *this;
  }
  KOKKOS_FUNCTION RandomAccessIterator operator++(int) {
    auto t = *this;
    m_ptr += m_stride;
    return t;
  }
  KOKKOS_FUNCTION RandomAccessIterator& operator--() {
    m_ptr -= m_stride;
    return *this;
  }
  KOKKOS_FUNCTION RandomAccessIterator operator--(int) {
    auto t = *this;
    m_ptr -= m_stride;
    return t;
  }
  KOKKOS_FUNCTION RandomAccessIterator& operator+=(difference_type n) {
    m_ptr += n * m_stride;
    return *this;
  }
  KOKKOS_FUNCTION RandomAccessIterator& operator-=(difference_type n) {
    m_ptr -= n * m_stride;
    return *this;
  }
  KOKKOS_FUNCTION RandomAccessIterator operator+(difference_type n) const {
    auto t = *this;
    return t += n;
  }
  KOKKOS_FUNCTION RandomAccessIterator operator-(difference_type n) const {
    auto t = *this;
    return t -= n;
  }
  KOKKOS_FUNCTION difference_type operator-(const RandomAccessIterator& o) const {
    return m_stride ? (m_ptr - o.m_ptr) / m_stride : 0;
  }
  KOKKOS_FUNCTION bool operator==(const RandomAccessIterator& o) const { return m_ptr == o.m_ptr; }
  KOKKOS_FUNCTION bool operator!=(const RandomAccessIterator& o) const { return m_ptr != o.m_ptr; }
  KOKKOS_FUNCTION bool operator<(const RandomAccessIterator& o) const { return m_ptr < o.m_ptr; }
  KOKKOS_FUNCTION bool operator<=(const RandomAccessIterator& o) const { return m_ptr <= o.m_ptr; }
  KOKKOS_FUNCTION bool operator>(const RandomAccessIterator& o) const { return m_ptr > o.m_ptr; }
  KOKKOS_FUNCTION bool operator>=(const RandomAccessIterator& o) const { return m_ptr >= o.m_ptr; }
  KOKKOS_FUNCTION pointer data() const { return m_ptr; }
  KOKKOS_FUNCTION difference_type stride() const { return m_stride; }

 private:
  pointer m_ptr              = nullptr;
  difference_type m_stride   = 1;
};

template <class D, class... P>
KOKKOS_FUNCTION auto begin(const View<D, P...>& v) {
  return RandomAccessIterator<View<D, P...>>(v, 0);
}
template <class D, class... P>
KOKKOS_FUNCTION auto end(const View<D, P...>& v) {
  return RandomAccessIterator<View<D, P...>>(v, ptrdiff_t(v.extent(0)));
}
template <class D, class... P>
KOKKOS_FUNCTION auto cbegin(const View<D, P...>& v) {
  using CV = typename View<D, P...>::const_type;
  return RandomAccessIterator<CV>(CV(v), 0);
}
template <class D, class... P>
KOKKOS_FUNCTION auto cend(const View<D, P...>& v) {
  using CV = typename View<D, P...>::const_type;
  return RandomAccessIterator<CV>(CV(v), ptrdiff_t(v.extent(0)));
}
template <class It>
KOKKOS_FUNCTION ptrdiff_t distance(const It& first, const It& last) {
  return last - first;
}

namespace Impl {

template <class T>
struct is_std_iterator : std::false_type {};
template <class V>
struct is_std_iterator<RandomAccessIterator<V>> : std::true_type {};
template <class... I>
constexpr bool are_iterators_v = (is_std_iterator<I>::value && ...);

// {value, empty} for operators without a neutral element.
template <class T>
struct NoNeutral {
  T val;
  bool empty;
};

struct PlusOp {
  template <class A, class B>
  KOKKOS_FUNCTION auto operator()(const A& a, const B& b) const {
    return a + b;
  }
};
struct IdentityOp {
  template <class A>
  KOKKOS_FUNCTION A operator()(const A& a) const {
    return a;
  }
};
struct MultiplyOp {
  template <class A, class B>
  KOKKOS_FUNCTION auto operator()(const A& a, const B& b) const {
    return a * b;
  }
};

// Scan functor: out[i] = (init op) u(in[0]) op ... op u(in[i]) (inclusive)
// or (init op) u(in[0]) op ... op u(in[i-1]) (exclusive).
template <class T, class InIt, class OutIt, class Op, class UOp, bool Inclusive, bool HasInit>
struct ScanFunctor {
  using value_type = NoNeutral<T>;
  InIt in;
  OutIt out;
  Op op;
  UOp u;
  T seed;  // the initial value, leftmost operand of every output
  KOKKOS_FUNCTION void init(value_type& v) const { v.empty = true; }
  KOKKOS_FUNCTION void join(value_type& a, const value_type& b) const {
    if (b.empty) return;
    if (a.empty) {
      a = b;
    } else {
      a.val = op(a.val, b.val);
    }
  }
  KOKKOS_FUNCTION void operator()(const int64_t i, value_type& upd, const bool final) const {
    const T x = static_cast<T>(u(in[i]));  // read before the write: in-place safe
    if (Inclusive) {
      if (upd.empty) {
        upd.val   = x;
        upd.empty = false;
      } else {
        upd.val = op(upd.val, x);
      }
      if (final) out[i] = HasInit ? op(seed, upd.val) : upd.val;
    } else {
      if (final) out[i] = upd.empty ? seed : op(seed, upd.val);
      if (upd.empty) {
        upd.val   = x;
        upd.empty = false;
      } else {
        upd.val = op(upd.val, x);
      }
    }
  }
};
// Scans with + and no user operator: plain values, 0 is the identity.
template <class T, class InIt, class OutIt, class UOp, bool Inclusive>
struct PlusScanFunctor {
  using value_type = T;
  InIt in;
  OutIt out;
  UOp u;
  T seed;
  KOKKOS_FUNCTION void operator()(const int64_t i, T& upd, const bool final) const {
    const T x = static_cast<T>(u(in[i]));
    if (Inclusive) {
      upd += x;
      if (final) out[i] = seed + upd;
    } else {
      if (final) out[i] = seed + upd;
      upd += x;
    }
  }
};

template <class T, class It1, class It2, class Join, class Transform, bool Binary>
struct ReduceFunctor {
  using value_type = NoNeutral<T>;
  It1 a;
  It2 b;
  Join op;
  Transform t;
  KOKKOS_FUNCTION void init(value_type& v) const { v.empty = true; }
  KOKKOS_FUNCTION void join(value_type& x, const value_type& y) const {
    if (y.empty) return;
    if (x.empty) {
      x = y;
    } else {
      x.val = op(x.val, y.val);
    }
  }
  KOKKOS_FUNCTION void operator()(const int64_t i, value_type& v) const {
    T x;
    if constexpr (Binary) x = static_cast<T>(t(a[i], b[i]));
    else x = static_cast<T>(t(a[i]));
    if (v.empty) {
      v.val   = x;
      v.empty = false;
    } else {
      v.val = op(v.val, x);
    }
  }
};
template <class T, class It1, class It2, class Transform, bool Binary>
struct PlusReduceFunctor {
  using value_type = T;
  It1 a;
  It2 b;
  Transform t;
  KOKKOS_FUNCTION void operator()(const int64_t i, T& v) const {
    if constexpr (Binary) v += static_cast<T>(t(a[i], b[i]));
    else v += static_cast<T>(t(a[i]));
  }
};

template <class Ex>
using Range64 = RangePolicy<Ex, IndexType<int64_t>>;

// Input of a scan whose source range shares no memory with its destination:
// the kernel never writes it, so its loads go through the constant address
// space, like those of a const RandomAccess View, and the scan's final pass
// can reuse the first pass's values instead of reading the input again.
template <class T>
struct InvariantInput {
  const T* ptr;
  ptrdiff_t stride;
  KOKKOS_FUNCTION T operator[](ptrdiff_t i) const {
#if defined(__HIP_DEVICE_COMPILE__)
    using invariant_ptr = const __attribute__((address_space(4))) T*;
    return *(invariant_ptr)(ptr + i * stride);
#else
    return ptr[i * stride];
#endif
  }
};

template <class It>
struct is_device_view_iterator : std::false_type {};
template <class D, class... P>
struct is_device_view_iterator<RandomAccessIterator<View<D, P...>>>
    : std::bool_constant<std::is_same<typename View<D, P...>::memory_space, MI355XSpace>::value &&
                         std::is_arithmetic<typename View<D, P...>::non_const_value_type>::value> {};

// Whether [a, a + n*sa) and [b, b + n*sb) (element strides, bytes ea / eb)
// share memory.
inline bool ranges_overlap(const void* a, ptrdiff_t sa, size_t ea, const void* b, ptrdiff_t sb, size_t eb,
                           int64_t n) {
  const auto pa = reinterpret_cast<uintptr_t>(a), pb = reinterpret_cast<uintptr_t>(b);
  const uintptr_t la = uintptr_t(n - 1) * uintptr_t(sa < 0 ? -sa : sa) * ea + ea;
  const uintptr_t lb = uintptr_t(n - 1) * uintptr_t(sb < 0 ? -sb : sb) * eb + eb;
  if (sa < 0 || sb < 0) return true;
  return pa < pb + lb && pb < pa + la;
}

template <class T, bool Inclusive, bool HasInit, class Ex, class InIt, class OutIt, class Op, class UOp>
void scan_run(const std::string& label, const Ex& ex, InIt first, OutIt dest, int64_t n, Op op, UOp u, T init) {
  if constexpr (std::is_same<Op, PlusOp>::value) {
    parallel_scan(label, Range64<Ex>(ex, 0, n),
                  PlusScanFunctor<T, InIt, OutIt, UOp, Inclusive>{first, dest, u, HasInit ? init : T(0)});
  } else {
    parallel_scan(label, Range64<Ex>(ex, 0, n), ScanFunctor<T, InIt, OutIt, Op, UOp, Inclusive, HasInit>{
                                                    first, dest, op, u, init});
  }
}

template <class T, bool Inclusive, bool HasInit, class Ex, class InIt, class OutIt, class Op, class UOp>
OutIt scan_impl(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, Op op, UOp u, T init) {
  const int64_t n = static_cast<int64_t>(last - first);
  if (n <= 0) return dest;
  if constexpr (std::is_same<Ex, MI355X>::value && is_device_view_iterator<InIt>::value &&
                is_device_view_iterator<OutIt>::value) {
    using IT = typename InIt::value_type;
    using OT = typename OutIt::value_type;
    if (!ranges_overlap(first.data(), first.stride(), sizeof(IT), dest.data(), dest.stride(), sizeof(OT), n)) {
      scan_run<T, Inclusive, HasInit>(label, ex, InvariantInput<IT>{first.data(), first.stride()}, dest, n, op, u,
                                      init);
      return dest + n;
    }
  }
  scan_run<T, Inclusive, HasInit>(label, ex, first, dest, n, op, u, init);
  return dest + n;
}

template <class T, bool Binary, class Ex, class It1, class It2, class Join, class Transform>
T reduce_impl(const std::string& label, const Ex& ex, It1 first, It1 last, It2 first2, T init, Join op,
              Transform t) {
  const int64_t n = static_cast<int64_t>(last - first);
  if (n <= 0) return init;
  if constexpr (std::is_same<Join, PlusOp>::value) {
    T r{};
    parallel_reduce(label, Range64<Ex>(ex, 0, n), PlusReduceFunctor<T, It1, It2, Transform, Binary>{first, first2, t},
                    r);
    return init + r;
  } else {
    NoNeutral<T> r{};
    parallel_reduce(label, Range64<Ex>(ex, 0, n),
                    ReduceFunctor<T, It1, It2, Join, Transform, Binary>{first, first2, op, t}, r);
    return r.empty ? init : op(init, r.val);
  }
}

template <class V>
using it_value_t = typename V::value_type;

}  // namespace Impl

// ------------------------------------------------------------ inclusive_scan
template <class Ex, class InIt, class OutIt,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt inclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest) {
  using T = Impl::it_value_t<InIt>;
  return Impl::scan_impl<T, true, false>(label, ex, first, last, dest, Impl::PlusOp{}, Impl::IdentityOp{}, T{});
}
template <class Ex, class InIt, class OutIt,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt inclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest) {
  return inclusive_scan("Kokkos::inclusive_scan_default_functors_iterator_api", ex, first, last, dest);
}
template <class Ex, class InIt, class OutIt, class Op,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt inclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, Op op) {
  using T = Impl::it_value_t<InIt>;
  return Impl::scan_impl<T, true, false>(label, ex, first, last, dest, op, Impl::IdentityOp{}, T{});
}
template <class Ex, class InIt, class OutIt, class Op,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt inclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest, Op op) {
  return inclusive_scan("Kokkos::inclusive_scan_custom_functors_iterator_api", ex, first, last, dest, op);
}
template <class Ex, class InIt, class OutIt, class Op, class T,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt inclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, Op op, T init) {
  return Impl::scan_impl<T, true, true>(label, ex, first, last, dest, op, Impl::IdentityOp{}, init);
}
template <class Ex, class InIt, class OutIt, class Op, class T,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt inclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest, Op op, T init) {
  return inclusive_scan("Kokkos::inclusive_scan_custom_functors_iterator_api", ex, first, last, dest, op, init);
}
template <class Ex, class D1, class... P1, class D2, class... P2,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto inclusive_scan(const std::string& label, const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to) {
  return inclusive_scan(label, ex, cbegin(from), cend(from), begin(to));
}
template <class Ex, class D1, class... P1, class D2, class... P2,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto inclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to) {
  return inclusive_scan("Kokkos::inclusive_scan_default_functors_view_api", ex, from, to);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class Op,
          std::enable_if_t<is_execution_space<Ex>::value && !is_view<Op>::value, int> = 0>
auto inclusive_scan(const std::string& label, const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to,
                    Op op) {
  return inclusive_scan(label, ex, cbegin(from), cend(from), begin(to), op);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class Op,
          std::enable_if_t<is_execution_space<Ex>::value && !is_view<Op>::value, int> = 0>
auto inclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to, Op op) {
  return inclusive_scan("Kokkos::inclusive_scan_custom_functors_view_api", ex, from, to, op);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class Op, class T,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto inclusive_scan(const std::string& label, const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to,
                    Op op, T init) {
  return inclusive_scan(label, ex, cbegin(from), cend(from), begin(to), op, init);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class Op, class T,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto inclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to, Op op, T init) {
  return inclusive_scan("Kokkos::inclusive_scan_custom_functors_view_api", ex, from, to, op, init);
}

// ------------------------------------------------------------ exclusive_scan
template <class Ex, class InIt, class OutIt, class T,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt exclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, T init) {
  return Impl::scan_impl<T, false, true>(label, ex, first, last, dest, Impl::PlusOp{}, Impl::IdentityOp{}, init);
}
template <class Ex, class InIt, class OutIt, class T,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt exclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest, T init) {
  return exclusive_scan("Kokkos::exclusive_scan_default_functors_iterator_api", ex, first, last, dest, init);
}
template <class Ex, class InIt, class OutIt, class T, class Op,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt exclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, T init, Op op) {
  return Impl::scan_impl<T, false, true>(label, ex, first, last, dest, op, Impl::IdentityOp{}, init);
}
template <class Ex, class InIt, class OutIt, class T, class Op,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt exclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest, T init, Op op) {
  return exclusive_scan("Kokkos::exclusive_scan_custom_functors_iterator_api", ex, first, last, dest, init, op);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class T,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto exclusive_scan(const std::string& label, const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to,
                    T init) {
  return exclusive_scan(label, ex, cbegin(from), cend(from), begin(to), init);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class T,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto exclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to, T init) {
  return exclusive_scan("Kokkos::exclusive_scan_default_functors_view_api", ex, from, to, init);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class T, class Op,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto exclusive_scan(const std::string& label, const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to,
                    T init, Op op) {
  return exclusive_scan(label, ex, cbegin(from), cend(from), begin(to), init, op);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class T, class Op,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto exclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to, T init, Op op) {
  return exclusive_scan("Kokkos::exclusive_scan_custom_functors_view_api", ex, from, to, init, op);
}

// ------------------------------------------------------------ transform scans
template <class Ex, class InIt, class OutIt, class Op, class UOp,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt transform_inclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, Op op,
                               UOp u) {
  using T = Impl::it_value_t<OutIt>;
  return Impl::scan_impl<T, true, false>(label, ex, first, last, dest, op, u, T{});
}
template <class Ex, class InIt, class OutIt, class Op, class UOp,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt transform_inclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest, Op op, UOp u) {
  return transform_inclusive_scan("Kokkos::transform_inclusive_scan_custom_functors_iterator_api", ex, first, last,
                                  dest, op, u);
}
template <class Ex, class InIt, class OutIt, class Op, class UOp, class T,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt transform_inclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, Op op,
                               UOp u, T init) {
  return Impl::scan_impl<T, true, true>(label, ex, first, last, dest, op, u, init);
}
template <class Ex, class InIt, class OutIt, class Op, class UOp, class T,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt transform_inclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest, Op op, UOp u, T init) {
  return transform_inclusive_scan("Kokkos::transform_inclusive_scan_custom_functors_iterator_api", ex, first, last,
                                  dest, op, u, init);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class Op, class UOp,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto transform_inclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to, Op op, UOp u) {
  return transform_inclusive_scan("Kokkos::transform_inclusive_scan_custom_functors_view_api", ex, cbegin(from),
                                  cend(from), begin(to), op, u);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class Op, class UOp, class T,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto transform_inclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to, Op op, UOp u,
                              T init) {
  return transform_inclusive_scan("Kokkos::transform_inclusive_scan_custom_functors_view_api", ex, cbegin(from),
                                  cend(from), begin(to), op, u, init);
}
template <class Ex, class InIt, class OutIt, class T, class Op, class UOp,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt transform_exclusive_scan(const std::string& label, const Ex& ex, InIt first, InIt last, OutIt dest, T init,
                               Op op, UOp u) {
  return Impl::scan_impl<T, false, true>(label, ex, first, last, dest, op, u, init);
}
template <class Ex, class InIt, class OutIt, class T, class Op, class UOp,
          std::enable_if_t<Impl::are_iterators_v<InIt, OutIt> && is_execution_space<Ex>::value, int> = 0>
OutIt transform_exclusive_scan(const Ex& ex, InIt first, InIt last, OutIt dest, T init, Op op, UOp u) {
  return transform_exclusive_scan("Kokkos::transform_exclusive_scan_custom_functors_iterator_api", ex, first, last,
                                  dest, init, op, u);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class T, class Op, class UOp,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
auto transform_exclusive_scan(const Ex& ex, const View<D1, P1...>& from, const View<D2, P2...>& to, T init, Op op,
                              UOp u) {
  return transform_exclusive_scan("Kokkos::transform_exclusive_scan_custom_functors_view_api", ex, cbegin(from),
                                  cend(from), begin(to), init, op, u);
}

// ------------------------------------------------------------ reduce
template <class Ex, class It, std::enable_if_t<Impl::are_iterators_v<It> && is_execution_space<Ex>::value, int> = 0>
typename It::value_type reduce(const Ex& ex, It first, It last) {
  using T = typename It::value_type;
  return Impl::reduce_impl<T, false>("Kokkos::reduce_default_functors_iterator_api", ex, first, last, first, T{},
                                     Impl::PlusOp{}, Impl::IdentityOp{});
}
template <class Ex, class It, class T,
          std::enable_if_t<Impl::are_iterators_v<It> && is_execution_space<Ex>::value, int> = 0>
T reduce(const Ex& ex, It first, It last, T init) {
  return Impl::reduce_impl<T, false>("Kokkos::reduce_default_functors_iterator_api", ex, first, last, first, init,
                                     Impl::PlusOp{}, Impl::IdentityOp{});
}
template <class Ex, class It, class T, class Op,
          std::enable_if_t<Impl::are_iterators_v<It> && is_execution_space<Ex>::value, int> = 0>
T reduce(const Ex& ex, It first, It last, T init, Op op) {
  return Impl::reduce_impl<T, false>("Kokkos::reduce_custom_functors_iterator_api", ex, first, last, first, init, op,
                                     Impl::IdentityOp{});
}
template <class Ex, class D, class... P, std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
typename View<D, P...>::non_const_value_type reduce(const Ex& ex, const View<D, P...>& v) {
  return reduce(ex, cbegin(v), cend(v));
}
template <class Ex, class D, class... P, class T,
          std::enable_if_t<is_execution_space<Ex>::value && !is_view<T>::value, int> = 0>
T reduce(const Ex& ex, const View<D, P...>& v, T init) {
  return reduce(ex, cbegin(v), cend(v), init);
}
template <class Ex, class D, class... P, class T, class Op, std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
T reduce(const Ex& ex, const View<D, P...>& v, T init, Op op) {
  return reduce(ex, cbegin(v), cend(v), init, op);
}

// ------------------------------------------------------------ transform_reduce
// init + sum_i a[i] * b[i]
template <class Ex, class It1, class It2, class T,
          std::enable_if_t<Impl::are_iterators_v<It1, It2> && is_execution_space<Ex>::value, int> = 0>
T transform_reduce(const Ex& ex, It1 first1, It1 last1, It2 first2, T init) {
  return Impl::reduce_impl<T, true>("Kokkos::transform_reduce_default_functors_iterator_api", ex, first1, last1,
                                    first2, init, Impl::PlusOp{}, Impl::MultiplyOp{});
}
// init join sum_i transform(a[i], b[i])
template <class Ex, class It1, class It2, class T, class Join, class Transform,
          std::enable_if_t<Impl::are_iterators_v<It1, It2> && is_execution_space<Ex>::value, int> = 0>
T transform_reduce(const Ex& ex, It1 first1, It1 last1, It2 first2, T init, Join join, Transform t) {
  return Impl::reduce_impl<T, true>("Kokkos::transform_reduce_custom_functors_iterator_api", ex, first1, last1, first2,
                                    init, join, t);
}
// init join sum_i transform(a[i])
template <class Ex, class It, class T, class Join, class Transform,
          std::enable_if_t<Impl::are_iterators_v<It> && is_execution_space<Ex>::value, int> = 0>
T transform_reduce(const Ex& ex, It first, It last, T init, Join join, Transform t) {
  return Impl::reduce_impl<T, false>("Kokkos::transform_reduce_custom_functors_iterator_api", ex, first, last, first,
                                     init, join, t);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class T,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
T transform_reduce(const Ex& ex, const View<D1, P1...>& a, const View<D2, P2...>& b, T init) {
  return transform_reduce(ex, cbegin(a), cend(a), cbegin(b), init);
}
template <class Ex, class D1, class... P1, class D2, class... P2, class T, class Join, class Transform,
          std::enable_if_t<is_execution_space<Ex>::value, int> = 0>
T transform_reduce(const Ex& ex, const View<D1, P1...>& a, const View<D2, P2...>& b, T init, Join join,
                   Transform t) {
  return transform_reduce(ex, cbegin(a), cend(a), cbegin(b), init, join, t);
}
template <class Ex, class D, class... P, class T, class Join, class Transform,
          std::enable_if_t<is_execution_space<Ex>::value && !is_view<T>::value, int> = 0>
T transform_reduce(const Ex& ex, const View<D, P...>& v, T init, Join join, Transform t) {
  return transform_reduce(ex, cbegin(v), cend(v), init, join, t);
}

}  // namespace Experimental
}  // namespace Kokkos

#endif
