// Kokkos::pair, Kokkos::ALL and Kokkos::subview.
//
// Reference interfaces: core/src/Kokkos_Pair.hpp:35-210 (pair, make_pair),
// core/src/View/Kokkos_ViewTraits.hpp:40-64 (ALL_t, ALL),
// core/src/Kokkos_View.hpp:988-1016 (subview, Subview), and the layout rules
// of core/src/View/Kokkos_ViewMapping.hpp:45-170 and 3215-3290.
//
// A subview argument per source dimension is an integer (the dimension is
// dropped), Kokkos::ALL (kept whole) or a pair [first, second) of integers
// (kept, narrowed).  The result points into the source allocation with the
// source's strides for the kept dimensions and counts the allocation like a
// copy of the source.  Its layout is the source's when that layout still
// describes it -- LayoutLeft needs the kept dimensions to start at source
// dimension 0 (stride 1), LayoutRight to end at the last source dimension --
// following the reference's rules exactly, and LayoutStride otherwise.  The
// layout tag only promises the unit-stride dimension; the other strides are
// the parent's, which the View's offsets read.
#ifndef KKMX_SUBVIEW_HPP
#define KKMX_SUBVIEW_HPP

#include <kkmx/View.hpp>

#include <type_traits>
#include <utility>

namespace Kokkos {

// ------------------------------------------------------------------ pair
template <class T1, class T2>
struct pair {
  using first_type  = T1;
  using second_type = T2;
  T1 first;
  T2 second;

  KOKKOS_DEFAULTED_FUNCTION constexpr pair() = default;
  KOKKOS_FORCEINLINE_FUNCTION constexpr pair(const first_type& f, const second_type& s) : first(f), second(s) {}
  template <class U, class V>
  KOKKOS_FORCEINLINE_FUNCTION constexpr pair(const pair<U, V>& p) : first(p.first), second(p.second) {}
  template <class U, class V>
  constexpr pair(const std::pair<U, V>& p) : first(p.first), second(p.second) {}
  template <class U, class V>
  KOKKOS_FORCEINLINE_FUNCTION pair& operator=(const pair<U, V>& p) {
    first  = p.first;
    second = p.second;
    return *this;
  }
  std::pair<T1, T2> to_std_pair() const { return std::make_pair(first, second); }
};

template <class T1, class T2>
KOKKOS_FORCEINLINE_FUNCTION constexpr pair<T1, T2> make_pair(T1 x, T2 y) {
  return pair<T1, T2>(x, y);
}
template <class T1, class T2>
KOKKOS_FORCEINLINE_FUNCTION constexpr bool operator==(const pair<T1, T2>& a, const pair<T1, T2>& b) {
  return a.first == b.first && a.second == b.second;
}
template <class T1, class T2>
KOKKOS_FORCEINLINE_FUNCTION constexpr bool operator!=(const pair<T1, T2>& a, const pair<T1, T2>& b) {
  return !(a == b);
}
template <class T1, class T2>
KOKKOS_FORCEINLINE_FUNCTION constexpr bool operator<(const pair<T1, T2>& a, const pair<T1, T2>& b) {
  return a.first < b.first || (!(b.first < a.first) && a.second < b.second);
}
template <class T1, class T2>
KOKKOS_FORCEINLINE_FUNCTION constexpr bool operator<=(const pair<T1, T2>& a, const pair<T1, T2>& b) {
  return !(b < a);
}
template <class T1, class T2>
KOKKOS_FORCEINLINE_FUNCTION constexpr bool operator>(const pair<T1, T2>& a, const pair<T1, T2>& b) {
  return b < a;
}
template <class T1, class T2>
KOKKOS_FORCEINLINE_FUNCTION constexpr bool operator>=(const pair<T1, T2>& a, const pair<T1, T2>& b) {
  return !(a < b);
}

// ------------------------------------------------------------------ ALL
struct ALL_t {
  KOKKOS_FORCEINLINE_FUNCTION constexpr const ALL_t& operator()() const { return *this; }
  KOKKOS_FORCEINLINE_FUNCTION constexpr bool operator==(const ALL_t&) const { return true; }
};
inline constexpr ALL_t ALL{};

// ------------------------------------------------------------------ subview
namespace Impl {

template <class A>
struct subview_range : std::false_type {};
template <class I>
struct subview_range<std::pair<I, I>> : std::is_integral<I> {};
template <class I>
struct subview_range<Kokkos::pair<I, I>> : std::is_integral<I> {};

template <class A>
struct subview_arg {
  using T = std::remove_cv_t<std::remove_reference_t<A>>;
  static constexpr bool all   = std::is_same<T, ALL_t>::value;
  static constexpr bool range = subview_range<T>::value;
  static constexpr bool index = std::is_integral<T>::value;
  static constexpr bool kept  = all || range;  // "integral extent" in the reference
  static_assert(kept || index, "subview argument must be either integral or integral extent");
};

template <class T, unsigned R>
struct add_pointers {
  using type = typename add_pointers<T, R - 1>::type*;
};
template <class T>
struct add_pointers<T, 0> {
  using type = T;
};

// Whether the result keeps the source layout L.
template <class L, class... A>
constexpr bool subview_keeps_layout() {
  constexpr unsigned n   = sizeof...(A);
  constexpr bool kept[]  = {subview_arg<A>::kept..., false};
  constexpr bool all[]   = {subview_arg<A>::all..., false};
  constexpr bool index[] = {subview_arg<A>::index..., false};
  unsigned r             = 0;
  for (unsigned c = 0; c < n; ++c) r += kept[c] ? 1u : 0u;
  if (r == 0 || std::is_same<L, LayoutStride>::value) return true;
  if (std::is_same<L, LayoutLeft>::value) {
    if (r <= 2 && kept[0]) return true;
    for (unsigned c = 0; c < n; ++c) {
      const bool ok = (c + 1 == n) ? (c + 1 == r || index[c])
                                   : ((c + 1 == r && kept[c]) || (c >= r && index[c]) || (c < r && all[c]) ||
                                      (c == 0 && kept[c]));
      if (!ok) return false;
    }
    return true;
  }
  if (std::is_same<L, LayoutRight>::value) {
    if (r <= 2 && kept[n - 1]) return true;
    for (unsigned c = 0; c < n; ++c) {
      const bool ok = (c + 1 == n) ? all[c]
                                   : ((c == n - r && kept[c]) || (c < n - r && index[c]) || (c >= n - r && all[c]));
      if (!ok) return false;
    }
    return true;
  }
  return false;
}

template <class SV, class... A>
struct SubviewType {
  static_assert(SV::rank == sizeof...(A),
                "Subview mapping requires one argument for each dimension of source View");
  static constexpr unsigned rank = (0u + ... + unsigned(subview_arg<A>::kept));
  using src_layout               = typename SV::array_layout;
  using layout =
      std::conditional_t<subview_keeps_layout<src_layout, A...>(), src_layout, LayoutStride>;
  using data_type = typename add_pointers<typename SV::value_type, rank>::type;
  using type      = View<data_type, layout, typename SV::device_type, typename SV::memory_traits>;
};

// One source dimension: advance the offset, record a kept dimension.
template <class SV, class A>
KOKKOS_INLINE_FUNCTION void subview_dim(const SV& src, unsigned c, const A& a, size_t& off, size_t* ext, size_t* str,
                                        unsigned& k) {
  using S = subview_arg<A>;
  if constexpr (S::index) {
#if !defined(__HIP_DEVICE_COMPILE__)
    if (size_t(a) >= src.extent(c) && src.extent(c) > 0)
      Kokkos::abort("Kokkos::subview: index out of the source extent");
#endif
    off += size_t(a) * src.stride(c);
  } else if constexpr (S::all) {
    ext[k]   = src.extent(c);
    str[k++] = src.stride(c);
  } else {
    const size_t b = size_t(a.first), e = size_t(a.second);
#if !defined(__HIP_DEVICE_COMPILE__)
    if (b > e || e > src.extent(c)) Kokkos::abort("Kokkos::subview: range outside the source extent");
#endif
    off += b * src.stride(c);
    ext[k]   = e - b;
    str[k++] = src.stride(c);
  }
}

}  // namespace Impl

template <class D, class... P, class... Args>
KOKKOS_INLINE_FUNCTION auto subview(const View<D, P...>& src, Args... args) {
  using SV  = View<D, P...>;
  using Out = typename Impl::SubviewType<SV, Args...>::type;
  size_t off = 0, ext[8] = {0, 0, 0, 0, 0, 0, 0, 0}, str[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  unsigned k = 0, c = 0;
  (Impl::subview_dim(src, c++, args, off, ext, str, k), ...);
  return Out(Impl::SubviewCtor{}, src.data() + off, ext, str, src.impl_record());
}

template <class V, class... Args>
using Subview = decltype(subview(std::declval<V>(), std::declval<Args>()...));

}  // namespace Kokkos

#endif
