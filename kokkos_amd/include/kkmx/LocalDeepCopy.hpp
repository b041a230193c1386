// Kokkos::Experimental::local_deep_copy and local_deep_copy_contiguous:
// View copies and fills made by one team (or one thread) inside a kernel.
//
// Reference interface: core/src/Kokkos_CopyViews.hpp:1268-2070 (team and
// sequential copies for ranks 1-7, fills with a value, contiguous forms).
//
// The team forms run over TeamVectorRange, so every lane of every thread of
// the team moves elements, between two team barriers.  When both Views are
// contiguous the elements are moved as one flat run over the span
// (consecutive lanes, consecutive addresses); otherwise a flat index is
// split into a multi-index with the first index fastest, as the reference
// does, so lanes of a LayoutLeft View still touch neighbouring elements.
// The sequential forms are plain loops for one thread.  A View without data
// is left alone.
#ifndef KKMX_LOCAL_DEEP_COPY_HPP
#define KKMX_LOCAL_DEEP_COPY_HPP

#include <kkmx/View.hpp>

#include <cstddef>
#include <type_traits>
#include <utility>

namespace Kokkos {
namespace Impl {

template <class V, size_t... I>
KOKKOS_INLINE_FUNCTION typename V::reference_type local_copy_at(const V& v, const size_t* idx,
                                                                std::index_sequence<I...>) {
  return v(idx[I]...);
}

// Element i of v's elements in first-index-fastest order.
template <class V>
KOKKOS_INLINE_FUNCTION typename V::reference_type local_copy_element(const V& v, size_t i) {
  constexpr unsigned R = V::rank;
  if constexpr (R == 0) {
    (void)i;
    return v();
  } else {
    size_t idx[R];
    for (unsigned d = 0; d < R; ++d) {
      const size_t e = v.extent(d);
      idx[d]         = i % e;
      i /= e;
    }
    return local_copy_at(v, idx, std::make_index_sequence<R>{});
  }
}

}  // namespace Impl

namespace Experimental {

template <class TeamType, class DT, class... DP, class ST, class... SP>
KOKKOS_INLINE_FUNCTION void local_deep_copy_contiguous(const TeamType& team, const View<DT, DP...>& dst,
                                                       const View<ST, SP...>& src) {
  parallel_for(TeamVectorRange(team, src.span()), [&](const size_t i) { dst.data()[i] = src.data()[i]; });
}
template <class DT, class... DP, class ST, class... SP>
KOKKOS_INLINE_FUNCTION void local_deep_copy_contiguous(const View<DT, DP...>& dst, const View<ST, SP...>& src) {
  for (size_t i = 0; i < src.span(); ++i) dst.data()[i] = src.data()[i];
}
template <class TeamType, class DT, class... DP>
KOKKOS_INLINE_FUNCTION void local_deep_copy_contiguous(const TeamType& team, const View<DT, DP...>& dst,
                                                       typename View<DT, DP...>::const_value_type& value) {
  parallel_for(TeamVectorRange(team, dst.span()), [&](const size_t i) { dst.data()[i] = value; });
}
template <class DT, class... DP>
KOKKOS_INLINE_FUNCTION void local_deep_copy_contiguous(const View<DT, DP...>& dst,
                                                       typename View<DT, DP...>::const_value_type& value) {
  for (size_t i = 0; i < dst.span(); ++i) dst.data()[i] = value;
}

// Team copy, src and dst of the same rank >= 1.
template <class TeamType, class DT, class... DP, class ST, class... SP>
KOKKOS_INLINE_FUNCTION void local_deep_copy(
    const TeamType& team, const View<DT, DP...>& dst, const View<ST, SP...>& src,
    std::enable_if_t<(unsigned(View<DT, DP...>::rank) == unsigned(View<ST, SP...>::rank) &&
                      unsigned(View<DT, DP...>::rank) >= 1)>* = nullptr) {
  if (dst.data() == nullptr) return;
  team.team_barrier();
  if (View<DT, DP...>::rank > 1 && dst.span_is_contiguous() && src.span_is_contiguous()) {
    local_deep_copy_contiguous(team, dst, src);
  } else {
    parallel_for(TeamVectorRange(team, dst.size()), [&](const size_t i) {
      Impl::local_copy_element(dst, i) = Impl::local_copy_element(src, i);
    });
  }
  team.team_barrier();
}

// Team fill with one value, rank >= 1.
template <class TeamType, class DT, class... DP>
KOKKOS_INLINE_FUNCTION void local_deep_copy(
    const TeamType& team, const View<DT, DP...>& dst, typename View<DT, DP...>::const_value_type& value,
    std::enable_if_t<(unsigned(View<DT, DP...>::rank) >= 1)>* = nullptr) {
  if (dst.data() == nullptr) return;
  team.team_barrier();
  if (View<DT, DP...>::rank > 1 && dst.span_is_contiguous()) {
    local_deep_copy_contiguous(team, dst, value);
  } else {
    parallel_for(TeamVectorRange(team, dst.size()),
                 [&](const size_t i) { Impl::local_copy_element(dst, i) = value; });
  }
  team.team_barrier();
}

// Sequential copy by one thread, src and dst of the same rank >= 1.
template <class DT, class... DP, class ST, class... SP>
KOKKOS_INLINE_FUNCTION void local_deep_copy(
    const View<DT, DP...>& dst, const View<ST, SP...>& src,
    std::enable_if_t<(unsigned(View<DT, DP...>::rank) == unsigned(View<ST, SP...>::rank) &&
                      unsigned(View<DT, DP...>::rank) >= 1)>* = nullptr) {
  if (dst.data() == nullptr) return;
  if (View<DT, DP...>::rank > 1 && dst.span_is_contiguous() && src.span_is_contiguous()) {
    local_deep_copy_contiguous(dst, src);
    return;
  }
  const size_t n = dst.size();
  for (size_t i = 0; i < n; ++i) Impl::local_copy_element(dst, i) = Impl::local_copy_element(src, i);
}

// Sequential fill by one thread, rank >= 1.
template <class DT, class... DP>
KOKKOS_INLINE_FUNCTION void local_deep_copy(const View<DT, DP...>& dst,
                                            typename View<DT, DP...>::const_value_type& value,
                                            std::enable_if_t<(unsigned(View<DT, DP...>::rank) >= 1)>* = nullptr) {
  if (dst.data() == nullptr) return;
  if (View<DT, DP...>::rank > 1 && dst.span_is_contiguous()) {
    local_deep_copy_contiguous(dst, value);
    return;
  }
  const size_t n = dst.size();
  for (size_t i = 0; i < n; ++i) Impl::local_copy_element(dst, i) = value;
}

}  // namespace Experimental
}  // namespace Kokkos

#endif
