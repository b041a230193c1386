// Team- and thread-level sorts inside a kernel: Kokkos::Experimental::
// sort_team, sort_by_key_team, sort_thread, sort_by_key_thread.
//
// Reference interfaces: algorithms/src/sorting/Kokkos_NestedSortPublicAPI.hpp
// :27-100; algorithm: impl/Kokkos_NestedSortImpl.hpp:24-110.
//
// Both levels run Batcher's bitonic network over the View, padded to the
// next power of two (comparisons whose partner lies past the end are
// skipped, as if the padding held +infinity).  Stage i (blocks of 2^(i+1)
// elements) starts with a flip step that compares element o of a block with
// element 2^(i+1)-1-o, then half-cleaner steps compare elements 2^(i-j)
// apart.  The pairs of a step are spread over the team's (or the thread's
// vector) lanes, and a step ends with a team barrier (or, for a thread,
// whose vector lanes share one wavefront, a wavefront-scope fence).  The
// network and the swap rule (swap when comp(b, a)) are the reference's, so
// equal keys -- and, for sort_by_key, their values -- end in the same places.
#ifndef KKMX_NESTED_SORT_HPP
#define KKMX_NESTED_SORT_HPP

#include <Kokkos_Core.hpp>

#include <cstddef>
#include <type_traits>

namespace Kokkos {
namespace Impl {

struct NestedLess {
  template <class T>
  KOKKOS_INLINE_FUNCTION bool operator()(const T& a, const T& b) const {
    return a < b;
  }
};

template <bool Team>
struct NestedSortLanes;

template <>
struct NestedSortLanes<true> {
  template <class M, class I>
  KOKKOS_INLINE_FUNCTION static auto range(const M& m, I n) {
    return TeamVectorRange(m, n);
  }
  template <class M>
  KOKKOS_INLINE_FUNCTION static void step_done(const M& m) {
    m.team_barrier();
  }
};

template <>
struct NestedSortLanes<false> {
  template <class M, class I>
  KOKKOS_INLINE_FUNCTION static auto range(const M& m, I n) {
    return ThreadVectorRange(m, n);
  }
  template <class M>
  KOKKOS_INLINE_FUNCTION static void step_done(const M&) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#endif
  }
};

template <bool Team, class M, class KV, class VV, class Comp>
KOKKOS_INLINE_FUNCTION void bitonic_nested(const M& m, const KV& keys, const VV& values, const Comp& comp) {
  using Lanes     = NestedSortLanes<Team>;
  using K         = typename KV::non_const_value_type;
  const size_t n  = keys.extent(0);
  unsigned levels = 0;
  while ((size_t(1) << levels) < n) ++levels;
  const size_t pairs = (size_t(1) << levels) / 2;
  for (unsigned i = 0; i < levels; ++i) {
    for (unsigned j = 0; j <= i; ++j) {
      const unsigned sh = i - j;
      const size_t half = size_t(1) << sh;
      parallel_for(Lanes::range(m, pairs), [&](const size_t k) {
        // pair k: a zero bit inserted at position sh gives its lower element
        const size_t lo = ((k >> sh) << (sh + 1)) | (k & (half - 1));
        const size_t hi = j == 0 ? (lo ^ (2 * half - 1)) : (lo | half);
        if (hi < n) {
          const K a = keys(lo), b = keys(hi);
          if (comp(b, a)) {
            keys(lo) = b;
            keys(hi) = a;
            if constexpr (!std::is_same<VV, std::nullptr_t>::value) {
              const auto t = values(lo);
              values(lo)   = values(hi);
              values(hi)   = t;
            }
          }
        }
      });
      Lanes::step_done(m);
    }
  }
}

}  // namespace Impl

namespace Experimental {

template <class TeamMember, class ViewType>
KOKKOS_INLINE_FUNCTION void sort_team(const TeamMember& t, const ViewType& view) {
  Kokkos::Impl::bitonic_nested<true>(t, view, nullptr, Kokkos::Impl::NestedLess{});
}
template <class TeamMember, class ViewType, class Comparator>
KOKKOS_INLINE_FUNCTION void sort_team(const TeamMember& t, const ViewType& view, const Comparator& comp) {
  Kokkos::Impl::bitonic_nested<true>(t, view, nullptr, comp);
}
template <class TeamMember, class KeyViewType, class ValueViewType>
KOKKOS_INLINE_FUNCTION void sort_by_key_team(const TeamMember& t, const KeyViewType& keys,
                                             const ValueViewType& values) {
  Kokkos::Impl::bitonic_nested<true>(t, keys, values, Kokkos::Impl::NestedLess{});
}
template <class TeamMember, class KeyViewType, class ValueViewType, class Comparator>
KOKKOS_INLINE_FUNCTION void sort_by_key_team(const TeamMember& t, const KeyViewType& keys, const ValueViewType& values,
                                             const Comparator& comp) {
  Kokkos::Impl::bitonic_nested<true>(t, keys, values, comp);
}

template <class TeamMember, class ViewType>
KOKKOS_INLINE_FUNCTION void sort_thread(const TeamMember& t, const ViewType& view) {
  Kokkos::Impl::bitonic_nested<false>(t, view, nullptr, Kokkos::Impl::NestedLess{});
}
template <class TeamMember, class ViewType, class Comparator>
KOKKOS_INLINE_FUNCTION void sort_thread(const TeamMember& t, const ViewType& view, const Comparator& comp) {
  Kokkos::Impl::bitonic_nested<false>(t, view, nullptr, comp);
}
template <class TeamMember, class KeyViewType, class ValueViewType>
KOKKOS_INLINE_FUNCTION void sort_by_key_thread(const TeamMember& t, const KeyViewType& keys,
                                               const ValueViewType& values) {
  Kokkos::Impl::bitonic_nested<false>(t, keys, values, Kokkos::Impl::NestedLess{});
}
template <class TeamMember, class KeyViewType, class ValueViewType, class Comparator>
KOKKOS_INLINE_FUNCTION void sort_by_key_thread(const TeamMember& t, const KeyViewType& keys,
                                               const ValueViewType& values, const Comparator& comp) {
  Kokkos::Impl::bitonic_nested<false>(t, keys, values, comp);
}

}  // namespace Experimental
}  // namespace Kokkos

#endif
