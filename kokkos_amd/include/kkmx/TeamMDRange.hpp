// Multidimensional nested ranges inside a team: TeamThreadMDRange,
// TeamVectorMDRange, ThreadVectorMDRange with parallel_for and (sum)
// parallel_reduce.
//
// Reference interface: core/src/Kokkos_ExecPolicy.hpp:1010-1240 (the range
// types, their deduction guides, scalar-only reductions summed over the
// team / vector lanes).  The reference walks all but one or two nest levels
// serially and spreads one level over threads and the innermost over
// vector lanes (impl/Kokkos_TeamMDPolicy.hpp).  Here the whole box is one
// flat index space dealt out over the same lanes as the one-dimensional
// range of that level (TeamThreadRange / TeamVectorRange /
// ThreadVectorRange); a flat index becomes a multi-index with the fastest
// index of the iteration direction varying fastest, so consecutive lanes
// touch consecutive elements of Views in that layout and every lane gets
// work even when one extent is small.  Each point is visited once either
// way; floating-point sums may associate differently.
#ifndef KKMX_TEAM_MD_RANGE_HPP
#define KKMX_TEAM_MD_RANGE_HPP

#include <kkmx/KernelTeam.hpp>

#include <type_traits>
#include <utility>

namespace Kokkos {
namespace Impl {

template <class Space>
constexpr Iterate team_md_direction(Iterate d) {
  return d == Iterate::Default ? default_outer_direction<Space>::value : d;
}

template <class RankT, class TeamHandle, int Level>
struct TeamMDRangeBase {
  using NestLevelType  = int;
  using BoundaryType   = int;
  using TeamHandleType = TeamHandle;
  using ExecutionSpace = typename TeamHandle::execution_space;
  using ArrayLayout    = typename ExecutionSpace::array_layout;
  static constexpr NestLevelType total_nest_level = RankT::rank;
  static constexpr Iterate iter                   = RankT::outer_direction;
  static constexpr Iterate direction              = team_md_direction<ExecutionSpace>(iter);
  static constexpr int level                      = Level;  // 0 threads, 1 threads x vector, 2 vector

  template <class... Args>
  KOKKOS_INLINE_FUNCTION TeamMDRangeBase(const TeamHandle& t, Args... args)
      : team(t), boundaries{static_cast<BoundaryType>(args)...} {
    static_assert(sizeof...(Args) == total_nest_level, "one extent per dimension");
  }
  const TeamHandle& team;
  BoundaryType boundaries[total_nest_level];

  KOKKOS_INLINE_FUNCTION int total() const {
    int n = 1;
    for (int d = 0; d < total_nest_level; ++d) n *= boundaries[d];
    return n;
  }
  // flat index -> multi-index, the direction's fastest index fastest
  KOKKOS_INLINE_FUNCTION void split(int flat, int* idx) const {
    if constexpr (direction == Iterate::Right) {
      for (int d = total_nest_level - 1; d >= 0; --d) {
        idx[d] = flat % boundaries[d];
        flat /= boundaries[d];
      }
    } else {
      for (int d = 0; d < total_nest_level; ++d) {
        idx[d] = flat % boundaries[d];
        flat /= boundaries[d];
      }
    }
  }
  KOKKOS_INLINE_FUNCTION auto flat_range() const {
    if constexpr (Level == 0) return TeamThreadRange(team, total());
    else if constexpr (Level == 1) return TeamVectorRange(team, total());
    else return ThreadVectorRange(team, total());
  }
};

template <class F, size_t... I>
KOKKOS_INLINE_FUNCTION void call_md(const F& f, const int* idx, std::index_sequence<I...>) {
  f(idx[I]...);
}
template <class F, class V, size_t... I>
KOKKOS_INLINE_FUNCTION void call_md_reduce(const F& f, const int* idx, V& v, std::index_sequence<I...>) {
  f(idx[I]..., v);
}

}  // namespace Impl

template <class RankT, class TeamHandle>
struct TeamThreadMDRange : Impl::TeamMDRangeBase<RankT, TeamHandle, 0> {
  using Impl::TeamMDRangeBase<RankT, TeamHandle, 0>::TeamMDRangeBase;
};
template <class RankT, class TeamHandle>
struct TeamVectorMDRange : Impl::TeamMDRangeBase<RankT, TeamHandle, 1> {
  using Impl::TeamMDRangeBase<RankT, TeamHandle, 1>::TeamMDRangeBase;
};
template <class RankT, class TeamHandle>
struct ThreadVectorMDRange : Impl::TeamMDRangeBase<RankT, TeamHandle, 2> {
  using Impl::TeamMDRangeBase<RankT, TeamHandle, 2>::TeamMDRangeBase;
};

template <class TeamHandle, class... Args>
TeamThreadMDRange(const TeamHandle&, Args...) -> TeamThreadMDRange<Rank<sizeof...(Args)>, TeamHandle>;
template <class TeamHandle, class... Args>
TeamVectorMDRange(const TeamHandle&, Args...) -> TeamVectorMDRange<Rank<sizeof...(Args)>, TeamHandle>;
template <class TeamHandle, class... Args>
ThreadVectorMDRange(const TeamHandle&, Args...) -> ThreadVectorMDRange<Rank<sizeof...(Args)>, TeamHandle>;

namespace Impl {
template <class P, class Lambda>
KOKKOS_INLINE_FUNCTION void team_md_for(const P& policy, const Lambda& lambda) {
  constexpr int N = P::total_nest_level;
  parallel_for(policy.flat_range(), [&](const int flat) {
    int idx[N];
    policy.split(flat, idx);
    call_md(lambda, idx, std::make_index_sequence<N>{});
  });
}
// Scalar results only, summed over the participating lanes; every lane
// receives the sum (as the reference's team_reduce of a Sum).
template <class P, class Lambda, class V>
KOKKOS_INLINE_FUNCTION void team_md_reduce(const P& policy, const Lambda& lambda, V& val) {
  static_assert(!std::is_array<V>::value && !std::is_pointer<V>::value && !is_reducer<V>::value,
                "Only scalar return types are allowed!");
  constexpr int N = P::total_nest_level;
  V sum{};
  parallel_reduce(
      policy.flat_range(),
      [&](const int flat, V& u) {
        int idx[N];
        policy.split(flat, idx);
        call_md_reduce(lambda, idx, u, std::make_index_sequence<N>{});
      },
      sum);
  val = sum;
}
}  // namespace Impl

template <class R, class T, class Lambda>
KOKKOS_INLINE_FUNCTION void parallel_for(const TeamThreadMDRange<R, T>& p, const Lambda& l) {
  Impl::team_md_for(p, l);
}
template <class R, class T, class Lambda>
KOKKOS_INLINE_FUNCTION void parallel_for(const TeamVectorMDRange<R, T>& p, const Lambda& l) {
  Impl::team_md_for(p, l);
}
template <class R, class T, class Lambda>
KOKKOS_INLINE_FUNCTION void parallel_for(const ThreadVectorMDRange<R, T>& p, const Lambda& l) {
  Impl::team_md_for(p, l);
}
template <class R, class T, class Lambda, class V>
KOKKOS_INLINE_FUNCTION void parallel_reduce(const TeamThreadMDRange<R, T>& p, const Lambda& l, V& val) {
  Impl::team_md_reduce(p, l, val);
}
template <class R, class T, class Lambda, class V>
KOKKOS_INLINE_FUNCTION void parallel_reduce(const TeamVectorMDRange<R, T>& p, const Lambda& l, V& val) {
  Impl::team_md_reduce(p, l, val);
}
template <class R, class T, class Lambda, class V>
KOKKOS_INLINE_FUNCTION void parallel_reduce(const ThreadVectorMDRange<R, T>& p, const Lambda& l, V& val) {
  Impl::team_md_reduce(p, l, val);
}

}  // namespace Kokkos

#endif
