// TeamPolicy on MI355X: one team per workgroup, vector lanes contiguous
// within a wavefront, level-0 scratch in LDS, level-1 scratch in HBM.
//
// Reference behaviour:
//  * member API, team_barrier / team_broadcast / team_reduce / team_scan:
//    core/src/HIP/Kokkos_HIP_Team.hpp:67-345;
//  * nested TeamThreadRange / ThreadVectorRange / TeamVectorRange for, reduce
//    and scan, single(PerTeam/PerThread): core/src/HIP/Kokkos_HIP_Team.hpp:357-1006
//    and core/src/Kokkos_ExecPolicy.hpp:779-980;
//  * scratch bump allocation with per-thread offsets:
//    core/src/Kokkos_ScratchSpace.hpp:83-140;
//  * a team reduction counts one value per team thread (vector lane 0).
//
// Layout: flat lane id = team_rank * vector_length + vector_lane.  The
// workgroup is the team rounded up to a multiple of 64 lanes; lanes beyond
// the team never call the functor but take part in barriers and in the
// workgroup-wide combine (with the identity), so every DPP and barrier sees
// whole wavefronts.  Collectives inside the functor use ds_bpermute shuffles
// within a wavefront and a 17-slot LDS area across wavefronts, always
// combining in thread order.
#ifndef KKMX_KERNEL_TEAM_HPP
#define KKMX_KERNEL_TEAM_HPP

#include <kkmx/KernelRange.hpp>

namespace Kokkos {

struct ParallelForTag {};
struct ParallelReduceTag {};
struct ParallelScanTag {};

// ------------------------------------------------------------------ scratch
// A team's scratch: two arenas, level 0 (LDS) and level 1 (the instance's HBM
// pool), each carved front to back.  A request reserves one block per thread
// sharing the arena (`copies`, set by PerThread use) and hands this thread the
// block at position `slot`; an exhausted arena returns nullptr and keeps its
// cursor.  The cursor is mutable so a member that is passed by const
// reference can still carve, as the reference's interface requires
// (core/src/Kokkos_ScratchSpace.hpp:83-140).
template <class ExecSpace>
class ScratchMemorySpace {
 public:
  static constexpr bool kkmx_is_scratch = true;
  using execution_space = ExecSpace;
  using memory_space    = ScratchMemorySpace<ExecSpace>;
  using device_type     = Device<execution_space, memory_space>;
  using array_layout    = typename ExecSpace::array_layout;
  using size_type       = typename ExecSpace::size_type;

  KOKKOS_INLINE_FUNCTION ScratchMemorySpace() = default;
  KOKKOS_INLINE_FUNCTION ScratchMemorySpace(void* lds, size_t lds_bytes, void* pool = nullptr,
                                            size_t pool_bytes = 0)
      : m_lds{static_cast<char*>(lds), static_cast<char*>(lds) + lds_bytes},
        m_pool{static_cast<char*>(pool), static_cast<char*>(pool) + pool_bytes} {}

  template <class IntType>
  KOKKOS_INLINE_FUNCTION void* get_shmem(const IntType& size, int level = -1) const {
    return carve(static_cast<size_t>(size), 1, level);
  }
  template <class IntType>
  KOKKOS_INLINE_FUNCTION void* get_shmem_aligned(const IntType& size, const ptrdiff_t align,
                                                 int level = -1) const {
    return carve(static_cast<size_t>(size), static_cast<size_t>(align), level);
  }
  // Later requests without a level go to `level`; each reserves `copies`
  // blocks and returns block `slot`.
  KOKKOS_INLINE_FUNCTION const ScratchMemorySpace& set_team_thread_mode(int level, int copies, int slot) const {
    m_level  = level;
    m_copies = copies;
    m_slot   = slot;
    return *this;
  }

 private:
  struct Arena {
    char* next  = nullptr;
    char* limit = nullptr;
  };
  KOKKOS_INLINE_FUNCTION void* carve(size_t bytes, size_t align, int level) const {
    Arena& a             = ((level < 0 ? m_level : level) == 0) ? m_lds : m_pool;
    const uintptr_t at   = reinterpret_cast<uintptr_t>(a.next);
    const uintptr_t from = align > 1 ? (at + align - 1) / align * align : at;
    const size_t span    = bytes * static_cast<size_t>(m_copies);
    if (from + span > reinterpret_cast<uintptr_t>(a.limit)) return nullptr;
    a.next = reinterpret_cast<char*>(from + span);
    return reinterpret_cast<char*>(from) + static_cast<size_t>(m_slot) * bytes;
  }
  mutable Arena m_lds{}, m_pool{};
  mutable int m_level = 0, m_copies = 1, m_slot = 0;
};

namespace Impl {

// Bytes of LDS the runtime keeps per team for collectives (17 slots of up to
// 64-byte values: 16 wavefront partials + 1 broadcast slot).
constexpr int kTeamCollectiveBytes = 17 * 64;

// ------------------------------------------------------------------ MI355X member
template <>
class TeamMember<MI355X> {
 public:
  using execution_space      = MI355X;
  using scratch_memory_space = ScratchMemorySpace<MI355X>;
  using team_handle          = TeamMember<MI355X>;

  KOKKOS_INLINE_FUNCTION TeamMember(char* collective, void* l0, size_t l0_size, void* l1, size_t l1_size,
                                    int league_rank, int league_size, int team_size, int vector_length)
      : m_coll(collective),
        m_scratch(l0, l0_size, l1, l1_size),
        m_league_rank(league_rank),
        m_league_size(league_size),
        m_team_size(team_size),
        m_vl(vector_length) {
#if defined(__HIP_DEVICE_COMPILE__)
    m_team_rank   = static_cast<int>(threadIdx.x) / vector_length;
    m_vector_lane = static_cast<int>(threadIdx.x) % vector_length;
#endif
  }

  KOKKOS_INLINE_FUNCTION int league_rank() const { return m_league_rank; }
  KOKKOS_INLINE_FUNCTION int league_size() const { return m_league_size; }
  KOKKOS_INLINE_FUNCTION int team_rank() const { return m_team_rank; }
  KOKKOS_INLINE_FUNCTION int team_size() const { return m_team_size; }
  KOKKOS_INLINE_FUNCTION int impl_vector_lane() const { return m_vector_lane; }
  KOKKOS_INLINE_FUNCTION int impl_vector_length() const { return m_vl; }

  KOKKOS_INLINE_FUNCTION const scratch_memory_space& team_shmem() const {
    return m_scratch.set_team_thread_mode(0, 1, 0);
  }
  KOKKOS_INLINE_FUNCTION const scratch_memory_space& team_scratch(int level) const {
    return m_scratch.set_team_thread_mode(level, 1, 0);
  }
  KOKKOS_INLINE_FUNCTION const scratch_memory_space& thread_scratch(int level) const {
    return m_scratch.set_team_thread_mode(level, m_team_size, m_team_rank);
  }

  KOKKOS_INLINE_FUNCTION void team_barrier() const {
#if defined(__HIP_DEVICE_COMPILE__)
    __syncthreads();
#endif
  }

  template <class T>
  KOKKOS_INLINE_FUNCTION void team_broadcast(T& val, const int root) const {
#if defined(__HIP_DEVICE_COMPILE__)
    static_assert(sizeof(T) <= 64, "team_broadcast: value larger than 64 bytes");
    T* slot = reinterpret_cast<T*>(m_coll + 16 * 64);
    __syncthreads();
    if (m_team_rank == root && m_vector_lane == 0) *slot = val;
    __syncthreads();
    val = *slot;
#else
    (void)val;
    (void)root;
#endif
  }
  template <class Closure, class T>
  KOKKOS_INLINE_FUNCTION void team_broadcast(const Closure& f, T& val, const int root) const {
    f(val);
    team_broadcast(val, root);
  }

  // ---- ordered collectives (device only)
#if defined(__HIP_DEVICE_COMPILE__)
  // Inclusive scan across lanes [seg_start, seg_start + seg_len) of this
  // wavefront, lower lanes on the left; returns the inclusive value.
  template <class T, class Join>
  __device__ static T seg_inclusive(T v, int pos_in_seg, int seg_len, const Join& join) {
    const int l = wave::lane();
    for (int d = 1; d < seg_len; d <<= 1) {
      T t = wave::shuffle(v, l - d >= 0 ? l - d : l);
      if (pos_in_seg >= d) {
        join(t, v);
        v = t;
      }
    }
    return v;
  }
  // Reduce over the team threads (vector lane 0 values) in thread order;
  // every lane of the team receives the result.
  template <class T, class Join>
  __device__ void reduce_threads(T& v, const T& identity, const Join& join) const {
    static_assert(sizeof(T) <= 64, "team reduction: value larger than 64 bytes");
    const int flat   = threadIdx.x;
    const int nthr   = m_team_size * m_vl;
    const int w      = flat >> 6;
    const int l      = wave::lane();
    const int wstart = w * 64;
    int nact         = nthr - wstart;
    if (nact > 64) nact = 64;
    T x  = (m_vector_lane == 0) ? v : identity;
    // lanes past the team are inactive: the DPP scan reads the identity there
    T in = wave::inclusive_scan(x, identity, join);
    const int nw = (nthr + 63) >> 6;
    if (nw == 1) {  // one-wavefront team: no LDS, no barrier
      v = wave::readlane(in, nact - 1);
      return;
    }
    T* slots = reinterpret_cast<T*>(m_coll);
    __syncthreads();
    if (l == nact - 1) slots[w] = in;
    __syncthreads();
    T r = identity;
    for (int k = 0; k < nw; ++k) join(r, slots[k]);
    __syncthreads();
    v = r;
  }
  // Exclusive scan over team threads in thread order; returns this thread's
  // exclusive prefix and writes the team total.
  template <class T, class Join>
  __device__ T scan_threads(const T& v, const T& identity, const Join& join, T* total) const {
    const int flat   = threadIdx.x;
    const int nthr   = m_team_size * m_vl;
    const int w      = flat >> 6;
    const int l      = wave::lane();
    const int wstart = w * 64;
    int nact         = nthr - wstart;
    if (nact > 64) nact = 64;
    T x  = (m_vector_lane == 0) ? v : identity;
    T in = wave::inclusive_scan(x, identity, join);
    // exclusive within the wavefront: inclusive value of the previous lane of
    // this thread's leader
    const int leader = (m_team_rank * m_vl) - wstart;
    T ex             = wave::shuffle(in, leader > 0 ? leader - 1 : 0);
    if (leader == 0) ex = identity;
    if (nthr <= 64) {  // one-wavefront team: no LDS, no barrier
      if (total) *total = wave::readlane(in, nact - 1);
      return ex;
    }
    T* slots = reinterpret_cast<T*>(m_coll);
    __syncthreads();
    if (l == nact - 1) slots[w] = in;
    __syncthreads();
    const int nw = (nthr + 63) >> 6;
    T off        = identity;
    T tot        = identity;
    for (int k = 0; k < nw; ++k) {
      if (k < w) join(off, slots[k]);
      join(tot, slots[k]);
    }
    __syncthreads();
    join(off, ex);
    if (total) *total = tot;
    return off;
  }
  // Reduce over the vector lanes of this thread, result in every lane.
  template <class T, class Join>
  __device__ void reduce_vector(T& v, const Join& join) const {
    const int seg0 = wave::lane() - m_vector_lane;
    T in           = seg_inclusive(v, m_vector_lane, m_vl, join);
    v              = wave::shuffle(in, seg0 + m_vl - 1);
  }
  template <class T, class Join>
  __device__ T scan_vector(const T& v, const T& identity, const Join& join, T* total) const {
    const int seg0 = wave::lane() - m_vector_lane;
    T in           = seg_inclusive(v, m_vector_lane, m_vl, join);
    T ex           = wave::shuffle(in, m_vector_lane > 0 ? wave::lane() - 1 : wave::lane());
    if (m_vector_lane == 0) ex = identity;
    if (total) *total = wave::shuffle(in, seg0 + m_vl - 1);
    return ex;
  }
  // Reduce over every lane of the team (TeamVectorRange), lane order.
  template <class T, class Join>
  __device__ void reduce_lanes(T& v, const T& identity, const Join& join) const {
    const int flat   = threadIdx.x;
    const int nthr   = m_team_size * m_vl;
    const int w      = flat >> 6;
    const int l      = wave::lane();
    int nact         = nthr - w * 64;
    if (nact > 64) nact = 64;
    T in         = wave::inclusive_scan(v, identity, join);
    const int nw = (nthr + 63) >> 6;
    if (nw == 1) {  // one-wavefront team: no LDS, no barrier
      v = wave::readlane(in, nact - 1);
      return;
    }
    T* slots = reinterpret_cast<T*>(m_coll);
    __syncthreads();
    if (l == nact - 1) slots[w] = in;
    __syncthreads();
    T r = identity;
    for (int k = 0; k < nw; ++k) join(r, slots[k]);
    __syncthreads();
    v = r;
  }
#endif

  template <class ReducerType>
  KOKKOS_INLINE_FUNCTION std::enable_if_t<is_reducer<ReducerType>::value> team_reduce(
      const ReducerType& reducer) const {
    team_reduce(reducer, reducer.reference());
  }
  template <class ReducerType>
  KOKKOS_INLINE_FUNCTION std::enable_if_t<is_reducer<ReducerType>::value> team_reduce(
      const ReducerType& reducer, typename ReducerType::value_type& value) const {
#if defined(__HIP_DEVICE_COMPILE__)
    using T = typename ReducerType::value_type;
    T id;
    reducer.init(id);
    reduce_threads(value, id, [&reducer](T& a, const T& b) { reducer.join(a, b); });
    reducer.reference() = value;
#else
    (void)reducer;
    (void)value;
#endif
  }
  template <class T>
  KOKKOS_INLINE_FUNCTION T team_scan(const T& value, T* const global_accum = nullptr) const {
#if defined(__HIP_DEVICE_COMPILE__)
    T total;
    T ex = scan_threads(value, T(0), [](T& a, const T& b) { a += b; }, &total);
    if (global_accum) {
      T* slot = reinterpret_cast<T*>(m_coll + 16 * 64);
      if (m_team_rank == 0 && m_vector_lane == 0) *slot = atomic_fetch_add(global_accum, total);
      __syncthreads();
      ex += *slot;
      __syncthreads();
    }
    return ex;
#else
    (void)global_accum;
    return value;
#endif
  }

 private:
  char* m_coll;
  scratch_memory_space m_scratch;
  int m_league_rank, m_league_size, m_team_size, m_vl;
  int m_team_rank   = 0;
  int m_vector_lane = 0;
};

// ------------------------------------------------------------------ Serial member
template <>
class TeamMember<Serial> {
 public:
  using execution_space      = Serial;
  using scratch_memory_space = ScratchMemorySpace<Serial>;
  using team_handle          = TeamMember<Serial>;
  TeamMember(void* l0, size_t l0_size, void* l1, size_t l1_size, int league_rank, int league_size)
      : m_scratch(l0, l0_size, l1, l1_size), m_league_rank(league_rank), m_league_size(league_size) {}
  KOKKOS_INLINE_FUNCTION int league_rank() const { return m_league_rank; }
  KOKKOS_INLINE_FUNCTION int league_size() const { return m_league_size; }
  KOKKOS_INLINE_FUNCTION int team_rank() const { return 0; }
  KOKKOS_INLINE_FUNCTION int team_size() const { return 1; }
  KOKKOS_INLINE_FUNCTION int impl_vector_lane() const { return 0; }
  KOKKOS_INLINE_FUNCTION int impl_vector_length() const { return 1; }
  const scratch_memory_space& team_shmem() const { return m_scratch.set_team_thread_mode(0, 1, 0); }
  const scratch_memory_space& team_scratch(int l) const { return m_scratch.set_team_thread_mode(l, 1, 0); }
  const scratch_memory_space& thread_scratch(int l) const { return m_scratch.set_team_thread_mode(l, 1, 0); }
  void team_barrier() const {}
  template <class T>
  void team_broadcast(T&, const int) const {}
  template <class Closure, class T>
  void team_broadcast(const Closure& f, T& val, const int) const {
    f(val);
  }
  template <class ReducerType>
  std::enable_if_t<is_reducer<ReducerType>::value> team_reduce(const ReducerType&) const {}
  template <class ReducerType>
  std::enable_if_t<is_reducer<ReducerType>::value> team_reduce(const ReducerType& r,
                                                               typename ReducerType::value_type& v) const {
    r.reference() = v;
  }
  template <class T>
  T team_scan(const T& value, T* const global_accum = nullptr) const {
    if (global_accum) {
      T old = *global_accum;
      *global_accum += value;
      return old;
    }
    return T(0);
  }

 private:
  scratch_memory_space m_scratch;
  int m_league_rank, m_league_size;
};

// ------------------------------------------------------------------ nested range structs
template <class iType, class Member>
struct TeamThreadRangeBoundariesStruct {
  using index_type = iType;
  const Member& member;
  iType start, end;
  KOKKOS_INLINE_FUNCTION TeamThreadRangeBoundariesStruct(const Member& m, iType b, iType e)
      : member(m), start(b), end(e) {}
};
template <class iType, class Member>
struct ThreadVectorRangeBoundariesStruct {
  using index_type = iType;
  const Member& member;
  iType start, end;
  KOKKOS_INLINE_FUNCTION ThreadVectorRangeBoundariesStruct(const Member& m, iType b, iType e)
      : member(m), start(b), end(e) {}
};
template <class iType, class Member>
struct TeamVectorRangeBoundariesStruct {
  using index_type = iType;
  const Member& member;
  iType start, end;
  KOKKOS_INLINE_FUNCTION TeamVectorRangeBoundariesStruct(const Member& m, iType b, iType e)
      : member(m), start(b), end(e) {}
};
template <class iType, class Member>
struct ThreadVectorMDRangeDummy {};

template <class T>
struct value_of_result {
  using type = std::remove_cv_t<std::remove_reference_t<T>>;
};

// Join/identity helpers for nested reductions: reducer or scalar sum.
template <class R, class = void>
struct NestedOps {
  using value_type = R;
  KOKKOS_INLINE_FUNCTION static void init(const R&, value_type& v) { v = value_type(); }
  KOKKOS_INLINE_FUNCTION static void join(const R&, value_type& a, const value_type& b) { a += b; }
};
template <class R>
struct NestedOps<R, std::enable_if_t<is_reducer<R>::value>> {
  using value_type = typename R::value_type;
  KOKKOS_INLINE_FUNCTION static void init(const R& r, value_type& v) { r.init(v); }
  KOKKOS_INLINE_FUNCTION static void join(const R& r, value_type& a, const value_type& b) { r.join(a, b); }
};
// Single-result nested reductions follow the reference's analysis of the
// closure (core/src/HIP/Kokkos_HIP_Team.hpp:509-573): a reducer's init /
// join / final; for a plain result, the closure's own init / join / final
// when it defines join, else a sum.  `final` runs on the reduced value.
template <class R, class F, class = void>
struct NestedClosureOps {
  using value_type = typename NestedOps<R>::value_type;
  KOKKOS_INLINE_FUNCTION static void init(const R& r, const F&, value_type& v) { NestedOps<R>::init(r, v); }
  KOKKOS_INLINE_FUNCTION static void join(const R& r, const F&, value_type& a, const value_type& b) {
    NestedOps<R>::join(r, a, b);
  }
  KOKKOS_INLINE_FUNCTION static void final(const R& r, const F&, value_type& v) {
    if constexpr (is_reducer<R>::value && has_final<R, void, value_type>::value) r.final(v);
  }
};
template <class R, class F>
struct NestedClosureOps<R, F, std::enable_if_t<!is_reducer<R>::value && has_join<F, void, R>::value>> {
  using value_type = R;
  KOKKOS_INLINE_FUNCTION static void init(const R&, const F& f, value_type& v) {
    if constexpr (has_init<F, void, R>::value) f.init(v);
    else v = value_type();
  }
  KOKKOS_INLINE_FUNCTION static void join(const R&, const F& f, value_type& a, const value_type& b) { f.join(a, b); }
  KOKKOS_INLINE_FUNCTION static void final(const R&, const F& f, value_type& v) {
    if constexpr (has_final<F, void, R>::value) f.final(v);
  }
};

template <class R>
KOKKOS_INLINE_FUNCTION void nested_store(const R& r, const typename NestedOps<R>::value_type& v) {
  if constexpr (is_reducer<R>::value) r.reference() = v;
}

}  // namespace Impl

template <class Member, class iType>
KOKKOS_INLINE_FUNCTION Impl::TeamThreadRangeBoundariesStruct<iType, Member> TeamThreadRange(const Member& m,
                                                                                            const iType& count) {
  return {m, iType(0), count};
}
template <class Member, class iType1, class iType2>
KOKKOS_INLINE_FUNCTION Impl::TeamThreadRangeBoundariesStruct<std::common_type_t<iType1, iType2>, Member>
TeamThreadRange(const Member& m, const iType1& b, const iType2& e) {
  using I = std::common_type_t<iType1, iType2>;
  return {m, I(b), I(e)};
}
template <class Member, class iType>
KOKKOS_INLINE_FUNCTION Impl::ThreadVectorRangeBoundariesStruct<iType, Member> ThreadVectorRange(const Member& m,
                                                                                                const iType& count) {
  return {m, iType(0), count};
}
template <class Member, class iType1, class iType2>
KOKKOS_INLINE_FUNCTION Impl::ThreadVectorRangeBoundariesStruct<std::common_type_t<iType1, iType2>, Member>
ThreadVectorRange(const Member& m, const iType1& b, const iType2& e) {
  using I = std::common_type_t<iType1, iType2>;
  return {m, I(b), I(e)};
}
template <class Member, class iType>
KOKKOS_INLINE_FUNCTION Impl::TeamVectorRangeBoundariesStruct<iType, Member> TeamVectorRange(const Member& m,
                                                                                            const iType& count) {
  return {m, iType(0), count};
}
template <class Member, class iType1, class iType2>
KOKKOS_INLINE_FUNCTION Impl::TeamVectorRangeBoundariesStruct<std::common_type_t<iType1, iType2>, Member>
TeamVectorRange(const Member& m, const iType1& b, const iType2& e) {
  using I = std::common_type_t<iType1, iType2>;
  return {m, I(b), I(e)};
}

template <class Member, class = decltype(std::declval<const Member&>().team_rank())>
KOKKOS_INLINE_FUNCTION Impl::SinglePerTeam<Member> PerTeam(const Member& m) {
  return {m};
}
template <class Member, class = decltype(std::declval<const Member&>().team_rank())>
KOKKOS_INLINE_FUNCTION Impl::SinglePerThread<Member> PerThread(const Member& m) {
  return {m};
}

// ------------------------------------------------------------------ nested parallel_for
template <class iType, class Member, class F>
KOKKOS_INLINE_FUNCTION void parallel_for(const Impl::TeamThreadRangeBoundariesStruct<iType, Member>& r, const F& f) {
  const int ts = r.member.team_size();
  for (iType i = r.start + r.member.team_rank(); i < r.end; i += ts) f(i);
}
template <class iType, class Member, class F>
KOKKOS_INLINE_FUNCTION void parallel_for(const Impl::ThreadVectorRangeBoundariesStruct<iType, Member>& r,
                                         const F& f) {
  const int vl = r.member.impl_vector_length();
  for (iType i = r.start + r.member.impl_vector_lane(); i < r.end; i += vl) f(i);
}
template <class iType, class Member, class F>
KOKKOS_INLINE_FUNCTION void parallel_for(const Impl::TeamVectorRangeBoundariesStruct<iType, Member>& r, const F& f) {
  const int n  = r.member.team_size() * r.member.impl_vector_length();
  const int id = r.member.team_rank() * r.member.impl_vector_length() + r.member.impl_vector_lane();
  for (iType i = r.start + id; i < r.end; i += n) f(i);
}

// ------------------------------------------------------------------ nested parallel_reduce
// Result: a scalar (sum) or a reducer; every thread of the team (or of the
// vector group) receives the result.
template <class iType, class Member, class F, class R>
KOKKOS_INLINE_FUNCTION void parallel_reduce(const Impl::TeamThreadRangeBoundariesStruct<iType, Member>& r,
                                            const F& f, R&& result) {
  using RR  = Impl::remove_cvref_t<R>;
  using Ops = Impl::NestedClosureOps<RR, F>;
  using T   = typename Ops::value_type;
  const RR& red = result;
  T v;
  Ops::init(red, f, v);
  const int ts = r.member.team_size();
  for (iType i = r.start + r.member.team_rank(); i < r.end; i += ts) f(i, v);
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (std::is_same<Member, Impl::TeamMember<MI355X>>::value) {
    T id;
    Ops::init(red, f, id);
    r.member.reduce_threads(v, id, [&red, &f](T& a, const T& b) { Ops::join(red, f, a, b); });
  }
#endif
  Ops::final(red, f, v);
  if constexpr (is_reducer<RR>::value) Impl::nested_store(red, v);
  else result = v;
}
template <class iType, class Member, class F, class R>
KOKKOS_INLINE_FUNCTION void parallel_reduce(const Impl::ThreadVectorRangeBoundariesStruct<iType, Member>& r,
                                            const F& f, R&& result) {
  using RR  = Impl::remove_cvref_t<R>;
  using Ops = Impl::NestedClosureOps<RR, F>;
  using T   = typename Ops::value_type;
  const RR& red = result;
  T v;
  Ops::init(red, f, v);
  const int vl = r.member.impl_vector_length();
  for (iType i = r.start + r.member.impl_vector_lane(); i < r.end; i += vl) f(i, v);
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (std::is_same<Member, Impl::TeamMember<MI355X>>::value) {
    if (vl > 1) r.member.reduce_vector(v, [&red, &f](T& a, const T& b) { Ops::join(red, f, a, b); });
  }
#endif
  Ops::final(red, f, v);
  if constexpr (is_reducer<RR>::value) Impl::nested_store(red, v);
  else result = v;
}
template <class iType, class Member, class F, class R>
KOKKOS_INLINE_FUNCTION void parallel_reduce(const Impl::TeamVectorRangeBoundariesStruct<iType, Member>& r,
                                            const F& f, R&& result) {
  using RR  = Impl::remove_cvref_t<R>;
  using Ops = Impl::NestedClosureOps<RR, F>;
  using T   = typename Ops::value_type;
  const RR& red = result;
  T v;
  Ops::init(red, f, v);
  const int n  = r.member.team_size() * r.member.impl_vector_length();
  const int id = r.member.team_rank() * r.member.impl_vector_length() + r.member.impl_vector_lane();
  for (iType i = r.start + id; i < r.end; i += n) f(i, v);
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (std::is_same<Member, Impl::TeamMember<MI355X>>::value) {
    T idv;
    Ops::init(red, f, idv);
    r.member.reduce_lanes(v, idv, [&red, &f](T& a, const T& b) { Ops::join(red, f, a, b); });
  }
#endif
  Ops::final(red, f, v);
  if constexpr (is_reducer<RR>::value) Impl::nested_store(red, v);
  else result = v;
}

// ------------------------------------------------------------------ nested combined reductions
// parallel_reduce(range, f, result0, result1, ...): several results (sums
// into scalars, or reducers) from one pass, the functor taking one update
// per result (core/src/Kokkos_Parallel_Reduce.hpp combined reducers, used
// at team level by core/unit_test/TestTeamCombinedReducers.hpp).  The
// values travel as one struct through the single-result reduction above,
// joined member by member with each result's own init and join.
namespace Impl {

template <class R>
struct is_nested_range : std::false_type {};
template <class I, class M>
struct is_nested_range<TeamThreadRangeBoundariesStruct<I, M>> : std::true_type {};
template <class I, class M>
struct is_nested_range<ThreadVectorRangeBoundariesStruct<I, M>> : std::true_type {};
template <class I, class M>
struct is_nested_range<TeamVectorRangeBoundariesStruct<I, M>> : std::true_type {};

template <class... Rs>
struct NestedVals;
template <class R0>
struct NestedVals<R0> {
  typename NestedOps<R0>::value_type head;
};
template <class R0, class R1, class... Rs>
struct NestedVals<R0, R1, Rs...> {
  typename NestedOps<R0>::value_type head;
  NestedVals<R1, Rs...> tail;
};

template <size_t K, class V>
KOKKOS_INLINE_FUNCTION auto& nested_get(V& v) {
  if constexpr (K == 0) return v.head;
  else return nested_get<K - 1>(v.tail);
}
template <size_t K, class V>
KOKKOS_INLINE_FUNCTION const auto& nested_get(const V& v) {
  if constexpr (K == 0) return v.head;
  else return nested_get<K - 1>(v.tail);
}

template <class... Rs>
struct NestedCombined {
  using reducer    = NestedCombined;
  using value_type = NestedVals<Rs...>;

  template <size_t... K>
  KOKKOS_INLINE_FUNCTION void init_k(value_type& v, std::index_sequence<K...>) const {
    (NestedOps<Rs>::init(*static_cast<const Rs*>(res_at<K>()), nested_get<K>(v)), ...);
  }
  template <size_t... K>
  KOKKOS_INLINE_FUNCTION void join_k(value_type& a, const value_type& b, std::index_sequence<K...>) const {
    (NestedOps<Rs>::join(*static_cast<const Rs*>(res_at<K>()), nested_get<K>(a), nested_get<K>(b)), ...);
  }
  template <size_t... K>
  KOKKOS_INLINE_FUNCTION void store_k(const value_type& v, std::index_sequence<K...>) const {
    (store_one(*res_at<K>(), nested_get<K>(v)), ...);
  }
  template <class R, class T>
  KOKKOS_INLINE_FUNCTION static void store_one(R& r, const T& v) {
    if constexpr (is_reducer<R>::value) r.reference() = v;
    else r = v;
  }
  template <size_t K>
  KOKKOS_INLINE_FUNCTION auto* res_at() const {
    using R = std::tuple_element_t<K, std::tuple<Rs...>>;
    return static_cast<R*>(ptrs[K]);
  }
  void* ptrs[sizeof...(Rs)];

  KOKKOS_INLINE_FUNCTION void init(value_type& v) const { init_k(v, std::index_sequence_for<Rs...>{}); }
  KOKKOS_INLINE_FUNCTION void join(value_type& a, const value_type& b) const {
    join_k(a, b, std::index_sequence_for<Rs...>{});
  }
  struct Store {
    const NestedCombined* c;
    KOKKOS_INLINE_FUNCTION void operator=(const value_type& v) const {
      c->store_k(v, std::index_sequence_for<Rs...>{});
    }
  };
  KOKKOS_INLINE_FUNCTION Store reference() const { return Store{this}; }
};

template <class F, class V, size_t... K, class I>
KOKKOS_INLINE_FUNCTION void nested_call(const F& f, const I& i, V& v, std::index_sequence<K...>) {
  f(i, nested_get<K>(v)...);
}

}  // namespace Impl

template <class Range, class F, class R0, class R1, class... Rs>
KOKKOS_INLINE_FUNCTION std::enable_if_t<Impl::is_nested_range<Range>::value> parallel_reduce(const Range& r,
                                                                                           const F& f, R0&& r0,
                                                                                           R1&& r1, Rs&&... rs) {
  using C = Impl::NestedCombined<Impl::remove_cvref_t<R0>, Impl::remove_cvref_t<R1>, Impl::remove_cvref_t<Rs>...>;
  C c;
  void* p[] = {const_cast<void*>(static_cast<const void*>(&r0)), const_cast<void*>(static_cast<const void*>(&r1)),
               const_cast<void*>(static_cast<const void*>(&rs))...};
  for (size_t k = 0; k < 2 + sizeof...(Rs); ++k) c.ptrs[k] = p[k];
  using V = typename C::value_type;
  parallel_reduce(
      r,
      [&](const auto& i, V& v) { Impl::nested_call(f, i, v, std::make_index_sequence<2 + sizeof...(Rs)>{}); },
      c);
}

// ------------------------------------------------------------------ nested parallel_scan
namespace Impl {
template <class F, class iType, class T>
KOKKOS_INLINE_FUNCTION void nested_scan_value(const F& f, iType i, T& v) {
  f(i, v, false);
}
}  // namespace Impl

template <class iType, class Member, class F, class T>
KOKKOS_INLINE_FUNCTION void parallel_scan(const Impl::TeamThreadRangeBoundariesStruct<iType, Member>& r,
                                          const F& f, T& return_value) {
  const int ts = r.member.team_size();
  T running    = T();
  for (iType base = r.start; base < r.end; base += ts) {
    const iType i = base + r.member.team_rank();
    T v           = T();
    if (i < r.end) f(i, v, false);
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<Member, Impl::TeamMember<MI355X>>::value) {
      T total;
      T ex  = r.member.scan_threads(v, T(), [](T& a, const T& b) { a += b; }, &total);
      T upd = running + ex;
      if (i < r.end) f(i, upd, true);
      running += total;
    } else
#endif
    {
      T upd = running;
      if (i < r.end) f(i, upd, true);
      running = upd;
    }
  }
  return_value = running;
}
template <class iType, class Member, class F>
KOKKOS_INLINE_FUNCTION void parallel_scan(const Impl::TeamThreadRangeBoundariesStruct<iType, Member>& r,
                                          const F& f) {
  using T = typename Impl::arg_from_back<typename Impl::op_args<F>::type, 1>::type;
  T dummy;
  parallel_scan(r, f, dummy);
}
template <class iType, class Member, class F, class T,
          class = std::enable_if_t<!is_reducer<std::remove_const_t<T>>::value>>
KOKKOS_INLINE_FUNCTION void parallel_scan(const Impl::ThreadVectorRangeBoundariesStruct<iType, Member>& r,
                                          const F& f, T& return_value) {
  const int vl = r.member.impl_vector_length();
  T running    = T();
  for (iType base = r.start; base < r.end; base += vl) {
    const iType i = base + r.member.impl_vector_lane();
    T v           = T();
    if (i < r.end) f(i, v, false);
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (std::is_same<Member, Impl::TeamMember<MI355X>>::value) {
      T total;
      T ex  = r.member.scan_vector(v, T(), [](T& a, const T& b) { a += b; }, &total);
      T upd = running + ex;
      if (i < r.end) f(i, upd, true);
      running += total;
    } else
#endif
    {
      T upd = running;
      if (i < r.end) f(i, upd, true);
      running = upd;
    }
  }
  return_value = running;
}
template <class iType, class Member, class F>
KOKKOS_INLINE_FUNCTION void parallel_scan(const Impl::ThreadVectorRangeBoundariesStruct<iType, Member>& r,
                                          const F& f) {
  using T = typename Impl::arg_from_back<typename Impl::op_args<F>::type, 1>::type;
  T dummy;
  parallel_scan(r, f, dummy);
}
// Scan over a thread's vector lanes with a reducer's identity and join in
// place of 0 and +; the total lands in the reducer's result
// (core/src/HIP/Kokkos_HIP_Team.hpp, parallel_scan(ThreadVectorRange,
// closure, reducer)).
template <class iType, class Member, class F, class R, class = std::enable_if_t<is_reducer<R>::value>>
KOKKOS_INLINE_FUNCTION void parallel_scan(const Impl::ThreadVectorRangeBoundariesStruct<iType, Member>& r,
                                          const F& f, const R& reducer) {
  using T = typename R::value_type;
  T identity;
  reducer.init(identity);
  T running = identity;
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (std::is_same<Member, Impl::TeamMember<MI355X>>::value) {
    const int vl = r.member.impl_vector_length();
    auto join    = [&](T& a, const T& b) { reducer.join(a, b); };
    for (iType base = r.start; base < r.end; base += vl) {
      const iType i = base + r.member.impl_vector_lane();
      T v           = identity;
      if (i < r.end) f(i, v, false);
      T total;
      T upd = running;
      reducer.join(upd, r.member.scan_vector(v, identity, join, &total));
      if (i < r.end) f(i, upd, true);
      reducer.join(running, total);
    }
  } else
#endif
  {
    for (iType i = r.start; i < r.end; ++i) f(i, running, true);
  }
  reducer.reference() = running;
}

// ------------------------------------------------------------------ single
template <class Member, class F>
KOKKOS_INLINE_FUNCTION void single(const Impl::SinglePerTeam<Member>& s, const F& f) {
  if (s.team.team_rank() == 0 && s.team.impl_vector_lane() == 0) f();
}
template <class Member, class F>
KOKKOS_INLINE_FUNCTION void single(const Impl::SinglePerThread<Member>& s, const F& f) {
  if (s.team.impl_vector_lane() == 0) f();
}
template <class Member, class F, class T>
KOKKOS_INLINE_FUNCTION void single(const Impl::SinglePerTeam<Member>& s, const F& f, T& val) {
  if (s.team.team_rank() == 0 && s.team.impl_vector_lane() == 0) f(val);
  s.team.team_broadcast(val, 0);
}
template <class Member, class F, class T>
KOKKOS_INLINE_FUNCTION void single(const Impl::SinglePerThread<Member>& s, const F& f, T& val) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (std::is_same<Member, Impl::TeamMember<MI355X>>::value) {
    if (s.team.impl_vector_lane() == 0) f(val);
    val = Impl::wave::shuffle(val, Impl::wave::lane() - s.team.impl_vector_lane());
    return;
  }
#endif
  f(val);
}

// ================================================================== closures
namespace Impl {

template <class F, class = void>
struct functor_team_shmem {
  static size_t get(const F&, int) { return 0; }
};
template <class F>
struct functor_team_shmem<F, std::void_t<decltype(std::declval<const F&>().team_shmem_size(0))>> {
  static size_t get(const F& f, int ts) { return static_cast<size_t>(f.team_shmem_size(ts)); }
};
template <class F, class = void>
struct functor_shmem_size {
  static size_t get(const F&, int) { return 0; }
};
template <class F>
struct functor_shmem_size<F, std::void_t<decltype(std::declval<const F&>().shmem_size(0))>> {
  static size_t get(const F& f, int ts) { return static_cast<size_t>(f.shmem_size(ts)); }
};

struct TeamLaunch {
  int league_size;
  int team_size;
  int vector_length;
  int l0_offset;      // start of level-0 scratch within dynamic LDS
  int l0_bytes;       // level-0 scratch bytes per team
  size_t l1_bytes;    // level-1 scratch bytes per team
  char* l1_pool;      // gridDim.x slices of l1_bytes
};

template <class F, class Tag>
__global__ __launch_bounds__(1024) void k_team_for(const F f, const TeamLaunch p) {
  extern __shared__ __align__(16) char kkmx_team_lds[];
  const bool active = static_cast<int>(threadIdx.x) < p.team_size * p.vector_length;
  char* l1          = p.l1_pool ? p.l1_pool + static_cast<size_t>(blockIdx.x) * p.l1_bytes : nullptr;
  for (int lr = blockIdx.x; lr < p.league_size; lr += gridDim.x) {
    TeamMember<MI355X> m(kkmx_team_lds, kkmx_team_lds + p.l0_offset, p.l0_bytes, l1, p.l1_bytes, lr,
                         p.league_size, p.team_size, p.vector_length);
    if (active) call_functor<Tag>(f, m);
    __syncthreads();
  }
}

template <class CR>
__global__ __launch_bounds__(1024) void k_team_reduce(const CR cr, const TeamLaunch p,
                                                      typename CR::value_type* partials, unsigned* counter,
                                                      typename CR::value_type* result) {
  using V   = typename CR::value_type;
  using Tag = typename CR::work_tag;
  extern __shared__ __align__(16) char kkmx_team_lds[];
  __shared__ LdsSlots16<V> slots;
  __shared__ int s_last;
  const auto join   = [&cr](V& a, const V& b) { cr.join(a, b); };
  const bool active = static_cast<int>(threadIdx.x) < p.team_size * p.vector_length;
  char* l1          = p.l1_pool ? p.l1_pool + static_cast<size_t>(blockIdx.x) * p.l1_bytes : nullptr;
  V identity;
  cr.init(identity);
  V v = identity;
  for (int lr = blockIdx.x; lr < p.league_size; lr += gridDim.x) {
    TeamMember<MI355X> m(kkmx_team_lds, kkmx_team_lds + p.l0_offset, p.l0_bytes, l1, p.l1_bytes, lr,
                         p.league_size, p.team_size, p.vector_length);
    if (active) call_functor<Tag>(cr.functor(), m, v);
    __syncthreads();
  }
  // one value per team thread: vector lane 0 only
  const bool counts = active && (static_cast<int>(threadIdx.x) % p.vector_length) == 0;
  V x               = counts ? v : identity;
  x                 = block_reduce(x, identity, join, slots.get(), blockDim.x);
  if (gridDim.x == 1) {
    if (threadIdx.x == 0) {
      cr.final(x);
      *result = x;
    }
    return;
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = x;
    const unsigned prev  = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    s_last               = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  V w = identity;
  for (int j = threadIdx.x; j < static_cast<int>(gridDim.x); j += blockDim.x) join(w, partials[j]);
  w = block_reduce(w, identity, join, slots.get(), blockDim.x);
  if (threadIdx.x == 0) {
    cr.final(w);
    *result = w;
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Runtime-length results: every lane keeps value_count accumulators in
// `acc` (after the team's LDS when they fit, else a global slice per
// workgroup); lanes other than vector lane 0 are reset before the ordered
// pairwise combine, since the vector lanes of a team thread share one value.
template <class CR>
__global__ __launch_bounds__(1024) void k_team_reduce_array(const CR cr, const TeamLaunch p, const int acc_offset,
                                                            typename CR::value_type* gacc,
                                                            typename CR::value_type* partials, unsigned* counter,
                                                            typename CR::value_type* result) {
  using V   = typename CR::value_type;
  using Tag = typename CR::work_tag;
  extern __shared__ __align__(16) char kkmx_team_lds[];
  __shared__ int s_last;
  const unsigned n  = cr.value_count();
  const bool active = static_cast<int>(threadIdx.x) < p.team_size * p.vector_length;
  char* l1          = p.l1_pool ? p.l1_pool + static_cast<size_t>(blockIdx.x) * p.l1_bytes : nullptr;
  V* base = gacc ? gacc + static_cast<size_t>(blockIdx.x) * blockDim.x * n : reinterpret_cast<V*>(kkmx_team_lds + acc_offset);
  V* mine = base + static_cast<size_t>(threadIdx.x) * n;
  cr.init_array(mine);
  for (int lr = blockIdx.x; lr < p.league_size; lr += gridDim.x) {
    TeamMember<MI355X> m(kkmx_team_lds, kkmx_team_lds + p.l0_offset, p.l0_bytes, l1, p.l1_bytes, lr,
                         p.league_size, p.team_size, p.vector_length);
    if (active) call_functor<Tag>(cr.functor(), m, mine);
    __syncthreads();
  }
  const bool counts = active && (static_cast<int>(threadIdx.x) % p.vector_length) == 0;
  if (!counts) cr.init_array(mine);
  __syncthreads();
  for (unsigned s = 1; s < blockDim.x; s <<= 1) {
    if ((threadIdx.x % (2 * s)) == 0 && threadIdx.x + s < blockDim.x)
      cr.join_array(mine, base + static_cast<size_t>(threadIdx.x + s) * n);
    __syncthreads();
  }
  if (gridDim.x == 1) {
    if (threadIdx.x == 0) {
      cr.final_array(base);
      for (unsigned k = 0; k < n; ++k) result[k] = base[k];
    }
    return;
  }
  if (threadIdx.x == 0) {
    for (unsigned k = 0; k < n; ++k) partials[static_cast<size_t>(blockIdx.x) * n + k] = base[k];
    const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    s_last              = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (threadIdx.x == 0) {
    cr.init_array(base);
    for (unsigned b = 0; b < gridDim.x; ++b) cr.join_array(base, partials + static_cast<size_t>(b) * n);
    cr.final_array(base);
    for (unsigned k = 0; k < n; ++k) result[k] = base[k];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Team shapes a tuning tool chooses from (team_size, vector_length), all
// within 1024 lanes.  The tool's answer only replaces the parts the policy
// leaves to Kokkos::AUTO (reference TeamSizeTuner,
// core/src/impl/Kokkos_Tools_Generic.hpp:219-230).
constexpr int kTeamCandidates[16][2] = {{64, 1},  {128, 1}, {256, 1}, {512, 1}, {1024, 1}, {32, 2},
                                        {64, 2},  {256, 2}, {16, 4},  {32, 4},  {64, 4},   {128, 4},
                                        {16, 8},  {32, 8},  {64, 8},  {128, 8}};
inline int team_candidate_index(int ts, int vl) {
  for (int c = 0; c < 16; ++c)
    if (kTeamCandidates[c][0] == ts && kTeamCandidates[c][1] == vl) return c;
  return -1;
}

// Launch shape of a team dispatch.  pattern: 0 for, 1 reduce.
template <class Policy, class F>
TeamLaunch team_launch_shape(const Policy& policy, const F& f, MI355XInstance* inst, int& block, int& grid,
                             size_t& lds, int pattern = 0, size_t value_bytes = 0, bool floating = false) {
  TeamLaunch p{};
  p.vector_length = policy.impl_vector_length() > 0 ? policy.impl_vector_length() : 1;
  int ts          = policy.impl_team_size_raw();
  if (ts <= 0) ts = policy.impl_default_team_size();
  // AUTO team size of a parallel_for: one wavefront. Nested team reductions
  // and barriers then stay inside one wave, which the team sweeps measure at
  // up to 3x faster than 256-lane teams; reductions keep 256 lanes.
  if (policy.impl_auto_team_size() && pattern == 0) ts = std::max(1, 64 / p.vector_length);
  if ((policy.impl_auto_team_size() || policy.impl_auto_vector_length()) &&
      Tools::Experimental::have_tuning_tool()) {
    const int h = team_candidate_index(ts, p.vector_length);
    const int c = request_launch_candidate<F>(LaunchVar::Team, pattern, policy.league_size(), value_bytes, floating, h);
    if (c >= 0 && c != h) {
      if (policy.impl_auto_team_size()) ts = kTeamCandidates[c][0];
      if (policy.impl_auto_vector_length()) p.vector_length = kTeamCandidates[c][1];
      if (ts * p.vector_length > 1024) ts = 1024 / p.vector_length;
    }
  }
  if (ts * p.vector_length > 1024)
    throw_runtime_exception("Kokkos::MI355X TeamPolicy: team_size * vector_length exceeds 1024");
  p.team_size   = ts;
  p.league_size = policy.league_size();
  block         = ((ts * p.vector_length + 63) / 64) * 64;
  const size_t l0 = policy.team_scratch_size(0) + static_cast<size_t>(ts) * policy.thread_scratch_size(0) +
                    functor_team_shmem<F>::get(f, ts) + functor_shmem_size<F>::get(f, ts);
  p.l0_offset = kTeamCollectiveBytes;
  p.l0_bytes  = static_cast<int>(l0);
  lds         = kTeamCollectiveBytes + ((l0 + 15) / 16) * 16;
  if (lds > 160 * 1024)
    throw_runtime_exception("Kokkos::MI355X TeamPolicy: requested level-0 scratch exceeds the 160 KiB of LDS");
  const int cus = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
  int per_cu    = (2048 / block) > 0 ? 2048 / block : 1;
  if (lds > 0) {
    int by_lds = static_cast<int>((160 * 1024) / lds);
    if (by_lds < per_cu) per_cu = by_lds > 0 ? by_lds : 1;
  }
  // parallel_for: 8 waves of resident teams; reductions: 2, so the last
  // workgroup's pass over the per-workgroup partials stays short.
  int64_t g = std::min<int64_t>(p.league_size, static_cast<int64_t>(cus) * per_cu * (pattern == 1 ? 2 : 8));
  grid      = static_cast<int>(g > 0 ? g : 0);
  p.l1_bytes = policy.team_scratch_size(1) + static_cast<size_t>(ts) * policy.thread_scratch_size(1);
  p.l1_pool  = nullptr;
  if (p.l1_bytes > 0 && grid > 0)
    p.l1_pool = static_cast<char*>(inst->team_l1_space(p.l1_bytes * static_cast<size_t>(grid)));
  return p;
}

template <class F, class... Traits>
class ParallelFor<F, TeamPolicy<Traits...>, MI355X> {
 public:
  using Policy = TeamPolicy<Traits...>;
  ParallelFor(const F& f, const Policy& p) : m_f(f), m_p(p) {}
  void execute() const {
    MI355XInstance* inst = m_p.space().impl_internal_space_instance();
    std::lock_guard<std::mutex> lk(inst->mtx);
    int block = 0, grid = 0;
    size_t lds = 0;
    TeamLaunch tl = team_launch_shape(m_p, m_f, inst, block, grid, lds);
    if (grid == 0) return;
    F f = m_f;
    const void* k = reinterpret_cast<const void*>(&k_team_for<F, typename Policy::work_tag>);
    mx_launch_kernel(k, static_cast<unsigned>(grid), block, lds, inst->stream, f, tl);
  }

 private:
  F m_f;
  Policy m_p;
};

template <class CR, class... Traits>
class ParallelReduce<CR, TeamPolicy<Traits...>, MI355X> {
 public:
  using Policy = TeamPolicy<Traits...>;
  using V      = typename CR::value_type;
  template <class Mem>
  ParallelReduce(const CR& cr, const Policy& p, V* result, Mem)
      : m_cr(cr), m_p(p), m_result(result), m_on_device(exec_can_access<MI355X, Mem>::value) {}
  void execute() const {
    MI355XInstance* inst = m_p.space().impl_internal_space_instance();
    std::lock_guard<std::mutex> lk(inst->mtx);
    int block = 0, grid = 0;
    size_t lds = 0;
    TeamLaunch tl = team_launch_shape(m_p, m_cr.functor(), inst, block, grid, lds, 1, sizeof(V),
                                      std::is_floating_point<V>::value);
    if (grid == 0) grid = 1;
    const unsigned cnt = m_cr.value_count();
    const size_t bytes = sizeof(V) * cnt;
    V* dst         = m_result;
    bool copy_back = false;
    if (!m_on_device || !dst) {
      dst       = static_cast<V*>(inst->pinned_space(bytes));
      copy_back = (m_result != nullptr);
    }
    CR cr             = m_cr;
    unsigned* counter = inst->control;
    if constexpr (!CR::is_array) {
      V* partials   = static_cast<V*>(inst->partials_space(sizeof(V) * grid));
      const void* k = reinterpret_cast<const void*>(&k_team_reduce<CR>);
      mx_launch_kernel(k, static_cast<unsigned>(grid), block, lds, inst->stream, cr, tl, partials, counter, dst);
    } else {
      // Per-lane accumulators after the team's LDS while the workgroup stays
      // within 64 KiB, else in a global slice per workgroup.
      const size_t acc_bytes = bytes * block;
      int acc_offset         = static_cast<int>(((lds + 15) / 16) * 16);
      const bool in_lds      = acc_offset + acc_bytes <= 64 * 1024;
      const size_t part      = ((bytes * grid + 255) / 256) * 256;
      char* scratch          = static_cast<char*>(inst->partials_space(part + (in_lds ? 0 : acc_bytes * grid)));
      V* partials            = reinterpret_cast<V*>(scratch);
      V* gacc                = in_lds ? nullptr : reinterpret_cast<V*>(scratch + part);
      const size_t lds_total = in_lds ? acc_offset + acc_bytes : lds;
      const void* k          = reinterpret_cast<const void*>(&k_team_reduce_array<CR>);
      mx_launch_kernel(k, static_cast<unsigned>(grid), block, lds_total, inst->stream, cr, tl, acc_offset, gacc,
                       partials, counter, dst);
    }
    if (copy_back) {
      KKMX_SAFE_CALL(mx_stream_sync(inst->stream));
      std::memcpy(static_cast<void*>(m_result), dst, bytes);
    }
  }

 private:
  CR m_cr;
  Policy m_p;
  V* m_result;
  bool m_on_device;
};

// ------------------------------------------------------------------ Serial team closures
template <class F, class... Traits>
class ParallelFor<F, TeamPolicy<Traits...>, Serial> {
 public:
  using Policy = TeamPolicy<Traits...>;
  ParallelFor(const F& f, const Policy& p) : m_f(f), m_p(p) {}
  void execute() const {
    const size_t l0 = m_p.team_scratch_size(0) + m_p.thread_scratch_size(0) + functor_team_shmem<F>::get(m_f, 1) +
                      functor_shmem_size<F>::get(m_f, 1);
    const size_t l1 = m_p.team_scratch_size(1) + m_p.thread_scratch_size(1);
    std::vector<char> b0(l0 + 16), b1(l1 + 16);
    for (int lr = 0; lr < m_p.league_size(); ++lr) {
      TeamMember<Serial> m(b0.data(), l0, b1.data(), l1, lr, m_p.league_size());
      call_functor<typename Policy::work_tag>(m_f, m);
    }
  }

 private:
  F m_f;
  Policy m_p;
};
template <class CR, class... Traits>
class ParallelReduce<CR, TeamPolicy<Traits...>, Serial> {
 public:
  using Policy = TeamPolicy<Traits...>;
  using V      = typename CR::value_type;
  template <class Mem>
  ParallelReduce(const CR& cr, const Policy& p, V* result, Mem) : m_cr(cr), m_p(p), m_result(result) {}
  void execute() const {
    const size_t l0 = m_p.team_scratch_size(0) + m_p.thread_scratch_size(0);
    const size_t l1 = m_p.team_scratch_size(1) + m_p.thread_scratch_size(1);
    std::vector<char> b0(l0 + 16), b1(l1 + 16);
    if constexpr (CR::is_array) {
      std::vector<V> tmp(m_cr.value_count());
      V* v = m_result ? m_result : tmp.data();
      m_cr.init_array(v);
      for (int lr = 0; lr < m_p.league_size(); ++lr) {
        TeamMember<Serial> m(b0.data(), l0, b1.data(), l1, lr, m_p.league_size());
        call_functor<typename CR::work_tag>(m_cr.functor(), m, v);
      }
      m_cr.final_array(v);
    } else {
      V v;
      m_cr.init(v);
      for (int lr = 0; lr < m_p.league_size(); ++lr) {
        TeamMember<Serial> m(b0.data(), l0, b1.data(), l1, lr, m_p.league_size());
        call_functor<typename CR::work_tag>(m_cr.functor(), m, v);
      }
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
