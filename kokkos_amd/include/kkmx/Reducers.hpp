// Built-in reducers and reduction identities.
//
// Join rules and identities follow core/src/Kokkos_Parallel_Reduce.hpp:33-1351
// and core/src/Kokkos_ReductionIdentity.hpp:143-398 of the reference:
//  * min/max identities are the type's max/lowest finite values (FLT_MAX,
//    DBL_MAX for floating point, not infinities);
//  * the *Loc reducers keep the source on a strict improvement and on a tie
//    only fill the location while it is still the index identity (:440-447);
//  * the *FirstLoc reducers keep the smaller location on a tie (:708-716);
//  * FirstLoc / LastLoc keep the smallest / largest index (:1134, :1195).
// Each reducer either references a host scalar (references_scalar() true,
// the dispatch fences) or a rank-0 View in the given space (asynchronous).
#ifndef KKMX_REDUCERS_HPP
#define KKMX_REDUCERS_HPP

#include <kkmx/View.hpp>

#include <cfloat>
#include <climits>

namespace Kokkos {

template <class T, class = void>
struct reduction_identity {
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T sum() { return T(0); }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T prod() { return T(1); }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T max() { return std::numeric_limits<T>::lowest(); }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T min() { return std::numeric_limits<T>::max(); }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T bor() { return T(0); }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T band() { return ~T(0); }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T lor() { return T(0); }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static T land() { return T(1); }
};
template <>
struct reduction_identity<float> {
  KOKKOS_FORCEINLINE_FUNCTION constexpr static float sum() { return 0.0f; }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static float prod() { return 1.0f; }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static float max() { return -FLT_MAX; }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static float min() { return FLT_MAX; }
};
template <>
struct reduction_identity<double> {
  KOKKOS_FORCEINLINE_FUNCTION constexpr static double sum() { return 0.0; }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static double prod() { return 1.0; }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static double max() { return -DBL_MAX; }
  KOKKOS_FORCEINLINE_FUNCTION constexpr static double min() { return DBL_MAX; }
};

namespace Impl {
// Storage and accessors common to every built-in reducer.
template <class ValueType, class Space>
class ReducerStorage {
 public:
  using value_type       = ValueType;
  using result_view_type = View<value_type, Space>;
  KOKKOS_INLINE_FUNCTION ReducerStorage(value_type& v) : m_value(&v), m_scalar(true) {}
  KOKKOS_INLINE_FUNCTION ReducerStorage(const result_view_type& v) : m_value(v), m_scalar(false) {}
  KOKKOS_INLINE_FUNCTION value_type& reference() const { return *m_value.data(); }
  KOKKOS_INLINE_FUNCTION result_view_type view() const { return m_value; }
  KOKKOS_INLINE_FUNCTION bool references_scalar() const { return m_scalar; }

 private:
  result_view_type m_value;
  bool m_scalar;
};
}  // namespace Impl

#define KKMX_SIMPLE_REDUCER(NAME, JOIN, IDENT)                                         \
  template <class Scalar, class Space = HostSpace>                                      \
  struct NAME : Impl::ReducerStorage<std::remove_cv_t<Scalar>, Space> {                 \
    using reducer          = NAME<Scalar, Space>;                                       \
    using value_type       = std::remove_cv_t<Scalar>;                                  \
    using result_view_type = View<value_type, Space>;                                   \
    using base             = Impl::ReducerStorage<value_type, Space>;                   \
    KOKKOS_INLINE_FUNCTION NAME(value_type& v) : base(v) {}                             \
    KOKKOS_INLINE_FUNCTION NAME(const result_view_type& v) : base(v) {}                 \
    KOKKOS_INLINE_FUNCTION void join(value_type& dest, const value_type& src) const {   \
      JOIN;                                                                             \
    }                                                                                   \
    KOKKOS_INLINE_FUNCTION void init(value_type& v) const {                             \
      v = reduction_identity<value_type>::IDENT();                                      \
    }                                                                                   \
  };                                                                                    \
  template <class S, class... P>                                                        \
  NAME(const View<S, P...>&)->NAME<S, typename View<S, P...>::memory_space>;

KKMX_SIMPLE_REDUCER(Sum, dest += src, sum)
KKMX_SIMPLE_REDUCER(Prod, dest *= src, prod)
KKMX_SIMPLE_REDUCER(Min, if (src < dest) dest = src, min)
KKMX_SIMPLE_REDUCER(Max, if (src > dest) dest = src, max)
KKMX_SIMPLE_REDUCER(LAnd, dest = dest && src, land)
KKMX_SIMPLE_REDUCER(LOr, dest = dest || src, lor)
KKMX_SIMPLE_REDUCER(BAnd, dest = dest & src, band)
KKMX_SIMPLE_REDUCER(BOr, dest = dest | src, bor)
#undef KKMX_SIMPLE_REDUCER

template <class Scalar, class Index>
struct ValLocScalar {
  Scalar val;
  Index loc;
};
template <class Scalar>
struct MinMaxScalar {
  Scalar min_val;
  Scalar max_val;
};
template <class Scalar, class Index>
struct MinMaxLocScalar {
  Scalar min_val;
  Scalar max_val;
  Index min_loc;
  Index max_loc;
};
template <class Index>
struct FirstLocScalar {
  Index min_loc_true;
};
template <class Index>
struct LastLocScalar {
  Index max_loc_true;
};

#define KKMX_LOC_REDUCER_HEAD(NAME)                                           \
  using reducer          = NAME;                                              \
  using result_view_type = View<value_type, Space>;                           \
  using base             = Impl::ReducerStorage<value_type, Space>;           \
  KOKKOS_INLINE_FUNCTION NAME(value_type& v) : base(v) {}                     \
  KOKKOS_INLINE_FUNCTION NAME(const result_view_type& v) : base(v) {}

template <class Scalar, class Index, class Space = HostSpace>
struct MinLoc : Impl::ReducerStorage<ValLocScalar<std::remove_cv_t<Scalar>, std::remove_cv_t<Index>>, Space> {
  using scalar_type = std::remove_cv_t<Scalar>;
  using index_type  = std::remove_cv_t<Index>;
  using value_type = ValLocScalar<scalar_type, index_type>;
  KKMX_LOC_REDUCER_HEAD(MinLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (s.val < d.val) d = s;
    else if (s.val == d.val && d.loc == reduction_identity<index_type>::min()) d.loc = s.loc;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.val = reduction_identity<scalar_type>::min();
    v.loc = reduction_identity<index_type>::min();
  }
};
template <class Scalar, class Index, class Space = HostSpace>
struct MaxLoc : Impl::ReducerStorage<ValLocScalar<std::remove_cv_t<Scalar>, std::remove_cv_t<Index>>, Space> {
  using scalar_type = std::remove_cv_t<Scalar>;
  using index_type  = std::remove_cv_t<Index>;
  using value_type = ValLocScalar<scalar_type, index_type>;
  KKMX_LOC_REDUCER_HEAD(MaxLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (s.val > d.val) d = s;
    else if (s.val == d.val && d.loc == reduction_identity<index_type>::min()) d.loc = s.loc;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.val = reduction_identity<scalar_type>::max();
    v.loc = reduction_identity<index_type>::min();
  }
};
template <class Scalar, class Space = HostSpace>
struct MinMax : Impl::ReducerStorage<MinMaxScalar<std::remove_cv_t<Scalar>>, Space> {
  using scalar_type = std::remove_cv_t<Scalar>;
  using value_type = MinMaxScalar<scalar_type>;
  KKMX_LOC_REDUCER_HEAD(MinMax)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (s.min_val < d.min_val) d.min_val = s.min_val;
    if (s.max_val > d.max_val) d.max_val = s.max_val;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.max_val = reduction_identity<scalar_type>::max();
    v.min_val = reduction_identity<scalar_type>::min();
  }
};
template <class Scalar, class Index, class Space = HostSpace>
struct MinMaxLoc : Impl::ReducerStorage<MinMaxLocScalar<std::remove_cv_t<Scalar>, std::remove_cv_t<Index>>, Space> {
  using scalar_type = std::remove_cv_t<Scalar>;
  using index_type  = std::remove_cv_t<Index>;
  using value_type = MinMaxLocScalar<scalar_type, index_type>;
  KKMX_LOC_REDUCER_HEAD(MinMaxLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (s.min_val < d.min_val) { d.min_val = s.min_val; d.min_loc = s.min_loc; }
    else if (d.min_val == s.min_val && d.min_loc == reduction_identity<index_type>::min()) d.min_loc = s.min_loc;
    if (s.max_val > d.max_val) { d.max_val = s.max_val; d.max_loc = s.max_loc; }
    else if (d.max_val == s.max_val && d.max_loc == reduction_identity<index_type>::min()) d.max_loc = s.max_loc;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.max_val = reduction_identity<scalar_type>::max();
    v.min_val = reduction_identity<scalar_type>::min();
    v.max_loc = reduction_identity<index_type>::min();
    v.min_loc = reduction_identity<index_type>::min();
  }
};
template <class Scalar, class Index, class Space = HostSpace>
struct MaxFirstLoc : Impl::ReducerStorage<ValLocScalar<std::remove_cv_t<Scalar>, std::remove_cv_t<Index>>, Space> {
  using scalar_type = std::remove_cv_t<Scalar>;
  using index_type  = std::remove_cv_t<Index>;
  using value_type = ValLocScalar<scalar_type, index_type>;
  KKMX_LOC_REDUCER_HEAD(MaxFirstLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (d.val < s.val) d = s;
    else if (!(s.val < d.val)) d.loc = (s.loc < d.loc) ? s.loc : d.loc;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.val = reduction_identity<scalar_type>::max();
    v.loc = reduction_identity<index_type>::min();
  }
};
template <class Scalar, class Index, class Space = HostSpace>
struct MinFirstLoc : Impl::ReducerStorage<ValLocScalar<std::remove_cv_t<Scalar>, std::remove_cv_t<Index>>, Space> {
  using scalar_type = std::remove_cv_t<Scalar>;
  using index_type  = std::remove_cv_t<Index>;
  using value_type = ValLocScalar<scalar_type, index_type>;
  KKMX_LOC_REDUCER_HEAD(MinFirstLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (s.val < d.val) d = s;
    else if (!(d.val < s.val)) d.loc = (s.loc < d.loc) ? s.loc : d.loc;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.val = reduction_identity<scalar_type>::min();
    v.loc = reduction_identity<index_type>::min();
  }
};
// Minimum with its first location and maximum with its last location
// (core/src/Kokkos_Parallel_Reduce.hpp:946-1010).
template <class Scalar, class Index, class Space = HostSpace>
struct MinMaxFirstLastLoc
    : Impl::ReducerStorage<MinMaxLocScalar<std::remove_cv_t<Scalar>, std::remove_cv_t<Index>>, Space> {
  using scalar_type = std::remove_cv_t<Scalar>;
  using index_type  = std::remove_cv_t<Index>;
  using value_type  = MinMaxLocScalar<scalar_type, index_type>;
  KKMX_LOC_REDUCER_HEAD(MinMaxFirstLastLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    if (s.min_val < d.min_val) { d.min_val = s.min_val; d.min_loc = s.min_loc; }
    else if (!(d.min_val < s.min_val)) d.min_loc = (s.min_loc < d.min_loc) ? s.min_loc : d.min_loc;
    if (d.max_val < s.max_val) { d.max_val = s.max_val; d.max_loc = s.max_loc; }
    else if (!(s.max_val < d.max_val)) d.max_loc = (s.max_loc > d.max_loc) ? s.max_loc : d.max_loc;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const {
    v.max_val = reduction_identity<scalar_type>::max();
    v.min_val = reduction_identity<scalar_type>::min();
    v.max_loc = reduction_identity<index_type>::max();
    v.min_loc = reduction_identity<index_type>::min();
  }
};
template <class Index, class Space = HostSpace>
struct FirstLoc : Impl::ReducerStorage<FirstLocScalar<std::remove_cv_t<Index>>, Space> {
  using index_type = std::remove_cv_t<Index>;
  using value_type = FirstLocScalar<index_type>;
  KKMX_LOC_REDUCER_HEAD(FirstLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    d.min_loc_true = (s.min_loc_true < d.min_loc_true) ? s.min_loc_true : d.min_loc_true;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const { v.min_loc_true = reduction_identity<index_type>::min(); }
};
template <class Index, class Space = HostSpace>
struct LastLoc : Impl::ReducerStorage<LastLocScalar<std::remove_cv_t<Index>>, Space> {
  using index_type = std::remove_cv_t<Index>;
  using value_type = LastLocScalar<index_type>;
  KKMX_LOC_REDUCER_HEAD(LastLoc)
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const {
    d.max_loc_true = (s.max_loc_true > d.max_loc_true) ? s.max_loc_true : d.max_loc_true;
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const { v.max_loc_true = reduction_identity<index_type>::max(); }
};
#undef KKMX_LOC_REDUCER_HEAD

template <class T, class = void>
struct is_reducer : std::false_type {};
template <class T>
struct is_reducer<T, std::enable_if_t<std::is_same<T, typename T::reducer>::value>> : std::true_type {};
template <class T>
inline constexpr bool is_reducer_v = is_reducer<T>::value;

}  // namespace Kokkos

#endif
