// Functor analysis: work-tag dispatch, reduction/scan value-type deduction and
// the init / join / final adaptor used by every reduce and scan kernel.
//
// Reference semantics (core/src/impl/Kokkos_FunctorAnalysis.hpp:865-960 and
// core/src/Kokkos_Parallel_Reduce.hpp:1356-1394, CombinedFunctorReducer):
//  * value_type is the reducer's value_type when a reducer is given, else the
//    functor's `value_type` typedef (which may be a runtime array `T[]` with a
//    `value_count` member), else the type of the last (reduce) or
//    second-to-last (scan) argument of operator(), else the result type;
//  * init: reducer.init, else functor.init([tag,] v), else value-initialise;
//  * join: reducer.join, else functor.join([tag,] dst, src), else dst += src;
//  * final: reducer.final / functor.final([tag,] v) when present.
#ifndef KKMX_FUNCTOR_ANALYSIS_HPP
#define KKMX_FUNCTOR_ANALYSIS_HPP

#include <kkmx/Reducers.hpp>

#include <tuple>

namespace Kokkos {
namespace Impl {

// ------------------------------------------------------------------ tag calls
template <class Tag, class F, class... Args>
KOKKOS_FORCEINLINE_FUNCTION void call_functor(const F& f, Args&&... args) {
  if constexpr (std::is_void<Tag>::value)
    f(static_cast<Args&&>(args)...);
  else
    f(Tag{}, static_cast<Args&&>(args)...);
}

// ------------------------------------------------------------------ operator() introspection
template <class M>
struct memfn_args {
  using type = void;
};
template <class C, class R, class... A>
struct memfn_args<R (C::*)(A...) const> {
  using type = std::tuple<A...>;
};
template <class C, class R, class... A>
struct memfn_args<R (C::*)(A...)> {
  using type = std::tuple<A...>;
};
template <class F, class = void>
struct op_args {
  using type = void;
};
template <class F>
struct op_args<F, std::void_t<decltype(&F::operator())>> {
  using type = typename memfn_args<decltype(&F::operator())>::type;
};
// Argument counted from the back (0 = last).
template <class Tuple, size_t FromBack, class = void>
struct arg_from_back {
  using type = void;
};
template <class... A, size_t K>
struct arg_from_back<std::tuple<A...>, K, std::enable_if_t<(sizeof...(A) > K)>> {
  using type = remove_cvref_t<std::tuple_element_t<sizeof...(A) - 1 - K, std::tuple<A...>>>;
};

template <class F, class = void>
struct functor_value_type {
  using type = void;
};
template <class F>
struct functor_value_type<F, std::void_t<typename F::value_type>> {
  using type = typename F::value_type;
};

// Value type of a reduce (FromBack = 0) or a scan (FromBack = 1).
template <class F, class Reducer, class Hint, size_t FromBack>
struct deduce_value_type {
  using from_functor = typename functor_value_type<F>::type;
  using from_op      = typename arg_from_back<typename op_args<F>::type, FromBack>::type;
  using type         = std::conditional_t<
      !std::is_void<Reducer>::value, typename functor_value_type<Reducer>::type,
      std::conditional_t<!std::is_void<from_functor>::value, from_functor,
                         std::conditional_t<!std::is_void<from_op>::value, from_op, Hint>>>;
};

// ------------------------------------------------------------------ detection of init/join/final
#define KKMX_DETECT(NAME, EXPR)                                                    \
  template <class F, class Tag, class V, class = void>                             \
  struct NAME : std::false_type {};                                                \
  template <class F, class Tag, class V>                                           \
  struct NAME<F, Tag, V, std::void_t<decltype(EXPR)>> : std::true_type {};

KKMX_DETECT(has_init_tag, std::declval<const F&>().init(std::declval<Tag>(), std::declval<V&>()))
KKMX_DETECT(has_init, std::declval<const F&>().init(std::declval<V&>()))
KKMX_DETECT(has_join_tag, std::declval<const F&>().join(std::declval<Tag>(), std::declval<V&>(), std::declval<const V&>()))
KKMX_DETECT(has_join, std::declval<const F&>().join(std::declval<V&>(), std::declval<const V&>()))
KKMX_DETECT(has_final_tag, std::declval<const F&>().final(std::declval<Tag>(), std::declval<V&>()))
KKMX_DETECT(has_final, std::declval<const F&>().final(std::declval<V&>()))
// Array (T[]) forms take pointers.
KKMX_DETECT(has_init_arr_tag, std::declval<const F&>().init(std::declval<Tag>(), std::declval<V*>()))
KKMX_DETECT(has_init_arr, std::declval<const F&>().init(std::declval<V*>()))
KKMX_DETECT(has_join_arr_tag, std::declval<const F&>().join(std::declval<Tag>(), std::declval<V*>(), std::declval<const V*>()))
KKMX_DETECT(has_join_arr, std::declval<const F&>().join(std::declval<V*>(), std::declval<const V*>()))
KKMX_DETECT(has_final_arr_tag, std::declval<const F&>().final(std::declval<Tag>(), std::declval<V*>()))
KKMX_DETECT(has_final_arr, std::declval<const F&>().final(std::declval<V*>()))
#undef KKMX_DETECT

struct NoTag {};
template <class Tag>
using tag_or_notag = std::conditional_t<std::is_void<Tag>::value, NoTag, Tag>;

struct EmptyReducer {};

// ------------------------------------------------------------------ CombinedReducer
// Functor + optional reducer, with the value type resolved.  Scalar values
// (including structs) are handled by reference; array values by pointer with a
// runtime value_count.
template <class F, class Tag, class Reducer, class ValueType>
class CombinedReducer {
 public:
  using functor_type  = F;
  using reducer_type  = Reducer;
  using work_tag      = Tag;
  static constexpr bool is_array = std::is_array<ValueType>::value;
  using value_type    = std::remove_cv_t<std::remove_extent_t<ValueType>>;
  using reducer_store = std::conditional_t<std::is_void<Reducer>::value, EmptyReducer, Reducer>;
  using T             = tag_or_notag<Tag>;

  KOKKOS_INLINE_FUNCTION CombinedReducer(const F& f, const reducer_store& r, unsigned count)
      : m_functor(f), m_reducer(r), m_count(count) {}

  KOKKOS_FORCEINLINE_FUNCTION const F& functor() const { return m_functor; }
  KOKKOS_FORCEINLINE_FUNCTION const reducer_store& reducer() const { return m_reducer; }
  KOKKOS_FORCEINLINE_FUNCTION unsigned value_count() const { return is_array ? m_count : 1u; }
  KOKKOS_FORCEINLINE_FUNCTION size_t value_size() const { return sizeof(value_type) * value_count(); }

  // -------- scalar forms
  KOKKOS_FORCEINLINE_FUNCTION void init(value_type& v) const {
    if constexpr (!std::is_void<Reducer>::value) {
      m_reducer.init(v);
    } else if constexpr (!std::is_void<Tag>::value && has_init_tag<F, Tag, value_type>::value) {
      m_functor.init(Tag{}, v);
    } else if constexpr (has_init<F, T, value_type>::value) {
      m_functor.init(v);
    } else {
      v = value_type();
    }
  }
  KOKKOS_FORCEINLINE_FUNCTION void join(value_type& dst, const value_type& src) const {
    if constexpr (!std::is_void<Reducer>::value) {
      m_reducer.join(dst, src);
    } else if constexpr (!std::is_void<Tag>::value && has_join_tag<F, Tag, value_type>::value) {
      m_functor.join(Tag{}, dst, src);
    } else if constexpr (has_join<F, T, value_type>::value) {
      m_functor.join(dst, src);
    } else {
      dst += src;
    }
  }
  KOKKOS_FORCEINLINE_FUNCTION void final(value_type& v) const {
    if constexpr (!std::is_void<Reducer>::value) {
      if constexpr (has_final<Reducer, T, value_type>::value) m_reducer.final(v);
    } else if constexpr (!std::is_void<Tag>::value && has_final_tag<F, Tag, value_type>::value) {
      m_functor.final(Tag{}, v);
    } else if constexpr (has_final<F, T, value_type>::value) {
      m_functor.final(v);
    }
  }

  // -------- array forms (value_count entries)
  KOKKOS_FORCEINLINE_FUNCTION void init_array(value_type* v) const {
    if constexpr (!std::is_void<Tag>::value && has_init_arr_tag<F, Tag, value_type>::value) {
      m_functor.init(Tag{}, v);
    } else if constexpr (has_init_arr<F, T, value_type>::value) {
      m_functor.init(v);
    } else {
      for (unsigned k = 0; k < m_count; ++k) v[k] = value_type();
    }
  }
  KOKKOS_FORCEINLINE_FUNCTION void join_array(value_type* dst, const value_type* src) const {
    if constexpr (!std::is_void<Tag>::value && has_join_arr_tag<F, Tag, value_type>::value) {
      m_functor.join(Tag{}, dst, src);
    } else if constexpr (has_join_arr<F, T, value_type>::value) {
      m_functor.join(dst, src);
    } else {
      for (unsigned k = 0; k < m_count; ++k) dst[k] += src[k];
    }
  }
  KOKKOS_FORCEINLINE_FUNCTION void final_array(value_type* v) const {
    if constexpr (!std::is_void<Tag>::value && has_final_arr_tag<F, Tag, value_type>::value) {
      m_functor.final(Tag{}, v);
    } else if constexpr (has_final_arr<F, T, value_type>::value) {
      m_functor.final(v);
    }
  }

 private:
  F m_functor;
  reducer_store m_reducer;
  unsigned m_count;
};

template <class F, class = void>
struct functor_value_count {
  static unsigned get(const F&) { return 1; }
};
template <class F>
struct functor_value_count<F, std::void_t<decltype(std::declval<const F&>().value_count)>> {
  static unsigned get(const F& f) { return static_cast<unsigned>(f.value_count); }
};

}  // namespace Impl
}  // namespace Kokkos

#endif
