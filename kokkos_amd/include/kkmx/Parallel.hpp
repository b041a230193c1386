// parallel_for / parallel_reduce / parallel_scan front end.
//
// Overload set and result rules follow core/src/Kokkos_Parallel.hpp:130-173,348-454
// and core/src/Kokkos_Parallel_Reduce.hpp:1495-1833 of the reference:
//  * the policy may be an execution policy or an integral count (then a
//    RangePolicy on the functor's execution_space, or the default space);
//  * a reduction result may be a reducer, a View (asynchronous), a host
//    scalar or a host pointer for runtime-length values (these fence);
//  * parallel_scan may return the total into a scalar (fence) or a View.
// Each call brackets the dispatch with the Tools begin/end hooks.
#ifndef KKMX_PARALLEL_HPP
#define KKMX_PARALLEL_HPP

#include <kkmx/HostExec.hpp>
#include <kkmx/KernelRange.hpp>
#include <kkmx/Tools.hpp>

namespace Kokkos {

template <class T, class = void>
struct is_execution_policy : std::false_type {};
template <class T>
struct is_execution_policy<T, std::void_t<typename T::execution_policy>> : std::true_type {};
template <class T>
inline constexpr bool is_execution_policy_v = is_execution_policy<T>::value;

namespace Impl {

template <class F, class = void>
struct functor_exec_space {
  using type = DefaultExecutionSpace;
};
template <class F>
struct functor_exec_space<F, std::void_t<typename F::execution_space>> {
  using type = typename F::execution_space;
};

template <class F, class P>
auto to_policy(const P& p) {
  if constexpr (is_execution_policy<P>::value) {
    return p;
  } else {
    static_assert(std::is_integral<P>::value, "parallel dispatch needs a policy or an integral count");
    using Exec = typename functor_exec_space<F>::type;
    return RangePolicy<Exec>(0, p);
  }
}

// Tools bracket (cheap when no tool and no tuning).
template <class Exec>
struct ToolsKernel {
  uint64_t kid = 0;
  bool active;
  int kind;  // 0 for, 1 reduce, 2 scan
  ToolsKernel(const std::string& label, const Exec& space, int k)
      : active(Tools::profileLibraryLoaded() || Kokkos::Impl::tune_internals()), kind(k) {
    if (!active) return;
    const std::string& name = label.empty() ? default_label() : label;
    const uint32_t dev      = Tools::Experimental::device_id(space);
    if (kind == 0) Tools::beginParallelFor(name, dev, &kid);
    else if (kind == 1) Tools::beginParallelReduce(name, dev, &kid);
    else Tools::beginParallelScan(name, dev, &kid);
  }
  ~ToolsKernel() {
    if (!active) return;
    if (kind == 0) Tools::endParallelFor(kid);
    else if (kind == 1) Tools::endParallelReduce(kid);
    else Tools::endParallelScan(kid);
  }
  static const std::string& default_label() {
    static const std::string s("Kokkos::parallel_kernel");
    return s;
  }
};

// Result descriptor for parallel_reduce.
template <class R, class = void>
struct ReduceReturn;

// Reducer.
template <class R>
struct ReduceReturn<R, std::enable_if_t<is_reducer<remove_cvref_t<R>>::value>> {
  using reducer    = remove_cvref_t<R>;
  using value_type = typename reducer::value_type;
  using mem        = typename reducer::result_view_type::memory_space;
  static value_type* ptr(const reducer& r) { return r.view().data(); }
  static bool fences(const reducer& r) { return r.references_scalar(); }
  static unsigned count(const reducer&) { return 1; }
};
// View.
template <class R>
struct ReduceReturn<R, std::enable_if_t<is_view<remove_cvref_t<R>>::value>> {
  using view       = remove_cvref_t<R>;
  using reducer    = void;
  using value_type = typename view::non_const_value_type;
  using mem        = typename view::memory_space;
  static value_type* ptr(const view& v) { return const_cast<value_type*>(v.data()); }
  static bool fences(const view&) { return false; }
  static unsigned count(const view& v) { return static_cast<unsigned>(v.size()); }
};
// Host pointer (runtime-length values).
template <class R>
struct ReduceReturn<R, std::enable_if_t<std::is_pointer<remove_cvref_t<R>>::value>> {
  using reducer    = void;
  using value_type = std::remove_pointer_t<remove_cvref_t<R>>;
  using mem        = HostSpace;
  static value_type* ptr(value_type* p) { return p; }
  static bool fences(value_type*) { return true; }
  static unsigned count(value_type*) { return 0; }
};
// Host array (runtime-length values written into a C array, which the
// reference receives as a decayed pointer).
template <class R>
struct ReduceReturn<R, std::enable_if_t<std::is_array<remove_cvref_t<R>>::value>> {
  using reducer    = void;
  using value_type = std::remove_extent_t<remove_cvref_t<R>>;
  using mem        = HostSpace;
  static value_type* ptr(value_type* p) { return p; }
  static bool fences(value_type*) { return true; }
  static unsigned count(value_type*) { return unsigned(std::extent<remove_cvref_t<R>>::value); }
};
// Host scalar / struct.
template <class R>
struct ReduceReturn<R, std::enable_if_t<!is_reducer<remove_cvref_t<R>>::value &&
                                        !is_view<remove_cvref_t<R>>::value &&
                                        !std::is_pointer<remove_cvref_t<R>>::value &&
                                        !std::is_array<remove_cvref_t<R>>::value>> {
  using reducer    = void;
  using value_type = remove_cvref_t<R>;
  using mem        = HostSpace;
  static value_type* ptr(value_type& v) { return &v; }
  static bool fences(const value_type&) { return true; }
  static unsigned count(const value_type&) { return 1; }
};

template <class Policy, class F, class Ret>
void reduce_dispatch(const std::string& label, const Policy& policy, const F& functor, Ret&& ret) {
  using RR      = ReduceReturn<Ret>;
  using Exec    = typename Policy::execution_space;
  using Tag     = typename Policy::work_tag;
  using Reducer = typename RR::reducer;
  using VT      = typename deduce_value_type<F, Reducer, typename RR::value_type, 0>::type;
  using CR      = CombinedReducer<F, Tag, Reducer, VT>;
  using Store   = typename CR::reducer_store;
  unsigned count = 1;
  if constexpr (std::is_array<VT>::value) count = functor_value_count<F>::get(functor);
  Store store = [&]() -> Store {
    if constexpr (std::is_void<Reducer>::value) return Store{};
    else return Store(ret);
  }();
  CR cr(functor, store, count);
  auto* p = RR::ptr(ret);
  using V = typename CR::value_type;
  static_assert(std::is_same<std::remove_cv_t<typename RR::value_type>, V>::value ||
                    std::is_convertible<V, typename RR::value_type>::value,
                "parallel_reduce: result type does not match the functor value_type");
  ToolsKernel<Exec> tk(label, policy.space(), 1);
  if constexpr (std::is_same<std::remove_cv_t<typename RR::value_type>, V>::value) {
    ParallelReduce<CR, Policy, Exec> closure(cr, policy, reinterpret_cast<V*>(p), typename RR::mem{});
    closure.execute();
  } else {
    // Result of a different (convertible) scalar type: reduce into a V.
    V tmp{};
    ParallelReduce<CR, Policy, Exec> closure(cr, policy, &tmp, HostSpace{});
    closure.execute();
    *p = static_cast<typename RR::value_type>(tmp);
  }
}

}  // namespace Impl

// ------------------------------------------------------------------ parallel_for
template <class Policy, class F>
inline void parallel_for(const std::string& label, const Policy& policy_in, const F& functor) {
  auto policy = Impl::to_policy<F>(policy_in);
  using P     = decltype(policy);
  using Exec  = typename P::execution_space;
  Impl::ToolsKernel<Exec> tk(label, policy.space(), 0);
  Impl::TrackingDisabledScope nt;
  Impl::ParallelFor<F, P, Exec> closure(functor, policy);
  closure.execute();
}
template <class Policy, class F,
          class = std::enable_if_t<is_execution_policy<Policy>::value || std::is_integral<Policy>::value>>
inline void parallel_for(const Policy& policy, const F& functor) {
  parallel_for(std::string(), policy, functor);
}

// ------------------------------------------------------------------ parallel_reduce
template <class Policy, class F, class Ret>
inline std::enable_if_t<!std::is_same<Impl::remove_cvref_t<Policy>, std::string>::value &&
                        !std::is_convertible<Policy, std::string>::value>
parallel_reduce(const std::string& label, const Policy& policy_in, const F& functor, Ret&& ret) {
  auto policy = Impl::to_policy<F>(policy_in);
  Impl::TrackingDisabledScope nt;
  Impl::reduce_dispatch(label, policy, functor, ret);
}
template <class Policy, class F, class Ret,
          class = std::enable_if_t<is_execution_policy<Policy>::value || std::is_integral<Policy>::value>>
inline void parallel_reduce(const Policy& policy, const F& functor, Ret&& ret) {
  parallel_reduce(std::string(), policy, functor, static_cast<Ret&&>(ret));
}
// No result: the functor's value_type decides, the value is discarded.
template <class Policy, class F>
inline void parallel_reduce(const std::string& label, const Policy& policy_in, const F& functor) {
  using VT = typename Impl::functor_value_type<F>::type;
  static_assert(!std::is_void<VT>::value, "parallel_reduce without a result needs a functor value_type");
  if constexpr (std::is_array<VT>::value) {
    std::vector<std::remove_extent_t<VT>> tmp(Impl::functor_value_count<F>::get(functor));
    parallel_reduce(label, policy_in, functor, tmp.data());
  } else {
    VT tmp{};
    parallel_reduce(label, policy_in, functor, tmp);
  }
}
template <class Policy, class F,
          class = std::enable_if_t<is_execution_policy<Policy>::value || std::is_integral<Policy>::value>>
inline void parallel_reduce(const Policy& policy, const F& functor) {
  parallel_reduce(std::string(), policy, functor);
}

// ------------------------------------------------------------------ several results
// parallel_reduce(label, policy, f, r1, r2, ...): f's operator() takes one
// value reference per result (core/src/impl/Kokkos_CombinedReducer.hpp).
// Each result keeps its own join: a reducer's own, Sum for a scalar or a
// View.  The values travel as one struct through the ordinary reduction
// kernels and are written back to each result after a fence.
namespace Impl {

template <class... Ts>
struct CombinedValue;
template <>
struct CombinedValue<> {};
template <class T, class... Ts>
struct CombinedValue<T, Ts...> {
  T head;
  CombinedValue<Ts...> tail;
};
template <size_t I, class CV>
KOKKOS_FORCEINLINE_FUNCTION auto& cv_get(CV& v) {
  if constexpr (I == 0) return v.head;
  else return cv_get<I - 1>(v.tail);
}
template <size_t I, class CV>
KOKKOS_FORCEINLINE_FUNCTION const auto& cv_get(const CV& v) {
  if constexpr (I == 0) return v.head;
  else return cv_get<I - 1>(v.tail);
}

template <class R, class = void>
struct result_reducer {  // host scalar
  using type = Sum<remove_cvref_t<R>, HostSpace>;
};
template <class R>
struct result_reducer<R, std::enable_if_t<is_view<remove_cvref_t<R>>::value>> {
  using V    = remove_cvref_t<R>;
  using type = Sum<typename V::non_const_value_type, typename V::memory_space>;
};
template <class R>
struct result_reducer<R, std::enable_if_t<is_reducer<remove_cvref_t<R>>::value>> {
  using type = remove_cvref_t<R>;
};

template <class F, class... Red>
struct CombinedResultFunctor {
  using value_type = CombinedValue<typename Red::value_type...>;
  using Seq        = std::index_sequence_for<Red...>;
  F f;
  CombinedValue<Red...> reducers;

  template <size_t... I>
  KOKKOS_FORCEINLINE_FUNCTION void init_(value_type& v, std::index_sequence<I...>) const {
    (cv_get<I>(reducers).init(cv_get<I>(v)), ...);
  }
  template <size_t... I>
  KOKKOS_FORCEINLINE_FUNCTION void join_(value_type& d, const value_type& s, std::index_sequence<I...>) const {
    (cv_get<I>(reducers).join(cv_get<I>(d), cv_get<I>(s)), ...);
  }
  template <size_t... I, class... A>
  KOKKOS_FORCEINLINE_FUNCTION void call_(value_type& v, std::index_sequence<I...>, const A&... a) const {
    f(a..., cv_get<I>(v)...);
  }
  KOKKOS_INLINE_FUNCTION void init(value_type& v) const { init_(v, Seq{}); }
  KOKKOS_INLINE_FUNCTION void join(value_type& d, const value_type& s) const { join_(d, s, Seq{}); }

  template <class A0>
  KOKKOS_INLINE_FUNCTION void operator()(const A0& a0, value_type& v) const { call_(v, Seq{}, a0); }
  template <class A0, class A1>
  KOKKOS_INLINE_FUNCTION void operator()(const A0& a0, const A1& a1, value_type& v) const {
    call_(v, Seq{}, a0, a1);
  }
  template <class A0, class A1, class A2>
  KOKKOS_INLINE_FUNCTION void operator()(const A0& a0, const A1& a1, const A2& a2, value_type& v) const {
    call_(v, Seq{}, a0, a1, a2);
  }
  template <class A0, class A1, class A2, class A3>
  KOKKOS_INLINE_FUNCTION void operator()(const A0& a0, const A1& a1, const A2& a2, const A3& a3,
                                         value_type& v) const {
    call_(v, Seq{}, a0, a1, a2, a3);
  }
  template <class A0, class A1, class A2, class A3, class A4>
  KOKKOS_INLINE_FUNCTION void operator()(const A0& a0, const A1& a1, const A2& a2, const A3& a3, const A4& a4,
                                         value_type& v) const {
    call_(v, Seq{}, a0, a1, a2, a3, a4);
  }
  template <class A0, class A1, class A2, class A3, class A4, class A5>
  KOKKOS_INLINE_FUNCTION void operator()(const A0& a0, const A1& a1, const A2& a2, const A3& a3, const A4& a4,
                                         const A5& a5, value_type& v) const {
    call_(v, Seq{}, a0, a1, a2, a3, a4, a5);
  }
  template <class A0, class A1, class A2, class A3, class A4, class A5, class A6>
  KOKKOS_INLINE_FUNCTION void operator()(const A0& a0, const A1& a1, const A2& a2, const A3& a3, const A4& a4,
                                         const A5& a5, const A6& a6, value_type& v) const {
    call_(v, Seq{}, a0, a1, a2, a3, a4, a5, a6);
  }
};

template <class Red, class R>
Red make_result_reducer(R& r) {
  if constexpr (is_reducer<remove_cvref_t<R>>::value) return r;
  else return Red(r);
}
template <class T, class R>
void write_back_result(const T& value, R& r) {
  T* dst;
  if constexpr (is_reducer<remove_cvref_t<R>>::value) dst = r.view().data();
  else if constexpr (is_view<remove_cvref_t<R>>::value) dst = const_cast<T*>(r.data());
  else dst = &r;
  KKMX_SAFE_CALL(mx_memcpy(dst, &value, sizeof(T)));
}

inline CombinedValue<> make_combined_reducers() { return {}; }
template <class R, class... Rest>
auto make_combined_reducers(R& r, Rest&... rest) {
  using Red = typename result_reducer<R>::type;
  return CombinedValue<Red, typename result_reducer<Rest>::type...>{make_result_reducer<Red>(r),
                                                                     make_combined_reducers(rest...)};
}
template <class CV, size_t... I, class... R>
void write_back_results(const CV& v, std::index_sequence<I...>, R&... rs) {
  (write_back_result(cv_get<I>(v), rs), ...);
}

template <class Policy, class F, class... R>
void combined_reduce(const std::string& label, const Policy& policy, const F& f, R&... rs) {
  using CF = CombinedResultFunctor<F, typename result_reducer<R>::type...>;
  CF cf{f, make_combined_reducers(rs...)};
  typename CF::value_type result;
  reduce_dispatch(label, policy, cf, result);
  write_back_results(result, std::index_sequence_for<R...>{}, rs...);
}

}  // namespace Impl

template <class Policy, class F, class R1, class R2, class... Rs>
inline std::enable_if_t<!std::is_convertible<Policy, std::string>::value> parallel_reduce(
    const std::string& label, const Policy& policy_in, const F& functor, R1&& r1, R2&& r2, Rs&&... rs) {
  auto policy = Impl::to_policy<F>(policy_in);
  Impl::TrackingDisabledScope nt;
  Impl::combined_reduce(label, policy, functor, r1, r2, rs...);
}
template <class Policy, class F, class R1, class R2, class... Rs,
          class = std::enable_if_t<is_execution_policy<Policy>::value || std::is_integral<Policy>::value>>
inline void parallel_reduce(const Policy& policy, const F& functor, R1&& r1, R2&& r2, Rs&&... rs) {
  parallel_reduce(std::string(), policy, functor, r1, r2, rs...);
}

// ------------------------------------------------------------------ parallel_scan
namespace Impl {
template <class Policy, class F, class Hint>
struct ScanTypes {
  using Tag = typename Policy::work_tag;
  using VT  = typename deduce_value_type<F, void, Hint, 1>::type;
  using CR  = CombinedReducer<F, Tag, void, VT>;
};
}  // namespace Impl

template <class Policy, class F>
inline void parallel_scan(const std::string& label, const Policy& policy_in, const F& functor) {
  auto policy = Impl::to_policy<F>(policy_in);
  using P     = decltype(policy);
  using Exec  = typename P::execution_space;
  using CR    = typename Impl::ScanTypes<P, F, void>::CR;
  static_assert(!std::is_void<typename CR::value_type>::value, "parallel_scan: cannot deduce value_type");
  Impl::ToolsKernel<Exec> tk(label, policy.space(), 2);
  Impl::TrackingDisabledScope nt;
  CR cr(functor, typename CR::reducer_store{}, 1);
  Impl::ParallelScan<CR, P, Exec> closure(cr, policy);
  closure.execute();
}
template <class Policy, class F,
          class = std::enable_if_t<is_execution_policy<Policy>::value || std::is_integral<Policy>::value>>
inline void parallel_scan(const Policy& policy, const F& functor) {
  parallel_scan(std::string(), policy, functor);
}

template <class Policy, class F, class Ret>
inline void parallel_scan(const std::string& label, const Policy& policy_in, const F& functor, Ret& ret) {
  auto policy = Impl::to_policy<F>(policy_in);
  using P     = decltype(policy);
  using Exec  = typename P::execution_space;
  Impl::ToolsKernel<Exec> tk(label, policy.space(), 2);
  Impl::TrackingDisabledScope nt;
  if constexpr (is_view<Ret>::value) {
    using CR = typename Impl::ScanTypes<P, F, typename Ret::non_const_value_type>::CR;
    using V  = typename CR::value_type;
    CR cr(functor, typename CR::reducer_store{}, 1);
    constexpr bool on_dev = Impl::exec_can_access<Exec, typename Ret::memory_space>::value;
    Impl::ParallelScanWithTotal<CR, P, Ret, Exec> closure(cr, policy, reinterpret_cast<V*>(ret.data()), on_dev);
    closure.execute();
  } else {
    using CR = typename Impl::ScanTypes<P, F, Ret>::CR;
    using V  = typename CR::value_type;
    CR cr(functor, typename CR::reducer_store{}, 1);
    V tmp{};
    Impl::ParallelScanWithTotal<CR, P, V, Exec> closure(cr, policy, &tmp, false);
    closure.execute();
    ret = static_cast<Ret>(tmp);
  }
}
template <class Policy, class F, class Ret,
          class = std::enable_if_t<is_execution_policy<Policy>::value || std::is_integral<Policy>::value>>
inline void parallel_scan(const Policy& policy, const F& functor, Ret& ret) {
  parallel_scan(std::string(), policy, functor, ret);
}

}  // namespace Kokkos

#endif
