// deep_copy, create_mirror, create_mirror_view, create_mirror_view_and_copy,
// realloc/resize.
//
// Reference: core/src/Kokkos_CopyViews.hpp:58-420 (ViewFill / ViewCopy are
// parallel_for kernels; whole-allocation copies of contiguous Views go to the
// runtime memcpy), core/src/HIP/Kokkos_HIP_DeepCopy.cpp:23-51.
// deep_copy without an execution space fences before and after, as in the
// reference; deep_copy(exec, ...) is ordered on exec's stream only.
#ifndef KKMX_COPY_VIEWS_HPP
#define KKMX_COPY_VIEWS_HPP

#include <kkmx/Parallel.hpp>

namespace Kokkos {
namespace Impl {

template <class ViewT>
struct FillFunctor {
  using value_type = typename ViewT::non_const_value_type;
  typename ViewT::value_type* ptr;
  value_type val;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { ptr[i] = val; }
};

template <class DV, class SV>
struct ElementCopy {
  DV dst;
  SV src;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const {
    // Rank-generic element copy through a flat index (LayoutRight order).
    size_t idx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t r     = i;
    for (int d = int(DV::rank) - 1; d >= 0; --d) {
      const size_t e = dst.extent(d);
      idx[d]         = static_cast<size_t>(r % static_cast<int64_t>(e));
      r /= static_cast<int64_t>(e);
    }
    if constexpr (DV::rank == 1) dst(idx[0]) = src(idx[0]);
    else if constexpr (DV::rank == 2) dst(idx[0], idx[1]) = src(idx[0], idx[1]);
    else if constexpr (DV::rank == 3) dst(idx[0], idx[1], idx[2]) = src(idx[0], idx[1], idx[2]);
    else if constexpr (DV::rank == 4) dst(idx[0], idx[1], idx[2], idx[3]) = src(idx[0], idx[1], idx[2], idx[3]);
  }
};

template <class Mem>
constexpr bool host_accessible() {
  return exec_can_access<Serial, Mem>::value;
}
template <class Mem>
constexpr bool device_accessible() {
  return exec_can_access<MI355X, Mem>::value;
}

template <class DV, class SV>
void check_shapes(const DV& dst, const SV& src) {
  static_assert(unsigned(DV::rank) == unsigned(SV::rank), "deep_copy: rank mismatch");
  for (unsigned r = 0; r < DV::rank; ++r)
    if (dst.extent(r) != src.extent(r))
      throw_runtime_exception("Deprecation Error: Kokkos::deep_copy extents of views don't match: " +
                              dst.label() + "(" + std::to_string(dst.extent(r)) + ") " + src.label() + "(" +
                              std::to_string(src.extent(r)) + ")");
}
}  // namespace Impl

// View <- scalar
template <class D, class... P>
inline void deep_copy(const View<D, P...>& dst, const typename View<D, P...>::const_value_type& value) {
  using V   = View<D, P...>;
  using Mem = typename V::memory_space;
  static_assert(!std::is_const<typename V::value_type>::value, "deep_copy into a const View");
  if (dst.span() == 0) return;
  Kokkos::fence("Kokkos::deep_copy: scalar copy, pre view equality check");
  if constexpr (Impl::host_accessible<Mem>()) {
    auto* p = dst.data();
    for (size_t i = 0; i < dst.span(); ++i) p[i] = value;
  } else {
    Impl::FillFunctor<V> f{dst.data(), value};
    parallel_for("Kokkos::ViewFill", RangePolicy<MI355X, IndexType<int64_t>>(0, int64_t(dst.span())), f);
  }
  Kokkos::fence("Kokkos::deep_copy: scalar copy, post copy fence");
}
template <class E, class D, class... P,
          class = std::enable_if_t<is_execution_space<E>::value>>
inline void deep_copy(const E& exec, const View<D, P...>& dst,
                      const typename View<D, P...>::const_value_type& value) {
  using V   = View<D, P...>;
  using Mem = typename V::memory_space;
  if (dst.span() == 0) return;
  if constexpr (Impl::host_accessible<Mem>()) {
    exec.fence();
    auto* p = dst.data();
    for (size_t i = 0; i < dst.span(); ++i) p[i] = value;
  } else {
    Impl::FillFunctor<V> f{dst.data(), value};
    parallel_for("Kokkos::ViewFill", RangePolicy<MI355X, IndexType<int64_t>>(exec, 0, int64_t(dst.span())), f);
  }
}

// scalar <- rank-0 View
template <class T, class S, class... P>
inline std::enable_if_t<!is_view<T>::value> deep_copy(T& dst, const View<S, P...>& src) {
  static_assert(View<S, P...>::rank == 0, "deep_copy(scalar, View) needs a rank-0 View");
  Kokkos::fence("Kokkos::deep_copy: copy into scalar, pre copy fence");
  KKMX_SAFE_CALL(mx_memcpy(&dst, src.data(), sizeof(T)));
}
// scalar <- rank-0 View, ordered on exec's stream; the value is valid after
// exec.fence().
template <class E, class T, class S, class... P,
          class = std::enable_if_t<is_execution_space<E>::value && !is_view<T>::value>>
inline void deep_copy(const E& exec, T& dst, const View<S, P...>& src) {
  static_assert(View<S, P...>::rank == 0, "deep_copy(exec, scalar, View) needs a rank-0 View");
  if constexpr (std::is_same<E, MI355X>::value) {
    KKMX_SAFE_CALL(mx_memcpy_async(&dst, src.data(), sizeof(T), exec.stream()));
  } else {
    exec.fence();
    KKMX_SAFE_CALL(mx_memcpy(&dst, src.data(), sizeof(T)));
  }
}
// rank-0 View <- scalar handled by the scalar fill overload above.

// View <- View
template <class D1, class... P1, class D2, class... P2>
inline void deep_copy(const View<D1, P1...>& dst, const View<D2, P2...>& src) {
  using DV = View<D1, P1...>;
  using SV = View<D2, P2...>;
  static_assert(std::is_same<typename DV::non_const_value_type, typename SV::non_const_value_type>::value,
                "deep_copy requires matching value types");
  if (dst.data() == nullptr && src.data() == nullptr) return;
  Impl::check_shapes(dst, src);
  if (dst.data() == src.data() && dst.span() == src.span()) {
    Kokkos::fence("Kokkos::deep_copy: copy between contiguous views, fence due to same spans");
    return;
  }
  Kokkos::fence("Kokkos::deep_copy: copy between views, pre view equality check");
  const bool same_layout = std::is_same<typename DV::array_layout, typename SV::array_layout>::value || DV::rank <= 1;
  if (same_layout && dst.span_is_contiguous() && src.span_is_contiguous()) {
    uint64_t kid = 0;
    (void)kid;
    KKMX_SAFE_CALL(mx_memcpy(dst.data(), src.data(), dst.span() * sizeof(typename DV::value_type)));
  } else if constexpr (Impl::host_accessible<typename DV::memory_space>() &&
                       Impl::host_accessible<typename SV::memory_space>()) {
    Impl::ElementCopy<DV, SV> f{dst, src};
    for (int64_t i = 0; i < int64_t(dst.size()); ++i) f(i);
  } else if constexpr (Impl::device_accessible<typename DV::memory_space>() &&
                       Impl::device_accessible<typename SV::memory_space>()) {
    Impl::ElementCopy<DV, SV> f{dst, src};
    parallel_for("Kokkos::ViewCopy", RangePolicy<MI355X, IndexType<int64_t>>(0, int64_t(dst.size())), f);
  } else {
    // Different layouts across spaces: stage through a host copy of src in
    // its own layout, then copy element-wise on the host / device.
    typename SV::HostMirror tmp(view_alloc(WithoutInitializing, std::string("deep_copy_tmp")), src.layout());
    deep_copy(tmp, src);
    if constexpr (Impl::host_accessible<typename DV::memory_space>()) {
      Impl::ElementCopy<DV, typename SV::HostMirror> f{dst, tmp};
      for (int64_t i = 0; i < int64_t(dst.size()); ++i) f(i);
    } else {
      typename DV::HostMirror tmp2(view_alloc(WithoutInitializing, std::string("deep_copy_tmp2")), dst.layout());
      Impl::ElementCopy<typename DV::HostMirror, typename SV::HostMirror> f{tmp2, tmp};
      for (int64_t i = 0; i < int64_t(dst.size()); ++i) f(i);
      deep_copy(dst, tmp2);
    }
  }
  Kokkos::fence("Kokkos::deep_copy: copy between views, post deep copy fence");
}

template <class E, class D1, class... P1, class D2, class... P2,
          class = std::enable_if_t<is_execution_space<E>::value>>
inline void deep_copy(const E& exec, const View<D1, P1...>& dst, const View<D2, P2...>& src) {
  using DV = View<D1, P1...>;
  using SV = View<D2, P2...>;
  Impl::check_shapes(dst, src);
  const bool same_layout = std::is_same<typename DV::array_layout, typename SV::array_layout>::value || DV::rank <= 1;
  if (!(same_layout && dst.span_is_contiguous() && src.span_is_contiguous()))
    return deep_copy(dst, src);
  if constexpr (std::is_same<E, MI355X>::value) {
    KKMX_SAFE_CALL(mx_memcpy_async(dst.data(), src.data(), dst.span() * sizeof(typename DV::value_type),
                                   exec.stream()));
  } else {
    exec.fence();
    std::memcpy(static_cast<void*>(dst.data()), src.data(), dst.span() * sizeof(typename DV::value_type));
  }
}

// ------------------------------------------------------------------ mirrors
template <class D, class... P>
inline typename View<D, P...>::HostMirror create_mirror(const View<D, P...>& v) {
  using M = typename View<D, P...>::HostMirror;
  return M(view_alloc(v.label() + "_mirror"), v.layout());
}
template <class D, class... P>
inline typename View<D, P...>::HostMirror create_mirror(Impl::WithoutInitializing_t, const View<D, P...>& v) {
  using M = typename View<D, P...>::HostMirror;
  return M(view_alloc(WithoutInitializing, v.label() + "_mirror"), v.layout());
}
template <class Space, class D, class... P,
          class = std::enable_if_t<is_memory_space<Space>::value || is_execution_space<Space>::value>>
inline auto create_mirror(const Space&, const View<D, P...>& v) {
  using Mem = typename Space::memory_space;
  using M   = View<typename View<D, P...>::traits::non_const_data_type, typename View<D, P...>::array_layout, Mem>;
  return M(view_alloc(v.label() + "_mirror"), v.layout());
}

template <class D, class... P>
inline auto create_mirror_view(const View<D, P...>& v) {
  using V = View<D, P...>;
  if constexpr (Impl::host_accessible<typename V::memory_space>() &&
                std::is_same<typename V::HostMirror::memory_space, typename V::memory_space>::value &&
                std::is_same<typename V::value_type, typename V::non_const_value_type>::value) {
    return typename V::HostMirror(v);
  } else {
    return create_mirror(v);
  }
}
template <class D, class... P>
inline auto create_mirror_view(Impl::WithoutInitializing_t, const View<D, P...>& v) {
  using V = View<D, P...>;
  if constexpr (Impl::host_accessible<typename V::memory_space>() &&
                std::is_same<typename V::HostMirror::memory_space, typename V::memory_space>::value &&
                std::is_same<typename V::value_type, typename V::non_const_value_type>::value) {
    return typename V::HostMirror(v);
  } else {
    return create_mirror(WithoutInitializing, v);
  }
}
template <class Space, class D, class... P,
          class = std::enable_if_t<is_memory_space<Space>::value || is_execution_space<Space>::value>>
inline auto create_mirror_view(const Space& s, const View<D, P...>& v) {
  using Mem = typename Space::memory_space;
  using V   = View<D, P...>;
  using M   = View<typename V::traits::non_const_data_type, typename V::array_layout, Mem>;
  if constexpr (std::is_same<typename V::memory_space, Mem>::value &&
                std::is_same<typename V::value_type, typename V::non_const_value_type>::value) {
    return M(v);
  } else {
    return create_mirror(s, v);
  }
}
template <class Space, class D, class... P,
          class = std::enable_if_t<is_memory_space<Space>::value || is_execution_space<Space>::value>>
inline auto create_mirror_view_and_copy(const Space& s, const View<D, P...>& v) {
  using Mem = typename Space::memory_space;
  using V   = View<D, P...>;
  using M   = View<typename V::traits::non_const_data_type, typename V::array_layout, Mem>;
  if constexpr (std::is_same<typename V::memory_space, Mem>::value &&
                std::is_same<typename V::value_type, typename V::non_const_value_type>::value) {
    Kokkos::fence("Kokkos::create_mirror_view_and_copy: fence before returning src view");
    return M(v);
  } else {
    M m(view_alloc(WithoutInitializing, v.label() + "_mirror"), v.layout());
    deep_copy(m, v);
    return m;
  }
}

// ------------------------------------------------------------------ resize / realloc
template <class D, class... P>
inline void realloc(View<D, P...>& v, size_t n0 = 0, size_t n1 = 0, size_t n2 = 0, size_t n3 = 0) {
  using V     = View<D, P...>;
  std::string l = v.label();
  v           = V();
  v           = V(l, n0, n1, n2, n3);
}
template <class D, class... P>
inline void resize(View<D, P...>& v, size_t n0 = 0, size_t n1 = 0, size_t n2 = 0, size_t n3 = 0) {
  using V = View<D, P...>;
  V nv(v.label(), n0, n1, n2, n3);
  if (v.data()) {
    auto hv  = create_mirror_view_and_copy(HostSpace(), v);
    auto hnv = create_mirror_view(nv);
    for (size_t i = 0; i < hnv.span() && i < hv.span(); ++i) hnv.data()[i] = hv.data()[i];
    deep_copy(nv, hnv);
  }
  v = nv;
}

}  // namespace Kokkos

#endif
