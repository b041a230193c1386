// Kokkos::Experimental::ScatterView for the MI355X execution space and the
// host (Serial) space: scatter-add / product / min / max contributions from
// a parallel loop into a View, with the contribution strategy chosen per
// space.
//
// Reference: containers/src/Kokkos_ScatterView.hpp (ScatterSum / Prod / Max
// / Min, ScatterNonDuplicated / ScatterDuplicated, ScatterNonAtomic /
// ScatterAtomic, ScatterView::access / subview / contribute_into / reset /
// reset_except, ScatterAccess::operator() returning a ScatterValue with
// +=, -=, *=, /= and update, create_scatter_view, contribute).
//
// MI355X: contributions are native device-scope atomics (the reference's
// GPU default, non-duplicated + atomic).  The reference's duplicated mode is
// one copy per host thread; tens of thousands of resident lanes cannot each
// own a copy, so on MI355X ScatterDuplicated (rank 1) keeps D replicas
// (D = 64 MiB / the View's bytes, 1..256) and a workgroup updates replica
// blockIdx % D with atomics.  Updates to one hot element then spread over D
// addresses instead of serialising on one; contribute_into() folds the D
// replicas into the destination in one coalesced pass.  Higher ranks use the
// single atomic copy.  Serial: one thread, plain updates.
#ifndef KKMX_SCATTER_VIEW_HPP
#define KKMX_SCATTER_VIEW_HPP

#include <Kokkos_Core.hpp>

#include <limits>

namespace Kokkos {
namespace Experimental {

struct ScatterSum {};
struct ScatterProd {};
struct ScatterMax {};
struct ScatterMin {};

struct ScatterNonDuplicated {};
struct ScatterDuplicated {};

struct ScatterNonAtomic {};
struct ScatterAtomic {};

}  // namespace Experimental

namespace Impl {
namespace Experimental {

namespace KE = Kokkos::Experimental;

template <class ExecSpace>
struct DefaultDuplication {
  using type = KE::ScatterNonDuplicated;
};
template <class ExecSpace, class Duplication>
struct DefaultContribution {
  using type = std::conditional_t<std::is_same<ExecSpace, MI355X>::value, KE::ScatterAtomic, KE::ScatterNonAtomic>;
};

// Identity of each operation (what reset() writes).
template <class T, class Op>
struct ScatterIdentity;
template <class T>
struct ScatterIdentity<T, KE::ScatterSum> {
  KOKKOS_FUNCTION static T value() { return T(0); }
};
template <class T>
struct ScatterIdentity<T, KE::ScatterProd> {
  KOKKOS_FUNCTION static T value() { return T(1); }
};
template <class T>
struct ScatterIdentity<T, KE::ScatterMax> {
  KOKKOS_FUNCTION static T value() { return reduction_identity<T>::max(); }
};
template <class T>
struct ScatterIdentity<T, KE::ScatterMin> {
  KOKKOS_FUNCTION static T value() { return reduction_identity<T>::min(); }
};

// One contribution target: the proxy ScatterAccess::operator() returns.
template <class T, class Op, bool Atomic>
struct ScatterValue;

template <class T, bool Atomic>
struct ScatterValue<T, KE::ScatterSum, Atomic> {
  T& value;
  KOKKOS_FORCEINLINE_FUNCTION void operator+=(const T& rhs) const { update(rhs); }
  KOKKOS_FORCEINLINE_FUNCTION void operator-=(const T& rhs) const { update(-rhs); }
  KOKKOS_FORCEINLINE_FUNCTION void operator++() const { update(T(1)); }
  KOKKOS_FORCEINLINE_FUNCTION void operator++(int) const { update(T(1)); }
  KOKKOS_FORCEINLINE_FUNCTION void operator--() const { update(T(-1)); }
  KOKKOS_FORCEINLINE_FUNCTION void operator--(int) const { update(T(-1)); }
  KOKKOS_FORCEINLINE_FUNCTION void update(const T& rhs) const {
    if constexpr (Atomic) Kokkos::atomic_add(&value, rhs);
    else value += rhs;
  }
  KOKKOS_FORCEINLINE_FUNCTION void reset() const { value = T(0); }
};
template <class T, bool Atomic>
struct ScatterValue<T, KE::ScatterProd, Atomic> {
  T& value;
  KOKKOS_FORCEINLINE_FUNCTION void operator*=(const T& rhs) const { update(rhs); }
  KOKKOS_FORCEINLINE_FUNCTION void operator/=(const T& rhs) const {
    if constexpr (Atomic) Kokkos::atomic_div(&value, rhs);
    else value /= rhs;
  }
  KOKKOS_FORCEINLINE_FUNCTION void update(const T& rhs) const {
    if constexpr (Atomic) Kokkos::atomic_mul(&value, rhs);
    else value *= rhs;
  }
  KOKKOS_FORCEINLINE_FUNCTION void reset() const { value = T(1); }
};
template <class T, bool Atomic>
struct ScatterValue<T, KE::ScatterMax, Atomic> {
  T& value;
  KOKKOS_FORCEINLINE_FUNCTION void update(const T& rhs) const {
    if constexpr (Atomic) Kokkos::atomic_max(&value, rhs);
    else value = rhs > value ? rhs : value;
  }
  KOKKOS_FORCEINLINE_FUNCTION void reset() const { value = ScatterIdentity<T, KE::ScatterMax>::value(); }
};
template <class T, bool Atomic>
struct ScatterValue<T, KE::ScatterMin, Atomic> {
  T& value;
  KOKKOS_FORCEINLINE_FUNCTION void update(const T& rhs) const {
    if constexpr (Atomic) Kokkos::atomic_min(&value, rhs);
    else value = rhs < value ? rhs : value;
  }
  KOKKOS_FORCEINLINE_FUNCTION void reset() const { value = ScatterIdentity<T, KE::ScatterMin>::value(); }
};

template <class T, class Op>
KOKKOS_FUNCTION void scatter_join(T& a, const T& b) {
  if constexpr (std::is_same<Op, KE::ScatterSum>::value) a += b;
  else if constexpr (std::is_same<Op, KE::ScatterProd>::value) a *= b;
  else if constexpr (std::is_same<Op, KE::ScatterMax>::value) a = b > a ? b : a;
  else a = b < a ? b : a;
}
// dst[i] = dst[i] op src[0][i] op ... op src[dups-1][i] over a contiguous
// span (contribute_into); replica d starts at src + d * n.
template <class T, class Op>
struct ScatterCombine {
  T* dst;
  const T* src;
  int64_t n;
  int dups;
  KOKKOS_FUNCTION void operator()(const int64_t i) const {
    T acc = src[i];
    for (int d = 1; d < dups; ++d) scatter_join<T, Op>(acc, src[d * n + i]);
    scatter_join<T, Op>(dst[i], acc);
  }
};
// Replicas for the duplicated mode on MI355X: as many as fit in 64 MiB.
inline int scatter_replicas(size_t bytes) {
  const size_t cap = size_t(64) << 20;
  size_t d         = bytes ? cap / bytes : 256;
  return int(d < 1 ? 1 : d > 256 ? 256 : d);
}

}  // namespace Experimental
}  // namespace Impl

namespace Experimental {

template <class DataType, class Layout, class DeviceType, class Op, class Duplication, class Contribution,
          class OverrideContribution>
class ScatterAccess;

template <class DataType, class Layout = typename DefaultExecutionSpace::array_layout,
          class DeviceType = DefaultExecutionSpace, class Op = ScatterSum,
          class Duplication  = typename Kokkos::Impl::Experimental::DefaultDuplication<
              typename DeviceType::execution_space>::type,
          class Contribution = typename Kokkos::Impl::Experimental::DefaultContribution<
              typename DeviceType::execution_space, Duplication>::type>
class ScatterView {
 public:
  using execution_space         = typename DeviceType::execution_space;
  using memory_space            = typename DeviceType::memory_space;
  using original_view_type      = View<DataType, Layout, memory_space>;
  using original_value_type     = typename original_view_type::non_const_value_type;
  using original_reference_type = typename original_view_type::reference_type;
  // Replicated duplicated mode on MI355X (see the top of this file).
  static constexpr bool replicated = std::is_same<execution_space, MI355X>::value &&
                                     std::is_same<Duplication, ScatterDuplicated>::value &&
                                     original_view_type::rank == 1;
  using replica_view_type = View<original_value_type*, LayoutRight, memory_space>;

  ScatterView() = default;
  template <class RT, class... RP>
  ScatterView(const View<RT, RP...>& original) : internal_view(original) {
    if constexpr (replicated) make_replicas(original.extent(0));
  }
  template <class RT, class... RP>
  ScatterView(const execution_space&, const View<RT, RP...>& original) : ScatterView(original) {}
  template <class... Dims>
  ScatterView(const std::string& name, Dims... dims) : internal_view(name, dims...) {
    if constexpr (replicated) make_replicas(internal_view.extent(0));
  }

  template <class OverrideContribution = Contribution>
  KOKKOS_FORCEINLINE_FUNCTION
      ScatterAccess<DataType, Layout, DeviceType, Op, Duplication, Contribution, OverrideContribution>
      access() const {
    if constexpr (replicated)
      return ScatterAccess<DataType, Layout, DeviceType, Op, Duplication, Contribution, OverrideContribution>(
          internal_view, m_replicas.data(), m_dups);
    else
      return ScatterAccess<DataType, Layout, DeviceType, Op, Duplication, Contribution, OverrideContribution>(
          internal_view, nullptr, 1);
  }

  original_view_type subview() const { return internal_view; }
  KOKKOS_INLINE_FUNCTION bool is_allocated() const { return internal_view.data() != nullptr; }
  int replicas() const { return replicated ? m_dups : 1; }

  template <class DT, class... RP>
  void contribute_into(const View<DT, RP...>& dest) const {
    contribute_into(execution_space(), dest);
  }
  template <class DT, class... RP>
  void contribute_into(const execution_space& exec, const View<DT, RP...>& dest) const {
    static_assert(std::is_same<typename View<DT, RP...>::array_layout, Layout>::value ||
                      original_view_type::rank <= 1,
                  "ScatterView contribute destination has a different layout");
    if constexpr (replicated) {
      // dest op= the fold of every replica (dest may be the original View)
      const int64_t n = int64_t(internal_view.extent(0));
      Kokkos::Impl::Experimental::ScatterCombine<original_value_type, Op> f{dest.data(), m_replicas.data(), n,
                                                                            m_dups};
      parallel_for("Kokkos::ScatterView::contribute_into", RangePolicy<execution_space, IndexType<int64_t>>(exec, 0, n),
                   f);
    } else {
      if (dest.data() == internal_view.data()) return;
      Kokkos::Impl::Experimental::ScatterCombine<original_value_type, Op> f{dest.data(), internal_view.data(),
                                                                            int64_t(internal_view.span()), 1};
      parallel_for("Kokkos::ScatterView::contribute_into",
                   RangePolicy<execution_space, IndexType<int64_t>>(exec, 0, int64_t(internal_view.span())), f);
    }
  }

  void reset(const execution_space& exec = execution_space()) {
    const auto id = Kokkos::Impl::Experimental::ScatterIdentity<original_value_type, Op>::value();
    if constexpr (replicated) {
      if (m_replicas.span()) deep_copy(exec, m_replicas, id);
    } else {
      if (internal_view.span()) deep_copy(exec, internal_view, id);
    }
  }
  template <class DT, class... RP>
  void reset_except(const View<DT, RP...>& view) {
    reset_except(execution_space(), view);
  }
  template <class DT, class... RP>
  void reset_except(const execution_space& exec, const View<DT, RP...>& view) {
    if (replicated || view.data() != internal_view.data()) reset(exec);
  }

 private:
  void make_replicas(size_t n) {
    m_dups     = Kokkos::Impl::Experimental::scatter_replicas(n * sizeof(original_value_type));
    m_replicas = replica_view_type(std::string("ScatterView::replicas"), n * size_t(m_dups));
    reset();
  }
  original_view_type internal_view;
  replica_view_type m_replicas;
  int m_dups = 1;
};

// The per-loop handle: holds the View (device-copyable) and returns one
// ScatterValue per element.
template <class DataType, class Layout, class DeviceType, class Op, class Duplication, class Contribution,
          class OverrideContribution>
class ScatterAccess {
 public:
  using view_type  = ScatterView<DataType, Layout, DeviceType, Op, Duplication, Contribution>;
  using value_type = Kokkos::Impl::Experimental::ScatterValue<
      typename view_type::original_value_type, Op,
      std::is_same<OverrideContribution, ScatterAtomic>::value ||
          (std::is_same<typename view_type::execution_space, MI355X>::value &&
           std::is_same<Duplication, ScatterDuplicated>::value)>;
  using T = typename view_type::original_value_type;
  KOKKOS_INLINE_FUNCTION ScatterAccess(const typename view_type::original_view_type& v, T* replicas, int dups)
      : m_view(v), m_rep(replicas), m_dups(dups) {}
  template <class... Args>
  KOKKOS_FORCEINLINE_FUNCTION value_type operator()(Args... args) const {
    if constexpr (view_type::replicated) return value_type{replica_base()[(args, ...)]};
    else return value_type{m_view(args...)};
  }
  template <class Arg, class = std::enable_if_t<std::is_integral<Arg>::value && view_type::original_view_type::rank == 1>>
  KOKKOS_FORCEINLINE_FUNCTION value_type operator[](Arg a) const {
    return (*this)(a);
  }

 private:
  // this workgroup's replica
  KOKKOS_FORCEINLINE_FUNCTION T* replica_base() const {
#if defined(__HIP_DEVICE_COMPILE__)
    const int64_t d = int64_t(__builtin_amdgcn_workgroup_id_x()) % m_dups;
#else
    const int64_t d = 0;
#endif
    return m_rep + d * int64_t(m_view.extent(0));
  }
  typename view_type::original_view_type m_view;
  T* m_rep;
  int m_dups;
};

template <class Op = ScatterSum, class Duplication = void, class Contribution = void, class RT, class... RP>
auto create_scatter_view(const View<RT, RP...>& original) {
  using V      = View<RT, RP...>;
  using Exec   = typename V::execution_space;
  using Dup    = std::conditional_t<std::is_void<Duplication>::value,
                                 typename Kokkos::Impl::Experimental::DefaultDuplication<Exec>::type, Duplication>;
  using Contr  = std::conditional_t<std::is_void<Contribution>::value,
                                   typename Kokkos::Impl::Experimental::DefaultContribution<Exec, Dup>::type,
                                   Contribution>;
  return ScatterView<typename V::data_type, typename V::array_layout, Exec, Op, Dup, Contr>(original);
}
template <class Op, class RT, class... RP>
auto create_scatter_view(Op, const View<RT, RP...>& original) {
  return create_scatter_view<Op>(original);
}
template <class Op, class Duplication, class Contribution, class RT, class... RP>
auto create_scatter_view(Op, Duplication, Contribution, const View<RT, RP...>& original) {
  return create_scatter_view<Op, Duplication, Contribution>(original);
}

template <class DT1, class... VP, class DT2, class LY, class ES, class OP, class DU, class CO>
void contribute(const typename ES::execution_space& exec, const View<DT1, VP...>& dest,
                const ScatterView<DT2, LY, ES, OP, DU, CO>& src) {
  src.contribute_into(exec, dest);
}
template <class DT1, class... VP, class DT2, class LY, class ES, class OP, class DU, class CO>
void contribute(const View<DT1, VP...>& dest, const ScatterView<DT2, LY, ES, OP, DU, CO>& src) {
  src.contribute_into(dest);
}

}  // namespace Experimental
}  // namespace Kokkos

#endif
