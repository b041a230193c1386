// Kokkos::View for the MI355X execution space: ranks 0-8, static and dynamic
// extents, LayoutLeft / LayoutRight / LayoutStride, memory traits, reference
// counted allocations, deep_copy and the create_mirror family.
//
// Reference behaviour: core/src/Kokkos_View.hpp:151-1055 (class View),
// core/src/View/Kokkos_ViewMapping.hpp (offsets of LayoutLeft/Right),
// core/src/Kokkos_CopyViews.hpp:58-420 (deep_copy, ViewFill, ViewCopy,
// create_mirror_view[_and_copy]), core/src/impl/Kokkos_SharedAlloc.hpp
// (reference-counted allocation records that are not touched on the device).
//
// Layout in HBM: one contiguous allocation, no padding; LayoutLeft makes the
// first index the fastest (so consecutive lanes of a wavefront touch
// consecutive addresses when the first index follows the lane id), which is
// why LayoutLeft is the MI355X default (as for the reference HIP space).
#ifndef KKMX_VIEW_HPP
#define KKMX_VIEW_HPP

#include <kkmx/Atomic.hpp>
#include <kkmx/Spaces.hpp>

#include <array>
#include <initializer_list>
#include <string>

namespace Kokkos {

// ------------------------------------------------------------------ layouts
struct LayoutLeft {
  using array_layout = LayoutLeft;
  size_t dimension[8];
  enum : bool { is_extent_constructible = true };
  KOKKOS_INLINE_FUNCTION explicit constexpr LayoutLeft(size_t N0 = 0, size_t N1 = 0, size_t N2 = 0,
                                                       size_t N3 = 0, size_t N4 = 0, size_t N5 = 0,
                                                       size_t N6 = 0, size_t N7 = 0)
      : dimension{N0, N1, N2, N3, N4, N5, N6, N7} {}
};
struct LayoutRight {
  using array_layout = LayoutRight;
  size_t dimension[8];
  enum : bool { is_extent_constructible = true };
  KOKKOS_INLINE_FUNCTION explicit constexpr LayoutRight(size_t N0 = 0, size_t N1 = 0, size_t N2 = 0,
                                                        size_t N3 = 0, size_t N4 = 0, size_t N5 = 0,
                                                        size_t N6 = 0, size_t N7 = 0)
      : dimension{N0, N1, N2, N3, N4, N5, N6, N7} {}
};
struct LayoutStride {
  using array_layout = LayoutStride;
  size_t dimension[8];
  size_t stride[8];
  enum : bool { is_extent_constructible = false };
  KOKKOS_INLINE_FUNCTION explicit constexpr LayoutStride(size_t N0 = 0, size_t S0 = 0, size_t N1 = 0,
                                                         size_t S1 = 0, size_t N2 = 0, size_t S2 = 0,
                                                         size_t N3 = 0, size_t S3 = 0, size_t N4 = 0,
                                                         size_t S4 = 0, size_t N5 = 0, size_t S5 = 0,
                                                         size_t N6 = 0, size_t S6 = 0, size_t N7 = 0,
                                                         size_t S7 = 0)
      : dimension{N0, N1, N2, N3, N4, N5, N6, N7}, stride{S0, S1, S2, S3, S4, S5, S6, S7} {}
  // Compact strides for the dimension order order[0] (stride 1), order[1],
  // ...; an order that is not a permutation of 0..rank-1 leaves every
  // stride zero (core/src/Kokkos_Layout.hpp:168-192).
  template <class IO, class ID>
  KOKKOS_INLINE_FUNCTION static LayoutStride order_dimensions(int const rank, IO const* const order,
                                                              ID const* const dimen) {
    LayoutStride l;
    if (rank < 0 || rank > 8) return l;
    unsigned seen = 0;
    for (int r = 0; r < rank; ++r) {
      const int o = int(order[r]);
      if (o < 0 || o >= rank) return l;
      seen |= 1u << o;
    }
    if (seen != (1u << rank) - 1u) return l;
    size_t n = 1;
    for (int r = 0; r < rank; ++r) {
      l.stride[order[r]] = n;
      n *= size_t(dimen[order[r]]);
      l.dimension[r] = size_t(dimen[r]);
    }
    return l;
  }
};

template <class T>
struct is_array_layout : std::false_type {};
template <>
struct is_array_layout<LayoutLeft> : std::true_type {};
template <>
struct is_array_layout<LayoutRight> : std::true_type {};
template <>
struct is_array_layout<LayoutStride> : std::true_type {};

// ------------------------------------------------------------------ memory traits
enum MemoryTraitsFlags : unsigned {
  Unmanaged    = 0x01,
  RandomAccess = 0x02,
  Atomic       = 0x04,
  Restrict     = 0x08,
  Aligned      = 0x10
};
template <unsigned T = 0>
struct MemoryTraits {
  using memory_traits = MemoryTraits<T>;
  static constexpr unsigned impl_value = T;
  static constexpr bool is_unmanaged     = (T & Unmanaged) != 0;
  static constexpr bool is_random_access = (T & RandomAccess) != 0;
  static constexpr bool is_atomic        = (T & Atomic) != 0;
  static constexpr bool is_restrict      = (T & Restrict) != 0;
  static constexpr bool is_aligned       = (T & Aligned) != 0;
};
using MemoryManaged      = MemoryTraits<0>;
using MemoryUnmanaged    = MemoryTraits<Unmanaged>;
using MemoryRandomAccess = MemoryTraits<Unmanaged | RandomAccess>;

template <class T>
struct is_memory_traits : std::false_type {};
template <unsigned U>
struct is_memory_traits<MemoryTraits<U>> : std::true_type {};

// ------------------------------------------------------------------ ctor properties
namespace Impl {
struct WithoutInitializing_t {};
struct AllowPadding_t {};
}  // namespace Impl
inline constexpr Impl::WithoutInitializing_t WithoutInitializing{};
inline constexpr Impl::AllowPadding_t AllowPadding{};

namespace Impl {
struct ViewCtorProp {
  std::string label;
  bool initialize = true;
  bool has_exec   = false;
  std::shared_ptr<MI355X> exec;
};
inline void apply_prop(ViewCtorProp& p, const std::string& s) { p.label = s; }
inline void apply_prop(ViewCtorProp& p, const char* s) { p.label = s; }
inline void apply_prop(ViewCtorProp& p, WithoutInitializing_t) { p.initialize = false; }
inline void apply_prop(ViewCtorProp& p, AllowPadding_t) {}
inline void apply_prop(ViewCtorProp& p, const MI355X& e) {
  p.has_exec = true;
  p.exec     = std::make_shared<MI355X>(e);
}
inline void apply_prop(ViewCtorProp& p, const Serial&) {}
inline void apply_prop(ViewCtorProp& p, const HostSpace&) {}
inline void apply_prop(ViewCtorProp& p, const MI355XSpace&) {}
}  // namespace Impl

template <class... Args>
Impl::ViewCtorProp view_alloc(const Args&... args) {
  Impl::ViewCtorProp p;
  (Impl::apply_prop(p, args), ...);
  return p;
}
template <class... Args>
Impl::ViewCtorProp view_wrap(const Args&... args) {
  return view_alloc(args...);
}

// ------------------------------------------------------------------ allocation records
namespace Impl {
struct SubviewCtor {};

template <unsigned N>
struct ViewRankConstant {
  using value_type                  = size_t;
  static constexpr size_t value     = N;
  KOKKOS_INLINE_FUNCTION constexpr operator size_t() const noexcept { return N; }
  KOKKOS_INLINE_FUNCTION constexpr size_t operator()() const noexcept { return N; }
};

struct SharedAllocRecord {
  std::atomic<int> count{1};
  std::string label;
  void* ptr    = nullptr;
  size_t bytes = 0;
  void (*release)(void*, size_t, const char*) = nullptr;
  ~SharedAllocRecord() {
    if (release) release(ptr, bytes, label.c_str());
  }
};
inline void record_inc(SharedAllocRecord* r) {
  if (r) r->count.fetch_add(1, std::memory_order_relaxed);
}
inline void record_dec(SharedAllocRecord* r) {
  if (r && r->count.fetch_sub(1, std::memory_order_acq_rel) == 1) delete r;
}

// Parse DataType: array extents (outermost first) are static dimensions that
// follow the dynamic (pointer) dimensions.
template <class T>
struct count_ptrs {
  static constexpr unsigned n = 0;
  using base                  = T;
};
template <class T>
struct count_ptrs<T*> {
  static constexpr unsigned n = count_ptrs<T>::n + 1;
  using base                  = typename count_ptrs<T>::base;
};
template <class DT>
struct ViewDataAnalysis {
  using no_extents                         = std::remove_all_extents_t<DT>;
  static constexpr unsigned rank_dynamic   = count_ptrs<no_extents>::n;
  static constexpr unsigned rank_static    = std::rank<DT>::value;
  static constexpr unsigned rank           = rank_dynamic + rank_static;
  using value_type                         = typename count_ptrs<no_extents>::base;
  template <unsigned K>
  static constexpr size_t static_extent() {
    if constexpr (K < rank_dynamic) return 0;
    else return std::extent<DT, K - rank_dynamic>::value;
  }
};

// Replace the scalar type inside a DataType (for const / non-const variants).
template <class DT, class V>
struct rebind_data {
  using type = V;
};
template <class DT, class V>
struct rebind_data<DT*, V> {
  using type = typename rebind_data<DT, V>::type*;
};
template <class DT, class V, size_t N>
struct rebind_data<DT[N], V> {
  using type = typename rebind_data<DT, V>::type[N];
};
template <class DT, class V>
struct rebind_data<DT[], V> {
  using type = typename rebind_data<DT, V>::type*;
};

// View template-property parsing: layout, space (memory, execution or Device)
// and memory traits in any order.
template <class... P>
struct ViewProps;
template <>
struct ViewProps<> {
  using layout = void;
  using space  = void;
  using traits = void;
};
template <class H, class... T>
struct ViewProps<H, T...> {
  using rest   = ViewProps<T...>;
  using layout = std::conditional_t<is_array_layout<H>::value, H, typename rest::layout>;
  using traits = std::conditional_t<is_memory_traits<H>::value, H, typename rest::traits>;
  template <class X, class = void>
  struct sp {
    using type = typename rest::space;
  };
  template <class X>
  struct sp<X, std::void_t<typename X::memory_space, typename X::execution_space>> {
    using type = X;
  };
  using space = std::conditional_t<is_array_layout<H>::value || is_memory_traits<H>::value,
                                   typename rest::space, typename sp<H>::type>;
};

template <class S>
struct space_traits {
  // A Device<E,M>, an execution space or a memory space.
  using memory_space    = typename S::memory_space;
  using execution_space = typename S::execution_space;
};
template <>
struct space_traits<void> {
  using memory_space    = MI355XSpace;
  using execution_space = MI355X;
};
template <>
struct space_traits<HostSpace> {
  using memory_space    = HostSpace;
  using execution_space = Serial;
};
template <>
struct space_traits<MI355XHostPinnedSpace> {
  using memory_space    = MI355XHostPinnedSpace;
  using execution_space = Serial;
};

}  // namespace Impl

template <class DataType, class... Properties>
struct ViewTraits {
  using analysis               = Impl::ViewDataAnalysis<DataType>;
  using props                  = Impl::ViewProps<Properties...>;
  using space_tr               = Impl::space_traits<typename props::space>;
  using memory_space           = typename space_tr::memory_space;
  using execution_space        = typename space_tr::execution_space;
  using device_type            = Device<execution_space, memory_space>;
  using array_layout           = std::conditional_t<std::is_void<typename props::layout>::value,
                                          typename execution_space::array_layout,
                                          typename props::layout>;
  using memory_traits          = std::conditional_t<std::is_void<typename props::traits>::value,
                                           MemoryManaged, typename props::traits>;
  using value_type             = typename analysis::value_type;
  using const_value_type       = std::add_const_t<value_type>;
  using non_const_value_type   = std::remove_const_t<value_type>;
  using data_type              = DataType;
  using const_data_type        = typename Impl::rebind_data<DataType, const_value_type>::type;
  using non_const_data_type    = typename Impl::rebind_data<DataType, non_const_value_type>::type;
  using scalar_array_type      = data_type;
  using size_type              = typename memory_space::size_type;
  using host_mirror_space      = std::conditional_t<
      Impl::exec_can_access<Serial, memory_space>::value, memory_space, HostSpace>;
  static constexpr unsigned rank         = analysis::rank;
  static constexpr unsigned rank_dynamic = analysis::rank_dynamic;
  static constexpr bool is_managed       = !memory_traits::is_unmanaged;
};

template <class DataType, class... Properties>
class View;

template <class T>
struct is_view : std::false_type {};
template <class D, class... P>
struct is_view<View<D, P...>> : std::true_type {};
template <class T>
inline constexpr bool is_view_v = is_view<T>::value;

namespace Impl {
inline thread_local bool g_tracking_enabled = true;
struct TrackingDisabledScope {
  bool prev;
  TrackingDisabledScope() : prev(g_tracking_enabled) { g_tracking_enabled = false; }
  ~TrackingDisabledScope() { g_tracking_enabled = prev; }
};

template <class Mem>
struct MemOps;
template <>
struct MemOps<HostSpace> {
  static void* alloc(const char* l, size_t n) { return HostSpace().allocate(l, n); }
  static void release(void* p, size_t n, const char* l) { HostSpace().deallocate(l, p, n); }
  static void zero(void* p, size_t n) { std::memset(p, 0, n); }
};
template <>
struct MemOps<MI355XSpace> {
  static void* alloc(const char* l, size_t n) { return MI355XSpace().allocate(l, n); }
  static void release(void* p, size_t n, const char* l) { MI355XSpace().deallocate(l, p, n); }
  static void zero(void* p, size_t n) {
    mx_stream_t s = MI355X().stream();
    KKMX_SAFE_CALL(mx_memset_async(p, 0, n, s));
    KKMX_SAFE_CALL(mx_stream_sync(s));
  }
};
template <>
struct MemOps<MI355XHostPinnedSpace> {
  static void* alloc(const char* l, size_t n) { return MI355XHostPinnedSpace().allocate(l, n); }
  static void release(void* p, size_t n, const char* l) { MI355XHostPinnedSpace().deallocate(l, p, n); }
  static void zero(void* p, size_t n) { std::memset(p, 0, n); }
};
}  // namespace Impl

template <class DataType, class... Properties>
class View : public ViewTraits<DataType, Properties...> {
 public:
  using traits                 = ViewTraits<DataType, Properties...>;
  using value_type             = typename traits::value_type;
  using const_value_type       = typename traits::const_value_type;
  using non_const_value_type   = typename traits::non_const_value_type;
  using memory_space           = typename traits::memory_space;
  using execution_space        = typename traits::execution_space;
  using device_type            = typename traits::device_type;
  using array_layout           = typename traits::array_layout;
  using memory_traits          = typename traits::memory_traits;
  using size_type              = typename traits::size_type;
  using pointer_type           = value_type*;
  using element_type           = value_type;
  // Read-only device Views of arithmetic values declared RandomAccess or
  // Restrict promise that no thread writes their data during a kernel; their
  // elements are read by value through the constant address space, whose
  // loads the compiler may move across stores and merge
  // (core/src/View/Kokkos_ViewMapping.hpp:2402-2423 gives Restrict Views a
  // restrict-qualified handle; the reference's CUDA RandomAccess Views also
  // return elements by value).
  static constexpr bool kkmx_invariant_reads =
      std::is_const<value_type>::value && std::is_arithmetic<typename traits::non_const_value_type>::value &&
      !memory_traits::is_atomic && (memory_traits::is_random_access || memory_traits::is_restrict) &&
      std::is_same<memory_space, MI355XSpace>::value;
  using reference_type         = std::conditional_t<
      memory_traits::is_atomic, Impl::AtomicDataElement<value_type>,
      std::conditional_t<kkmx_invariant_reads, const_value_type, value_type&>>;
  using data_handle_type       = pointer_type;
  using type                   = View<typename traits::data_type, array_layout, device_type, memory_traits>;
  using const_type             = View<typename traits::const_data_type, array_layout, device_type, memory_traits>;
  using non_const_type         = View<typename traits::non_const_data_type, array_layout, device_type, memory_traits>;
  using host_mirror_type       = View<typename traits::non_const_data_type, array_layout,
                                Device<Serial, typename traits::host_mirror_space>>;
  using HostMirror             = host_mirror_type;
  using uniform_type           = View<typename traits::data_type, array_layout, device_type, memory_traits>;

  // Integral constants: usable as values (V::rank) and as calls (V::rank()),
  // as in the reference (core/src/View/Kokkos_ViewLegacy.hpp:315-319).
  static constexpr Impl::ViewRankConstant<traits::rank> rank                 = {};
  static constexpr Impl::ViewRankConstant<traits::rank_dynamic> rank_dynamic = {};
  static constexpr unsigned Rank         = rank;
  static constexpr bool reference_type_is_lvalue_reference = !memory_traits::is_atomic && !kkmx_invariant_reads;

  // ------------------------------------------------------------ constructors
  KOKKOS_INLINE_FUNCTION View() : m_ptr(nullptr), m_rec(nullptr) {
    for (unsigned r = 0; r < kR; ++r) m_ext[r] = m_str[r] = 0;
    init_static();
  }

  KOKKOS_INLINE_FUNCTION View(const View& o) : m_ptr(o.m_ptr), m_rec(o.m_rec) {
    for (unsigned r = 0; r < kR; ++r) { m_ext[r] = o.m_ext[r]; m_str[r] = o.m_str[r]; }
#if !defined(__HIP_DEVICE_COMPILE__)
    if (Impl::g_tracking_enabled) Impl::record_inc(m_rec); else m_rec = nullptr;
#endif
  }
  KOKKOS_INLINE_FUNCTION View(View&& o) : m_ptr(o.m_ptr), m_rec(o.m_rec) {
    for (unsigned r = 0; r < kR; ++r) { m_ext[r] = o.m_ext[r]; m_str[r] = o.m_str[r]; }
    o.m_rec = nullptr;
  }
  KOKKOS_INLINE_FUNCTION View& operator=(const View& o) {
    if (this == &o) return *this;
#if !defined(__HIP_DEVICE_COMPILE__)
    Impl::record_inc(o.m_rec);
    Impl::record_dec(m_rec);
#endif
    m_ptr = o.m_ptr;
    m_rec = o.m_rec;
    for (unsigned r = 0; r < kR; ++r) { m_ext[r] = o.m_ext[r]; m_str[r] = o.m_str[r]; }
    return *this;
  }
  KOKKOS_INLINE_FUNCTION View& operator=(View&& o) {
    if (this == &o) return *this;
#if !defined(__HIP_DEVICE_COMPILE__)
    Impl::record_dec(m_rec);
#endif
    m_ptr   = o.m_ptr;
    m_rec   = o.m_rec;
    o.m_rec = nullptr;
    for (unsigned r = 0; r < kR; ++r) { m_ext[r] = o.m_ext[r]; m_str[r] = o.m_str[r]; }
    return *this;
  }
  KOKKOS_INLINE_FUNCTION ~View() {
#if !defined(__HIP_DEVICE_COMPILE__)
    Impl::record_dec(m_rec);
#endif
  }

  // Conversion from a compatible View (non-const -> const, other traits).
  template <class D2, class... P2,
            class = std::enable_if_t<
                !std::is_same<View<D2, P2...>, View>::value &&
                std::is_same<typename View<D2, P2...>::memory_space, memory_space>::value &&
                std::is_convertible<typename View<D2, P2...>::value_type*, value_type*>::value &&
                View<D2, P2...>::rank == rank>>
  KOKKOS_INLINE_FUNCTION View(const View<D2, P2...>& o) : m_ptr(o.data()), m_rec(nullptr) {
    // Any layout goes to LayoutStride; LayoutStride goes to LayoutLeft or
    // LayoutRight when its unit-stride dimension really has stride 1
    // (core/src/View/Kokkos_ViewMapping.hpp, ViewMapping assignment rules).
    using L2 = typename View<D2, P2...>::array_layout;
    static_assert(std::is_same<L2, array_layout>::value || std::is_same<array_layout, LayoutStride>::value ||
                      std::is_same<L2, LayoutStride>::value || rank <= 1,
                  "View assignment requires matching layouts");
    for (unsigned r = 0; r < kR; ++r) { m_ext[r] = o.impl_ext(r); m_str[r] = o.impl_str(r); }
    if constexpr (std::is_same<L2, LayoutStride>::value && !std::is_same<array_layout, LayoutStride>::value &&
                  rank >= 1) {
      const unsigned unit = std::is_same<array_layout, LayoutLeft>::value ? 0 : rank - 1;
#if !defined(__HIP_DEVICE_COMPILE__)
      if (m_ext[unit] > 1 && m_str[unit] != 1)
        Kokkos::abort("View assignment: the source is not unit-stride where the layout needs it");
#endif
      m_str[unit] = 1;
    }
#if !defined(__HIP_DEVICE_COMPILE__)
    if (Impl::g_tracking_enabled && !memory_traits::is_unmanaged) {
      m_rec = o.impl_record();
      Impl::record_inc(m_rec);
    }
#endif
  }

  // Allocating constructors.
  explicit View(const std::string& label, size_t n0 = KOKKOS_IMPL_NOARG, size_t n1 = KOKKOS_IMPL_NOARG,
                size_t n2 = KOKKOS_IMPL_NOARG, size_t n3 = KOKKOS_IMPL_NOARG, size_t n4 = KOKKOS_IMPL_NOARG,
                size_t n5 = KOKKOS_IMPL_NOARG, size_t n6 = KOKKOS_IMPL_NOARG, size_t n7 = KOKKOS_IMPL_NOARG)
      : View(view_alloc(label), n0, n1, n2, n3, n4, n5, n6, n7) {}
  // String-literal labels (char arrays, as the reference's is_view_label,
  // core/src/View/Kokkos_ViewCtor.hpp:59-68): a char pointer stays free for
  // the unmanaged constructor of char Views.
  template <size_t N>
  explicit View(const char (&label)[N], size_t n0 = KOKKOS_IMPL_NOARG, size_t n1 = KOKKOS_IMPL_NOARG,
                size_t n2 = KOKKOS_IMPL_NOARG, size_t n3 = KOKKOS_IMPL_NOARG, size_t n4 = KOKKOS_IMPL_NOARG,
                size_t n5 = KOKKOS_IMPL_NOARG, size_t n6 = KOKKOS_IMPL_NOARG, size_t n7 = KOKKOS_IMPL_NOARG)
      : View(view_alloc(std::string(label)), n0, n1, n2, n3, n4, n5, n6, n7) {}
  explicit View(const Impl::ViewCtorProp& prop, size_t n0 = KOKKOS_IMPL_NOARG, size_t n1 = KOKKOS_IMPL_NOARG,
                size_t n2 = KOKKOS_IMPL_NOARG, size_t n3 = KOKKOS_IMPL_NOARG, size_t n4 = KOKKOS_IMPL_NOARG,
                size_t n5 = KOKKOS_IMPL_NOARG, size_t n6 = KOKKOS_IMPL_NOARG, size_t n7 = KOKKOS_IMPL_NOARG)
      : m_ptr(nullptr), m_rec(nullptr) {
    size_t dyn[8] = {n0, n1, n2, n3, n4, n5, n6, n7};
    set_extents(dyn);
    allocate(prop);
  }
  template <class L, class = std::enable_if_t<is_array_layout<L>::value>>
  View(const std::string& label, const L& layout) : View(view_alloc(label), layout) {}
  template <class L, class = std::enable_if_t<is_array_layout<L>::value>>
  View(const Impl::ViewCtorProp& prop, const L& layout) : m_ptr(nullptr), m_rec(nullptr) {
    set_layout(layout);
    allocate(prop);
  }

  // Unmanaged wrap of an existing pointer.
  KOKKOS_INLINE_FUNCTION explicit View(pointer_type p, size_t n0 = KOKKOS_IMPL_NOARG,
                                       size_t n1 = KOKKOS_IMPL_NOARG, size_t n2 = KOKKOS_IMPL_NOARG,
                                       size_t n3 = KOKKOS_IMPL_NOARG, size_t n4 = KOKKOS_IMPL_NOARG,
                                       size_t n5 = KOKKOS_IMPL_NOARG, size_t n6 = KOKKOS_IMPL_NOARG,
                                       size_t n7 = KOKKOS_IMPL_NOARG)
      : m_ptr(p), m_rec(nullptr) {
    size_t dyn[8] = {n0, n1, n2, n3, n4, n5, n6, n7};
    set_extents(dyn);
  }
  template <class L, class = std::enable_if_t<is_array_layout<L>::value>>
  KOKKOS_INLINE_FUNCTION View(pointer_type p, const L& layout) : m_ptr(p), m_rec(nullptr) {
    set_layout(layout);
  }

  // A window of another View's allocation (Kokkos::subview): pointer,
  // extents and strides given; shares and counts the allocation record.
  KOKKOS_INLINE_FUNCTION View(Impl::SubviewCtor, pointer_type p, const size_t* ext, const size_t* str,
                              Impl::SharedAllocRecord* rec)
      : m_ptr(p), m_rec(nullptr) {
    for (unsigned r = 0; r < kR; ++r) {
      m_ext[r] = r < rank ? ext[r] : 0;
      m_str[r] = r < rank ? str[r] : 1;
    }
#if !defined(__HIP_DEVICE_COMPILE__)
    if (rec && Impl::g_tracking_enabled && !memory_traits::is_unmanaged) {
      m_rec = rec;
      Impl::record_inc(m_rec);
    }
#else
    (void)rec;
#endif
  }

  // Scratch-space constructor: carve from a team scratch pad.
  template <class ScratchSpace,
            class = std::enable_if_t<std::is_same<typename ScratchSpace::memory_space,
                                                  typename ScratchSpace::memory_space>::value &&
                                     ScratchSpace::kkmx_is_scratch>>
  KOKKOS_INLINE_FUNCTION View(const ScratchSpace& space, size_t n0 = KOKKOS_IMPL_NOARG,
                              size_t n1 = KOKKOS_IMPL_NOARG, size_t n2 = KOKKOS_IMPL_NOARG,
                              size_t n3 = KOKKOS_IMPL_NOARG, size_t n4 = KOKKOS_IMPL_NOARG,
                              size_t n5 = KOKKOS_IMPL_NOARG, size_t n6 = KOKKOS_IMPL_NOARG,
                              size_t n7 = KOKKOS_IMPL_NOARG)
      : m_ptr(nullptr), m_rec(nullptr) {
    size_t dyn[8] = {n0, n1, n2, n3, n4, n5, n6, n7};
    set_extents(dyn);
    m_ptr = static_cast<pointer_type>(space.get_shmem(span() * sizeof(value_type)));
  }

  // Bytes an allocation of these extents (or this layout) takes
  // (core/src/View/Kokkos_ViewLegacy.hpp required_allocation_size).
  static size_t required_allocation_size(size_t n0 = KOKKOS_IMPL_NOARG, size_t n1 = KOKKOS_IMPL_NOARG,
                                         size_t n2 = KOKKOS_IMPL_NOARG, size_t n3 = KOKKOS_IMPL_NOARG,
                                         size_t n4 = KOKKOS_IMPL_NOARG, size_t n5 = KOKKOS_IMPL_NOARG,
                                         size_t n6 = KOKKOS_IMPL_NOARG, size_t n7 = KOKKOS_IMPL_NOARG) {
    View v(pointer_type(nullptr), n0, n1, n2, n3, n4, n5, n6, n7);
    return v.span() * sizeof(value_type);
  }
  template <class L, class = std::enable_if_t<is_array_layout<L>::value>>
  static size_t required_allocation_size(const L& layout) {
    View v(pointer_type(nullptr), layout);
    return v.span() * sizeof(value_type);
  }
  template <class L, class = std::enable_if_t<is_array_layout<L>::value>>
  static size_t shmem_size(const L& layout) {
    View v(pointer_type(nullptr), layout);
    return ((v.span() * sizeof(value_type) + 7) / 8) * 8 + 8;
  }

  static constexpr size_t shmem_size(size_t n0 = KOKKOS_IMPL_NOARG, size_t n1 = KOKKOS_IMPL_NOARG,
                                     size_t n2 = KOKKOS_IMPL_NOARG, size_t n3 = KOKKOS_IMPL_NOARG,
                                     size_t n4 = KOKKOS_IMPL_NOARG, size_t n5 = KOKKOS_IMPL_NOARG,
                                     size_t n6 = KOKKOS_IMPL_NOARG, size_t n7 = KOKKOS_IMPL_NOARG) {
    size_t dyn[8] = {n0, n1, n2, n3, n4, n5, n6, n7};
    size_t total  = 1;
    unsigned d    = 0;
    for (unsigned r = 0; r < rank; ++r) {
      size_t se = static_ext(r);
      total *= se ? se : dyn[d++];
    }
    return ((total * sizeof(value_type) + 7) / 8) * 8 + 8;
  }

  // ------------------------------------------------------------ queries
  KOKKOS_INLINE_FUNCTION constexpr size_t extent(unsigned r) const { return r < rank ? m_ext[r] : 1; }
  KOKKOS_INLINE_FUNCTION constexpr int extent_int(unsigned r) const { return static_cast<int>(extent(r)); }
  KOKKOS_INLINE_FUNCTION constexpr size_t stride(unsigned r) const { return r < rank ? m_str[r] : 0; }
  template <class iType>
  KOKKOS_INLINE_FUNCTION void stride(iType* s) const {
    for (unsigned r = 0; r < rank; ++r) s[r] = static_cast<iType>(m_str[r]);
    if (rank) s[rank] = static_cast<iType>(span());
  }
  KOKKOS_INLINE_FUNCTION size_t stride_0() const { return stride(0); }
  KOKKOS_INLINE_FUNCTION size_t stride_1() const { return stride(1); }
  KOKKOS_INLINE_FUNCTION size_t stride_2() const { return stride(2); }
  KOKKOS_INLINE_FUNCTION constexpr size_t size() const {
    size_t s = 1;
    for (unsigned r = 0; r < rank; ++r) s *= m_ext[r];
    return s;
  }
  // Strides may exceed the extents below them in every layout (a subview
  // keeps its parent's strides), so the span comes from the strides.
  KOKKOS_INLINE_FUNCTION constexpr size_t span() const {
    if constexpr (rank == 0) {
      return 1;
    } else {
      size_t s = 0;
      for (unsigned r = 0; r < rank; ++r) if (m_ext[r]) s += (m_ext[r] - 1) * m_str[r];
      return size() ? s + 1 : 0;
    }
  }
  KOKKOS_INLINE_FUNCTION constexpr bool span_is_contiguous() const { return span() == size(); }
  KOKKOS_INLINE_FUNCTION constexpr pointer_type data() const { return m_ptr; }
  KOKKOS_INLINE_FUNCTION constexpr bool is_allocated() const { return m_ptr != nullptr; }
  std::string label() const { return m_rec ? m_rec->label : std::string(); }
  int use_count() const { return m_rec ? m_rec->count.load() : 0; }
  array_layout layout() const {
    array_layout l{};
    for (unsigned r = 0; r < rank; ++r) l.dimension[r] = m_ext[r];
    if constexpr (std::is_same<array_layout, LayoutStride>::value)
      for (unsigned r = 0; r < rank; ++r) l.stride[r] = m_str[r];
    return l;
  }
  KOKKOS_INLINE_FUNCTION size_t impl_ext(unsigned r) const { return m_ext[r < kR ? r : 0]; }
  KOKKOS_INLINE_FUNCTION size_t impl_str(unsigned r) const { return m_str[r < kR ? r : 0]; }
  KOKKOS_INLINE_FUNCTION Impl::SharedAllocRecord* impl_record() const { return m_rec; }
  static constexpr size_t static_ext(unsigned r) {
    constexpr size_t se[9] = {Impl::ViewDataAnalysis<DataType>::template static_extent<0>(),
                              Impl::ViewDataAnalysis<DataType>::template static_extent<1>(),
                              Impl::ViewDataAnalysis<DataType>::template static_extent<2>(),
                              Impl::ViewDataAnalysis<DataType>::template static_extent<3>(),
                              Impl::ViewDataAnalysis<DataType>::template static_extent<4>(),
                              Impl::ViewDataAnalysis<DataType>::template static_extent<5>(),
                              Impl::ViewDataAnalysis<DataType>::template static_extent<6>(),
                              Impl::ViewDataAnalysis<DataType>::template static_extent<7>(), 0};
    return r < 8 ? se[r] : 0;
  }
  static constexpr size_t static_extent(unsigned r) { return static_ext(r); }

  // ------------------------------------------------------------ element access
  template <class... Is>
  KOKKOS_FORCEINLINE_FUNCTION reference_type operator()(Is... is) const {
    static_assert(sizeof...(Is) == rank || (rank == 0 && sizeof...(Is) == 0),
                  "View::operator(): number of indices must equal the rank");
    return ref(offset(static_cast<size_t>(is)...));
  }
  template <class I0>
  KOKKOS_FORCEINLINE_FUNCTION reference_type operator[](I0 i0) const {
    static_assert(rank == 1, "View::operator[] requires rank 1");
    return ref(offset(static_cast<size_t>(i0)));
  }
  template <class... Is>
  KOKKOS_FORCEINLINE_FUNCTION reference_type access(Is... is) const {
    return ref(offset(static_cast<size_t>(is)...));
  }

 private:
  static constexpr unsigned kR = rank ? rank : 1;
  static constexpr size_t KOKKOS_IMPL_NOARG = ~size_t(0);
  pointer_type m_ptr;
  size_t m_ext[kR];
  size_t m_str[kR];
  Impl::SharedAllocRecord* m_rec;

  template <class D2, class... P2>
  friend class View;

  KOKKOS_FORCEINLINE_FUNCTION reference_type ref(size_t off) const {
    if constexpr (memory_traits::is_atomic) {
      return reference_type(m_ptr + off);
    } else if constexpr (kkmx_invariant_reads) {
#if defined(__HIP_DEVICE_COMPILE__)
      using invariant_ptr = const __attribute__((address_space(4))) typename traits::non_const_value_type*;
      return *(invariant_ptr)(m_ptr + off);
#else
      return m_ptr[off];
#endif
    } else {
      return m_ptr[off];
    }
  }
  KOKKOS_INLINE_FUNCTION void init_static() {
    for (unsigned r = 0; r < rank; ++r) if (static_ext(r)) m_ext[r] = static_ext(r);
  }
  // Extents from a layout object; a LayoutStride brings its strides too.
  template <class L>
  KOKKOS_INLINE_FUNCTION void set_layout(const L& layout) {
    if constexpr (std::is_same<L, LayoutStride>::value) {
      for (unsigned r = 0; r < kR; ++r) {
        m_ext[r] = r < rank ? layout.dimension[r] : 0;
        m_str[r] = r < rank ? layout.stride[r] : 1;
      }
      if constexpr (rank == 0) m_str[0] = 1;
    } else {
      size_t dyn[8];
      unsigned d = 0;
      // Layout objects list every extent, static ones included.
      for (unsigned r = 0; r < rank; ++r)
        if (!static_ext(r)) dyn[d++] = layout.dimension[r];
      for (; d < 8; ++d) dyn[d] = KOKKOS_IMPL_NOARG;
      set_extents(dyn);
    }
  }
  KOKKOS_INLINE_FUNCTION void set_extents(const size_t* dyn) {
    unsigned d = 0;
    for (unsigned r = 0; r < kR; ++r) m_ext[r] = 0;
    for (unsigned r = 0; r < rank; ++r) {
      size_t se = static_ext(r);
      if (se) {
        m_ext[r] = se;
      } else {
        size_t v = dyn[d++];
        m_ext[r] = (v == KOKKOS_IMPL_NOARG) ? 0 : v;
      }
    }
    compute_strides();
  }
  KOKKOS_INLINE_FUNCTION void compute_strides() {
    if constexpr (std::is_same<array_layout, LayoutRight>::value) {
      size_t s = 1;
      for (int r = int(rank) - 1; r >= 0; --r) { m_str[r] = s; s *= m_ext[r]; }
    } else if constexpr (std::is_same<array_layout, LayoutLeft>::value) {
      size_t s = 1;
      for (unsigned r = 0; r < rank; ++r) { m_str[r] = s; s *= m_ext[r]; }
    } else {
      size_t s = 1;
      for (int r = int(rank) - 1; r >= 0; --r) { m_str[r] = s; s *= m_ext[r]; }
    }
    if constexpr (rank == 0) m_str[0] = 1;
  }
  // Offsets: rank 1 is the bare index for LayoutLeft/LayoutRight (their
  // unit-stride dimension); higher ranks multiply the other indices by their
  // strides, which a subview may have inherited from a larger parent.
  KOKKOS_FORCEINLINE_FUNCTION size_t offset() const { return 0; }
  KOKKOS_FORCEINLINE_FUNCTION size_t offset(size_t i0) const {
    if constexpr (std::is_same<array_layout, LayoutStride>::value) return i0 * m_str[0];
    else return i0;
  }
  template <class... Is>
  KOKKOS_FORCEINLINE_FUNCTION size_t offset(size_t i0, size_t i1, Is... rest) const {
    const size_t idx[] = {i0, i1, static_cast<size_t>(rest)...};
    constexpr unsigned n = 2 + sizeof...(Is);
    if constexpr (std::is_same<array_layout, LayoutRight>::value) {
      size_t o = idx[n - 1];
      for (unsigned r = 0; r + 1 < n; ++r) o += idx[r] * m_str[r];
      return o;
    } else if constexpr (std::is_same<array_layout, LayoutLeft>::value) {
      size_t o = idx[0];
      for (unsigned r = 1; r < n; ++r) o += idx[r] * m_str[r];
      return o;
    } else {
      size_t o = 0;
      for (unsigned r = 0; r < n; ++r) o += idx[r] * m_str[r];
      return o;
    }
  }

  void allocate(const Impl::ViewCtorProp& prop) {
    static_assert(!memory_traits::is_unmanaged, "Unmanaged Views cannot allocate");
    size_t bytes = span() * sizeof(value_type);
    auto* rec    = new Impl::SharedAllocRecord();
    rec->label   = prop.label;
    rec->bytes   = bytes;
    try {
      rec->ptr = Impl::MemOps<memory_space>::alloc(prop.label.c_str(), bytes);
    } catch (...) {
      delete rec;
      throw;
    }
    rec->release = &Impl::MemOps<memory_space>::release;
    m_rec        = rec;
    m_ptr        = static_cast<pointer_type>(rec->ptr);
    if (prop.initialize && bytes) Impl::MemOps<memory_space>::zero(rec->ptr, bytes);
  }
};

template <class V>
KOKKOS_INLINE_FUNCTION constexpr unsigned rank(const V&) {
  return V::rank;
}

template <class D1, class... P1, class D2, class... P2>
KOKKOS_INLINE_FUNCTION bool operator==(const View<D1, P1...>& a, const View<D2, P2...>& b) {
  if (a.data() != b.data() || View<D1, P1...>::rank != View<D2, P2...>::rank) return false;
  for (unsigned r = 0; r < View<D1, P1...>::rank; ++r)
    if (a.extent(r) != b.extent(r)) return false;
  return true;
}
template <class D1, class... P1, class D2, class... P2>
KOKKOS_INLINE_FUNCTION bool operator!=(const View<D1, P1...>& a, const View<D2, P2...>& b) {
  return !(a == b);
}

}  // namespace Kokkos

#endif
