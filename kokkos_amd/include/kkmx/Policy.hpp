// Execution policies: RangePolicy, MDRangePolicy<Rank<N>>, TeamPolicy, and the
// nested TeamThreadRange / ThreadVectorRange / TeamVectorRange / single
// helpers.
//
// Reference behaviour:
//  * RangePolicy: core/src/Kokkos_ExecPolicy.hpp:76-354 (bounds check :235-250
//    aborts when end < begin, chunk must be a power of two :215-218, WorkRange
//    split :303-338); default index type is execution_space::size_type
//    (core/src/impl/Kokkos_AnalyzePolicy.hpp:173-176).
//  * MDRangePolicy: core/src/KokkosExp_MDRangePolicy.hpp:158-426 (tile
//    defaults, m_tile_end, m_num_tiles; a tile product above the space's limit
//    is an error :418-424).
//  * TeamPolicy and nested ranges: core/src/Kokkos_ExecPolicy.hpp:365-980,
//    ScratchRequest :549-583.
#ifndef KKMX_POLICY_HPP
#define KKMX_POLICY_HPP

#include <kkmx/Spaces.hpp>

#include <array>
#include <tuple>
#include <initializer_list>

namespace Kokkos {

// ------------------------------------------------------------------ trait tags
struct Static {};
struct Dynamic {};
template <class T>
struct Schedule {
  using schedule_type = Schedule<T>;
  using type          = T;
};
template <class T>
struct IndexType {
  static_assert(std::is_integral<T>::value, "IndexType<T> requires an integral T");
  using index_type = IndexType<T>;
  using type       = T;
};
template <unsigned MaxT = 0, unsigned MinB = 0>
struct LaunchBounds {
  using launch_bounds                 = LaunchBounds<MaxT, MinB>;
  static constexpr unsigned maxTperB  = MaxT;
  static constexpr unsigned minBperSM = MinB;
};
struct ChunkSize {
  int value;
  explicit ChunkSize(int v) : value(v) {}
};

namespace Experimental {
struct WorkItemProperty {
  template <unsigned long P>
  struct ImplWorkItemProperty {
    static constexpr unsigned value = P;
    using work_item_property        = ImplWorkItemProperty<P>;
  };
  static constexpr ImplWorkItemProperty<0> None{};
  static constexpr ImplWorkItemProperty<1> HintLightWeight{};
  static constexpr ImplWorkItemProperty<2> HintHeavyWeight{};
  static constexpr ImplWorkItemProperty<4> HintRegular{};
  static constexpr ImplWorkItemProperty<8> HintIrregular{};
  static constexpr ImplWorkItemProperty<16> ImplForceGlobalLaunch{};
  using None_t                  = ImplWorkItemProperty<0>;
  using HintLightWeight_t       = ImplWorkItemProperty<1>;
  using HintHeavyWeight_t       = ImplWorkItemProperty<2>;
  using HintRegular_t           = ImplWorkItemProperty<4>;
  using HintIrregular_t         = ImplWorkItemProperty<8>;
  using ImplForceGlobalLaunch_t = ImplWorkItemProperty<16>;
};
template <class Policy, class Prop>
Policy require(const Policy& p, const Prop&) {
  return p;
}
struct DesiredOccupancy {
  int value;
  explicit DesiredOccupancy(int v) : value(v) {}
};
struct MaximizeOccupancy {};
template <class Policy>
Policy prefer(const Policy& p, const DesiredOccupancy&) { return p; }
template <class Policy>
Policy prefer(const Policy& p, const MaximizeOccupancy&) { return p; }
}  // namespace Experimental

enum class Iterate { Default, Left, Right };
template <unsigned N, Iterate OuterDir = Iterate::Default, Iterate InnerDir = Iterate::Default>
struct Rank {
  static_assert(N >= 2 && N <= 8, "Kokkos::Rank supports ranks 2 to 8 (MDRangePolicy: 2 to 6)");
  static constexpr unsigned rank           = N;
  static constexpr Iterate outer_direction = OuterDir;
  static constexpr Iterate inner_direction = InnerDir;
  using iteration_pattern                  = Rank<N, OuterDir, InnerDir>;
};

// Fixed-size array usable in kernels (core/src/Kokkos_Array.hpp: aggregate
// initialisation, size / empty / max_size / data, element-wise equality, free
// begin / end / get and to_array).
template <class T, size_t N>
struct Array {
  using value_type      = T;
  using size_type       = size_t;
  using difference_type = ptrdiff_t;
  using reference       = T&;
  using const_reference = const T&;
  using pointer         = T*;
  using const_pointer   = const T*;
  T m_internal_implementation_private_member_data[N ? N : 1];
  KOKKOS_INLINE_FUNCTION constexpr T& operator[](size_t i) { return m_internal_implementation_private_member_data[i]; }
  KOKKOS_INLINE_FUNCTION constexpr const T& operator[](size_t i) const { return m_internal_implementation_private_member_data[i]; }
  KOKKOS_INLINE_FUNCTION static constexpr size_t size() { return N; }
  KOKKOS_INLINE_FUNCTION static constexpr bool empty() { return N == 0; }
  KOKKOS_INLINE_FUNCTION constexpr size_t max_size() const { return N; }
  KOKKOS_INLINE_FUNCTION constexpr T* data() { return N ? m_internal_implementation_private_member_data : nullptr; }
  KOKKOS_INLINE_FUNCTION constexpr const T* data() const {
    return N ? m_internal_implementation_private_member_data : nullptr;
  }
  friend KOKKOS_INLINE_FUNCTION constexpr bool operator==(const Array& a, const Array& b) {
    for (size_t i = 0; i < N; ++i)
      if (!(a[i] == b[i])) return false;
    return true;
  }
  friend KOKKOS_INLINE_FUNCTION constexpr bool operator!=(const Array& a, const Array& b) { return !(a == b); }
};
// The empty array holds nothing, so a const one needs no initialiser.
template <class T>
struct Array<T, 0> {
  using value_type      = T;
  using size_type       = size_t;
  using difference_type = ptrdiff_t;
  using reference       = T&;
  using const_reference = const T&;
  using pointer         = T*;
  using const_pointer   = const T*;
  template <class I>
  KOKKOS_INLINE_FUNCTION T& operator[](const I&) {
    Kokkos::abort("Kokkos::Array<T, 0>: no elements");
    return *reinterpret_cast<T*>(this);
  }
  template <class I>
  KOKKOS_INLINE_FUNCTION const T& operator[](const I&) const {
    Kokkos::abort("Kokkos::Array<T, 0>: no elements");
    return *reinterpret_cast<const T*>(this);
  }
  KOKKOS_INLINE_FUNCTION static constexpr size_t size() { return 0; }
  KOKKOS_INLINE_FUNCTION static constexpr bool empty() { return true; }
  KOKKOS_INLINE_FUNCTION constexpr size_t max_size() const { return 0; }
  KOKKOS_INLINE_FUNCTION constexpr T* data() { return nullptr; }
  KOKKOS_INLINE_FUNCTION constexpr const T* data() const { return nullptr; }
  friend KOKKOS_INLINE_FUNCTION constexpr bool operator==(const Array&, const Array&) { return true; }
  friend KOKKOS_INLINE_FUNCTION constexpr bool operator!=(const Array&, const Array&) { return false; }
};
template <class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr T* begin(Array<T, N>& a) noexcept { return a.data(); }
template <class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr const T* begin(const Array<T, N>& a) noexcept { return a.data(); }
template <class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr T* end(Array<T, N>& a) noexcept { return a.data() + N; }
template <class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr const T* end(const Array<T, N>& a) noexcept { return a.data() + N; }
template <size_t I, class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr T& get(Array<T, N>& a) noexcept {
  static_assert(I < N, "Kokkos::get: index out of range");
  return a[I];
}
template <size_t I, class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr const T& get(const Array<T, N>& a) noexcept {
  static_assert(I < N, "Kokkos::get: index out of range");
  return a[I];
}
template <size_t I, class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr T&& get(Array<T, N>&& a) noexcept {
  static_assert(I < N, "Kokkos::get: index out of range");
  return static_cast<T&&>(a[I]);
}
template <class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr Array<std::remove_cv_t<T>, N> to_array(T (&a)[N]) {
  Array<std::remove_cv_t<T>, N> r{};
  for (size_t i = 0; i < N; ++i) r[i] = a[i];
  return r;
}
template <class T, size_t N>
KOKKOS_INLINE_FUNCTION constexpr Array<std::remove_cv_t<T>, N> to_array(T (&&a)[N]) {
  Array<std::remove_cv_t<T>, N> r{};
  for (size_t i = 0; i < N; ++i) r[i] = static_cast<T&&>(a[i]);
  return r;
}

namespace Impl {
// Trait a graph adds to the policy type of each kernel node it records
// (core/src/impl/Kokkos_GraphImpl_fwd.hpp:65-67); it changes nothing in how
// the kernel runs.
struct IsGraphKernelTag {};
struct KernelInGraphProperty {};

template <class T>
struct is_schedule : std::false_type {};
template <class T>
struct is_schedule<Schedule<T>> : std::true_type {};
template <class T>
struct is_index_type : std::false_type {};
template <class T>
struct is_index_type<IndexType<T>> : std::true_type {};
template <class T>
struct is_launch_bounds : std::false_type {};
template <unsigned A, unsigned B>
struct is_launch_bounds<LaunchBounds<A, B>> : std::true_type {};
template <class T>
struct is_rank : std::false_type {};
template <unsigned N, Iterate O, Iterate I>
struct is_rank<Rank<N, O, I>> : std::true_type {};
template <class T, class = void>
struct is_work_item_property : std::false_type {};
template <class T>
struct is_work_item_property<T, std::void_t<typename T::work_item_property>> : std::true_type {};

// Analyse the property pack of a policy (any order, each at most once).
template <class... P>
struct PolicyTraits;
template <>
struct PolicyTraits<> {
  using execution_space = void;
  using schedule_type   = void;
  using index_type      = void;
  using work_tag        = void;
  using launch_bounds   = void;
  using iteration_pattern = void;
};
template <class H, class... T>
struct PolicyTraits<H, T...> {
  using rest = PolicyTraits<T...>;
  static constexpr bool h_exec  = is_execution_space<H>::value;
  static constexpr bool h_sched = is_schedule<H>::value;
  static constexpr bool h_index = is_index_type<H>::value || std::is_integral<H>::value;
  static constexpr bool h_lb    = is_launch_bounds<H>::value;
  static constexpr bool h_rank  = is_rank<H>::value;
  static constexpr bool h_wip   = is_work_item_property<H>::value;
  static constexpr bool h_graph = std::is_same<H, IsGraphKernelTag>::value;
  static constexpr bool h_tag   = !(h_exec || h_sched || h_index || h_lb || h_rank || h_wip || h_graph);
  template <class X, bool B = std::is_integral<X>::value>
  struct idx {
    using type = IndexType<X>;
  };
  template <class X>
  struct idx<X, false> {
    using type = X;
  };
  using execution_space   = std::conditional_t<h_exec, H, typename rest::execution_space>;
  using schedule_type     = std::conditional_t<h_sched, H, typename rest::schedule_type>;
  using index_type        = std::conditional_t<h_index, typename idx<H>::type, typename rest::index_type>;
  using work_tag          = std::conditional_t<h_tag, H, typename rest::work_tag>;
  using launch_bounds     = std::conditional_t<h_lb, H, typename rest::launch_bounds>;
  using iteration_pattern = std::conditional_t<h_rank, H, typename rest::iteration_pattern>;
};

template <class... P>
struct PolicyTraitsWithDefaults {
  using raw             = PolicyTraits<P...>;
  using execution_space = std::conditional_t<std::is_void<typename raw::execution_space>::value,
                                             DefaultExecutionSpace, typename raw::execution_space>;
  using schedule_type   = std::conditional_t<std::is_void<typename raw::schedule_type>::value,
                                           Schedule<Static>, typename raw::schedule_type>;
  using index_type      = typename std::conditional_t<std::is_void<typename raw::index_type>::value,
                                                 IndexType<typename execution_space::size_type>,
                                                 typename raw::index_type>::type;
  using work_tag        = typename raw::work_tag;
  using launch_bounds   = std::conditional_t<std::is_void<typename raw::launch_bounds>::value,
                                           LaunchBounds<>, typename raw::launch_bounds>;
  using iteration_pattern = typename raw::iteration_pattern;
};
}  // namespace Impl

// ------------------------------------------------------------------ RangePolicy
template <class... Properties>
class RangePolicy : public Impl::PolicyTraitsWithDefaults<Properties...> {
 public:
  using traits          = Impl::PolicyTraitsWithDefaults<Properties...>;
  using execution_space = typename traits::execution_space;
  using work_tag        = typename traits::work_tag;
  using schedule_type   = typename traits::schedule_type;
  using launch_bounds   = typename traits::launch_bounds;
  using index_type      = typename traits::index_type;
  using member_type     = index_type;
  using execution_policy = RangePolicy<Properties...>;
  using impl_policy_kind = RangePolicy<>;

  RangePolicy() : m_space(), m_begin(0), m_end(0), m_chunk(0) {}

  template <class I1, class I2,
            class = std::enable_if_t<std::is_convertible<I1, member_type>::value &&
                                     std::is_convertible<I2, member_type>::value>>
  RangePolicy(const I1 b, const I2 e) : RangePolicy(execution_space(), b, e) {}

  template <class I1, class I2,
            class = std::enable_if_t<std::is_convertible<I1, member_type>::value &&
                                     std::is_convertible<I2, member_type>::value>>
  RangePolicy(const execution_space& space, const I1 b, const I2 e)
      : m_space(space), m_begin(static_cast<member_type>(b)), m_end(static_cast<member_type>(e)), m_chunk(0) {
    check_conversion(b);
    check_conversion(e);
    check_bounds();
  }
  template <class I1, class I2>
  RangePolicy(const I1 b, const I2 e, const ChunkSize c) : RangePolicy(execution_space(), b, e) {
    set_chunk_size(c.value);
  }
  template <class I1, class I2>
  RangePolicy(const execution_space& space, const I1 b, const I2 e, const ChunkSize c)
      : RangePolicy(space, b, e) {
    set_chunk_size(c.value);
  }

  KOKKOS_INLINE_FUNCTION member_type begin() const { return m_begin; }
  KOKKOS_INLINE_FUNCTION member_type end() const { return m_end; }
  const execution_space& space() const { return m_space; }
  // Rebind to another instance of the same space (used by graph nodes).
  void impl_set_space(const execution_space& s) { m_space = s; }
  member_type chunk_size() const { return m_chunk; }
  RangePolicy& set_chunk_size(int c) {
    if (c > 0 && !Impl::is_pow2(static_cast<uint64_t>(c)))
      Kokkos::abort("RangePolicy blocking granularity must be power of two");
    m_chunk = c;
    return *this;
  }
  template <class... OtherProps>
  RangePolicy(const RangePolicy<OtherProps...>& o)
      : m_space(o.space()), m_begin(o.begin()), m_end(o.end()), m_chunk(o.chunk_size()),
        m_occupancy(o.impl_get_desired_occupancy().value) {}
  // Desired occupancy in percent, as RangePolicyOccupancyTuner sets it.  The
  // MI355X launches size their grids from the tile count, so the value is
  // kept for the tuner's bookkeeping and does not change the launch.
  RangePolicy& impl_set_desired_occupancy(const Experimental::DesiredOccupancy& o) {
    m_occupancy = o.value;
    return *this;
  }
  Experimental::DesiredOccupancy impl_get_desired_occupancy() const {
    return Experimental::DesiredOccupancy{m_occupancy};
  }

 private:
  execution_space m_space;
  member_type m_begin, m_end, m_chunk;
  int m_occupancy = 100;

  void check_bounds() {
    if (m_end < m_begin) {
      std::string msg = "Kokkos::RangePolicy bounds error: The lower bound (" +
                        std::to_string(m_begin) + ") is greater than the upper bound (" +
                        std::to_string(m_end) + ").\n";
      Kokkos::abort(msg.c_str());
    }
  }
  template <class I>
  static void check_conversion(const I bound) {
    if constexpr (std::is_arithmetic<I>::value) {
      bool bad = static_cast<I>(static_cast<member_type>(bound)) != bound;
      if constexpr (std::is_signed<I>::value && !std::is_signed<member_type>::value) bad |= bound < 0;
      if constexpr (!std::is_signed<I>::value && std::is_signed<member_type>::value)
        bad |= bound > static_cast<I>(std::numeric_limits<member_type>::max());
      if (bad) {
        std::string msg =
            "Kokkos::RangePolicy bound type error: an unsafe implicit conversion is performed on a bound (" +
            std::to_string(bound) + "), which may not preserve its original value.\n";
        Kokkos::abort(msg.c_str());
      }
    }
  }
};

// ------------------------------------------------------------------ MDRangePolicy
// Iteration directions an execution space uses when Rank<N> leaves them to it
// (core/src/KokkosExp_MDRangePolicy.hpp:49-60).  On MI355X both are Left:
// the first index is fastest, as in LayoutLeft Views, and it runs along the
// 64 lanes of a wavefront.
template <class ExecSpace>
struct default_outer_direction {
  using type                     = Iterate;
  static constexpr Iterate value = Iterate::Right;
};
template <class ExecSpace>
struct default_inner_direction {
  using type                     = Iterate;
  static constexpr Iterate value = Iterate::Right;
};
template <>
struct default_outer_direction<MI355X> {
  using type                     = Iterate;
  static constexpr Iterate value = Iterate::Left;
};
template <>
struct default_inner_direction<MI355X> {
  using type                     = Iterate;
  static constexpr Iterate value = Iterate::Left;
};

namespace Impl {
// Tile limits and defaults of an execution space
// (core/src/KokkosExp_MDRangePolicy.hpp:137-156): the largest tile product a
// workgroup runs, the extent given to the fastest dimension and to each other
// one when the caller leaves the tile to the implementation.  Host spaces
// take the whole fastest extent.
struct TileSizeProperties {
  int max_threads;
  int default_largest_tile_size;
  int default_tile_size;
  int max_total_tile_size;
};
template <class ExecutionSpace>
TileSizeProperties get_tile_size_properties(const ExecutionSpace&) {
  TileSizeProperties t;
  t.max_threads               = std::numeric_limits<int>::max();
  t.default_largest_tile_size = 0;
  t.default_tile_size         = 2;
  t.max_total_tile_size       = std::numeric_limits<int>::max();
  return t;
}
// MI355X: one lane per tile point, at most 1024 lanes per workgroup; a full
// wavefront along the fastest dimension.
template <>
inline TileSizeProperties get_tile_size_properties<MI355X>(const MI355X&) {
  TileSizeProperties t;
  t.max_threads               = 1024;
  t.default_largest_tile_size = 64;
  t.default_tile_size         = 2;
  t.max_total_tile_size       = 1024;
  return t;
}
}  // namespace Impl

template <class... Properties>
class MDRangePolicy : public Impl::PolicyTraitsWithDefaults<Properties...> {
 public:
  using traits            = Impl::PolicyTraitsWithDefaults<Properties...>;
  using execution_space   = typename traits::execution_space;
  using work_tag          = typename traits::work_tag;
  using launch_bounds     = typename traits::launch_bounds;
  using iteration_pattern = typename traits::iteration_pattern;
  static_assert(!std::is_void<iteration_pattern>::value, "MDRangePolicy requires a Kokkos::Rank<N>");
  static constexpr unsigned rank = iteration_pattern::rank;
  static_assert(rank <= 6, "Kokkos::MDRangePolicy supports ranks 2 to 6");
  using index_type        = typename traits::index_type;
  using array_index_type  = std::int64_t;  // as the reference, whatever index_type is
  using point_type        = Kokkos::Array<array_index_type, rank>;
  using tile_type         = Kokkos::Array<array_index_type, rank>;
  using execution_policy  = MDRangePolicy<Properties...>;
  using impl_policy_kind  = MDRangePolicy<>;
  static constexpr Iterate outer_direction = iteration_pattern::outer_direction == Iterate::Default
                                                 ? default_outer_direction<execution_space>::value
                                                 : iteration_pattern::outer_direction;
  static constexpr Iterate inner_direction = iteration_pattern::inner_direction == Iterate::Default
                                                 ? default_inner_direction<execution_space>::value
                                                 : iteration_pattern::inner_direction;

  MDRangePolicy() = default;

  template <class LT, class UT, class TT = array_index_type>
  MDRangePolicy(std::initializer_list<LT> lower, std::initializer_list<UT> upper,
                std::initializer_list<TT> tile = {})
      : MDRangePolicy(execution_space(), lower, upper, tile) {}
  template <class LT, class UT, class TT = array_index_type>
  MDRangePolicy(const execution_space& space, std::initializer_list<LT> lower,
                std::initializer_list<UT> upper, std::initializer_list<TT> tile = {})
      : m_space(space) {
    if (lower.size() != rank || upper.size() != rank || (tile.size() != 0 && tile.size() != rank))
      Kokkos::abort("Kokkos::MDRangePolicy: the number of bounds/tiles must match the rank");
    unsigned r = 0;
    for (auto v : lower) m_lower[r++] = static_cast<array_index_type>(v);
    r = 0;
    for (auto v : upper) m_upper[r++] = static_cast<array_index_type>(v);
    r = 0;
    for (auto v : tile) m_tile[r++] = static_cast<array_index_type>(v);
    for (; r < rank; ++r) m_tile[r] = 0;
    init_helper();
  }
  // Double-braced bounds: MDRangePolicy({{0, 0}}, {{n, m}}).
  MDRangePolicy(const point_type& lower, const point_type& upper, const tile_type& tile = tile_type{})
      : MDRangePolicy(execution_space(), lower, upper, tile) {}
  MDRangePolicy(const execution_space& space, const point_type& lower, const point_type& upper,
                const tile_type& tile = tile_type{})
      : m_space(space), m_lower(lower), m_upper(upper), m_tile(tile) {
    init_helper();
  }
  // Bounds as Kokkos::Array of rank entries; the tile may list fewer
  // dimensions (the rest are chosen at launch) and use another index type.
  template <class AT, size_t N, class TT = AT, size_t M = N>
  MDRangePolicy(const Kokkos::Array<AT, N>& lower, const Kokkos::Array<AT, N>& upper,
                const Kokkos::Array<TT, M>& tile = Kokkos::Array<TT, M>{})
      : MDRangePolicy(execution_space(), lower, upper, tile) {}
  template <class AT, size_t N, class TT = AT, size_t M = N>
  MDRangePolicy(const execution_space& space, const Kokkos::Array<AT, N>& lower,
                const Kokkos::Array<AT, N>& upper, const Kokkos::Array<TT, M>& tile = Kokkos::Array<TT, M>{})
      : m_space(space) {
    static_assert(N == rank, "Kokkos::MDRangePolicy: bounds must have rank entries");
    static_assert(M <= rank, "Kokkos::MDRangePolicy: the tile has more entries than the rank");
    for (unsigned r = 0; r < rank; ++r) {
      m_lower[r] = lower[r];
      m_upper[r] = upper[r];
      m_tile[r]  = r < M ? tile[r] : 0;
    }
    init_helper();
  }

  const execution_space& space() const { return m_space; }
  void impl_set_space(const execution_space& s) { m_space = s; }
  KOKKOS_INLINE_FUNCTION const point_type& lower() const { return m_lower; }
  KOKKOS_INLINE_FUNCTION const point_type& upper() const { return m_upper; }
  KOKKOS_INLINE_FUNCTION const tile_type& tile_size() const { return m_tile; }
  KOKKOS_INLINE_FUNCTION array_index_type num_tiles(unsigned r) const { return m_tile_end[r]; }
  KOKKOS_INLINE_FUNCTION int64_t m_num_tiles_total() const { return m_num_tiles; }
  int64_t num_tiles() const { return m_num_tiles; }
  bool impl_tune_tile_size() const { return m_tune; }
  // Bit r set: tile[r] was left to the implementation.
  unsigned impl_unset_tile_mask() const { return m_tile_unset; }
  // Replace the tile shape (used by the tuning hooks when tiles were left to
  // the implementation).
  void impl_change_tile_size(const tile_type& t) {
    m_tile = t;
    recompute();
  }
  int64_t impl_tile_product() const {
    int64_t p = 1;
    for (unsigned r = 0; r < rank; ++r) p *= m_tile[r];
    return p;
  }
  // The tile shape the space recommends for these bounds, from the fastest
  // dimension outwards: the fastest takes default_largest_tile_size (the
  // whole extent when that is 0), every other one default_tile_size while the
  // product stays below max_total_tile_size, else 1.
  tile_type tile_size_recommended() const {
    const Impl::TileSizeProperties pr = Impl::get_tile_size_properties(m_space);
    tile_type t{};
    const bool left = inner_direction == Iterate::Left;
    int64_t prod    = 1;
    for (unsigned k = 0; k < rank; ++k) {
      const unsigned r = left ? k : rank - 1 - k;
      if (k == 0) {
        const int64_t ext = m_upper[r] - m_lower[r];
        t[r] = pr.default_largest_tile_size == 0 ? (ext > 1 ? ext : 1) : pr.default_largest_tile_size;
      } else {
        t[r] = prod * pr.default_tile_size < pr.max_total_tile_size ? pr.default_tile_size : 1;
      }
      prod *= t[r];
    }
    return t;
  }
  int max_total_tile_size() const { return Impl::get_tile_size_properties(m_space).max_total_tile_size; }

  point_type m_lower{};
  point_type m_upper{};
  tile_type m_tile{};
  point_type m_tile_end{};
  int64_t m_num_tiles = 0;
  int64_t m_prod_tile_dims = 0;
  bool m_tune = false;
  unsigned m_tile_unset = 0;  // bit r: the caller left tile[r] to the implementation

 private:
  execution_space m_space;

  void init_helper() {
    for (unsigned r = 0; r < rank; ++r)
      if (m_upper[r] < m_lower[r]) {
        std::string msg = "Kokkos::MDRangePolicy bounds error: The lower bound (" +
                          std::to_string(m_lower[r]) + ") is greater than its upper bound (" +
                          std::to_string(m_upper[r]) + ") in dimension " + std::to_string(r) + ".\n";
        Kokkos::abort(msg.c_str());
      }
    // Tiles left to the implementation (any entry <= 0) take the
    // recommended shape of the space's tile properties; the launch may
    // still pick another shape for them (the tuning hooks, or the MI355X
    // heuristic per pattern), which m_tune records.
    for (unsigned r = 0; r < rank; ++r)
      if (m_tile[r] <= 0) {
        m_tune = true;
        m_tile_unset |= 1u << r;
      }
    if (m_tune) {
      const tile_type rec = tile_size_recommended();
      for (unsigned r = 0; r < rank; ++r)
        if (m_tile[r] <= 0) m_tile[r] = rec[r];
    }
    recompute();
  }
  void recompute() {
    m_num_tiles      = 1;
    m_prod_tile_dims = 1;
    for (unsigned r = 0; r < rank; ++r) {
      int64_t ext    = m_upper[r] - m_lower[r];
      m_tile_end[r]  = static_cast<array_index_type>((ext + m_tile[r] - 1) / m_tile[r]);
      m_num_tiles   *= m_tile_end[r];
      m_prod_tile_dims *= m_tile[r];
    }
    const int max_total = Impl::get_tile_size_properties(m_space).max_total_tile_size;
    if (m_prod_tile_dims > max_total) {
      Kokkos::Impl::throw_runtime_exception(
          std::string("Kokkos::MDRange Error: Exceeded rank bounds with ") + execution_space::name() +
          " execution space: tile dimensions product " + std::to_string(m_prod_tile_dims) +
          " is larger than the maximum of " + std::to_string(max_total));
    }
  }
};

// ------------------------------------------------------------------ scratch requests
namespace Impl {
struct PerTeamValue {
  size_t value;
  explicit PerTeamValue(size_t v) : value(v) {}
};
struct PerThreadValue {
  size_t value;
  explicit PerThreadValue(size_t v) : value(v) {}
};
}  // namespace Impl
inline Impl::PerTeamValue PerTeam(const size_t v) { return Impl::PerTeamValue(v); }
inline Impl::PerThreadValue PerThread(const size_t v) { return Impl::PerThreadValue(v); }

template <class Member>
struct PerTeamHandle {
  const Member& m;
};
template <class Member>
struct PerThreadHandle {
  const Member& m;
};
namespace Impl {
template <class Member>
struct SinglePerTeam {
  const Member& team;
};
template <class Member>
struct SinglePerThread {
  const Member& team;
};
}  // namespace Impl

// ------------------------------------------------------------------ TeamPolicy
namespace Impl {
template <class Space>
class TeamMember;  // specialised per execution space
}

template <class... Properties>
class TeamPolicy : public Impl::PolicyTraitsWithDefaults<Properties...> {
 public:
  using traits          = Impl::PolicyTraitsWithDefaults<Properties...>;
  using execution_space = typename traits::execution_space;
  using work_tag        = typename traits::work_tag;
  using schedule_type   = typename traits::schedule_type;
  using launch_bounds   = typename traits::launch_bounds;
  using index_type      = typename traits::index_type;
  using member_type     = Impl::TeamMember<execution_space>;
  using execution_policy = TeamPolicy<Properties...>;
  using impl_policy_kind = TeamPolicy<>;

  TeamPolicy() = default;
  TeamPolicy(const execution_space& space, int league, int team, int vector = 1)
      : m_space(space), m_league(league), m_team(team), m_vector(vector) { check(); }
  TeamPolicy(const execution_space& space, int league, const AUTO_t&, int vector = 1)
      : m_space(space), m_league(league), m_team(-1), m_vector(vector), m_team_auto(true) { check(); }
  TeamPolicy(const execution_space& space, int league, const AUTO_t&, const AUTO_t&)
      : m_space(space), m_league(league), m_team(-1), m_vector(1), m_team_auto(true), m_vector_auto(true) { check(); }
  TeamPolicy(const execution_space& space, int league, int team, const AUTO_t&)
      : m_space(space), m_league(league), m_team(team), m_vector(1), m_vector_auto(true) { check(); }
  TeamPolicy(int league, int team, int vector = 1) : TeamPolicy(execution_space(), league, team, vector) {}
  TeamPolicy(int league, const AUTO_t& a, int vector = 1) : TeamPolicy(execution_space(), league, a, vector) {}
  TeamPolicy(int league, const AUTO_t& a, const AUTO_t& b) : TeamPolicy(execution_space(), league, a, b) {}
  TeamPolicy(int league, int team, const AUTO_t& a) : TeamPolicy(execution_space(), league, team, a) {}
  template <class... OtherProps>
  TeamPolicy(const TeamPolicy<OtherProps...>& o)
      : m_space(o.space()), m_league(o.league_size()), m_team(o.impl_team_size_raw()),
        m_vector(o.impl_vector_length()), m_team_auto(o.impl_auto_team_size()),
        m_vector_auto(o.impl_auto_vector_length()), m_chunk(o.chunk_size()) {
    for (int l = 0; l < 2; ++l) {
      m_team_scratch[l]   = o.team_scratch_size(l);
      m_thread_scratch[l] = o.thread_scratch_size(l);
    }
  }

  const execution_space& space() const { return m_space; }
  void impl_set_space(const execution_space& s) { m_space = s; }
  int league_size() const { return m_league; }
  int team_size() const { return m_team > 0 ? m_team : impl_default_team_size(); }
  int impl_team_size_raw() const { return m_team; }
  int impl_vector_length() const { return m_vector; }
  int vector_length() const { return m_vector; }
  bool impl_auto_team_size() const { return m_team_auto; }
  bool impl_auto_vector_length() const { return m_vector_auto; }
  int chunk_size() const { return m_chunk; }
  TeamPolicy& set_chunk_size(int c) { m_chunk = c; return *this; }
  TeamPolicy& impl_set_team_size(int t) { m_team = t; return *this; }
  TeamPolicy& impl_set_vector_length(int v) { m_vector = v; return *this; }

  TeamPolicy& set_scratch_size(int level, const Impl::PerTeamValue& t) {
    check_level(level);
    m_team_scratch[level] = t.value;
    return *this;
  }
  TeamPolicy& set_scratch_size(int level, const Impl::PerThreadValue& t) {
    check_level(level);
    m_thread_scratch[level] = t.value;
    return *this;
  }
  TeamPolicy& set_scratch_size(int level, const Impl::PerTeamValue& t, const Impl::PerThreadValue& th) {
    check_level(level);
    m_team_scratch[level]   = t.value;
    m_thread_scratch[level] = th.value;
    return *this;
  }
  TeamPolicy& set_scratch_size(int level, const Impl::PerThreadValue& th, const Impl::PerTeamValue& t) {
    return set_scratch_size(level, t, th);
  }
  size_t team_scratch_size(int level) const { return m_team_scratch[level]; }
  size_t thread_scratch_size(int level) const { return m_thread_scratch[level]; }
  size_t scratch_size(int level, int team = -1) const {
    if (team < 0) team = team_size();
    return m_team_scratch[level] + size_t(team) * m_thread_scratch[level];
  }

  // Limits of the MI355X team (one workgroup): 1024 lanes, vector lanes
  // within one wavefront, 160 KiB of LDS per CU minus what the runtime keeps.
  template <class F, class Tag>
  int team_size_max(const F&, const Tag&) const { return impl_max_threads() / m_vector; }
  template <class F, class Tag>
  int team_size_recommended(const F&, const Tag&) const { return impl_recommended() / m_vector; }
  template <class F>
  int team_size_max(const F& f) const { return team_size_max(f, 0); }
  template <class F>
  int team_size_recommended(const F& f) const { return team_size_recommended(f, 0); }
  static int vector_length_max() { return std::is_same<execution_space, MI355X>::value ? 64 : 1; }
  static int scratch_size_max(int level) {
    if (std::is_same<execution_space, MI355X>::value)
      return level == 0 ? (160 * 1024 - 4096) : (1 << 30);
    return 1 << 30;
  }
  int impl_default_team_size() const { return impl_recommended() / (m_vector > 0 ? m_vector : 1); }
  static int impl_max_threads() { return std::is_same<execution_space, MI355X>::value ? 1024 : 1; }
  static int impl_recommended() { return std::is_same<execution_space, MI355X>::value ? 256 : 1; }

 private:
  execution_space m_space{};
  int m_league = 0, m_team = 0, m_vector = 1;
  bool m_team_auto = false, m_vector_auto = false;
  int m_chunk = 0;
  size_t m_team_scratch[2]   = {0, 0};
  size_t m_thread_scratch[2] = {0, 0};

  void check() {
    if (m_league < 0) Kokkos::abort("Kokkos::TeamPolicy: league size must be non-negative");
    if (m_vector <= 0 || !Impl::is_pow2(static_cast<uint64_t>(m_vector)) || m_vector > vector_length_max()) {
      if (std::is_same<execution_space, MI355X>::value)
        Kokkos::Impl::throw_runtime_exception(
            "Kokkos::TeamPolicy: the vector length must be a power of two no larger than 64 on MI355X");
      m_vector = 1;
    }
    if (m_team > 0 && m_team * m_vector > impl_max_threads() && std::is_same<execution_space, MI355X>::value)
      Kokkos::Impl::throw_runtime_exception(
          "Kokkos::TeamPolicy: team_size * vector_length exceeds 1024 lanes on MI355X");
  }
  static void check_level(int level) {
    if (level < 0 || level > 1) Kokkos::abort("TeamPolicy::set_scratch_size: invalid level");
  }
};

}  // namespace Kokkos

// Structured bindings of Kokkos::Array.
template <class T, size_t N>
struct std::tuple_size<Kokkos::Array<T, N>> : std::integral_constant<size_t, N> {};
template <size_t I, class T, size_t N>
struct std::tuple_element<I, Kokkos::Array<T, N>> {
  using type = T;
};


#endif
