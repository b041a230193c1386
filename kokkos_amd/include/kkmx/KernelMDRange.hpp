// MDRangePolicy kernels for MI355X.
//
// One workgroup per tile (tile product <= 1024 lanes, rounded up to whole
// wavefronts); the tile's fastest dimension (Iterate::Left: the first index,
// matching LayoutLeft) runs along the lanes of a wavefront, so each functor
// call of a wavefront touches consecutive addresses.
//
// XCD-aware tile order: the hardware hands consecutive workgroup ids to the 8
// XCDs in turn; the kernel maps workgroup b to logical tile
// start(b % 8) + b / 8, so every XCD walks one contiguous block of tiles and
// the halo rows/planes neighbouring tiles share stay in that XCD's L2.
//
// Reference: core/src/HIP/Kokkos_HIP_ParallelFor_MDRange.hpp:29-196,
// Kokkos_HIP_ParallelReduce_MDRange.hpp:33-247, impl/KokkosExp_IterateTileGPU.hpp.
#ifndef KKMX_KERNEL_MDRANGE_HPP
#define KKMX_KERNEL_MDRANGE_HPP

#include <kkmx/KernelRange.hpp>

namespace Kokkos {
namespace Impl {

struct MDLaunch {
  int64_t lower[6];
  int64_t upper[6];
  int64_t tile[6];
  int64_t tile_end[6];  // number of tiles per dimension
  int64_t ntiles;
  int tile_points;      // product of tile extents
  int xcds;             // XCD count used by the tile remap (1 = identity)
  // Coarsening of parallel_for: each workgroup runs `depth` consecutive tiles
  // along the slowest tile dimension (loop_dim), so the planes a tile shares
  // with its successor are still in the CU's L1 and the XCD's L2 when it is
  // reached.  groups = number of workgroup tiles.
  int depth;
  int loop_dim;
  int64_t groups;
};

__device__ __forceinline__ int64_t md_logical_tile(int64_t b, int64_t ntiles, int xcds) {
  if (xcds <= 1) return b;
  const int64_t q = ntiles / xcds, r = ntiles % xcds;
  const int64_t x = b % xcds, k = b / xcds;
  return x * q + (x < r ? x : r) + k;
}

// Decompose tile index and lane into the point's indices.
template <unsigned R, bool InnerLeft, bool OuterLeft>
__device__ __forceinline__ bool md_point(const MDLaunch& p, int64_t t, int lane, int64_t* idx) {
  int64_t tc[6];
  if (OuterLeft) {
    for (unsigned d = 0; d < R; ++d) {
      tc[d] = t % p.tile_end[d];
      t /= p.tile_end[d];
    }
  } else {
    for (int d = int(R) - 1; d >= 0; --d) {
      tc[d] = t % p.tile_end[d];
      t /= p.tile_end[d];
    }
  }
  int64_t loc = lane;
  bool in     = true;
  if (InnerLeft) {
    for (unsigned d = 0; d < R; ++d) {
      const int64_t o = loc % p.tile[d];
      loc /= p.tile[d];
      idx[d] = p.lower[d] + tc[d] * p.tile[d] + o;
      in &= idx[d] < p.upper[d];
    }
  } else {
    for (int d = int(R) - 1; d >= 0; --d) {
      const int64_t o = loc % p.tile[d];
      loc /= p.tile[d];
      idx[d] = p.lower[d] + tc[d] * p.tile[d] + o;
      in &= idx[d] < p.upper[d];
    }
  }
  return in;
}

// Point of lane `lane` in repeat q of workgroup tile g (coarsened launch).
template <unsigned R, bool InnerLeft, bool OuterLeft>
__device__ __forceinline__ bool md_point_group(const MDLaunch& p, int64_t g, int q, int lane, int64_t* idx) {
  int64_t tc[6];
  const int ld = p.loop_dim;
  if (OuterLeft) {
    for (unsigned d = 0; d < R; ++d) {
      const int64_t te = (int(d) == ld) ? (p.tile_end[d] + p.depth - 1) / p.depth : p.tile_end[d];
      tc[d]            = g % te;
      g /= te;
    }
  } else {
    for (int d = int(R) - 1; d >= 0; --d) {
      const int64_t te = (d == ld) ? (p.tile_end[d] + p.depth - 1) / p.depth : p.tile_end[d];
      tc[d]            = g % te;
      g /= te;
    }
  }
  tc[ld] = tc[ld] * p.depth + q;
  if (tc[ld] >= p.tile_end[ld]) return false;
  int64_t loc = lane;
  bool in     = true;
  if (InnerLeft) {
    for (unsigned d = 0; d < R; ++d) {
      const int64_t o = loc % p.tile[d];
      loc /= p.tile[d];
      idx[d] = p.lower[d] + tc[d] * p.tile[d] + o;
      in &= idx[d] < p.upper[d];
    }
  } else {
    for (int d = int(R) - 1; d >= 0; --d) {
      const int64_t o = loc % p.tile[d];
      loc /= p.tile[d];
      idx[d] = p.lower[d] + tc[d] * p.tile[d] + o;
      in &= idx[d] < p.upper[d];
    }
  }
  return in;
}

template <class Tag, class Index, unsigned R, class F, class... Extra>
__device__ __forceinline__ void md_invoke(const F& f, const int64_t* i, Extra&... extra) {
  if constexpr (R == 2) call_functor<Tag>(f, Index(i[0]), Index(i[1]), extra...);
  else if constexpr (R == 3) call_functor<Tag>(f, Index(i[0]), Index(i[1]), Index(i[2]), extra...);
  else if constexpr (R == 4) call_functor<Tag>(f, Index(i[0]), Index(i[1]), Index(i[2]), Index(i[3]), extra...);
  else if constexpr (R == 5)
    call_functor<Tag>(f, Index(i[0]), Index(i[1]), Index(i[2]), Index(i[3]), Index(i[4]), extra...);
  else
    call_functor<Tag>(f, Index(i[0]), Index(i[1]), Index(i[2]), Index(i[3]), Index(i[4]), Index(i[5]), extra...);
}

template <class F, class Tag, class Index, unsigned R, bool InnerLeft, bool OuterLeft>
__global__ __launch_bounds__(1024) void k_md_for(const F f, const MDLaunch p) {
  if (static_cast<int>(threadIdx.x) >= p.tile_points) return;
  for (int64_t b = blockIdx.x; b < p.groups; b += gridDim.x) {
    const int64_t g = (gridDim.x >= p.groups) ? md_logical_tile(b, p.groups, p.xcds) : b;
    for (int q = 0; q < p.depth; ++q) {
      int64_t idx[6];
      if (md_point_group<R, InnerLeft, OuterLeft>(p, g, q, threadIdx.x, idx)) md_invoke<Tag, Index, R>(f, idx);
    }
  }
}

template <class CR, class Index, unsigned R, bool InnerLeft, bool OuterLeft>
__global__ __launch_bounds__(1024) void k_md_reduce(const CR cr, const MDLaunch p,
                                                    typename CR::value_type* partials, unsigned* counter,
                                                    typename CR::value_type* result) {
  using V   = typename CR::value_type;
  using Tag = typename CR::work_tag;
  __shared__ LdsSlots16<V> slots;
  __shared__ int s_last;
  const auto join = [&cr](V& a, const V& b) { cr.join(a, b); };
  V identity;
  cr.init(identity);
  V v = identity;
  if (static_cast<int>(threadIdx.x) < p.tile_points) {
    for (int64_t b = blockIdx.x; b < p.ntiles; b += gridDim.x) {
      int64_t idx[6];
      if (md_point<R, InnerLeft, OuterLeft>(p, b, threadIdx.x, idx)) md_invoke<Tag, Index, R>(cr.functor(), idx, v);
    }
  }
  v = block_reduce(v, identity, join, slots.get(), blockDim.x);
  if (gridDim.x == 1) {
    if (threadIdx.x == 0) {
      cr.final(v);
      *result = v;
    }
    return;
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = v;
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

template <class Policy>
MDLaunch md_launch(const Policy& p) {
  MDLaunch l{};
  constexpr unsigned R = Policy::rank;
  l.ntiles             = 1;
  l.tile_points        = 1;
  for (unsigned d = 0; d < R; ++d) {
    l.lower[d]    = p.m_lower[d];
    l.upper[d]    = p.m_upper[d];
    l.tile[d]     = p.m_tile[d];
    l.tile_end[d] = p.m_tile_end[d];
    l.ntiles *= p.m_tile_end[d];
    l.tile_points *= static_cast<int>(p.m_tile[d]);
  }
  for (unsigned d = R; d < 6; ++d) {
    l.lower[d] = 0;
    l.upper[d] = 1;
    l.tile[d] = l.tile_end[d] = 1;
  }
  for (unsigned d = 0; d < R; ++d)
    if (l.upper[d] <= l.lower[d]) l.ntiles = 0;
  l.xcds     = MI355X::impl_props().xcc_count > 0 ? MI355X::impl_props().xcc_count : 1;
  l.loop_dim = (Policy::outer_direction == Iterate::Left) ? int(R) - 1 : 0;
  l.depth    = 1;
  l.groups   = l.ntiles;
  return l;
}

// Tiles per workgroup along the slowest tile dimension for parallel_for.
// Heuristic: 8 while at least 16 workgroup tiles per resident workgroup slot
// remain; a tuning override (md_depth_override() > 0) replaces it.
inline int& md_depth_override() {
  static int d = 0;
  return d;
}
inline void md_set_depth(MDLaunch& l, int block) {
  const int64_t te = l.tile_end[l.loop_dim];
  int depth        = md_depth_override();
  if (depth <= 0) {
    const int cus       = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
    const int64_t slots = int64_t(cus) * (2048 / (block > 0 ? block : 64));
    depth               = 8;
    while (depth > 1 && l.ntiles / depth < 16 * slots) depth /= 2;
  }
  if (depth > te) depth = int(te > 0 ? te : 1);
  if (depth < 1) depth = 1;
  l.depth  = depth;
  l.groups = l.ntiles / te * ((te + depth - 1) / depth);
}

// MDRange tiles a tuning tool chooses from when the policy leaves the tile
// to the implementation: the fastest dimension takes 16 << (c % 4) points,
// the next 1 << (c / 4), the others 1 (reference MDRangeTuner,
// core/src/impl/Kokkos_Tools_Generic.hpp:106-122,340-353).
template <class Policy, class F>
Policy md_tuned_policy(const Policy& p, int pattern, size_t value_bytes, bool floating) {
  if (!p.impl_tune_tile_size() || !Tools::Experimental::have_tuning_tool()) return p;
  constexpr unsigned R = Policy::rank;
  const bool left      = Policy::inner_direction == Iterate::Left;
  const int fast = left ? 0 : int(R) - 1, next = left ? 1 : int(R) - 2;
  int64_t points = 1;
  for (unsigned r = 0; r < R; ++r) points *= (p.upper()[r] - p.lower()[r]);
  int h = -1;
  for (int c = 0; c < 16; ++c) {
    bool match = p.tile_size()[fast] == (16 << (c % 4)) && p.tile_size()[next] == (1 << (c / 4));
    for (int r = 0; r < int(R); ++r)
      if (r != fast && r != next && p.tile_size()[r] != 1) match = false;
    if (match) h = c;
  }
  const int c = request_launch_candidate<F>(LaunchVar::MDRange, pattern, points, value_bytes, floating, h);
  if (c < 0 || c == h) return p;
  typename Policy::tile_type t;
  for (unsigned r = 0; r < R; ++r) t[r] = 1;
  t[fast] = 16 << (c % 4);
  t[next] = 1 << (c / 4);
  Policy q = p;
  q.impl_change_tile_size(t);
  return q;
}

template <class F, class... Traits>
class ParallelFor<F, MDRangePolicy<Traits...>, MI355X> {
 public:
  using Policy             = MDRangePolicy<Traits...>;
  static constexpr unsigned R = Policy::rank;
  static constexpr bool IL = Policy::inner_direction == Iterate::Left;
  static constexpr bool OL = Policy::outer_direction == Iterate::Left;
  ParallelFor(const F& f, const Policy& p) : m_f(f), m_p(md_tuned_policy<Policy, F>(p, 0, 0, false)) {}
  void execute() const {
    MDLaunch l = md_launch(m_p);
    if (l.ntiles == 0) return;
    const int block = ((l.tile_points + 63) / 64) * 64;
    md_set_depth(l, block);
    int64_t grid    = l.groups > 0x7fffffff ? 0x7fffffff : l.groups;
    MI355XInstance* inst = m_p.space().impl_internal_space_instance();
    F f = m_f;
    const void* k = reinterpret_cast<const void*>(
        &k_md_for<F, typename Policy::work_tag, typename Policy::index_type, R, IL, OL>);
    mx_launch_kernel(k, static_cast<unsigned>(grid), block, 0, inst->stream, f, l);
  }
  // Largest tile (in points) the kernel accepts
  // (reference: ParallelFor<MDRange,HIP>::max_tile_size_product).
  template <class Pol, class Fn>
  static int max_tile_size_product(const Pol&, const Fn&) {
    return 1024;
  }

 private:
  F m_f;
  Policy m_p;
};

template <class CR, class... Traits>
class ParallelReduce<CR, MDRangePolicy<Traits...>, MI355X> {
 public:
  using Policy             = MDRangePolicy<Traits...>;
  using V                  = typename CR::value_type;
  static constexpr unsigned R = Policy::rank;
  static constexpr bool IL = Policy::inner_direction == Iterate::Left;
  static constexpr bool OL = Policy::outer_direction == Iterate::Left;
  template <class Mem>
  ParallelReduce(const CR& cr, const Policy& p, V* result, Mem)
      : m_cr(cr),
        m_p(md_tuned_policy<Policy, typename CR::functor_type>(p, 1, sizeof(V), std::is_floating_point<V>::value)),
        m_result(result),
        m_on_device(exec_can_access<MI355X, Mem>::value) {}
  void execute() const {
    static_assert(!CR::is_array, "MDRange reductions with runtime-length value types are not supported");
    MI355XInstance* inst = m_p.space().impl_internal_space_instance();
    std::lock_guard<std::mutex> lk(inst->mtx);
    MDLaunch l      = md_launch(m_p);
    const int block = ((l.tile_points + 63) / 64) * 64;
    const int cus   = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
    int64_t grid    = std::min<int64_t>(l.ntiles, static_cast<int64_t>(cus) * (2048 / block) * 2);
    if (grid < 1) grid = 1;
    V* dst         = m_result;
    bool copy_back = false;
    if (!m_on_device || !dst) {
      dst       = static_cast<V*>(inst->pinned_space(sizeof(V)));
      copy_back = (m_result != nullptr);
    }
    V* partials       = static_cast<V*>(inst->partials_space(sizeof(V) * grid));
    CR cr             = m_cr;
    unsigned* counter = inst->control;
    const void* k     = reinterpret_cast<const void*>(&k_md_reduce<CR, typename Policy::index_type, R, IL, OL>);
    mx_launch_kernel(k, static_cast<unsigned>(grid), block, 0, inst->stream, cr, l, partials, counter, dst);
    if (copy_back) {
      KKMX_SAFE_CALL(mx_stream_sync(inst->stream));
      std::memcpy(static_cast<void*>(m_result), dst, sizeof(V));
    }
  }

 private:
  CR m_cr;
  Policy m_p;
  V* m_result;
  bool m_on_device;
};

}  // namespace Impl
}  // namespace Kokkos

#endif
