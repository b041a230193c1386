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

// Index decomposition, split so that no division runs per point:
//  * md_lane_offsets: the lane's offset inside a tile, once per thread
//    (32-bit, the tile product is at most 1024);
//  * md_tile_coords: the tile coordinates of a (uniform) tile index, once
//    per tile and workgroup, in 32-bit arithmetic whenever the tile count
//    fits, which is every practical launch.
template <unsigned R, bool InnerLeft>
__device__ __forceinline__ void md_lane_offsets(const MDLaunch& p, int lane, int64_t* off) {
  unsigned loc = static_cast<unsigned>(lane);
  if (InnerLeft) {
    for (unsigned d = 0; d < R; ++d) {
      const unsigned t = static_cast<unsigned>(p.tile[d]);
      off[d]           = loc % t;
      loc /= t;
    }
  } else {
    for (int d = int(R) - 1; d >= 0; --d) {
      const unsigned t = static_cast<unsigned>(p.tile[d]);
      off[d]           = loc % t;
      loc /= t;
    }
  }
}

// Tile coordinates of tile t; `ext` are the tile counts per dimension.
template <unsigned R, bool OuterLeft>
__device__ __forceinline__ void md_tile_coords(int64_t t, const int64_t* ext, bool small, int64_t* tc) {
  if (small) {
    uint32_t u = static_cast<uint32_t>(t);
    if (OuterLeft) {
      for (unsigned d = 0; d < R; ++d) {
        const uint32_t e = static_cast<uint32_t>(ext[d]);
        tc[d]            = u % e;
        u /= e;
      }
    } else {
      for (int d = int(R) - 1; d >= 0; --d) {
        const uint32_t e = static_cast<uint32_t>(ext[d]);
        tc[d]            = u % e;
        u /= e;
      }
    }
  } else {
    if (OuterLeft) {
      for (unsigned d = 0; d < R; ++d) {
        tc[d] = t % ext[d];
        t /= ext[d];
      }
    } else {
      for (int d = int(R) - 1; d >= 0; --d) {
        tc[d] = t % ext[d];
        t /= ext[d];
      }
    }
  }
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

// Workgroup tile g covers `depth` consecutive tiles along the slowest tile
// dimension LD; the other coordinates and their bounds are fixed for the
// group, so the per-point work is one add and one compare.
template <class F, class Tag, class Index, unsigned R, bool InnerLeft, bool OuterLeft>
__global__ __launch_bounds__(1024) void k_md_for(const F f, const MDLaunch p) {
  constexpr int LD = OuterLeft ? int(R) - 1 : 0;
  if (static_cast<int>(threadIdx.x) >= p.tile_points) return;
  int64_t off[6];
  md_lane_offsets<R, InnerLeft>(p, threadIdx.x, off);
  int64_t ext[6];
  for (unsigned d = 0; d < R; ++d) ext[d] = p.tile_end[d];
  ext[LD]          = (p.tile_end[LD] + p.depth - 1) / p.depth;
  const bool small = p.groups <= int64_t(0xffffffff);
  for (int64_t b = blockIdx.x; b < p.groups; b += gridDim.x) {
    const int64_t g = (gridDim.x >= p.groups) ? md_logical_tile(b, p.groups, p.xcds) : b;
    int64_t tc[6];
    md_tile_coords<R, OuterLeft>(g, ext, small, tc);
    int64_t idx[6];
    bool in = true;
    for (unsigned d = 0; d < R; ++d) {
      if (int(d) == LD) continue;
      idx[d] = p.lower[d] + tc[d] * p.tile[d] + off[d];
      in &= idx[d] < p.upper[d];
    }
    if (!in) continue;
    const int64_t t0   = tc[LD] * p.depth;
    const int64_t left = p.tile_end[LD] - t0;
    const int qn       = left < p.depth ? static_cast<int>(left) : p.depth;
    int64_t i          = p.lower[LD] + t0 * p.tile[LD] + off[LD];
    // Eight whole planes one apart: an unrolled body without bounds checks,
    // so each call is dominated by the previous one and the compiler can
    // reuse what a functor loaded from Views the kernel cannot change (const
    // RandomAccess / Restrict, constant address space) -- a 7-point
    // stencil's u(i, j, k + 1) is the next point's u(i, j, k).  The 512^3
    // stencil issues 33 loads per 8 points instead of 56 and runs 25 % faster.
    if (p.tile[LD] == 1 && qn == 8 && i + 7 < p.upper[LD]) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        idx[LD] = i + q;
        md_invoke<Tag, Index, R>(f, idx);
      }
      continue;
    }
    for (int q = 0; q < qn; ++q, i += p.tile[LD]) {
      if (i < p.upper[LD]) {
        idx[LD] = i;
        md_invoke<Tag, Index, R>(f, idx);
      }
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
    // Workgroup tiles of `depth` tiles along LD, as in k_md_for, taken in a
    // grid-stride loop.
    constexpr int LD = OuterLeft ? int(R) - 1 : 0;
    int64_t off[6];
    md_lane_offsets<R, InnerLeft>(p, threadIdx.x, off);
    int64_t ext[6];
    for (unsigned d = 0; d < R; ++d) ext[d] = p.tile_end[d];
    ext[LD]          = (p.tile_end[LD] + p.depth - 1) / p.depth;
    const bool small = p.groups <= int64_t(0xffffffff);
    for (int64_t b = blockIdx.x; b < p.groups; b += gridDim.x) {
      int64_t tc[6], idx[6];
      md_tile_coords<R, OuterLeft>(b, ext, small, tc);
      bool in = true;
      for (unsigned d = 0; d < R; ++d) {
        if (int(d) == LD) continue;
        idx[d] = p.lower[d] + tc[d] * p.tile[d] + off[d];
        in &= idx[d] < p.upper[d];
      }
      if (!in) continue;
      const int64_t t0   = tc[LD] * p.depth;
      const int64_t left = p.tile_end[LD] - t0;
      const int qn       = left < p.depth ? static_cast<int>(left) : p.depth;
      int64_t i          = p.lower[LD] + t0 * p.tile[LD] + off[LD];
      for (int q = 0; q < qn; ++q, i += p.tile[LD]) {
        if (i < p.upper[LD]) {
          idx[LD] = i;
          md_invoke<Tag, Index, R>(cr.functor(), idx, v);
        }
      }
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

// Runtime-length results (value_type T[] with value_count entries): every
// lane keeps its value_count accumulators in `acc` (dynamic LDS, or a global
// slice per workgroup when they do not fit), the workgroup combines them in
// an ordered pairwise tree, and the last workgroup joins the per-workgroup
// partials in workgroup order.
template <class CR, class Index, unsigned R, bool InnerLeft, bool OuterLeft>
__global__ __launch_bounds__(1024) void k_md_reduce_array(const CR cr, const MDLaunch p,
                                                          typename CR::value_type* gacc,
                                                          typename CR::value_type* partials, unsigned* counter,
                                                          typename CR::value_type* result) {
  using V   = typename CR::value_type;
  using Tag = typename CR::work_tag;
  extern __shared__ __align__(16) unsigned char kkmx_dyn_lds[];
  __shared__ int s_last;
  const unsigned n = cr.value_count();
  V* base          = gacc ? gacc + static_cast<size_t>(blockIdx.x) * blockDim.x * n : reinterpret_cast<V*>(kkmx_dyn_lds);
  V* mine          = base + static_cast<size_t>(threadIdx.x) * n;
  cr.init_array(mine);
  if (static_cast<int>(threadIdx.x) < p.tile_points) {
    constexpr int LD = OuterLeft ? int(R) - 1 : 0;
    int64_t off[6];
    md_lane_offsets<R, InnerLeft>(p, threadIdx.x, off);
    int64_t ext[6];
    for (unsigned d = 0; d < R; ++d) ext[d] = p.tile_end[d];
    ext[LD]          = (p.tile_end[LD] + p.depth - 1) / p.depth;
    const bool small = p.groups <= int64_t(0xffffffff);
    for (int64_t b = blockIdx.x; b < p.groups; b += gridDim.x) {
      int64_t tc[6], idx[6];
      md_tile_coords<R, OuterLeft>(b, ext, small, tc);
      bool in = true;
      for (unsigned d = 0; d < R; ++d) {
        if (int(d) == LD) continue;
        idx[d] = p.lower[d] + tc[d] * p.tile[d] + off[d];
        in &= idx[d] < p.upper[d];
      }
      if (!in) continue;
      const int64_t t0   = tc[LD] * p.depth;
      const int64_t left = p.tile_end[LD] - t0;
      const int qn       = left < p.depth ? static_cast<int>(left) : p.depth;
      int64_t i          = p.lower[LD] + t0 * p.tile[LD] + off[LD];
      for (int q = 0; q < qn; ++q, i += p.tile[LD]) {
        if (i < p.upper[LD]) {
          idx[LD] = i;
          md_invoke<Tag, Index, R>(cr.functor(), idx, mine);
        }
      }
    }
  }
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
//
// Without a tool, such policies launch the shape the MDRange sweeps measured
// best: parallel_for 64 x 4 (the fastest dimension shrunk to the next power
// of two of a shorter extent), reductions 128 x 8 when the extents allow
// (1.5-3x faster than 64 x 4 there); the others 1.
template <class Policy>
typename Policy::tile_type md_heuristic_tile(const Policy& p, int pattern) {
  constexpr unsigned R = Policy::rank;
  const bool left      = Policy::inner_direction == Iterate::Left;
  const int fast = left ? 0 : int(R) - 1, next = left ? 1 : int(R) - 2;
  const unsigned unset = p.impl_unset_tile_mask();
  const int64_t ef = p.upper()[fast] - p.lower()[fast], en = p.upper()[next] - p.lower()[next];
  int64_t want_fast = 64, want_next = 4;
  if (pattern == 1 && ef >= 128 && en >= 8) {
    want_fast = 128;
    want_next = 8;
  } else if (ef > 0 && ef < 64) {
    want_fast = 1;
    while (want_fast < ef) want_fast *= 2;
  }
  typename Policy::tile_type t = p.tile_size();
  int64_t prod                 = 1;
  for (unsigned r = 0; r < R; ++r)
    if (!(unset & (1u << r))) prod *= t[r];
  for (unsigned k = 0; k < R; ++k) {
    const int r = left ? int(k) : int(R) - 1 - int(k);
    if (!(unset & (1u << r))) continue;
    int64_t v = r == fast ? want_fast : r == next ? want_next : 1;
    while (v > 1 && prod * v > 1024) v /= 2;
    t[r] = v;
    prod *= v;
  }
  return t;
}
template <class Policy, class F>
Policy md_tuned_policy(const Policy& p, int pattern, size_t value_bytes, bool floating) {
  if (!p.impl_tune_tile_size()) return p;
  constexpr unsigned R = Policy::rank;
  const bool left      = Policy::inner_direction == Iterate::Left;
  const int fast = left ? 0 : int(R) - 1, next = left ? 1 : int(R) - 2;
  Policy q = p;
  q.impl_change_tile_size(md_heuristic_tile(p, pattern));
  // The tool chooses only whole shapes: tiles the caller fixed in part stay.
  if (!Tools::Experimental::have_tuning_tool() || p.impl_unset_tile_mask() != (1u << R) - 1u) return q;
  int64_t points = 1;
  for (unsigned r = 0; r < R; ++r) points *= (p.upper()[r] - p.lower()[r]);
  int h = -1;
  for (int c = 0; c < 16; ++c) {
    bool match = q.tile_size()[fast] == (16 << (c % 4)) && q.tile_size()[next] == (1 << (c / 4));
    for (int r = 0; r < int(R); ++r)
      if (r != fast && r != next && q.tile_size()[r] != 1) match = false;
    if (match) h = c;
  }
  const int c = request_launch_candidate<F>(LaunchVar::MDRange, pattern, points, value_bytes, floating, h);
  if (c < 0 || c == h) return q;
  typename Policy::tile_type t;
  for (unsigned r = 0; r < R; ++r) t[r] = 1;
  t[fast] = 16 << (c % 4);
  t[next] = 1 << (c / 4);
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
    MI355XInstance* inst = m_p.space().impl_internal_space_instance();
    std::lock_guard<std::mutex> lk(inst->mtx);
    MDLaunch l      = md_launch(m_p);
    const int block = ((l.tile_points + 63) / 64) * 64;
    const int cus   = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
    if (l.ntiles > 0) md_set_depth(l, block);
    int64_t grid    = std::min<int64_t>(l.groups, static_cast<int64_t>(cus) * (2048 / block) * 2);
    if (grid < 1) grid = 1;
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
      const void* k = reinterpret_cast<const void*>(&k_md_reduce<CR, typename Policy::index_type, R, IL, OL>);
      mx_launch_kernel(k, static_cast<unsigned>(grid), block, 0, inst->stream, cr, l, partials, counter, dst);
    } else {
      // Per-lane accumulators in LDS up to 64 KiB per workgroup, else in a
      // global slice per workgroup placed after the partials.
      const size_t acc_bytes = bytes * block;
      const bool in_lds      = acc_bytes <= 64 * 1024;
      const size_t part      = ((bytes * grid + 255) / 256) * 256;
      char* scratch          = static_cast<char*>(inst->partials_space(part + (in_lds ? 0 : acc_bytes * grid)));
      V* partials            = reinterpret_cast<V*>(scratch);
      V* gacc                = in_lds ? nullptr : reinterpret_cast<V*>(scratch + part);
      const void* k =
          reinterpret_cast<const void*>(&k_md_reduce_array<CR, typename Policy::index_type, R, IL, OL>);
      mx_launch_kernel(k, static_cast<unsigned>(grid), block, in_lds ? acc_bytes : 0, inst->stream, cr, l, gacc,
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

}  // namespace Impl
}  // namespace Kokkos

#endif
