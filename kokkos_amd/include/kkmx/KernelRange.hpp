// RangePolicy kernels for MI355X (gfx950) and the dispatch closures that
// launch them.
//
//  K1 k_range_for     grid-stride loop over tiles of blockDim*U consecutive
//                     indices; within a tile, round u covers blockDim
//                     consecutive indices, so every functor call of a
//                     wavefront touches 64 consecutive elements (coalesced).
//  K2 k_range_reduce  same traversal with U independent loads in flight per
//                     lane, ordered DPP wave reduction, LDS combine of the
//                     wavefronts, one partial per workgroup, and a last-block
//                     finish behind an agent-scope acq_rel counter that the
//                     last block resets to zero (no memset between launches).
//                     The grid size is a function of the range only, so the
//                     floating-point combination order is reproducible.
//  K3 k_range_scan    single-pass decoupled look-back scan (Merrill and
//                     Garland, "Single-pass Parallel Prefix Scan with
//                     Decoupled Look-back", NVIDIA NVR-2016-002).  Tiles are
//                     claimed from an atomic counter in launch order, so every
//                     predecessor of a tile is already running (forward
//                     progress).  Each wavefront owns 64*IPT consecutive
//                     indices, handled in IPT coalesced rounds; pass 1 calls
//                     f(i, v, false) once per index and keeps the per-round
//                     inclusive wave scans in registers, pass 2 calls
//                     f(i, prefix, true) once per index.  Tile status words
//                     carry a launch epoch, so they are cleared only once
//                     every 2^29 launches (kScanClearPeriod).
//
// Reference closures these specialise: core/src/HIP/Kokkos_HIP_ParallelFor_Range.hpp:28-101,
// Kokkos_HIP_ParallelReduce_Range.hpp:30-337, Kokkos_HIP_ParallelScan_Range.hpp:29-388.
#ifndef KKMX_KERNEL_RANGE_HPP
#define KKMX_KERNEL_RANGE_HPP

#include <kkmx/FunctorAnalysis.hpp>
#include <kkmx/Policy.hpp>
#include <kkmx/Tools.hpp>
#include <kkmx/Wave.hpp>

namespace Kokkos {
namespace Impl {

// ------------------------------------------------------------------ launch helpers
// When set, dispatches size and allocate their scratch but launch nothing:
// graph construction runs every node once this way before stream capture, so
// no allocation or memset is ever recorded into a graph.
inline thread_local bool g_size_only_dispatch = false;

template <class... Args>
inline void mx_launch_kernel(const void* fn, unsigned grid, unsigned block, size_t lds,
                             mx_stream_t stream, Args&... args) {
  if (g_size_only_dispatch) return;
  void* argv[] = {static_cast<void*>(&args)..., nullptr};
  KKMX_SAFE_CALL(mx_launch(fn, grid, 1, 1, block, 1, 1, lds, argv, stream));
}

// Launch shape of a Range dispatch.  block: lanes per workgroup (multiple of
// 64, <= 1024); items: indices per lane per tile; max_blocks: 0 = one
// workgroup per tile, else a grid-stride cap.
struct RangeConfig {
  int block      = 256;
  int items      = 1;
  int max_blocks = 0;
};

inline int normalise_items(int u) { return u >= 8 ? 8 : u >= 4 ? 4 : u >= 2 ? 2 : 1; }
inline int normalise_block(int b) {
  if (b < 64) b = 64;
  if (b > 1024) b = 1024;
  return (b / 64) * 64;
}

template <class T>
struct alignas(16) LdsSlots16 {
  unsigned char bytes[16 * sizeof(T)];
  __device__ T* get() { return reinterpret_cast<T*>(bytes); }
};

// ------------------------------------------------------------------ K1
template <class F, class Tag, class Index, int U>
__global__ __launch_bounds__(1024) void k_range_for(const F f, const int64_t begin,
                                                    const int64_t end) {
  const int64_t bs     = blockDim.x;
  const int64_t tile   = bs * U;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * tile;
  for (int64_t base = begin + static_cast<int64_t>(blockIdx.x) * tile; base < end; base += stride) {
    const int64_t i0 = base + threadIdx.x;
    if (base + tile <= end) {
#pragma unroll
      for (int u = 0; u < U; ++u) call_functor<Tag>(f, static_cast<Index>(i0 + u * bs));
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * bs;
        if (i < end) call_functor<Tag>(f, static_cast<Index>(i));
      }
    }
  }
}

// ------------------------------------------------------------------ K2
template <class CR, class Index, int U>
__global__ __launch_bounds__(1024) void k_range_reduce(
    const CR cr, const int64_t begin, const int64_t end, typename CR::value_type* partials,
    unsigned* counter, typename CR::value_type* result) {
  using V   = typename CR::value_type;
  using Tag = typename CR::work_tag;
  __shared__ LdsSlots16<V> slots;
  __shared__ int s_last;
  const auto join = [&cr](V& a, const V& b) { cr.join(a, b); };

  V identity;
  cr.init(identity);
  V v = identity;

  const int64_t bs     = blockDim.x;
  const int64_t tile   = bs * U;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * tile;
  for (int64_t base = begin + static_cast<int64_t>(blockIdx.x) * tile; base < end; base += stride) {
    const int64_t i0 = base + threadIdx.x;
    if (base + tile <= end) {
#pragma unroll
      for (int u = 0; u < U; ++u) call_functor<Tag>(cr.functor(), static_cast<Index>(i0 + u * bs), v);
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * bs;
        if (i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), v);
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
    const unsigned prev =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  V w = identity;
  // Thread t folds partials t, t+bs, ... in order; the block reduction keeps
  // thread order, so the combination tree depends on the grid size only.
  for (int64_t j = threadIdx.x; j < gridDim.x; j += bs) join(w, partials[j]);
  w = block_reduce(w, identity, join, slots.get(), blockDim.x);
  if (threadIdx.x == 0) {
    cr.final(w);
    *result = w;
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Runtime-length (value_type[]) reductions: per-lane partials live in LDS.
template <class CR, class Index>
__global__ __launch_bounds__(256) void k_range_reduce_array(
    const CR cr, const int64_t begin, const int64_t end, typename CR::value_type* partials,
    unsigned* counter, typename CR::value_type* result) {
  using V   = typename CR::value_type;
  using Tag = typename CR::work_tag;
  extern __shared__ __align__(16) unsigned char kkmx_dyn_lds[];
  __shared__ int s_last;
  const unsigned n = cr.value_count();
  V* mine          = reinterpret_cast<V*>(kkmx_dyn_lds) + threadIdx.x * n;
  cr.init_array(mine);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = begin + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < end;
       i += stride)
    call_functor<Tag>(cr.functor(), static_cast<Index>(i), mine);
  __syncthreads();
  // Pairwise tree in LDS, lower thread on the left.
  for (unsigned s = 1; s < blockDim.x; s <<= 1) {
    if ((threadIdx.x % (2 * s)) == 0 && threadIdx.x + s < blockDim.x)
      cr.join_array(mine, reinterpret_cast<V*>(kkmx_dyn_lds) + (threadIdx.x + s) * n);
    __syncthreads();
  }
  V* block_v = reinterpret_cast<V*>(kkmx_dyn_lds);
  if (gridDim.x == 1) {
    if (threadIdx.x == 0) {
      cr.final_array(block_v);
      for (unsigned k = 0; k < n; ++k) result[k] = block_v[k];
    }
    return;
  }
  if (threadIdx.x == 0) {
    for (unsigned k = 0; k < n; ++k) partials[blockIdx.x * n + k] = block_v[k];
    const unsigned prev =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (threadIdx.x == 0) {
    cr.init_array(block_v);
    for (unsigned b = 0; b < gridDim.x; ++b) cr.join_array(block_v, partials + b * n);
    cr.final_array(block_v);
    for (unsigned k = 0; k < n; ++k) result[k] = block_v[k];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------------ K3 tile state
constexpr uint32_t kScanAgg  = 1;
constexpr uint32_t kScanIncl = 2;
KOKKOS_INLINE_FUNCTION uint32_t scan_flag(uint32_t epoch, uint32_t st) {
  return ((epoch & 0x3fffffffu) << 2) | st;
}

// A status word keeps the low 30 bits of its launch's epoch, so a word left
// 2^30 launches earlier would pass for a current one.  The host counts the
// launches that use one status array and zeroes the array, in stream order
// and with a kernel (so the step may be stream-captured), before that count
// since the last clear exceeds 2^29: every word a launch can then see was
// written by one of the last 2^29 launches, whose epochs differ modulo 2^30.
inline constexpr uint64_t kScanClearPeriod = uint64_t(1) << 29;

template <class Unused = void>
__global__ __launch_bounds__(256) void k_clear_words(uint64_t* __restrict__ p, const int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// Zeroes `bytes` (a multiple of 8) at p on stream s.
inline void clear_words_async(void* p, size_t bytes, mx_stream_t s) {
  if (!p || bytes < 8) return;
  uint64_t* w     = static_cast<uint64_t*>(p);
  int64_t n       = static_cast<int64_t>(bytes / 8);
  const int64_t g = (n + 255) / 256;
  mx_launch_kernel(reinterpret_cast<const void*>(&k_clear_words<>), static_cast<unsigned>(g < 2048 ? g : 2048),
                   256, 0, s, w, n);
}

// Accounts for `k` scan launches about to be queued on stream s that use
// inst's status arrays, zeroing the arrays first when the period is due.
inline void note_scan_launches(MI355XInstance* inst, uint64_t k, mx_stream_t s) {
  if (inst->scan_since_clear + k > kScanClearPeriod) {
    clear_words_async(inst->scan_status, inst->scan_status_bytes, s);
    clear_words_async(inst->scan_flags, inst->scan_flags_bytes, s);
    inst->scan_since_clear = 0;
  }
  inst->scan_since_clear += k;
}

template <class T, int Kind = (sizeof(T) == 4 ? 1 : (sizeof(T) == 8 ? 2 : 0))>
struct ScanTileState;

// Status words are published with a buffer store carrying the agent-scope
// coherence bit (sc1), the same memory operation a relaxed agent-scope
// atomic store of the word compiles to.  As an intrinsic call rather than an
// IR atomic, the compiler can see that it does not write memory that is
// invariant for the kernel, so a functor's loads through const RandomAccess /
// Restrict Views (constant address space) are not clobbered by a publish and
// can be reused across the look-back.  Word offsets are 32-bit: the host
// keeps the tile count below 2^27.
typedef unsigned int scan_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void scan_publish_word(uint64_t* words, int64_t w, uint32_t flag, uint32_t bits) {
  scan_u32x2 d;
  d.x                        = bits;
  d.y                        = flag;
  __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(words, 0, 0x7fffffff, 0x00020000);
  __builtin_amdgcn_raw_buffer_store_b64(d, rsrc, static_cast<int>(w * 8), 0, /*sc1*/ 16);
}
inline constexpr int64_t kScanMaxTiles = int64_t(1) << 27;

// 4-byte values: one 64-bit word {flag, value}.
template <class T>
struct ScanTileState<T, 1> {
  uint64_t* words;
  static size_t bytes(int64_t ntiles) { return static_cast<size_t>(ntiles) * 8; }
  void bind(void* status, void*) { words = static_cast<uint64_t*>(status); }
  __device__ void publish(int64_t t, uint32_t flag, const T& v) const {
    uint32_t bits;
    __builtin_memcpy(&bits, &v, 4);
    scan_publish_word(words, t, flag, bits);
  }
  __device__ uint32_t read(int64_t t, T& v) const {
    const uint64_t w = __hip_atomic_load(&words[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t bits = static_cast<uint32_t>(w);
    __builtin_memcpy(&v, &bits, 4);
    return static_cast<uint32_t>(w >> 32);
  }
};

// 8-byte values: two 64-bit words, each {flag, half}; a read is valid only
// when both halves carry the same flag.
template <class T>
struct ScanTileState<T, 2> {
  uint64_t* words;
  static size_t bytes(int64_t ntiles) { return static_cast<size_t>(ntiles) * 16; }
  void bind(void* status, void*) { words = static_cast<uint64_t*>(status); }
  __device__ void publish(int64_t t, uint32_t flag, const T& v) const {
    uint64_t bits;
    __builtin_memcpy(&bits, &v, 8);
    scan_publish_word(words, 2 * t, flag, static_cast<uint32_t>(bits));
    scan_publish_word(words, 2 * t + 1, flag, static_cast<uint32_t>(bits >> 32));
  }
  __device__ uint32_t read(int64_t t, T& v) const {
    const uint64_t a = __hip_atomic_load(&words[2 * t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t b = __hip_atomic_load(&words[2 * t + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t fa = static_cast<uint32_t>(a >> 32), fb = static_cast<uint32_t>(b >> 32);
    if (fa != fb) return 0;
    const uint64_t bits = (a & 0xffffffffu) | (b << 32);
    __builtin_memcpy(&v, &bits, 8);
    return fa;
  }
};

// Any other value: flag word plus separate aggregate / inclusive slots,
// ordered with agent-scope release/acquire fences.
template <class T>
struct ScanTileState<T, 0> {
  uint32_t* flags;
  T* values;  // 2 per tile: [2t] aggregate, [2t+1] inclusive prefix
  static size_t bytes(int64_t ntiles) { return static_cast<size_t>(ntiles) * 4; }
  static size_t value_bytes(int64_t ntiles) { return static_cast<size_t>(ntiles) * 2 * sizeof(T); }
  void bind(void* status, void* vals) {
    flags  = static_cast<uint32_t*>(status);
    values = static_cast<T*>(vals);
  }
  __device__ void publish(int64_t t, uint32_t flag, const T& v) const {
    values[2 * t + ((flag & 3) == kScanIncl ? 1 : 0)] = v;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(&flags[t], flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ uint32_t read(int64_t t, T& v) const {
    const uint32_t f = __hip_atomic_load(&flags[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((f & 3) != 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      v = values[2 * t + ((f & 3) == kScanIncl ? 1 : 0)];
    }
    return f;
  }
};

template <class CR>
__device__ __forceinline__ typename CR::value_type cr_identity(const CR& cr) {
  typename CR::value_type v;
  cr.init(v);
  return v;
}

template <class T>
struct scan_state_kind {
  static constexpr int value = (sizeof(T) == 4 ? 1 : (sizeof(T) == 8 ? 2 : 0));
};

// Exclusive prefix of tile `tile` from its predecessors, computed by the
// whole workgroup: thread j < W of a window reads the status of tile
// hi-(W-1)+j (W = 32, or the block when smaller), so one round trip covers
// W predecessors and the ordered block reduction keeps tile order.  The
// nearest tile that already holds an inclusive prefix ends the walk, which
// the immediate predecessor usually does; a window of 32 rather than one
// word per lane of a 1024-lane block reads 32 times fewer status words per
// round (2^28 int32 scan about 2 % faster; with 64 1.2 %, with 128 0.3 %).
// Tiles before `first` (another sequence's tiles) count as the identity.
// Returns the same value in every thread.
template <class State, class V, class Join>
__device__ V scan_lookback(const State& st, int64_t tile, uint32_t epoch, const V& identity,
                           const Join& join, V* lds_vals, int* lds_pos, int64_t first = 0) {
  const int T         = blockDim.x;
  const int tid       = threadIdx.x;
  const uint32_t want = (epoch & 0x3fffffffu);
  V excl              = identity;
  int64_t hi          = tile - 1;
  const int W         = T < 32 ? T : 32;
  while (true) {
    const int64_t t = hi - (W - 1) + tid;
    V v             = identity;
    uint32_t s      = tid < W ? kScanIncl : kScanAgg;
    if (tid < W && t >= first) {
      uint32_t f = st.read(t, v);
      int spins  = 0;
      while ((f >> 2) != want || (f & 3) == 0) {
        if (++spins > 16) __builtin_amdgcn_s_sleep(1);
        f = st.read(t, v);
      }
      s = f & 3;
    }
    if (tid == 0) *lds_pos = -1;
    __syncthreads();
    const uint64_t m = wave::ballot(s == kScanIncl);
    if (wave::lane() == 0 && m) atomicMax(lds_pos, (tid & ~63) + 63 - __builtin_clzll(m));
    __syncthreads();
    const int k = *lds_pos;
    if (k >= 0 && tid < k) v = identity;
    const V agg = block_reduce(v, identity, join, lds_vals, T);
    V r         = agg;
    join(r, excl);
    excl = r;
    if (k >= 0) return excl;
    hi -= W;
    __syncthreads();  // lds_pos is reset at the top of the next round
  }
}

// One tile per workgroup, claimed in launch order.  Wavefront w owns the
// 64*IPT consecutive indices starting at tile_base + w*64*IPT and walks them
// in rounds of 64 (one index per lane, coalesced), B rounds at a time so B
// loads per lane are in flight together.
//   pass 1: f(i, v, false) per index, ordered wave reduction per round ->
//           wave totals -> tile aggregate (published), nothing kept in
//           registers, so occupancy stays high;
//   look-back for the tile's exclusive prefix (whole workgroup);
//   pass 2: the same rounds again (the tile is still in L2 / the memory-side
//           cache), per-round inclusive wave scans, then f(i, prefix, true).
// The functor sees final=false twice and final=true once per index, which
// the reference allows (core/src/HIP/Kokkos_HIP_ParallelScan_Range.hpp calls
// it three times per index).
template <class CR, class Index, int IPT, bool WithTotal>
__global__ __launch_bounds__(1024) void k_range_scan(
    const CR cr, const int64_t begin, const int64_t end, const int64_t ntiles,
    ScanTileState<typename CR::value_type> st, unsigned* control,
    typename CR::value_type* total_out) {
  using V          = typename CR::value_type;
  using Tag        = typename CR::work_tag;
  constexpr int B  = IPT < 8 ? IPT : 8;
  __shared__ LdsSlots16<V> s_wave;  // per-wave totals
  __shared__ LdsSlots16<V> s_red;   // look-back reduction scratch
  __shared__ uint64_t s_claim;
  __shared__ int s_pos;
  const auto join = [&cr](V& a, const V& b) { cr.join(a, b); };
  // control[2..3] is one 64-bit word {tile counter (low), epoch (high)}: a
  // single relaxed fetch_add returns a consistent (epoch, tile) pair, so the
  // claim needs no fence (an acq_rel claim would write back and invalidate
  // the L2 at every tile).
  uint64_t* claim_word = reinterpret_cast<uint64_t*>(control + 2);
  if (threadIdx.x == 0)
    s_claim = __hip_atomic_fetch_add(claim_word, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const uint64_t claim = s_claim;
  const int64_t tile   = static_cast<int64_t>(static_cast<uint32_t>(claim));
  const uint32_t epoch = static_cast<uint32_t>(claim >> 32);
  const int w          = threadIdx.x >> 6;
  const int nw         = blockDim.x >> 6;
  const int l          = wave::lane();
  V identity;
  cr.init(identity);
  const int64_t base = begin + tile * (static_cast<int64_t>(blockDim.x) * IPT) +
                       static_cast<int64_t>(w) * (64 * IPT) + l;

  // Pass 1: ordered reduction of the wavefront's indices.
  V wave_total = identity;
#pragma unroll
  for (int r0 = 0; r0 < IPT; r0 += B) {
    V v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int64_t i = base + (r0 + b) * 64;
      v[b]            = identity;
      if (i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), v[b], false);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) join(wave_total, wave::reduce(v[b], identity, join));
  }
  if (l == 0) s_wave.get()[w] = wave_total;
  __syncthreads();
  const V x    = (l < nw) ? s_wave.get()[l] : identity;
  const V xs   = wave::inclusive_scan(x, identity, join);
  const V tagg = wave::readlane(xs, nw - 1);
  const V wpre = (w == 0) ? identity : wave::readlane(xs, w - 1);

  V prefix = identity;
  if (tile == 0) {
    if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanIncl), tagg);
  } else {
    if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanAgg), tagg);
    prefix = scan_lookback(st, tile, epoch, identity, join, s_red.get(), &s_pos);
    if (threadIdx.x == 0) {
      V incl_t = prefix;
      join(incl_t, tagg);
      st.publish(tile, scan_flag(epoch, kScanIncl), incl_t);
    }
  }
  if (threadIdx.x == 0 && tile == ntiles - 1) {
    if constexpr (WithTotal) {
      V tot = prefix;
      join(tot, tagg);
      *total_out = tot;
    }
    // Last tile: every workgroup has claimed; the next launch starts at tile
    // 0 with the next epoch.
    __hip_atomic_store(claim_word, static_cast<uint64_t>(epoch + 1u) << 32, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }

  // Pass 2: values again, per-round scans, final calls.
  V running = prefix;
  join(running, wpre);
#pragma unroll
  for (int r0 = 0; r0 < IPT; r0 += B) {
    V v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int64_t i = base + (r0 + b) * 64;
      v[b]            = identity;
      if (i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), v[b], false);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const V inc = wave::inclusive_scan(v[b], identity, join);
      V upd       = running;
      join(upd, wave::shift_up(inc, identity));
      join(running, wave::readlane(inc, 63));
      v[b] = upd;
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int64_t i = base + (r0 + b) * 64;
      if (i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), v[b], true);
    }
  }
}

// Register-resident variant: each lane keeps its IPT contributions from the
// first functor pass in registers, so the tile is read from memory once; the
// second (final) pass only re-reads what the functor itself loads, from
// cache.  Lane l of wavefront w holds index base + r*64 for r < IPT (one
// contiguous 64-index run per round).  Used while IPT values fit in about
// 256 bytes of registers per lane; larger IPT takes the two-pass tile above.
// One tile of the register-resident scan, run by the whole workgroup: pass 1
// with the values kept in registers, aggregate published, look-back,
// inclusive prefix published, final calls.  The thread that runs tile
// ntiles-1 writes the total.  Uses s_wave, s_red and s_pos; the caller puts a
// barrier between two calls.
template <class CR, class Index, int IPT, bool WithTotal>
__device__ __forceinline__ void scan_tile_rr(const CR& cr, const int64_t begin, const int64_t end,
                                             const int64_t ntiles, const ScanTileState<typename CR::value_type>& st,
                                             const int64_t tile, const uint32_t epoch,
                                             typename CR::value_type* total_out,
                                             typename CR::value_type* s_wave, typename CR::value_type* s_red,
                                             int* s_pos) {
  using V         = typename CR::value_type;
  using Tag       = typename CR::work_tag;
  const auto join = [&cr](V& a, const V& b) { cr.join(a, b); };
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6, l = wave::lane();
  V identity;
  cr.init(identity);
  const int64_t tile_elems = static_cast<int64_t>(blockDim.x) * IPT;
  const int64_t tbase      = begin + tile * tile_elems;
  const int64_t base       = tbase + static_cast<int64_t>(w) * (64 * IPT) + l;
  const bool full          = tbase + tile_elems <= end;

  // Separate unguarded loop for full tiles: all IPT loads issue back to back.
  V v[IPT];
  if (full) {
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
      v[r] = identity;
      call_functor<Tag>(cr.functor(), static_cast<Index>(base + r * 64), v[r], false);
    }
  } else {
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
      const int64_t i = base + r * 64;
      v[r]            = identity;
      if (i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), v[r], false);
    }
  }
  V inc[IPT];
  V wave_total = identity;
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    inc[r] = wave::inclusive_scan(v[r], identity, join);
    join(wave_total, wave::readlane(inc[r], 63));
  }
  if (l == 0) s_wave[w] = wave_total;
  __syncthreads();
  const V x    = (l < nw) ? s_wave[l] : identity;
  const V xs   = wave::inclusive_scan(x, identity, join);
  const V tagg = wave::readlane(xs, nw - 1);
  const V wpre = (w == 0) ? identity : wave::readlane(xs, w - 1);

  V prefix = identity;
  if (tile == 0) {
    if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanIncl), tagg);
  } else {
    if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanAgg), tagg);
    prefix = scan_lookback(st, tile, epoch, identity, join, s_red, s_pos);
    if (threadIdx.x == 0) {
      V t = prefix;
      join(t, tagg);
      st.publish(tile, scan_flag(epoch, kScanIncl), t);
    }
  }
  if constexpr (WithTotal) {
    if (threadIdx.x == 0 && tile == ntiles - 1) {
      V tot = prefix;
      join(tot, tagg);
      *total_out = tot;
    }
  }
  V running = prefix;
  join(running, wpre);
  if (full) {
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
      V u = running;
      join(u, wave::shift_up(inc[r], identity));
      join(running, wave::readlane(inc[r], 63));
      call_functor<Tag>(cr.functor(), static_cast<Index>(base + r * 64), u, true);
    }
  } else {
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
      V u = running;
      join(u, wave::shift_up(inc[r], identity));
      join(running, wave::readlane(inc[r], 63));
      const int64_t i = base + r * 64;
      if (i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), u, true);
    }
  }
}

// The same tile with the functor's contributions kept in registers instead
// of the per-round wave scans; the final pass recomputes each round's wave
// scan from them (the same DPP steps in the same order, so the results are
// bitwise those of scan_tile_rr).  What it buys: when the functor reads
// through Views whose data the kernel cannot change (const RandomAccess /
// Restrict, loaded through the constant address space) and the library is
// built with the larger memory-dependence limits, the compiler replaces the
// final pass's loads with the first pass's values, so the tile is read from
// HBM once.  For that the first pass must dominate the final pass, so full
// and ragged tiles run separate copies of the body (Full).  Index bases are
// made opaque per chunk of 8 rounds, so every round's address is a chunk base
// plus an immediate offset instead of a separately materialised 64-bit
// address; the ragged copy re-derives its bases before the final pass, so its
// bounds masks are not carried across the look-back.
//   K > 1 (value types that are not arithmetic scalars, whose joins cost
// many instructions): each round gives a lane K consecutive indices, which
// it runs through the functor in order with one running update, as a serial
// scan does; the wave scans then combine one value per lane and round, K
// times fewer than with one index per lane.  A lane's K indices are
// adjacent, so its K loads of a unit-stride input fall in one or two cache
// lines.  Floating-point joins then associate differently from K = 1
// (integer results are identical).
template <class CR, class Index, int IPT, bool WithTotal, bool Full, int K = 1>
__device__ __forceinline__ void scan_tile_keep_body(const CR& cr, const int64_t end, const int64_t ntiles,
                                                    const ScanTileState<typename CR::value_type>& st,
                                                    const int64_t tile, const uint32_t epoch,
                                                    typename CR::value_type* total_out,
                                                    typename CR::value_type* s_wave, typename CR::value_type* s_red,
                                                    int* s_pos, const int64_t base) {
  using V         = typename CR::value_type;
  using Tag       = typename CR::work_tag;
  const auto join = [&cr](V& a, const V& b) { cr.join(a, b); };
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6, l = wave::lane();
  V identity;
  cr.init(identity);
  static_assert(K >= 1 && IPT % K == 0, "K must divide the values per lane");
  constexpr int kRounds = IPT / K;
  constexpr int kChunk  = 8;
  constexpr int kChunks = (IPT + kChunk - 1) / kChunk;
  int64_t chunk_base[kChunks];
#pragma unroll
  for (int c = 0; c < kChunks; ++c) {
    chunk_base[c] = base + c * (kChunk * 64);
    if constexpr (K == 1) asm volatile("" : "+v"(chunk_base[c]));
  }
  // Index of round r's k-th call in this lane.
  const auto index = [&](int r, int k) -> int64_t {
    if constexpr (K == 1) return chunk_base[r / kChunk] + (r % kChunk) * 64;
    else return base + static_cast<int64_t>(r) * (64 * K) + k;
  };
  V v[kRounds];
  V wave_total = identity;
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    v[r] = identity;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t i = index(r, k);
      if (Full || i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), v[r], false);
    }
  }
#pragma unroll
  for (int r = 0; r < kRounds; ++r) join(wave_total, wave::reduce(v[r], identity, join));
  if (l == 0) s_wave[w] = wave_total;
  __syncthreads();
  const V x    = (l < nw) ? s_wave[l] : identity;
  const V xs   = wave::inclusive_scan(x, identity, join);
  const V tagg = wave::readlane(xs, nw - 1);
  const V wpre = (w == 0) ? identity : wave::readlane(xs, w - 1);

  V prefix = identity;
  if (tile == 0) {
    if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanIncl), tagg);
  } else {
    if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanAgg), tagg);
    prefix = scan_lookback(st, tile, epoch, identity, join, s_red, s_pos);
    if (threadIdx.x == 0) {
      V t = prefix;
      join(t, tagg);
      st.publish(tile, scan_flag(epoch, kScanIncl), t);
    }
  }
  if constexpr (WithTotal) {
    if (threadIdx.x == 0 && tile == ntiles - 1) {
      V tot = prefix;
      join(tot, tagg);
      *total_out = tot;
    }
  }
  V running = prefix;
  join(running, wpre);
  if constexpr (!Full && K == 1) {
#pragma unroll
    for (int c = 0; c < kChunks; ++c) asm volatile("" : "+v"(chunk_base[c]));
  }
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const V inc = wave::inclusive_scan(v[r], identity, join);
    V u         = running;
    join(u, wave::shift_up(inc, identity));
    join(running, wave::readlane(inc, 63));
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t i = index(r, k);
      if (Full || i < end) call_functor<Tag>(cr.functor(), static_cast<Index>(i), u, true);
    }
  }
}

// Consecutive indices per lane and round in the contribution-keeping body:
// 1 for arithmetic values, 4 for other value types (structs, wrapped values
// with an emptiness flag), whose joins dominate the tile's time.
template <class V>
inline constexpr int scan_keep_k = std::is_arithmetic<V>::value ? 1 : 4;

template <class CR, class Index, int IPT, bool WithTotal>
__device__ __forceinline__ void scan_tile_keep(const CR& cr, const int64_t begin, const int64_t end,
                                               const int64_t ntiles, const ScanTileState<typename CR::value_type>& st,
                                               const int64_t tile, const uint32_t epoch,
                                               typename CR::value_type* total_out, typename CR::value_type* s_wave,
                                               typename CR::value_type* s_red, int* s_pos) {
  constexpr int K          = scan_keep_k<typename CR::value_type>;
  const int64_t tile_elems = static_cast<int64_t>(blockDim.x) * IPT;
  const int64_t tbase      = begin + tile * tile_elems;
  const int64_t base = tbase + static_cast<int64_t>(threadIdx.x >> 6) * (64 * IPT) + int64_t(wave::lane()) * K;
  if (tbase + tile_elems <= end)
    scan_tile_keep_body<CR, Index, IPT, WithTotal, true, K>(cr, end, ntiles, st, tile, epoch, total_out, s_wave,
                                                            s_red, s_pos, base);
  else
    scan_tile_keep_body<CR, Index, IPT, WithTotal, false, K>(cr, end, ntiles, st, tile, epoch, total_out, s_wave,
                                                             s_red, s_pos, base);
}

// One tile per workgroup, claimed in launch order from control[2..3] (one
// 64-bit word {tile counter (low), epoch (high)}: a single relaxed fetch_add
// returns a consistent pair, so the claim needs no fence).  The workgroup of
// the last tile returns the word to {0, epoch + 1}: every workgroup has
// claimed by then.
// Keep selects the tile body: 0 = scan_tile_rr, 1 = scan_tile_keep.
template <class CR, class Index, int IPT, bool WithTotal, int Keep = 0>
__global__ __launch_bounds__(1024) void k_range_scan_rr(
    const CR cr, const int64_t begin, const int64_t end, const int64_t ntiles,
    ScanTileState<typename CR::value_type> st, unsigned* control,
    typename CR::value_type* total_out) {
  using V = typename CR::value_type;
  __shared__ LdsSlots16<V> s_wave;
  __shared__ LdsSlots16<V> s_red;
  __shared__ uint64_t s_claim;
  __shared__ int s_pos;
  uint64_t* claim_word = reinterpret_cast<uint64_t*>(control + 2);
  if (threadIdx.x == 0)
    s_claim = __hip_atomic_fetch_add(claim_word, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const uint64_t claim = s_claim;
  const int64_t tile   = static_cast<int64_t>(static_cast<uint32_t>(claim));
  const uint32_t epoch = static_cast<uint32_t>(claim >> 32);
  if (threadIdx.x == 0 && tile == ntiles - 1)
    __hip_atomic_store(claim_word, static_cast<uint64_t>(epoch + 1u) << 32, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if constexpr (Keep == 1)
    scan_tile_keep<CR, Index, IPT, WithTotal>(cr, begin, end, ntiles, st, tile, epoch, total_out, s_wave.get(),
                                              s_red.get(), &s_pos);
  else
    scan_tile_rr<CR, Index, IPT, WithTotal>(cr, begin, end, ntiles, st, tile, epoch, total_out, s_wave.get(),
                                            s_red.get(), &s_pos);
}

// The contribution-keeping body in a loop: a grid of resident workgroups,
// each claiming its next tile from the same word once its current tile's
// final calls are issued, so a workgroup's end and a new one's start (the
// dispatch between them) are paid once per workgroup instead of once per tile
// (3 % at 2^28 int32 elements with 4096 tiles).
//   * Claims are counted past the last tile: every workgroup stops on its
//     first id >= ntiles, so the launch makes exactly ntiles + gridDim.x
//     claims and the one that receives the last id returns the word to
//     {0, epoch + 1} after every other claim of the launch.
//   * The claim is taken after the look-back of the current tile, so a
//     workgroup holding an unstarted tile only has its own final pass to
//     finish: the lowest unfinished tile is always running or about to run.
//     A workgroup that is not resident holds no tile, so the launch needs no
//     co-residency.
//   * The claim is made after the last final call, so the compiler still
//     replaces the final pass's loads with the first pass's values.
template <class CR, class Index, int IPT, bool WithTotal>
__global__ __launch_bounds__(1024) void k_range_scan_keep_loop(
    const CR cr, const int64_t begin, const int64_t end, const int64_t ntiles,
    ScanTileState<typename CR::value_type> st, unsigned* control,
    typename CR::value_type* total_out) {
  using V = typename CR::value_type;
  __shared__ LdsSlots16<V> s_wave;
  __shared__ LdsSlots16<V> s_red;
  __shared__ uint64_t s_claim;
  __shared__ int s_pos;
  uint64_t* claim_word   = reinterpret_cast<uint64_t*>(control + 2);
  const uint32_t last_id = static_cast<uint32_t>(ntiles + static_cast<int64_t>(gridDim.x) - 1);
  if (threadIdx.x == 0)
    s_claim = __hip_atomic_fetch_add(claim_word, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  uint64_t claim = s_claim;
  while (true) {
    const uint32_t id    = static_cast<uint32_t>(claim);
    const uint32_t epoch = static_cast<uint32_t>(claim >> 32);
    if (threadIdx.x == 0 && id == last_id)
      __hip_atomic_store(claim_word, static_cast<uint64_t>(epoch + 1u) << 32, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (static_cast<int64_t>(id) >= ntiles) break;
    scan_tile_keep<CR, Index, IPT, WithTotal>(cr, begin, end, ntiles, st, static_cast<int64_t>(id), epoch, total_out,
                                              s_wave.get(), s_red.get(), &s_pos);
    uint64_t next = 0;
    if (threadIdx.x == 0)
      next = __hip_atomic_fetch_add(claim_word, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();  // the tile's LDS reads are done before s_claim changes
    if (threadIdx.x == 0) s_claim = next;
    __syncthreads();
    claim = s_claim;
  }
}

// ------------------------------------------------------------------ host helpers
template <class V>
struct ResultSink {
  // Where a reduction/scan total lands: the caller's pointer when the device
  // can write it, else the instance's pinned staging buffer.
  V* device_ptr   = nullptr;
  V* host_target  = nullptr;
  size_t bytes    = 0;
  bool needs_copy = false;
};

template <class V, class ResultMemSpace>
ResultSink<V> make_sink(MI355XInstance* inst, V* result, size_t count) {
  ResultSink<V> s;
  s.bytes = sizeof(V) * count;
  if constexpr (exec_can_access<MI355X, ResultMemSpace>::value) {
    s.device_ptr = result;
  } else {
    s.device_ptr  = static_cast<V*>(inst->pinned_space(s.bytes));
    s.host_target = result;
    s.needs_copy  = (result != nullptr);
  }
  return s;
}
template <class V>
void finish_sink(MI355XInstance* inst, const ResultSink<V>& s) {
  if (s.needs_copy) {
    KKMX_SAFE_CALL(mx_stream_sync(inst->stream));
    std::memcpy(static_cast<void*>(s.host_target), s.device_ptr, s.bytes);
  }
}

template <class F, class Tag, class Index>
const void* range_for_kernel(int items) {
  switch (normalise_items(items)) {
    case 8: return reinterpret_cast<const void*>(&k_range_for<F, Tag, Index, 8>);
    case 4: return reinterpret_cast<const void*>(&k_range_for<F, Tag, Index, 4>);
    case 2: return reinterpret_cast<const void*>(&k_range_for<F, Tag, Index, 2>);
    default: return reinterpret_cast<const void*>(&k_range_for<F, Tag, Index, 1>);
  }
}

template <class CR, class Index>
const void* range_reduce_kernel(int items) {
  switch (normalise_items(items)) {
    case 8: return reinterpret_cast<const void*>(&k_range_reduce<CR, Index, 8>);
    case 4: return reinterpret_cast<const void*>(&k_range_reduce<CR, Index, 4>);
    case 2: return reinterpret_cast<const void*>(&k_range_reduce<CR, Index, 2>);
    default: return reinterpret_cast<const void*>(&k_range_reduce<CR, Index, 1>);
  }
}

inline int64_t range_grid(int64_t n, const RangeConfig& c) {
  const int64_t tile = static_cast<int64_t>(c.block) * c.items;
  int64_t g          = ceil_div(n, tile);
  if (c.max_blocks > 0 && g > c.max_blocks) g = c.max_blocks;
  if (g > 0x7fffffff) g = 0x7fffffff;
  return g;
}

// Process-wide launch-shape overrides per pattern (0 for, 1 reduce, 2 scan):
// set by kkmx_set_range_config for configuration sweeps; a tuning tool's
// answer for one dispatch takes precedence.
struct RangeOverrides {
  RangeConfig cfg[3];
  bool set[3]   = {false, false, false};
  int scan_body = -1;  // >= 0: every scan runs this tile body (no trials)
};
inline RangeOverrides& range_overrides() {
  static RangeOverrides o;
  return o;
}

// Heuristic launch shapes (the learned tuner may replace them through the
// tuning hooks).
// Default launch shapes, per octave of the extent, from the launch-shape
// sweep on MI355X (20 dispatch families x 160 extents x 16 candidates): the
// candidate with the least mean slowdown against the per-case best.
// Large parallel_for: functors of 96 bytes or more (three or more captured
// 1-D Views of 32 bytes, e.g. STREAM add / triad) already have enough bytes
// in flight with 2 indices per lane and run ~3 % faster with them; smaller
// ones (copies, scales, fills) keep 4.
inline RangeConfig default_for_config(int64_t n, size_t functor_bytes = 0) {
  if (n >= (int64_t(1) << 23)) return RangeConfig{128, functor_bytes >= 96 ? 2 : 4, 0};
  return RangeConfig{256, 2, 0};
}
inline RangeConfig default_reduce_config(int64_t n) {
  RangeConfig c = n < (int64_t(1) << 15)   ? RangeConfig{256, 4, 0}
                  : n < (int64_t(1) << 18) ? RangeConfig{256, 8, 0}
                  : n < (int64_t(1) << 21) ? RangeConfig{1024, 8, 0}
                                           : RangeConfig{1024, 4, 0};
  // At most as many resident lanes as four 256-lane workgroups per CU.
  const int cus = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
  c.max_blocks  = cus * 4 * 256 / c.block;
  if (n < static_cast<int64_t>(c.block) * c.items * c.max_blocks) c.max_blocks = 0;
  return c;
}
inline RangeConfig default_scan_config(int64_t n) {
  return n < (int64_t(1) << 10)   ? RangeConfig{128, 8, 0}
         : n < (int64_t(1) << 17) ? RangeConfig{256, 8, 0}
         : n < (int64_t(1) << 20) ? RangeConfig{512, 8, 0}
         : n < (int64_t(1) << 21) ? RangeConfig{1024, 8, 0}
         : n < (int64_t(1) << 22) ? RangeConfig{1024, 16, 0}
                                  : RangeConfig{1024, 32, 0};
}

// ------------------------------------------------------------------ tuning hooks
// When a Kokkos Tools tuning tool is loaded, Range dispatches without an
// explicit launch shape ask it for one of 16 candidates through the
// reference's tuning protocol (core/src/impl/Kokkos_Profiling.cpp:1032-1211:
// declare_input_type / declare_output_type, set_input_values,
// request_output_values inside a context).  The request carries the
// heuristic's own candidate as the default value, so a tool that does not
// answer leaves the heuristic in charge.
//   for / reduce: block = 128 << (c / 4), items per lane = 1 << (c % 4);
//   scan:         block = 128 << (c / 4), items per lane = 8 << (c % 4).
constexpr int kRangeCandidates = 16;
inline RangeConfig range_candidate(int pattern, int c, int64_t n) {
  const int block = 128 << (c >> 2);
  if (pattern == 2) return RangeConfig{block, 8 << (c & 3), 0};
  RangeConfig r{block, 1 << (c & 3), 0};
  if (pattern == 1) {  // same resident-thread cap as the heuristic
    const int cus = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
    r.max_blocks  = cus * 4 * 256 / block;
    if (n < static_cast<int64_t>(r.block) * r.items * r.max_blocks) r.max_blocks = 0;
  }
  return r;
}
inline int range_candidate_index(int pattern, const RangeConfig& c) {
  int bi = -1;
  for (int k = 0; k < 4; ++k)
    if (c.block == (128 << k)) bi = k;
  int ii = -1;
  for (int k = 0; k < 4; ++k)
    if (c.items == (pattern == 2 ? 8 << k : 1 << k)) ii = k;
  return (bi < 0 || ii < 0) ? -1 : bi * 4 + ii;
}

// Optional cost hints a functor may declare for the tuner:
//   static constexpr double kkmx_bytes_per_index, kkmx_flops_per_index;
template <class F, class = void>
struct functor_cost_hint {
  static constexpr double bytes = 0.0, flops = 0.0;
};
template <class F>
struct functor_cost_hint<F, std::void_t<decltype(F::kkmx_bytes_per_index), decltype(F::kkmx_flops_per_index)>> {
  static constexpr double bytes = F::kkmx_bytes_per_index, flops = F::kkmx_flops_per_index;
};

// The launch-shape variables (Range launch shape, Team shape, MDRange tile,
// tile shape of the fused batch kernel) share one set of declared inputs;
// policy_kind tells them apart.
enum class LaunchVar : int { Range = 0, Team = 1, MDRange = 2, Batch = 3 };
constexpr int kLaunchVars          = 4;
constexpr int kBatchShapeCandidates = 8;
constexpr int launch_var_candidates(LaunchVar v) {
  return v == LaunchVar::Batch ? kBatchShapeCandidates : kRangeCandidates;
}

struct LaunchTuningIds {
  size_t out[kLaunchVars] = {0, 0, 0, 0};
  size_t pattern = 0, extent = 0, value_bytes = 0, functor_bytes = 0, floating = 0, bytes = 0, flops = 0,
         cus = 0, kind = 0;
};
inline const LaunchTuningIds& launch_tuning_ids() {
  static const LaunchTuningIds ids = [] {
    using namespace Tools::Experimental;
    LaunchTuningIds r;
    static int64_t cand[kRangeCandidates];
    for (int k = 0; k < kRangeCandidates; ++k) cand[k] = k;
    const char* names[kLaunchVars] = {"kokkos.mi355x.range_launch", "kokkos.mi355x.team_shape",
                                      "kokkos.mi355x.mdrange_tile", "kokkos.mi355x.batch_shape"};
    for (int v = 0; v < kLaunchVars; ++v) {
      VariableInfo o{};
      o.type          = kokkos_value_int64;
      o.category      = kokkos_value_categorical;
      o.valueQuantity = kokkos_value_set;
      o.candidates    = make_candidate_set(size_t(launch_var_candidates(LaunchVar(v))), cand);
      r.out[v]        = declare_output_type(names[v], o);
    }
    auto in = [](const char* name, ValueType t, StatisticalCategory cat) {
      VariableInfo i{};
      i.type          = t;
      i.category      = cat;
      i.valueQuantity = kokkos_value_unbounded;
      return declare_input_type(name, i);
    };
    r.pattern       = in("kokkos.mi355x.pattern", kokkos_value_int64, kokkos_value_categorical);
    r.extent        = in("kokkos.mi355x.extent", kokkos_value_int64, kokkos_value_ratio);
    r.value_bytes   = in("kokkos.mi355x.value_bytes", kokkos_value_int64, kokkos_value_ratio);
    r.functor_bytes = in("kokkos.mi355x.functor_bytes", kokkos_value_int64, kokkos_value_ratio);
    r.floating      = in("kokkos.mi355x.floating", kokkos_value_int64, kokkos_value_categorical);
    r.bytes         = in("kokkos.mi355x.bytes_per_index", kokkos_value_double, kokkos_value_ratio);
    r.flops         = in("kokkos.mi355x.flops_per_index", kokkos_value_double, kokkos_value_ratio);
    r.cus           = in("kokkos.mi355x.cu_count", kokkos_value_int64, kokkos_value_ratio);
    r.kind          = in("kokkos.mi355x.policy_kind", kokkos_value_int64, kokkos_value_categorical);
    return r;
  }();
  return ids;
}

// Inputs of one launch-shape request.  pattern: 0 for, 1 reduce, 2 scan,
// 3 batch; functor_bytes is sizeof(functor) (for a batch: the mean op
// extent); bytes / flops are per index (0 when unknown).
struct LaunchFeatures {
  int pattern;
  int64_t extent;
  int64_t value_bytes;
  int64_t functor_bytes;
  bool floating;
  double bytes, flops;
};

// Asks the tuning tool for one of the candidates of `var`; returns the
// answer, or `dflt` when no tool is loaded or it leaves the default.
inline int request_launch_candidate(LaunchVar var, const LaunchFeatures& x, int dflt) {
  using namespace Tools::Experimental;
  if (!have_tuning_tool()) return dflt;
  const LaunchTuningIds& id = launch_tuning_ids();
  const size_t ctx          = get_new_context_id();
  begin_context(ctx);
  VariableValue in[9] = {make_variable_value(id.pattern, int64_t(x.pattern)),
                         make_variable_value(id.extent, x.extent),
                         make_variable_value(id.value_bytes, x.value_bytes),
                         make_variable_value(id.functor_bytes, x.functor_bytes),
                         make_variable_value(id.floating, int64_t(x.floating ? 1 : 0)),
                         make_variable_value(id.bytes, x.bytes),
                         make_variable_value(id.flops, x.flops),
                         make_variable_value(id.cus, int64_t(MI355X::impl_cu_count())),
                         make_variable_value(id.kind, int64_t(var))};
  set_input_values(ctx, 9, in);
  VariableValue out = make_variable_value(id.out[int(var)], int64_t(dflt));
  request_output_values(ctx, 1, &out);
  end_context(ctx);
  const int64_t c = out.value.int_value;
  return (c >= 0 && c < launch_var_candidates(var)) ? static_cast<int>(c) : dflt;
}

template <class F>
int request_launch_candidate(LaunchVar var, int pattern, int64_t n, size_t value_bytes, bool floating, int dflt) {
  if (!Tools::Experimental::have_tuning_tool()) return dflt;
  const LaunchFeatures x{pattern, n, int64_t(value_bytes), int64_t(sizeof(F)), floating, functor_cost_hint<F>::bytes,
                         functor_cost_hint<F>::flops};
  return request_launch_candidate(var, x, dflt);
}

template <class F>
RangeConfig tuned_range_config(int pattern, int64_t n, size_t value_bytes, bool floating,
                               const RangeConfig& heuristic) {
  if (!Tools::Experimental::have_tuning_tool()) return heuristic;
  const int h = range_candidate_index(pattern, heuristic);
  const int c = request_launch_candidate<F>(LaunchVar::Range, pattern, n, value_bytes, floating, h);
  if (c < 0 || c == h) return heuristic;
  return range_candidate(pattern, c, n);
}

}  // namespace Impl

// ================================================================== closures
namespace Impl {

template <class FunctorType, class... Traits>
class ParallelFor<FunctorType, RangePolicy<Traits...>, MI355X> {
 public:
  using Policy = RangePolicy<Traits...>;
  using Tag    = typename Policy::work_tag;
  using Index  = typename Policy::member_type;

  ParallelFor(const FunctorType& f, const Policy& p) : m_functor(f), m_policy(p) {}
  void set_config(const RangeConfig& c) { m_config = c; m_has_config = true; }

  void execute() const {
    const int64_t b = static_cast<int64_t>(m_policy.begin());
    const int64_t e = static_cast<int64_t>(m_policy.end());
    if (e <= b) return;
    RangeConfig c = m_has_config ? m_config
                    : range_overrides().set[0]
                        ? range_overrides().cfg[0]
                        : tuned_range_config<FunctorType>(0, e - b, 0, false, default_for_config(e - b, sizeof(FunctorType)));
    c.block       = normalise_block(c.block);
    c.items       = normalise_items(c.items);
    const void* k = range_for_kernel<FunctorType, Tag, Index>(c.items);
    MI355XInstance* inst = m_policy.space().impl_internal_space_instance();
    static_assert(sizeof(FunctorType) <= 4000, "functor exceeds the kernel-argument limit");
    FunctorType f = m_functor;
    int64_t bb = b, ee = e;
    mx_launch_kernel(k, static_cast<unsigned>(range_grid(e - b, c)), c.block, 0, inst->stream, f,
                     bb, ee);
  }

 private:
  FunctorType m_functor;
  Policy m_policy;
  RangeConfig m_config{};
  bool m_has_config = false;
};

// CR = CombinedReducer<F, Tag, Reducer, ValueType>
template <class CR, class... Traits>
class ParallelReduce<CR, RangePolicy<Traits...>, MI355X> {
 public:
  using Policy = RangePolicy<Traits...>;
  using Index  = typename Policy::member_type;
  using V      = typename CR::value_type;

  template <class ResultMem>
  ParallelReduce(const CR& cr, const Policy& p, V* result, ResultMem)
      : m_cr(cr), m_policy(p), m_result(result),
        m_result_on_device(exec_can_access<MI355X, ResultMem>::value) {}
  void set_config(const RangeConfig& c) { m_config = c; m_has_config = true; }

  void execute() const {
    MI355XInstance* inst = m_policy.space().impl_internal_space_instance();
    std::lock_guard<std::mutex> lk(inst->mtx);
    const int64_t b   = static_cast<int64_t>(m_policy.begin());
    const int64_t e   = static_cast<int64_t>(m_policy.end());
    const unsigned cnt = m_cr.value_count();
    const size_t bytes = sizeof(V) * cnt;
    V* dst = m_result;
    bool copy_back = false;
    if (!m_result_on_device) {
      dst       = static_cast<V*>(inst->pinned_space(bytes));
      copy_back = (m_result != nullptr);
    }
    if (dst == nullptr) {  // no result requested: still run (final may have effects)
      dst = static_cast<V*>(inst->pinned_space(bytes));
    }
    CR cr     = m_cr;
    int64_t bb = b, ee = e;
    unsigned* counter = inst->control;
    if constexpr (!CR::is_array) {
      RangeConfig c = m_has_config ? m_config
                      : range_overrides().set[1]
                          ? range_overrides().cfg[1]
                          : tuned_range_config<typename CR::functor_type>(
                                1, e - b, sizeof(V), std::is_floating_point<V>::value, default_reduce_config(e - b));
      c.block       = normalise_block(c.block);
      c.items       = normalise_items(c.items);
      if (e <= b) {
        c.max_blocks = 1;
      }
      int64_t grid = (e <= b) ? 1 : range_grid(e - b, c);
      V* partials  = static_cast<V*>(inst->partials_space(sizeof(V) * grid));
      const void* k = range_reduce_kernel<CR, Index>(c.items);
      mx_launch_kernel(k, static_cast<unsigned>(grid), c.block, 0, inst->stream, cr, bb, ee,
                       partials, counter, dst);
    } else {
      const int block = 256;
      int64_t grid    = (e <= b) ? 1 : ceil_div(e - b, block);
      const int cus   = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
      if (grid > cus * 4) grid = cus * 4;
      V* partials = static_cast<V*>(inst->partials_space(bytes * grid));
      const size_t lds = bytes * block;
      if (lds > 64 * 1024)
        throw_runtime_exception("Kokkos::MI355X parallel_reduce: value_count too large for LDS");
      const void* k = reinterpret_cast<const void*>(&k_range_reduce_array<CR, Index>);
      mx_launch_kernel(k, static_cast<unsigned>(grid), block, lds, inst->stream, cr, bb, ee,
                       partials, counter, dst);
    }
    if (copy_back) {
      KKMX_SAFE_CALL(mx_stream_sync(inst->stream));
      std::memcpy(static_cast<void*>(m_result), dst, bytes);
    }
  }

 private:
  CR m_cr;
  Policy m_policy;
  V* m_result;
  bool m_result_on_device;
  RangeConfig m_config{};
  bool m_has_config = false;
};

// ------------------------------------------------------------------ scan tile body choice
// Which tile body a scan functor type runs (k_range_scan_rr's Keep, for 16
// and 32 values per lane).  Both give bitwise-identical results, so the
// choice is made from the runtime's own timings: the first 2 * kScanTrials
// launches of a functor type over at least kScanTrialElems indices alternate
// the two bodies, each bracketed by events on the launching stream; at a
// later launch, once every trial has completed (checked without blocking),
// the body with the lower best time per index is kept for the type.
// Launches captured into a graph, size-only passes and smaller launches use
// the choice made so far (body 0 until then).
inline constexpr int kScanTrials         = 3;
inline constexpr int64_t kScanTrialElems = int64_t(1) << 21;
struct ScanBodyChoice {
  std::mutex m;
  int chosen   = -1;
  int launched = 0;
  mx_event_t ev[2 * kScanTrials][2] = {};
  int64_t elems[2 * kScanTrials]   = {};
};
template <class CR>
ScanBodyChoice& scan_body_choice() {
  static ScanBodyChoice c;
  return c;
}
// The body for a launch of n indices on stream s; for a trial launch,
// *start / *stop are the events to record around it (else null).
inline int pick_scan_body(ScanBodyChoice& c, int64_t n, mx_stream_t s, mx_event_t* start, mx_event_t* stop) {
  *start = *stop = nullptr;
  if (range_overrides().scan_body >= 0) return range_overrides().scan_body;
  std::lock_guard<std::mutex> lk(c.m);
  if (c.chosen >= 0) return c.chosen;
  if (n < kScanTrialElems || g_size_only_dispatch) return 0;
  int capturing = 0;
  if (mx_stream_capturing(s, &capturing) != 0 || capturing) return 0;
  if (c.launched < 2 * kScanTrials) {
    const int k = c.launched++;
    for (int e = 0; e < 2; ++e)
      if (!c.ev[k][e]) KKMX_SAFE_CALL(mx_event_create(&c.ev[k][e]));
    c.elems[k] = n;
    *start     = c.ev[k][0];
    *stop      = c.ev[k][1];
    return k & 1;
  }
  for (int k = 0; k < 2 * kScanTrials; ++k)
    if (mx_event_query(c.ev[k][1]) != 0) return 0;
  double best[2] = {1e300, 1e300};
  for (int k = 0; k < 2 * kScanTrials; ++k) {
    float ms = 0.f;
    if (mx_event_elapsed_ms(&ms, c.ev[k][0], c.ev[k][1]) != 0) continue;
    best[k & 1] = std::min(best[k & 1], double(ms) / double(c.elems[k]));
  }
  c.chosen = best[1] < best[0] ? 1 : 0;
  return c.chosen;
}

template <class CR, class Index, bool WithTotal>
const void* range_scan_kernel(int ipt, int keep = 0) {
  // Register-resident while IPT values take at most 256 bytes per lane
  // (measured: 32 int64 values per lane at 1024 lanes still run at 5 waves
  // per SIMD and beat the two-pass tile by 2x).
  constexpr int rr_max = int(256 / sizeof(typename CR::value_type));
  switch (ipt >= 64 ? 64 : ipt >= 32 ? 32 : ipt >= 16 ? 16 : 8) {
    case 64:
      if constexpr (rr_max >= 64)
        if (keep) return reinterpret_cast<const void*>(&k_range_scan_rr<CR, Index, 64, WithTotal, 1>);
      return reinterpret_cast<const void*>(&k_range_scan<CR, Index, 64, WithTotal>);
    case 32:
      if constexpr (rr_max >= 32)
        return keep ? reinterpret_cast<const void*>(&k_range_scan_rr<CR, Index, 32, WithTotal, 1>)
                    : reinterpret_cast<const void*>(&k_range_scan_rr<CR, Index, 32, WithTotal>);
      else return reinterpret_cast<const void*>(&k_range_scan<CR, Index, 32, WithTotal>);
    case 16:
      if constexpr (rr_max >= 16)
        return keep ? reinterpret_cast<const void*>(&k_range_scan_rr<CR, Index, 16, WithTotal, 1>)
                    : reinterpret_cast<const void*>(&k_range_scan_rr<CR, Index, 16, WithTotal>);
      else return reinterpret_cast<const void*>(&k_range_scan<CR, Index, 16, WithTotal>);
    default:
      if constexpr (rr_max >= 8) return reinterpret_cast<const void*>(&k_range_scan_rr<CR, Index, 8, WithTotal>);
      else return reinterpret_cast<const void*>(&k_range_scan<CR, Index, 8, WithTotal>);
  }
}

// Shared body of ParallelScan and ParallelScanWithTotal.
template <class CR, class Policy, bool WithTotal>
void range_scan_execute(const CR& cr_in, const Policy& policy, typename CR::value_type* total,
                        bool total_on_device, const RangeConfig* cfg) {
  using V     = typename CR::value_type;
  using Index = typename Policy::member_type;
  static_assert(!CR::is_array, "parallel_scan does not support runtime-length value types");
  MI355XInstance* inst = policy.space().impl_internal_space_instance();
  std::lock_guard<std::mutex> lk(inst->mtx);
  const int64_t b = static_cast<int64_t>(policy.begin());
  const int64_t e = static_cast<int64_t>(policy.end());
  CR cr           = cr_in;
  if (e <= b) {
    if (WithTotal && total) {
      V id;
      cr.init(id);
      if (total_on_device) {
        KKMX_SAFE_CALL(mx_memcpy_async(total, &id, sizeof(V), inst->stream));
        KKMX_SAFE_CALL(mx_stream_sync(inst->stream));
      } else {
        *total = id;
      }
    }
    return;
  }
  RangeConfig c = cfg ? *cfg
                  : range_overrides().set[2]
                      ? range_overrides().cfg[2]
                      : tuned_range_config<typename CR::functor_type>(2, e - b, sizeof(V), std::is_floating_point<V>::value,
                                                                     default_scan_config(e - b));
  c.block = normalise_block(c.block);
  int ipt = c.items >= 64 ? 64 : c.items >= 32 ? 32 : c.items >= 16 ? 16 : 8;
  // The tile body (trial launches alternate the two), then the shape that
  // suits it: at the default shape for large extents (1024 lanes x 32) the
  // contribution-keeping body takes 64 values per lane while they fit its
  // register budget, since it no longer relies on the L2 to serve a re-read
  // and larger tiles halve the claims and look-backs -- once there are at
  // least as many 64-wide tiles as CUs (at 2^23 indices the 32-wide tiles,
  // which fill every CU, take 19.2 us against 22.6 us).
  mx_event_t ev_start = nullptr, ev_stop = nullptr;
  const int keep = ipt >= 16 ? pick_scan_body(scan_body_choice<CR>(), e - b, inst->stream, &ev_start, &ev_stop) : 0;
  constexpr bool keep64_fits = sizeof(V) <= 4 && scan_state_kind<V>::value != 0;
  if (keep64_fits && keep == 1 && !cfg && !range_overrides().set[2] && c.block == 1024 && ipt == 32) {
    const RangeConfig d = default_scan_config(e - b);
    const int64_t cus   = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
    if (d.block == 1024 && d.items == 32 && ceil_div(e - b, int64_t(1024) * 64) >= cus) ipt = 64;
  }
  const int64_t tile   = static_cast<int64_t>(c.block) * ipt;
  const int64_t ntiles = ceil_div(e - b, tile);
  if (ntiles >= kScanMaxTiles) throw_runtime_exception("Kokkos::parallel_scan: extent needs too many tiles");
  ScanTileState<V> st;
  void* status = scan_state_kind<V>::value == 0 ? inst->scan_flags_space(ScanTileState<V>::bytes(ntiles))
                                                : inst->scan_status_space(ScanTileState<V>::bytes(ntiles));
  void* vals   = nullptr;
  if constexpr (scan_state_kind<V>::value == 0)
    vals = inst->scan_values_space(ScanTileState<V>::value_bytes(ntiles));
  st.bind(status, vals);
  V* tot_dst     = total;
  bool copy_back = false;
  if (WithTotal && (!total_on_device || total == nullptr)) {
    tot_dst   = static_cast<V*>(inst->pinned_space(sizeof(V)));
    copy_back = (total != nullptr);
  }
  int64_t bb = b, ee = e, nt = ntiles;
  unsigned* control = inst->control;
  const void* k     = range_scan_kernel<CR, Index, WithTotal>(ipt, keep);
  unsigned grid     = static_cast<unsigned>(ntiles);
  if constexpr (keep64_fits) {
    // 1024 lanes holding 64 values each need more than 64 registers per lane
    // (118 for an int32 sum), so one workgroup fills a CU, and a grid of one
    // workgroup per CU loops over the tiles.
    if (keep == 1 && ipt == 64 && c.block == 1024) {
      k                 = reinterpret_cast<const void*>(&k_range_scan_keep_loop<CR, Index, 64, WithTotal>);
      const int64_t cus = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
      grid              = static_cast<unsigned>(ntiles < cus ? ntiles : cus);
    }
  }
  if (!g_size_only_dispatch) note_scan_launches(inst, 1, inst->stream);
  if (ev_start) KKMX_SAFE_CALL(mx_event_record(ev_start, inst->stream));
  mx_launch_kernel(k, grid, c.block, 0, inst->stream, cr, bb, ee, nt, st, control, tot_dst);
  if (ev_stop) KKMX_SAFE_CALL(mx_event_record(ev_stop, inst->stream));
  if (copy_back) {
    KKMX_SAFE_CALL(mx_stream_sync(inst->stream));
    std::memcpy(static_cast<void*>(total), tot_dst, sizeof(V));
  }
}

template <class CR, class... Traits>
class ParallelScan<CR, RangePolicy<Traits...>, MI355X> {
 public:
  using Policy = RangePolicy<Traits...>;
  ParallelScan(const CR& cr, const Policy& p) : m_cr(cr), m_policy(p) {}
  void set_config(const RangeConfig& c) { m_config = c; m_has_config = true; }
  void execute() const {
    range_scan_execute<CR, Policy, false>(m_cr, m_policy, nullptr, true,
                                          m_has_config ? &m_config : nullptr);
  }

 private:
  CR m_cr;
  Policy m_policy;
  RangeConfig m_config{};
  bool m_has_config = false;
};

template <class CR, class ReturnType, class... Traits>
class ParallelScanWithTotal<CR, RangePolicy<Traits...>, ReturnType, MI355X> {
 public:
  using Policy = RangePolicy<Traits...>;
  using V      = typename CR::value_type;
  ParallelScanWithTotal(const CR& cr, const Policy& p, V* total, bool total_on_device)
      : m_cr(cr), m_policy(p), m_total(total), m_on_device(total_on_device) {}
  void set_config(const RangeConfig& c) { m_config = c; m_has_config = true; }
  void execute() const {
    range_scan_execute<CR, Policy, true>(m_cr, m_policy, m_total, m_on_device,
                                         m_has_config ? &m_config : nullptr);
  }

 private:
  CR m_cr;
  Policy m_policy;
  V* m_total;
  bool m_on_device;
  RangeConfig m_config{};
  bool m_has_config = false;
};

}  // namespace Impl
}  // namespace Kokkos

#endif
