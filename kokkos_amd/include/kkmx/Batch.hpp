// Batched uneven dispatch: many independent reductions, scans and
// element-wise ops over contiguous arrays, executed by one persistent fused
// kernel that takes fixed-size tiles from a device work queue.
//
// The reference runs each such op as its own parallel_reduce /
// parallel_scan / parallel_for (core/src/Kokkos_Parallel_Reduce.hpp:1679-1833,
// core/src/Kokkos_Parallel.hpp:130-173,348-454), one or two kernel launches
// each, or as one hipGraph kernel node per op
// (core/src/HIP/Kokkos_HIP_Graph_Impl.hpp:82-221).  For thousands of small
// ops launch latency dominates; here the whole batch is one launch:
//
//  * plan(): every op is cut into tiles (BatchShape); the tiles of
//    all ops form one global sequence (op 0's tiles, then op 1's, ...) and a
//    tile -> op table is uploaded with the op descriptors.  All device
//    memory is allocated here; submit() allocates and clears nothing, so it
//    can be captured into a hipGraph and replayed;
//  * the kernel is persistent (grid = resident workgroups); each workgroup
//    claims the next tile with one relaxed 64-bit fetch_add on a
//    {epoch, counter} word, so tiles are processed in sequence order;
//  * reduce tiles publish their aggregate; the op's last tile combines the
//    aggregates of its predecessors with the decoupled look-back of the
//    range scan (tiles of other ops count as the identity) and writes the
//    result.  Scan tiles publish aggregate, look back, publish the inclusive
//    prefix (only for scans longer than kBatchDetTiles tiles) and write their
//    outputs.  Reductions and scans of up to kBatchDetTiles tiles combine in
//    a fixed order, so their floating-point results are bit-reproducible run
//    to run; integer results are exact in every case;
//  * the last workgroup to leave resets the queue and advances the epoch
//    carried by every tile status word: the next replay needs no clearing.
#ifndef KKMX_BATCH_HPP
#define KKMX_BATCH_HPP

#include <kkmx/Graph.hpp>

namespace Kokkos {
namespace Experimental {

enum class BatchOp : int {
  ReduceSum     = 0,
  ReduceMin     = 1,
  ReduceMax     = 2,
  ScanExclusive = 3,
  ScanInclusive = 4,
  Axpb          = 5,  // out[i] = alpha * in[i] + beta
};
// Input element type; the accumulator/output type is int64 for both integer
// inputs and double for double.
enum class BatchType : int { Int32 = 0, Int64 = 1, Float64 = 2 };

}  // namespace Experimental

namespace Impl {

struct alignas(16) BatchDesc {
  const void* in;
  void* out;     // scan / axpb output array (accumulator type)
  void* result;  // reduction result or scan total (accumulator type); may be null
  int64_t n;
  int64_t tile0;   // first tile of this op in the global sequence
  int64_t ntiles;  // tiles of this op
  int64_t alpha_bits, beta_bits;
  int32_t op, type;
};

// Tile shapes of the fused kernel: `block` lanes, each holding `k`
// consecutive elements in each of `r` rounds, so a tile is block*k*r
// elements.  k = 4 lets the compiler move each lane's run with one 16-byte
// (int32) or two 16-byte (8-byte types) accesses.
struct BatchShape {
  int block, k, r;
  constexpr int64_t tile() const { return int64_t(block) * k * r; }
};
constexpr int kBatchShapes = 8;
constexpr BatchShape kBatchShape[kBatchShapes] = {{256, 1, 16}, {256, 4, 4}, {256, 4, 8},  {512, 4, 4},
                                                  {1024, 1, 16}, {512, 4, 8}, {1024, 4, 4}, {512, 1, 32}};
constexpr int kBatchDefaultShape = 5;
static_assert(kBatchShapes == kBatchShapeCandidates, "the batch_shape tuning variable lists every tile shape");
constexpr int64_t kBatchDetTiles = 2048;

template <class Acc>
__device__ __forceinline__ Acc batch_identity(int op) {
  if (op == int(Experimental::BatchOp::ReduceMin)) return reduction_identity<Acc>::min();
  if (op == int(Experimental::BatchOp::ReduceMax)) return reduction_identity<Acc>::max();
  return Acc(0);
}

struct BatchJoin {
  int op;
  template <class Acc>
  __device__ __forceinline__ void operator()(Acc& a, const Acc& b) const {
    if (op == int(Experimental::BatchOp::ReduceMin)) a = b < a ? b : a;
    else if (op == int(Experimental::BatchOp::ReduceMax)) a = a < b ? b : a;
    else a += b;
  }
};

template <class Acc>
KOKKOS_FORCEINLINE_FUNCTION Acc batch_bits(int64_t b) {
  Acc a;
  __builtin_memcpy(&a, &b, sizeof(Acc));
  return a;
}

template <class T>
using GlobalPtr = __attribute__((address_space(1))) T*;

template <int K, class T>
__device__ __forceinline__ void batch_store(GlobalPtr<T> p, const T& v) {
  if constexpr (K == 1) __builtin_nontemporal_store(v, p);
  else *p = v;
}

struct BatchLds {
  LdsSlots16<int64_t> wave;  // per-wave totals
  LdsSlots16<int64_t> red;   // look-back / block reduction scratch
  int pos;
};

// One tile of one op.  Wavefront w owns R rounds of 64*K consecutive
// elements; in round r lane l holds elements
//   tile_base + w*64*K*R + r*64*K + l*K + [0, K),
// so each round is one contiguous run and each lane's part of it is K
// consecutive elements.
template <class In, class Acc, int BLOCK, int K, int R>
__device__ __forceinline__ void batch_tile(const BatchDesc& d, int64_t tile, uint32_t epoch, uint64_t* status,
                                           BatchLds& lds) {
  using Experimental::BatchOp;
  constexpr int64_t TILE = int64_t(BLOCK) * K * R;
  constexpr int nw       = BLOCK / 64;
  const int op           = d.op;
  const BatchJoin join{op};
  const Acc identity  = batch_identity<Acc>(op);
  const int64_t local = tile - d.tile0;
  const int w         = threadIdx.x >> 6;
  const int l         = wave::lane();
  const int64_t base  = local * TILE + int64_t(w) * (64 * K * R) + int64_t(l) * K;
  // The descriptors hold generic pointers; the arrays are device (global)
  // memory, so address them as such (global_load/store rather than flat).
  const GlobalPtr<const In> in = (GlobalPtr<const In>)(d.in);
  const int64_t n              = d.n;
  const bool full              = (local + 1) * TILE <= n;

  // Inputs are read once and outputs never re-read: single-element lanes
  // use non-temporal accesses; K-element lanes use plain ones, which the
  // compiler merges into 16-byte accesses.
  auto ld = [&](int64_t i) -> Acc {
    if constexpr (K == 1) return static_cast<Acc>(__builtin_nontemporal_load(in + i));
    else return static_cast<Acc>(in[i]);
  };
  Acc v[R][K];
  if (full) {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < K; ++k) v[r][k] = ld(base + r * 64 * K + k);
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int64_t i = base + r * 64 * K + k;
        v[r][k]         = i < n ? ld(i) : identity;
      }
  }

  if (op == int(BatchOp::Axpb)) {
    const Acc a = batch_bits<Acc>(d.alpha_bits), b = batch_bits<Acc>(d.beta_bits);
    const GlobalPtr<Acc> out = (GlobalPtr<Acc>)(d.out);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int64_t i = base + r * 64 * K + k;
        if (full || i < n) batch_store<K>(out + i, a * v[r][k] + b);
      }
    return;
  }

  ScanTileState<Acc> st;
  st.words           = status;
  Acc* s_wave        = reinterpret_cast<Acc*>(lds.wave.get());
  Acc* s_red         = reinterpret_cast<Acc*>(lds.red.get());
  const bool is_scan = op == int(BatchOp::ScanExclusive) || op == int(BatchOp::ScanInclusive);

  // Lane totals per round; for scans their wave-inclusive scans are kept for
  // the output pass.  The wavefront total combines the rounds in order.
  Acc inc[R];
  Acc wave_total = identity;
  if (is_scan) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      Acc s = v[r][0];
#pragma unroll
      for (int k = 1; k < K; ++k) join(s, v[r][k]);
      inc[r] = wave::inclusive_scan(s, identity, join);
      join(wave_total, wave::readlane(inc[r], 63));
    }
  } else {
    Acc mine = v[0][0];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (r | k) join(mine, v[r][k]);
    wave_total = wave::reduce(mine, identity, join);
  }
  if (l == 0) s_wave[w] = wave_total;
  __syncthreads();
  const Acc x    = (l < nw) ? s_wave[l] : identity;
  const Acc xs   = wave::inclusive_scan(x, identity, join);
  const Acc tagg = wave::readlane(xs, nw - 1);
  const Acc wpre = (w == 0) ? identity : wave::readlane(xs, w - 1);
  __syncthreads();  // s_wave is reused by the next tile

  // Ops of up to kBatchDetTiles tiles never publish inclusive prefixes past
  // their first tile, so every look-back sums all predecessor aggregates in
  // the same fixed order: results are bit-reproducible run to run (a few
  // look-back rounds per tile at most).  Longer scans also publish
  // inclusive prefixes, which shortens the walk but makes the floating-point
  // combination order depend on timing.
  const bool last = local == d.ntiles - 1;
  const bool det  = d.ntiles <= kBatchDetTiles;
  Acc prefix      = identity;
  if (d.ntiles > 1) {
    if (local == 0) {
      if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanIncl), tagg);
    } else if (is_scan || last) {
      if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanAgg), tagg);
      prefix = scan_lookback(st, tile, epoch, identity, join, s_red, &lds.pos, d.tile0);
      if (is_scan && !det && threadIdx.x == 0) {
        Acc incl = prefix;
        join(incl, tagg);
        st.publish(tile, scan_flag(epoch, kScanIncl), incl);
      }
    } else {
      if (threadIdx.x == 0) st.publish(tile, scan_flag(epoch, kScanAgg), tagg);
    }
  }
  if (last && threadIdx.x == 0 && d.result) {
    Acc tot = prefix;
    join(tot, tagg);
    *static_cast<Acc*>(d.result) = tot;
  }
  if (!is_scan) return;

  const bool exclusive     = op == int(BatchOp::ScanExclusive);
  const GlobalPtr<Acc> out = (GlobalPtr<Acc>)(d.out);
  Acc running           = prefix;
  join(running, wpre);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    Acc e = running;  // exclusive prefix of the lane's first element
    join(e, wave::shift_up(inc[r], identity));
    join(running, wave::readlane(inc[r], 63));
    Acc o[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (exclusive) {
        o[k] = e;
        join(e, v[r][k]);
      } else {
        join(e, v[r][k]);
        o[k] = e;
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t i = base + r * 64 * K + k;
      if (full || i < n) batch_store<K>(out + i, o[k]);
    }
  }
}

// control[0..1]: {tile counter (low 32), epoch (high 32)}; control[2]: number
// of workgroups that have left.  A workgroup claims its next tile only when
// it is ready to start it: a claim made earlier (prefetched) would hold a
// tile whose aggregate later tiles' look-backs wait for, until the current
// tile is done, and measured 8x slower.
template <int BLOCK, int K, int R>
__global__ __launch_bounds__(BLOCK) void k_batch(const BatchDesc* __restrict__ ops,
                                                 const int32_t* __restrict__ tile_op, int64_t total_tiles,
                                                 uint64_t* status, unsigned* control) {
  __shared__ BatchLds lds;
  __shared__ uint64_t s_claim;
  uint64_t* claim_word = reinterpret_cast<uint64_t*>(control);
  uint32_t epoch       = 0;
  while (true) {
    if (threadIdx.x == 0)
      s_claim = __hip_atomic_fetch_add(claim_word, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint64_t claim = s_claim;
    __syncthreads();  // every thread has read s_claim before it is rewritten
    const int64_t tile = static_cast<int64_t>(static_cast<uint32_t>(claim));
    epoch              = static_cast<uint32_t>(claim >> 32);
    if (tile >= total_tiles) break;
    const BatchDesc d = ops[tile_op[tile]];
    switch (d.type) {
      case int(Experimental::BatchType::Int32):
        batch_tile<int32_t, int64_t, BLOCK, K, R>(d, tile, epoch, status, lds);
        break;
      case int(Experimental::BatchType::Int64):
        batch_tile<int64_t, int64_t, BLOCK, K, R>(d, tile, epoch, status, lds);
        break;
      default: batch_tile<double, double, BLOCK, K, R>(d, tile, epoch, status, lds); break;
    }
  }
  if (threadIdx.x == 0) {
    const unsigned left = __hip_atomic_fetch_add(control + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (left == gridDim.x - 1) {
      // Every workgroup has made its last claim: restart the queue at tile 0
      // with the next epoch.
      __hip_atomic_store(control + 2, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(claim_word, static_cast<uint64_t>(epoch + 1u) << 32, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

template <int S>
inline const void* batch_kernel() {
  constexpr BatchShape sh = kBatchShape[S];
  return reinterpret_cast<const void*>(&k_batch<sh.block, sh.k, sh.r>);
}
inline const void* batch_kernel(int shape) {
  switch (shape) {
    case 0: return batch_kernel<0>();
    case 1: return batch_kernel<1>();
    case 2: return batch_kernel<2>();
    case 3: return batch_kernel<3>();
    case 4: return batch_kernel<4>();
    case 5: return batch_kernel<5>();
    case 6: return batch_kernel<6>();
    default: return batch_kernel<7>();
  }
}

// Per-call (plain) execution of one op through the public dispatch API.
template <class In, class Acc>
struct BatchReduceF {
  static constexpr double kkmx_bytes_per_index = sizeof(In), kkmx_flops_per_index = 1;
  View<const In*, MI355XSpace, MemoryTraits<Unmanaged>> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, Acc& u) const { u += static_cast<Acc>(x(i)); }
};
template <class In, class Acc>
struct BatchMinF {
  static constexpr double kkmx_bytes_per_index = sizeof(In), kkmx_flops_per_index = 1;
  View<const In*, MI355XSpace, MemoryTraits<Unmanaged>> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, Acc& u) const {
    const Acc a = static_cast<Acc>(x(i));
    if (a < u) u = a;
  }
};
template <class In, class Acc>
struct BatchMaxF {
  static constexpr double kkmx_bytes_per_index = sizeof(In), kkmx_flops_per_index = 1;
  View<const In*, MI355XSpace, MemoryTraits<Unmanaged>> x;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, Acc& u) const {
    const Acc a = static_cast<Acc>(x(i));
    if (a > u) u = a;
  }
};
template <class In, class Acc, bool Exclusive>
struct BatchScanF {
  static constexpr double kkmx_bytes_per_index = sizeof(In) + sizeof(Acc), kkmx_flops_per_index = 1;
  View<const In*, MI355XSpace, MemoryTraits<Unmanaged>> x;
  View<Acc*, MI355XSpace, MemoryTraits<Unmanaged>> y;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, Acc& u, const bool final) const {
    const Acc a = static_cast<Acc>(x(i));
    if (Exclusive && final) y(i) = u;
    u += a;
    if (!Exclusive && final) y(i) = u;
  }
};
template <class In, class Acc>
struct BatchAxpbF {
  static constexpr double kkmx_bytes_per_index = sizeof(In) + sizeof(Acc), kkmx_flops_per_index = 2;
  View<const In*, MI355XSpace, MemoryTraits<Unmanaged>> x;
  View<Acc*, MI355XSpace, MemoryTraits<Unmanaged>> y;
  Acc a, b;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { y(i) = a * static_cast<Acc>(x(i)) + b; }
};

template <class In, class Acc>
void batch_plain_one(const BatchDesc& d, const MI355X& ex) {
  using Experimental::BatchOp;
  using P = RangePolicy<MI355X, IndexType<int64_t>>;
  const P p(ex, 0, d.n);
  View<const In*, MI355XSpace, MemoryTraits<Unmanaged>> x(static_cast<const In*>(d.in), d.n);
  View<Acc*, MI355XSpace, MemoryTraits<Unmanaged>> y(static_cast<Acc*>(d.out), d.out ? d.n : 0);
  View<Acc, MI355XSpace, MemoryTraits<Unmanaged>> r(static_cast<Acc*>(d.result));
  switch (BatchOp(d.op)) {
    case BatchOp::ReduceSum: parallel_reduce("batch.sum", p, BatchReduceF<In, Acc>{x}, r); break;
    case BatchOp::ReduceMin:
      parallel_reduce("batch.min", p, BatchMinF<In, Acc>{x}, Min<Acc, MI355XSpace>(r));
      break;
    case BatchOp::ReduceMax:
      parallel_reduce("batch.max", p, BatchMaxF<In, Acc>{x}, Max<Acc, MI355XSpace>(r));
      break;
    case BatchOp::ScanExclusive:
      if (d.result) parallel_scan("batch.scan", p, BatchScanF<In, Acc, true>{x, y}, r);
      else parallel_scan("batch.scan", p, BatchScanF<In, Acc, true>{x, y});
      break;
    case BatchOp::ScanInclusive:
      if (d.result) parallel_scan("batch.scan", p, BatchScanF<In, Acc, false>{x, y}, r);
      else parallel_scan("batch.scan", p, BatchScanF<In, Acc, false>{x, y});
      break;
    case BatchOp::Axpb:
      parallel_for("batch.axpb", p,
                   BatchAxpbF<In, Acc>{x, y, batch_bits<Acc>(d.alpha_bits), batch_bits<Acc>(d.beta_bits)});
      break;
  }
}

// Scan total for graph nodes (then_parallel_scan has no total argument):
// total = out[n-1] (+ in[n-1] for an exclusive scan), identity for n == 0.
template <class In, class Acc, bool Exclusive>
struct BatchScanTotalF {
  const In* x;
  const Acc* y;
  Acc* total;
  int64_t n;
  KOKKOS_INLINE_FUNCTION void operator()() const {
    Acc t = Acc(0);
    if (n > 0) {
      t = y[n - 1];
      if (Exclusive) t += static_cast<Acc>(x[n - 1]);
    }
    *total = t;
  }
};

template <class In, class Acc>
void batch_graph_one(const BatchDesc& d, const Experimental::GraphNodeRef<MI355X>& root) {
  using Experimental::BatchOp;
  using P = RangePolicy<MI355X, IndexType<int64_t>>;
  const P p(0, d.n);
  View<const In*, MI355XSpace, MemoryTraits<Unmanaged>> x(static_cast<const In*>(d.in), d.n);
  View<Acc*, MI355XSpace, MemoryTraits<Unmanaged>> y(static_cast<Acc*>(d.out), d.out ? d.n : 0);
  View<Acc, MI355XSpace, MemoryTraits<Unmanaged>> r(static_cast<Acc*>(d.result));
  switch (BatchOp(d.op)) {
    case BatchOp::ReduceSum: root.then_parallel_reduce("batch.sum", p, BatchReduceF<In, Acc>{x}, r); break;
    case BatchOp::ReduceMin:
      root.then_parallel_reduce("batch.min", p, BatchMinF<In, Acc>{x}, Min<Acc, MI355XSpace>(r));
      break;
    case BatchOp::ReduceMax:
      root.then_parallel_reduce("batch.max", p, BatchMaxF<In, Acc>{x}, Max<Acc, MI355XSpace>(r));
      break;
    case BatchOp::ScanExclusive: {
      auto s = root.then_parallel_scan("batch.scan", p, BatchScanF<In, Acc, true>{x, y});
      if (d.result)
        s.then("batch.total", BatchScanTotalF<In, Acc, true>{x.data(), y.data(), r.data(), d.n});
      break;
    }
    case BatchOp::ScanInclusive: {
      auto s = root.then_parallel_scan("batch.scan", p, BatchScanF<In, Acc, false>{x, y});
      if (d.result)
        s.then("batch.total", BatchScanTotalF<In, Acc, false>{x.data(), y.data(), r.data(), d.n});
      break;
    }
    case BatchOp::Axpb:
      root.then_parallel_for("batch.axpb", p,
                             BatchAxpbF<In, Acc>{x, y, batch_bits<Acc>(d.alpha_bits), batch_bits<Acc>(d.beta_bits)});
      break;
  }
}

}  // namespace Impl

namespace Experimental {

// A fixed set of independent ops over device arrays, planned once and
// submitted many times.  Pointers are device memory the caller owns; they
// must stay valid while the batch is submitted.
class BatchDispatch {
 public:
  explicit BatchDispatch(const MI355X& ex = MI355X()) : m_exec(ex) {}
  ~BatchDispatch() { release(); }
  BatchDispatch(const BatchDispatch&)            = delete;
  BatchDispatch& operator=(const BatchDispatch&) = delete;

  // Adds one op; returns its index.  `out` is required for scans and Axpb,
  // `result` for reductions (optional for scans: the total).
  int add(BatchOp op, BatchType type, const void* in, int64_t n, void* out, void* result, double alpha = 1.0,
          double beta = 0.0) {
    if (m_planned) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: add() after plan()");
    if (n < 0) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: negative extent");
    const int o = int(op);
    if (o < 0 || o > int(BatchOp::Axpb)) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: bad op");
    if (int(type) < 0 || int(type) > int(BatchType::Float64))
      Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: bad element type");
    const bool needs_out = op == BatchOp::ScanExclusive || op == BatchOp::ScanInclusive || op == BatchOp::Axpb;
    const bool is_reduce = o <= int(BatchOp::ReduceMax);
    if (n > 0 && !in) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: null input");
    if (needs_out && n > 0 && !out) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: null output");
    if (is_reduce && !result) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: null result");
    Impl::BatchDesc d{};
    d.in     = in;
    d.out    = needs_out ? out : nullptr;
    d.result = (op == BatchOp::Axpb) ? nullptr : result;
    d.n      = n;
    d.op     = o;
    d.type   = int(type);
    if (type == BatchType::Float64) {
      std::memcpy(&d.alpha_bits, &alpha, 8);
      std::memcpy(&d.beta_bits, &beta, 8);
    } else {
      d.alpha_bits = static_cast<int64_t>(alpha);
      d.beta_bits  = static_cast<int64_t>(beta);
    }
    m_ops.push_back(d);
    return static_cast<int>(m_ops.size()) - 1;
  }

  // Builds the tile sequence and the device tables.  Zero-length ops keep
  // one tile so their result (the identity) is still written.
  void plan() {
    if (m_planned) return;
    if (!m_shape_set) m_shape = tuned_shape();
    const Impl::BatchShape shape = Impl::kBatchShape[m_shape];
    const int64_t tile_elems     = shape.tile();
    int64_t t                    = 0;
    std::vector<int32_t> tile_op;
    for (size_t o = 0; o < m_ops.size(); ++o) {
      auto& d  = m_ops[o];
      d.ntiles = d.n > 0 ? (d.n + tile_elems - 1) / tile_elems : 1;
      d.tile0  = t;
      for (int64_t k = 0; k < d.ntiles; ++k) tile_op.push_back(static_cast<int32_t>(o));
      t += d.ntiles;
      m_elements += d.n;
    }
    if (t >= Impl::kScanMaxTiles) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: too many tiles");
    m_tiles = t;
    if (m_ops.empty()) {
      m_planned = true;
      return;
    }
    mx_stream_t s = m_exec.stream();
    KKMX_SAFE_CALL(mx_malloc(&m_desc, sizeof(Impl::BatchDesc) * m_ops.size(), nullptr));
    KKMX_SAFE_CALL(mx_malloc(&m_tile_op, sizeof(int32_t) * tile_op.size(), nullptr));
    KKMX_SAFE_CALL(mx_malloc(&m_status, 16 * size_t(m_tiles), nullptr));
    KKMX_SAFE_CALL(mx_malloc(&m_control, 64, nullptr));
    KKMX_SAFE_CALL(mx_memcpy_async(m_desc, m_ops.data(), sizeof(Impl::BatchDesc) * m_ops.size(), s));
    KKMX_SAFE_CALL(mx_memcpy_async(m_tile_op, tile_op.data(), sizeof(int32_t) * tile_op.size(), s));
    KKMX_SAFE_CALL(mx_memset_async(m_status, 0, 16 * size_t(m_tiles), s));
    unsigned ctrl[16] = {};
    ctrl[1]           = 1;  // epoch 1: zeroed status words are never valid
    KKMX_SAFE_CALL(mx_memcpy_async(m_control, ctrl, sizeof(ctrl), s));
    const void* k = Impl::batch_kernel(m_shape);
    int per_cu    = 0;
    KKMX_SAFE_CALL(mx_occupancy(k, shape.block, 0, &per_cu));
    if (per_cu < 1) per_cu = 1;
    const int64_t cap = int64_t(MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256) * per_cu;
    m_grid            = static_cast<unsigned>(std::min<int64_t>(cap, m_tiles));
    KKMX_SAFE_CALL(mx_stream_sync(s));
    m_planned = true;
  }

  // The whole batch as one persistent fused kernel on exec's stream
  // (asynchronous; allocates nothing, so it may be stream-captured).
  void submit(const MI355X& ex) const {
    if (!m_planned) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: submit() before plan()");
    if (m_ops.empty()) return;
    const Impl::BatchDesc* ops = static_cast<const Impl::BatchDesc*>(m_desc);
    const int32_t* tile_op     = static_cast<const int32_t*>(m_tile_op);
    int64_t total              = m_tiles;
    uint64_t* status           = static_cast<uint64_t*>(m_status);
    unsigned* control          = static_cast<unsigned*>(m_control);
    if (!Impl::g_size_only_dispatch) impl_note_replays(ex, 1);
    Impl::mx_launch_kernel(Impl::batch_kernel(m_shape), m_grid, Impl::kBatchShape[m_shape].block, 0,
                           ex.stream(), ops, tile_op, total, status, control);
  }
  void submit() const { submit(m_exec); }

  // Counts `k` runs of the fused kernel about to be queued on ex's stream
  // that submit() did not launch itself (replays of a graph that captured
  // submit()); the tile status words, which keep 30 bits of the launch
  // epoch, are zeroed first when 2^29 runs have gone by since the last clear.
  void impl_note_replays(const MI355X& ex, uint64_t k) const {
    if (m_ops.empty()) return;
    if (m_since_clear + k > Impl::kScanClearPeriod) {
      Impl::clear_words_async(m_status, 16 * size_t(m_tiles), ex.stream());
      m_since_clear = 0;
    }
    m_since_clear += k;
  }

  // The same ops as one parallel_reduce / parallel_scan / parallel_for each.
  void submit_plain(const MI355X& ex) const {
    for (const auto& d : m_ops) {
      switch (BatchType(d.type)) {
        case BatchType::Int32: Impl::batch_plain_one<int32_t, int64_t>(d, ex); break;
        case BatchType::Int64: Impl::batch_plain_one<int64_t, int64_t>(d, ex); break;
        case BatchType::Float64: Impl::batch_plain_one<double, double>(d, ex); break;
      }
    }
  }
  void submit_plain() const { submit_plain(m_exec); }

  // The per-op dispatches as an Experimental::Graph: one node (or two, for
  // a scan with a total) per op, all independent.  This is what the
  // reference's hipGraph backend would run for the same batch.
  Graph<MI355X> plain_graph(const MI355X& ex) const {
    return create_graph(ex, [&](const GraphNodeRef<MI355X>& root) {
      for (const auto& d : m_ops) {
        switch (BatchType(d.type)) {
          case BatchType::Int32: Impl::batch_graph_one<int32_t, int64_t>(d, root); break;
          case BatchType::Int64: Impl::batch_graph_one<int64_t, int64_t>(d, root); break;
          case BatchType::Float64: Impl::batch_graph_one<double, double>(d, root); break;
        }
      }
    });
  }

  // Tile shape of the fused kernel (Impl::kBatchShape), before plan().
  void set_shape(int shape) {
    if (m_planned) Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: set_shape() after plan()");
    if (shape < 0 || shape >= Impl::kBatchShapes)
      Impl::throw_runtime_exception("Kokkos::Experimental::BatchDispatch: bad tile shape");
    m_shape     = shape;
    m_shape_set = true;
  }
  int shape() const { return m_shape; }

  // Features the tile-shape request describes the batch with (Impl::
  // LaunchFeatures, pattern 3): total elements, mean element bytes, mean op
  // extent, all-float64 flag and algorithmic bytes per element.
  Impl::LaunchFeatures tuning_features() const {
    int64_t total = 0;
    double bytes = 0, vbytes = 0, f64 = 0;
    for (const auto& d : m_ops) {
      const double esz = d.type == int(BatchType::Int32) ? 4.0 : 8.0;
      const bool outs  = d.op == int(BatchOp::ScanExclusive) || d.op == int(BatchOp::ScanInclusive) ||
                        d.op == int(BatchOp::Axpb);
      total += d.n;
      vbytes += esz * double(d.n);
      bytes += (esz + (outs ? 8.0 : 0.0)) * double(d.n);
      if (d.type == int(BatchType::Float64)) f64 += double(d.n);
    }
    const double t = total > 0 ? double(total) : 1.0;
    const size_t m = m_ops.empty() ? 1 : m_ops.size();
    return Impl::LaunchFeatures{3, total, int64_t(vbytes / t + 0.5), int64_t(double(total) / double(m) + 0.5),
                                f64 >= t, bytes / t, 1.0};
  }

  size_t size() const { return m_ops.size(); }
  int64_t tiles() const { return m_tiles; }
  int64_t elements() const { return m_elements; }
  unsigned grid() const { return m_grid; }
  const std::vector<Impl::BatchDesc>& ops() const { return m_ops; }

 private:
  void release() {
    if (!m_desc) return;
    (void)mx_stream_sync(m_exec.stream());
    (void)mx_free(m_desc, nullptr);
    (void)mx_free(m_tile_op, nullptr);
    (void)mx_free(m_status, nullptr);
    (void)mx_free(m_control, nullptr);
    m_desc = nullptr;
  }
  // The tile shape: the learned tuner's answer when a tuning tool is
  // loaded, else the default.
  int tuned_shape() const {
    if (m_ops.empty() || !Tools::Experimental::have_tuning_tool()) return Impl::kBatchDefaultShape;
    const int c = Impl::request_launch_candidate(Impl::LaunchVar::Batch, tuning_features(), Impl::kBatchDefaultShape);
    return (c >= 0 && c < Impl::kBatchShapes) ? c : Impl::kBatchDefaultShape;
  }
  MI355X m_exec;
  std::vector<Impl::BatchDesc> m_ops;
  bool m_planned     = false;
  bool m_shape_set   = false;
  int m_shape        = Impl::kBatchDefaultShape;
  int64_t m_tiles    = 0;
  int64_t m_elements = 0;
  unsigned m_grid    = 0;
  void* m_desc       = nullptr;
  void* m_tile_op    = nullptr;
  void* m_status     = nullptr;
  void* m_control    = nullptr;
  mutable uint64_t m_since_clear = 0;
};

}  // namespace Experimental
}  // namespace Kokkos

#endif
