// Sorting for the MI355X execution space: Kokkos::BinSort with BinOp1D /
// BinOp3D, and Kokkos::sort.
//
// Reference: algorithms/src/sorting/Kokkos_BinSortPublicAPI.hpp (BinSort:
// count keys per bin, exclusive scan of the counts into bin offsets, place
// every index at its bin's next slot, optionally sort within bins, then
// permute values through the permutation vector),
// algorithms/src/sorting/Kokkos_BinOpsPublicAPI.hpp (BinOp1D / BinOp3D:
// bin(keys, i), max_bins(), and the in-bin comparison), and
// algorithms/src/sorting/Kokkos_SortPublicAPI.hpp (sort(view),
// sort(exec, view), sort(view, comparator), sort(view, begin, end)).
//
// Kokkos::sort of a contiguous rank-1 device View of 4- or 8-byte
// arithmetic keys is an LSD radix sort written for gfx950: 8-bit digits,
// one histogram pass and one scatter pass per digit over 4096-key tiles of
// 256 lanes, grouped into one segment of consecutive tiles per block (about
// four blocks per CU).  The scatter ranks keys with equal digits inside
// each wavefront from eight ballots, combines the wavefronts through LDS
// into a digit-sorted copy of the tile in LDS, and writes that copy out so
// each digit's keys leave as one contiguous run, at offsets from a
// digit-major exclusive scan of the per-segment histograms: every pass
// reads its keys twice and writes them once, stably.  The first histogram
// also folds the AND and OR of all keys, and passes whose digit is the same
// in every key (the high digits of small-range keys) are skipped.  Signed
// and floating-point keys are mapped to unsigned order on the first pass
// that runs and back on the last.  As in the reference's HIP path (thrust::sort, a radix
// sort for arithmetic keys: algorithms/src/sorting/impl/
// Kokkos_SortImpl.hpp:222-240), -0.0 and +0.0 compare equal and keep their
// input order, and NaNs go below -inf (sign bit set) or above +inf.  Other
// element types, comparators and host Views sort on the host with
// std::sort (results identical to an ascending sort).
#ifndef KKMX_SORT_HPP
#define KKMX_SORT_HPP

#include <Kokkos_Core.hpp>

#include <algorithm>
#include <mutex>
#include <vector>

namespace Kokkos {

// ------------------------------------------------------------------ bin ops
template <class KeyViewType>
struct BinOp1D {
  int max_bins_ = 0;
  double mul_   = 0;
  double min_   = 0;
  BinOp1D() = delete;
  // max_bins bins over [min, max] (one extra bin holds max itself); integer
  // keys whose range is at most max_bins get one value per bin.
  BinOp1D(int max_bins, typename KeyViewType::const_value_type min, typename KeyViewType::const_value_type max)
      : max_bins_(max_bins + 1),
        mul_(static_cast<double>(max_bins) / (static_cast<double>(max) - static_cast<double>(min))),
        min_(static_cast<double>(min)) {
    if (std::is_integral<typename KeyViewType::non_const_value_type>::value &&
        static_cast<double>(max) - static_cast<double>(min) <= static_cast<double>(max_bins))
      mul_ = 1.0;
  }
  template <class ViewType>
  KOKKOS_INLINE_FUNCTION int bin(ViewType& keys, const int& i) const {
    return static_cast<int>(mul_ * (static_cast<double>(keys(i)) - min_));
  }
  KOKKOS_INLINE_FUNCTION int max_bins() const { return max_bins_; }
  // true: key i1 goes before key i2 inside a bin
  template <class ViewType, class I1, class I2>
  KOKKOS_INLINE_FUNCTION bool operator()(ViewType& keys, I1& i1, I2& i2) const {
    return keys(i1) < keys(i2);
  }
};

template <class KeyViewType>
struct BinOp3D {
  int max_bins_[3] = {0, 0, 0};
  double mul_[3]   = {0, 0, 0};
  double min_[3]   = {0, 0, 0};
  BinOp3D() = delete;
  BinOp3D(int max_bins[], typename KeyViewType::const_value_type min[],
          typename KeyViewType::const_value_type max[]) {
    for (int d = 0; d < 3; ++d) {
      max_bins_[d] = max_bins[d];
      mul_[d]      = static_cast<double>(max_bins[d]) / (static_cast<double>(max[d]) - static_cast<double>(min[d]));
      min_[d]      = static_cast<double>(min[d]);
    }
  }
  template <class ViewType>
  KOKKOS_INLINE_FUNCTION int bin(ViewType& keys, const int& i) const {
    const int b0 = int(mul_[0] * (keys(i, 0) - min_[0]));
    const int b1 = int(mul_[1] * (keys(i, 1) - min_[1]));
    const int b2 = int(mul_[2] * (keys(i, 2) - min_[2]));
    return (b0 * max_bins_[1] + b1) * max_bins_[2] + b2;
  }
  KOKKOS_INLINE_FUNCTION int max_bins() const { return max_bins_[0] * max_bins_[1] * max_bins_[2]; }
  // lexicographically greater first (the reference's in-bin order)
  template <class ViewType, class I1, class I2>
  KOKKOS_INLINE_FUNCTION bool operator()(ViewType& keys, I1& i1, I2& i2) const {
    for (int d = 0; d < 3; ++d) {
      if (keys(i1, d) > keys(i2, d)) return true;
      if (keys(i1, d) != keys(i2, d)) return false;
    }
    return false;
  }
};

namespace Impl {

template <class Keys, class Op>
struct BinCountF {
  Keys keys;
  Op op;
  View<int*, MI355XSpace, MemoryTraits<Unmanaged>> count;
  int begin;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const { atomic_add(&count(op.bin(keys, begin + i)), 1); }
};
struct BinOffsetF {
  View<const int*, MI355XSpace, MemoryTraits<Unmanaged>> count;
  View<int*, MI355XSpace, MemoryTraits<Unmanaged>> offsets;
  KOKKOS_INLINE_FUNCTION void operator()(const int i, int& upd, const bool final) const {
    if (final) offsets(i) = upd;
    upd += count(i);
  }
};
template <class Keys, class Op>
struct BinPlaceF {
  Keys keys;
  Op op;
  View<int*, MI355XSpace, MemoryTraits<Unmanaged>> fill;
  View<const int*, MI355XSpace, MemoryTraits<Unmanaged>> offsets;
  View<int*, MI355XSpace, MemoryTraits<Unmanaged>> order;
  int begin;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    const int j = begin + i;
    const int b = op.bin(keys, j);
    order(offsets(b) + atomic_fetch_add(&fill(b), 1)) = j;
  }
};
template <class T>
struct is_bin_op_1d : std::false_type {};
template <class K>
struct is_bin_op_1d<BinOp1D<K>> : std::true_type {};

template <class Dst, class Keys>
struct BinKeyCopyF {
  Dst dst;
  Keys keys;
  int begin;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const { dst(i) = keys(begin + i); }
};
struct BinShiftF {
  View<int*, MI355XSpace, MemoryTraits<Unmanaged>> order;
  int begin;
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const { order(i) += begin; }
};

// Insertion sort of one bin's indices by one thread, as the reference does on
// devices (quadratic in the bin size: meant for the small bins BinSort is
// used with); equal keys keep index order, so the result does not depend on
// the order the binning atomics ran in.
template <class Keys, class Op>
struct BinSortWithinF {
  Keys keys;
  Op op;
  View<const int*, MI355XSpace, MemoryTraits<Unmanaged>> count;
  View<const int*, MI355XSpace, MemoryTraits<Unmanaged>> offsets;
  View<int*, MI355XSpace, MemoryTraits<Unmanaged>> order;
  KOKKOS_INLINE_FUNCTION void operator()(const int b) const {
    const int lo = offsets(b), hi = lo + count(b);
    for (int k = lo + 1; k < hi; ++k) {
      int cur = order(k);
      int j   = k - 1;
      while (j >= lo) {
        int prev       = order(j);
        const bool before = op(keys, cur, prev) || (!op(keys, prev, cur) && cur < prev);
        if (!before) break;
        order(j + 1) = prev;
        --j;
      }
      order(j + 1) = cur;
    }
  }
};
template <class Dst, class Src>
struct PermuteRowsF {
  Dst dst;  // rows [0, len)
  Src src;
  View<const int*, MI355XSpace, MemoryTraits<Unmanaged>> order;
  int shift;  // values index = order(i) + shift
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    const int s = order(i) + shift;
    if constexpr (Dst::rank == 1) {
      dst(i) = src(s);
    } else {
      for (size_t c = 0; c < dst.extent(1); ++c) dst(i, c) = src(s, c);
    }
  }
};
template <class Dst, class Src>
struct CopyRowsF {
  Dst dst;
  Src src;
  int begin;  // dst rows [begin, begin + len)
  KOKKOS_INLINE_FUNCTION void operator()(const int i) const {
    if constexpr (Dst::rank == 1) {
      dst(begin + i) = src(i);
    } else {
      for (size_t c = 0; c < dst.extent(1); ++c) dst(begin + i, c) = src(i, c);
    }
  }
};

}  // namespace Impl

// ------------------------------------------------------------------ radix sort
namespace Impl {

template <class T>
struct RadixKey {
  using U = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
  static constexpr U top = U(1) << (8 * sizeof(T) - 1);
  KOKKOS_FORCEINLINE_FUNCTION static U to(T v) {
    U u;
    __builtin_memcpy(&u, &v, sizeof(T));
    if constexpr (std::is_floating_point<T>::value) return (u & top) ? ~u : (u | top);
    else if constexpr (std::is_signed<T>::value) return u ^ top;
    else return u;
  }
  // Digit of an encoded key.  -0.0 (encoded ~top) takes the digits of +0.0
  // (encoded top), so the two compare equal and keep their input order, as
  // under operator<; the key itself keeps its bits.
  KOKKOS_FORCEINLINE_FUNCTION static U canon(U u) {
    if constexpr (std::is_floating_point<T>::value) return u == U(~top) ? top : u;
    else return u;
  }
  KOKKOS_FORCEINLINE_FUNCTION static unsigned digit(U u, int shift) { return unsigned(canon(u) >> shift) & 255u; }
  KOKKOS_FORCEINLINE_FUNCTION static T from(U u) {
    if constexpr (std::is_floating_point<T>::value) u = (u & top) ? (u ^ top) : ~u;
    else if constexpr (std::is_signed<T>::value) u ^= top;
    T v;
    __builtin_memcpy(&v, &u, sizeof(T));
    return v;
  }
};

constexpr int kRadixLanes = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixTile  = kRadixLanes * kRadixItems;

// Each block owns a segment of `tpb` consecutive tiles of kRadixTile keys
// (the last segment may be short).  k_radix_hist counts digit
// R::digit(key, shift) over its segment, in one LDS histogram per wavefront,
// and stores the counts digit-major: hist[d * nblocks + block].
//
// The First variant also folds the bitwise AND and OR of every (canonical)
// key into summary[0] and summary[1]: digits where the two agree are the
// same in every key, and their passes move nothing.
template <class T, bool First>
__global__ __launch_bounds__(256) void k_radix_hist(const void* in, const int64_t n, const int shift, uint32_t* hist,
                                                    const int64_t nblocks, const int64_t tpb,
                                                    typename RadixKey<T>::U* summary) {
  using R = RadixKey<T>;
  using U = typename R::U;
  constexpr int NW = kRadixLanes / 64;
  __shared__ uint32_t h[NW][256];
  __shared__ U s_and, s_or;
  const int t = threadIdx.x, w = t >> 6;
#pragma unroll
  for (int k = 0; k < NW; ++k) h[k][t] = 0;
  if (First && t == 0) {
    s_and = ~U(0);
    s_or  = U(0);
  }
  U all_and = ~U(0), any_or = U(0);
  __syncthreads();
  const int64_t first = int64_t(blockIdx.x) * tpb * kRadixTile;
  const int64_t last  = first + tpb * kRadixTile < n ? first + tpb * kRadixTile : n;
  for (int64_t tb = first; tb < last; tb += kRadixTile) {
    U u[kRadixItems];
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
      const int64_t i = tb + int64_t(r) * kRadixLanes + t;
      u[r]            = 0;
      if (i < last) u[r] = First ? R::to(static_cast<const T*>(in)[i]) : static_cast<const U*>(in)[i];
    }
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
      if (tb + int64_t(r) * kRadixLanes + t < last) {
        atomicAdd(&h[w][R::digit(u[r], shift)], 1u);
        if (First) {
          all_and &= R::canon(u[r]);
          any_or |= R::canon(u[r]);
        }
      }
    }
  }
  if (First) {
    atomicAnd(&s_and, all_and);
    atomicOr(&s_or, any_or);
  }
  __syncthreads();
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) c += h[k][t];
  hist[int64_t(t) * nblocks + blockIdx.x] = c;
  if (First && t == 0) {
    atomicAnd(&summary[0], s_and);
    atomicOr(&summary[1], s_or);
  }
}

// Stable placement of a segment's keys at their digit's offsets, one tile at
// a time.  Wavefront w ranks keys [w * 1024, (w + 1) * 1024) of the tile with
// eight ballots per 64 keys and its own LDS digit counters (no block barrier
// per round), the block turns the per-wavefront counts into slots of a
// digit-sorted copy of the tile in LDS, and writes that copy out so each
// digit's keys leave as one contiguous run.  Lane t carries digit t's next
// output position from tile to tile in a register.
//
// A payload word PT (4 or 8 bytes) can travel with each key through the same
// slots: the values of a sort by key, or with Iota the key's input position,
// generated on the first pass, so the last pass leaves the sorting
// permutation.  NoPayload sorts keys alone.
struct NoPayload {};

template <class T, bool First, bool Last, class PT, bool Iota>
__global__ __launch_bounds__(256) void k_radix_scatter(const void* in, void* out, const PT* in_p, PT* out_p,
                                                       const int64_t n, const int shift, const uint32_t* offsets,
                                                       const int64_t nblocks, const int64_t tpb) {
  constexpr bool Pairs = !std::is_same<PT, NoPayload>::value;
  using R = RadixKey<T>;
  using U = typename R::U;
  constexpr int NW   = kRadixLanes / 64;
  constexpr int WKEY = kRadixTile / NW;  // keys per wavefront per tile
  static_assert(WKEY == 64 * kRadixItems, "each lane ranks kRadixItems keys per tile");
  __shared__ uint32_t gshift[256];  // output position of tile slot j of digit d: gshift[d] + j
  __shared__ uint32_t wcnt[NW][256];
  __shared__ uint32_t wsum[NW];
  __shared__ U tile[kRadixTile];
  __shared__ PT ptile[Pairs ? kRadixTile : 1];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const uint64_t lt = (uint64_t(1) << l) - 1;
  uint32_t gnext    = offsets[int64_t(t) * nblocks + blockIdx.x];  // digit t's next output position
  const int64_t seg = int64_t(blockIdx.x) * tpb * kRadixTile;
  for (int64_t k = 0; k < tpb; ++k) {
    const int64_t tfirst = seg + k * kRadixTile;
    if (tfirst >= n) break;
#pragma unroll
    for (int q = 0; q < NW; ++q) wcnt[q][t] = 0;
    __syncthreads();  // counters cleared; the previous tile is fully written out
    const int64_t wfirst = tfirst + int64_t(w) * WKEY;
    U u[kRadixItems];
    PT ix[Pairs ? kRadixItems : 1];
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
      const int64_t i = wfirst + int64_t(r) * 64 + l;
      u[r]            = 0;
      if (i < n) u[r] = First ? R::to(static_cast<const T*>(in)[i]) : static_cast<const U*>(in)[i];
      if constexpr (Pairs) {
        if constexpr (First && Iota) ix[r] = PT(i);
        else if (i < n) ix[r] = in_p[i];
      }
    }
    uint32_t rk[kRadixItems];
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
      const bool valid = wfirst + int64_t(r) * 64 + l < n;
      const unsigned d = R::digit(u[r], shift);
      // lanes differing from this lane in some digit bit; the ballot mask of
      // a bit doubles as the per-lane select between it and its complement
      const uint64_t vm = __builtin_amdgcn_ballot_w64(valid);
      uint32_t dlo = ~uint32_t(vm), dhi = ~uint32_t(vm >> 32);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int32_t xb = int32_t(d << (31 - b)) >> 31;  // 0 or all ones
        const uint64_t m = __builtin_amdgcn_ballot_w64(xb != 0);
        dlo |= uint32_t(m) ^ uint32_t(xb);
        dhi |= uint32_t(m >> 32) ^ uint32_t(xb);
      }
      const uint64_t peers = ~((uint64_t(dhi) << 32) | dlo);
      const uint32_t below = __popcll(peers & lt);
      const uint32_t prior = wcnt[w][d];
      rk[r]                = prior + below;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (valid && below == 0) wcnt[w][d] = prior + uint32_t(__popcll(peers));
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    __syncthreads();  // per-wavefront counts complete
    uint32_t total = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const uint32_t c = wcnt[q][t];
      wcnt[q][t]       = total;
      total += c;
    }
    {
      const auto plus    = [](uint32_t& a, const uint32_t& b) { a += b; };
      const uint32_t inc = wave::inclusive_scan(total, 0u, plus);
      if (l == 63) wsum[w] = inc;
      __syncthreads();
      uint32_t before = 0;
#pragma unroll
      for (int q = 0; q < NW; ++q) before += q < w ? wsum[q] : 0u;
      // digit t's first slot in the tile; fold it into the per-wavefront
      // prefixes and into the output shift, so placement and write-out
      // read one LDS word per key each
      const uint32_t first = before + inc - total;
#pragma unroll
      for (int q = 0; q < NW; ++q) wcnt[q][t] += first;
      gshift[t] = gnext - first;
      gnext += total;
    }
    __syncthreads();  // slot bases and output shifts visible
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
      if (wfirst + int64_t(r) * 64 + l < n) {
        const uint32_t slot = wcnt[w][R::digit(u[r], shift)] + rk[r];
        tile[slot]          = u[r];
        if constexpr (Pairs) ptile[slot] = ix[r];
      }
    }
    __syncthreads();  // tile sorted by digit
    const int64_t left = n - tfirst;
    const int count    = left < kRadixTile ? int(left) : kRadixTile;
    for (int j = t; j < count; j += kRadixLanes) {
      const U v          = tile[j];
      const uint32_t pos = gshift[R::digit(v, shift)] + uint32_t(j);
      if (Last) static_cast<T*>(out)[pos] = R::from(v);
      else static_cast<U*>(out)[pos] = v;
      if constexpr (Pairs) out_p[pos] = ptile[j];
    }
    // the next tile writes tile, wcnt and gshift only after its first
    // barrier, which every lane reaches after this write-out
  }
}

struct ExclusiveScanU32 {
  const uint32_t* in;
  uint32_t* out;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i, uint32_t& upd, const bool final) const {
    const uint32_t x = in[i];
    if (final) out[i] = upd;
    upd += x;
  }
};

// Device workspace of the radix sort (ping-pong keys and the two digit
// tables), kept between calls and grown on demand: a fresh allocation per
// call costs more than the smaller sorts themselves and slows the kernels
// right after it.  Calls hold the mutex until their passes have run, so
// sorts on different instances take turns; Kokkos::finalize frees it.
struct RadixWorkspace {
  std::mutex mtx;
  void* buf    = nullptr;
  size_t bytes = 0;
  bool hooked  = false;
};

inline RadixWorkspace& radix_workspace() {
  static RadixWorkspace w;
  return w;
}

inline void radix_workspace_release() {
  RadixWorkspace& w = radix_workspace();
  std::lock_guard<std::mutex> lock(w.mtx);
  if (w.buf) KKMX_SAFE_CALL(mx_free(w.buf, nullptr));
  w.buf    = nullptr;
  w.bytes  = 0;
  w.hooked = false;
}

struct IotaU32 {
  uint32_t* p;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { p[i] = uint32_t(i); }
};

template <class T, class PT, bool Iota>
const void* radix_scatter_kernel(bool first, bool last) {
  if (first) return last ? reinterpret_cast<const void*>(&k_radix_scatter<T, true, true, PT, Iota>)
                         : reinterpret_cast<const void*>(&k_radix_scatter<T, true, false, PT, Iota>);
  return last ? reinterpret_cast<const void*>(&k_radix_scatter<T, false, true, PT, Iota>)
              : reinterpret_cast<const void*>(&k_radix_scatter<T, false, false, PT, Iota>);
}

// Ascending LSD radix sort of n keys in place (device memory), on exec.
// With a payload (n device words of PT), the sort is of (key, payload)
// pairs and the payload is permuted with the keys, in place; with iota the
// payload's input is ignored and it receives the sorting permutation
// (sorted key i came from position payload[i]).
template <class T, class PT = NoPayload, bool Iota = false>
void radix_sort(const MI355X& exec, T* keys, int64_t n, PT* payload = nullptr) {
  static_assert(std::is_same<PT, NoPayload>::value || sizeof(PT) == 4 || sizeof(PT) == 8,
                "radix sort payloads are 4- or 8-byte words");
  static_assert(!Iota || std::is_same<PT, uint32_t>::value, "a permutation payload is 32-bit");
  if (n > int64_t(0x7fffffff)) throw_runtime_exception("Kokkos::sort: more than 2^31 - 1 keys");
  if (n < 2) {
    if constexpr (Iota) {
      if (n == 1) {
        const uint32_t zero = 0;
        KKMX_SAFE_CALL(mx_memcpy_async(payload, &zero, sizeof(zero), exec.stream()));
        exec.fence("Kokkos::sort: one key");
      }
    }
    return;
  }
  using U              = typename RadixKey<T>::U;
  const int64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
  // segments of whole tiles, about four blocks per CU
  const int cus        = MI355X::impl_cu_count() > 0 ? MI355X::impl_cu_count() : 256;
  const int64_t want   = int64_t(cus) * 4;
  const int64_t tpb    = (ntiles + want - 1) / want;
  const int64_t nblk   = (ntiles + tpb - 1) / tpb;
  const int passes     = int(sizeof(T));
  const size_t key_bytes   = (size_t(n) * sizeof(U) + 255) & ~size_t(255);
  const size_t table_bytes = size_t(256 * nblk) * sizeof(uint32_t);
  constexpr bool pairs     = !std::is_same<PT, NoPayload>::value;
  const size_t idx_bytes   = pairs ? (size_t(n) * sizeof(PT) + 255) & ~size_t(255) : 0;
  const size_t need        = key_bytes + 2 * table_bytes + idx_bytes + 256;
  RadixWorkspace& ws       = radix_workspace();
  std::lock_guard<std::mutex> lock(ws.mtx);
  if (ws.bytes < need) {
    if (ws.buf) KKMX_SAFE_CALL(mx_free(ws.buf, nullptr));
    ws.buf   = nullptr;
    ws.bytes = 0;
    KKMX_SAFE_CALL(mx_malloc(&ws.buf, need, nullptr));
    ws.bytes = need;
  }
  if (!ws.hooked) {
    push_finalize_hook(&radix_workspace_release);
    ws.hooked = true;
  }
  char* base   = static_cast<char*>(ws.buf);
  void* buf[2] = {static_cast<void*>(keys), static_cast<void*>(base)};
  uint32_t* hp = reinterpret_cast<uint32_t*>(base + key_bytes);
  uint32_t* op = reinterpret_cast<uint32_t*>(base + key_bytes + table_bytes);
  PT* ibuf[2]   = {payload, pairs ? reinterpret_cast<PT*>(base + key_bytes + 2 * table_bytes) : nullptr};
  U* summary    = reinterpret_cast<U*>(base + key_bytes + 2 * table_bytes + idx_bytes);
  int64_t nb = nblk, tp = tpb;

  // Pass 0's histogram, with the AND / OR summary of all keys; passes whose
  // digit is the same in every key are identity permutations and are skipped.
  const U init[2] = {~U(0), U(0)};
  KKMX_SAFE_CALL(mx_memcpy_async(summary, init, sizeof(init), exec.stream()));
  const void* keys_in = keys;
  int shift0          = 0;
  mx_launch_kernel(reinterpret_cast<const void*>(&k_radix_hist<T, true>), unsigned(nblk), kRadixLanes, 0,
                   exec.stream(), keys_in, n, shift0, hp, nb, tp, summary);
  U got[2] = {0, 0};
  KKMX_SAFE_CALL(mx_memcpy_async(got, summary, sizeof(got), exec.stream()));
  exec.fence("Kokkos::sort: key summary");
  const U varying = got[0] ^ got[1];
  int run[8], nrun = 0;
  for (int p = 0; p < passes; ++p)
    if ((varying >> (8 * p)) & U(255)) run[nrun++] = p;

  if (nrun == 0) {  // every key has the same digits: the keys are sorted
    if constexpr (Iota)
      parallel_for("Kokkos::sort::identity_permutation", RangePolicy<MI355X, IndexType<int64_t>>(exec, 0, n),
                   IotaU32{payload});
    exec.fence("Kokkos::sort: nothing to move");
    return;
  }
  for (int k = 0; k < nrun; ++k) {
    const int p      = run[k];
    const void* in   = buf[k & 1];
    void* out        = buf[(k + 1) & 1];
    int shift        = 8 * p;
    const bool first = k == 0, last = k == nrun - 1;
    if (!(first && p == 0)) {  // pass 0's histogram is already there
      const void* kh = first ? reinterpret_cast<const void*>(&k_radix_hist<T, true>)
                             : reinterpret_cast<const void*>(&k_radix_hist<T, false>);
      mx_launch_kernel(kh, unsigned(nblk), kRadixLanes, 0, exec.stream(), in, n, shift, hp, nb, tp, summary);
    }
    parallel_scan("Kokkos::sort::radix_offsets", RangePolicy<MI355X, IndexType<int64_t>>(exec, 0, 256 * nblk),
                  ExclusiveScanU32{hp, op});
    const void* ks      = radix_scatter_kernel<T, PT, Iota>(first, last);
    const uint32_t* cop = op;
    const PT* iin       = ibuf[k & 1];
    PT* iout            = ibuf[(k + 1) & 1];
    mx_launch_kernel(ks, unsigned(nblk), kRadixLanes, 0, exec.stream(), in, out, iin, iout, n, shift, cop, nb, tp);
  }
  if (nrun & 1) {  // an odd number of passes left the result in the workspace
    parallel_for("Kokkos::sort::copy_back", RangePolicy<MI355X, IndexType<int64_t>>(exec, 0, n),
                 CopyElems<U>{static_cast<const U*>(buf[1]), reinterpret_cast<U*>(keys)});
    if constexpr (pairs)
      parallel_for("Kokkos::sort::copy_back_payload", RangePolicy<MI355X, IndexType<int64_t>>(exec, 0, n),
                   CopyElems<PT>{ibuf[1], payload});
  }
  // the workspace is free for the next call once the stream has run these
  exec.fence("Kokkos::sort: radix passes");
}

template <class V>
constexpr bool radix_sortable() {
  using T = typename V::non_const_value_type;
  return V::rank == 1 && std::is_arithmetic<T>::value && !std::is_same<T, bool>::value &&
         (sizeof(T) == 4 || sizeof(T) == 8) && exec_can_access<MI355X, typename V::memory_space>::value;
}

template <class V, class Comp>
void host_sort(const V& view, size_t begin, size_t end, Comp comp) {
  using T = typename V::non_const_value_type;
  const size_t n = end - begin;
  std::vector<T> h(n);
  if constexpr (exec_can_access<Serial, typename V::memory_space>::value) {
    for (size_t i = 0; i < n; ++i) h[i] = view(begin + i);
    std::sort(h.begin(), h.end(), comp);
    for (size_t i = 0; i < n; ++i) view(begin + i) = h[i];
  } else {
    static_assert(V::rank == 1, "Kokkos::sort sorts rank-1 Views");
    Kokkos::fence("Kokkos::sort: host fallback");
    if (view.span_is_contiguous()) {
      KKMX_SAFE_CALL(mx_memcpy(h.data(), view.data() + begin, n * sizeof(T)));
      std::sort(h.begin(), h.end(), comp);
      KKMX_SAFE_CALL(mx_memcpy(view.data() + begin, h.data(), n * sizeof(T)));
    } else {
      auto m = create_mirror_view_and_copy(HostSpace(), view);
      for (size_t i = 0; i < n; ++i) h[i] = m(begin + i);
      std::sort(h.begin(), h.end(), comp);
      for (size_t i = 0; i < n; ++i) m(begin + i) = h[i];
      deep_copy(view, m);
    }
  }
}

}  // namespace Impl

// ------------------------------------------------------------------ BinSort
template <class KeyViewType, class BinSortOp, class Space = typename KeyViewType::device_type,
          class SizeType = typename KeyViewType::size_type>
class BinSort {
 public:
  using exec_space       = MI355X;
  using offset_type      = View<int*, MI355XSpace>;
  using bin_count_type   = View<const int*, MI355XSpace>;
  using const_key_view_type = typename KeyViewType::const_type;

  BinSort() = delete;
  template <class ExecutionSpace>
  BinSort(const ExecutionSpace& exec, const_key_view_type keys_, int range_begin_, int range_end_, BinSortOp bin_op_,
          bool sort_within_bins_ = false)
      : keys(keys_),
        range_begin(range_begin_),
        range_end(range_end_),
        bin_op(bin_op_),
        sort_within_bins(sort_within_bins_) {
    static_assert(std::is_same<ExecutionSpace, MI355X>::value, "BinSort runs on the MI355X execution space");
    if (range_begin > range_end) Kokkos::abort("BinSort: range_begin > range_end");
    sort_order  = offset_type("Kokkos::BinSort::sort_order", size_t(range_end - range_begin));
    bin_count   = View<int*, MI355XSpace>("Kokkos::BinSort::bin_count", size_t(bin_op.max_bins()));
    bin_fill    = View<int*, MI355XSpace>("Kokkos::BinSort::bin_fill", size_t(bin_op.max_bins()));
    bin_offsets = offset_type("Kokkos::BinSort::bin_offsets", size_t(bin_op.max_bins()));
    (void)exec;
  }
  BinSort(const_key_view_type keys_, int range_begin_, int range_end_, BinSortOp bin_op_, bool sort_within_bins_ = false)
      : BinSort(exec_space{}, keys_, range_begin_, range_end_, bin_op_, sort_within_bins_) {}
  template <class ExecutionSpace>
  BinSort(const ExecutionSpace& exec, const_key_view_type keys_, BinSortOp bin_op_, bool sort_within_bins_ = false)
      : BinSort(exec, keys_, 0, int(keys_.extent(0)), bin_op_, sort_within_bins_) {}
  BinSort(const_key_view_type keys_, BinSortOp bin_op_, bool sort_within_bins_ = false)
      : BinSort(exec_space{}, keys_, bin_op_, sort_within_bins_) {}

  // Counts, offsets and the permutation; call again when the keys change.
  template <class ExecutionSpace>
  void create_permute_vector(const ExecutionSpace& exec) {
    using U        = View<int*, MI355XSpace, MemoryTraits<Unmanaged>>;
    using CU       = View<const int*, MI355XSpace, MemoryTraits<Unmanaged>>;
    const int len  = range_end - range_begin;
    const int bins = bin_op.max_bins();
    deep_copy(exec, bin_count, 0);
    deep_copy(exec, bin_fill, 0);
    U cnt(bin_count.data(), size_t(bins)), fill(bin_fill.data(), size_t(bins));
    U off(bin_offsets.data(), size_t(bins)), order(sort_order.data(), size_t(len));
    parallel_for("Kokkos::Sort::BinCount", RangePolicy<MI355X>(exec, 0, len),
                 Impl::BinCountF<const_key_view_type, BinSortOp>{keys, bin_op, cnt, range_begin});
    parallel_scan("Kokkos::Sort::BinOffset", RangePolicy<MI355X>(exec, 0, bins), Impl::BinOffsetF{CU(cnt), off});
    if constexpr (Impl::is_bin_op_1d<BinSortOp>::value && KeyViewType::rank == 1 &&
                  Impl::radix_sortable<KeyViewType>()) {
      // BinOp1D's bin never decreases as the key grows, so "by bin, then by
      // key, equal keys in index order" is the stable order of the keys: the
      // permutation of a radix sort of (key, position) pairs
      if (sort_within_bins && len > 0) {
        using T = typename KeyViewType::non_const_value_type;
        View<T*, MI355XSpace> copy(view_alloc(WithoutInitializing, std::string("Kokkos::BinSort::key_copy")),
                                   size_t(len));
        parallel_for("Kokkos::Sort::BinKeyCopy", RangePolicy<MI355X>(exec, 0, len),
                     Impl::BinKeyCopyF<View<T*, MI355XSpace>, const_key_view_type>{copy, keys, range_begin});
        Impl::radix_sort<T, uint32_t, true>(exec, copy.data(), int64_t(len),
                                            reinterpret_cast<uint32_t*>(sort_order.data()));
        if (range_begin != 0)
          parallel_for("Kokkos::Sort::BinShift", RangePolicy<MI355X>(exec, 0, len), Impl::BinShiftF{order, range_begin});
        return;
      }
    }
    parallel_for("Kokkos::Sort::BinBinning", RangePolicy<MI355X>(exec, 0, len),
                 Impl::BinPlaceF<const_key_view_type, BinSortOp>{keys, bin_op, fill, CU(off), order, range_begin});
    if (sort_within_bins)
      parallel_for("Kokkos::Sort::BinSort", RangePolicy<MI355X>(exec, 0, bins),
                   Impl::BinSortWithinF<const_key_view_type, BinSortOp>{keys, bin_op, CU(cnt), CU(off), order});
  }
  void create_permute_vector() {
    Kokkos::fence("Kokkos::BinSort::create_permute_vector: before");
    exec_space e{};
    create_permute_vector(e);
    e.fence("Kokkos::BinSort::create_permute_vector: after");
  }

  // values(range_begin + i) = values(sort_order(i) + values_range_begin - range_begin)
  template <class ExecutionSpace, class ValuesViewType>
  void sort(const ExecutionSpace& exec, const ValuesViewType& values, int values_range_begin,
            int values_range_end) const {
    if (values.extent(0) == 0) return;
    const int len = range_end - range_begin;
    if (values_range_end - values_range_begin != len)
      Kokkos::abort("BinSort::sort: values range length != permutation vector length");
    using Tmp = View<typename ValuesViewType::non_const_data_type, typename ValuesViewType::array_layout, MI355XSpace>;
    Tmp tmp(std::string("Kokkos::SortImpl::BinSortFunctor::sorted_values"), size_t(len),
            ValuesViewType::rank > 1 ? values.extent(1) : ~size_t(0));  // ~0: no extent
    View<const int*, MI355XSpace, MemoryTraits<Unmanaged>> order(sort_order.data(), size_t(len));
    parallel_for("Kokkos::Sort::CopyPermute", RangePolicy<MI355X>(exec, 0, len),
                 Impl::PermuteRowsF<Tmp, ValuesViewType>{tmp, values, order, values_range_begin - range_begin});
    parallel_for("Kokkos::Sort::Copy", RangePolicy<MI355X>(exec, 0, len),
                 Impl::CopyRowsF<ValuesViewType, Tmp>{values, tmp, range_begin});
  }
  template <class ValuesViewType>
  void sort(const ValuesViewType& values, int values_range_begin, int values_range_end) const {
    Kokkos::fence("Kokkos::BinSort::sort: before");
    exec_space exec;
    sort(exec, values, values_range_begin, values_range_end);
    exec.fence("Kokkos::BinSort::sort: after");
  }
  template <class ExecutionSpace, class ValuesViewType, class = std::enable_if_t<is_execution_space<ExecutionSpace>::value>>
  void sort(const ExecutionSpace& exec, const ValuesViewType& values) const {
    sort(exec, values, 0, range_end - range_begin);
  }
  template <class ValuesViewType>
  void sort(const ValuesViewType& values) const {
    sort(values, 0, range_end - range_begin);
  }

  offset_type get_permute_vector() const { return sort_order; }
  offset_type get_bin_offsets() const { return bin_offsets; }
  bin_count_type get_bin_count() const { return bin_count; }

 private:
  const_key_view_type keys;
  int range_begin = 0, range_end = 0;
  BinSortOp bin_op;
  bool sort_within_bins = false;
  offset_type sort_order;
  View<int*, MI355XSpace> bin_count, bin_fill;
  offset_type bin_offsets;
};

// ------------------------------------------------------------------ sort
template <class ExecutionSpace, class DataType, class... Properties,
          class = std::enable_if_t<is_execution_space<ExecutionSpace>::value>>
void sort(const ExecutionSpace& exec, const View<DataType, Properties...>& view) {
  using V = View<DataType, Properties...>;
  using T = typename V::non_const_value_type;
  static_assert(V::rank == 1, "Kokkos::sort sorts rank-1 Views");
  if (view.extent(0) < 2) return;
  if constexpr (Impl::radix_sortable<V>() && std::is_same<ExecutionSpace, MI355X>::value) {
    if (view.span_is_contiguous()) {
      Impl::radix_sort<T>(exec, view.data(), int64_t(view.extent(0)));
    } else {
      // strided keys: gather into a contiguous copy, sort it, scatter back
      View<T*, MI355XSpace> packed(view_alloc(WithoutInitializing, std::string("Kokkos::sort::packed")),
                                   view.extent(0));
      deep_copy(exec, packed, view);
      Impl::radix_sort<T>(exec, packed.data(), int64_t(packed.extent(0)));
      deep_copy(exec, view, packed);
      exec.fence("Kokkos::sort: strided keys written back");
    }
    return;
  }
  exec.fence();
  Impl::host_sort(view, 0, view.extent(0), std::less<T>());
}
template <class DataType, class... Properties>
void sort(const View<DataType, Properties...>& view) {
  Kokkos::fence("Kokkos::sort: before");
  sort(MI355X(), view);
  Kokkos::fence("Kokkos::sort: after");
}
template <class ExecutionSpace, class Comparator, class DataType, class... Properties,
          class = std::enable_if_t<is_execution_space<ExecutionSpace>::value>>
void sort(const ExecutionSpace& exec, const View<DataType, Properties...>& view, const Comparator& comp) {
  exec.fence();
  Impl::host_sort(view, 0, view.extent(0), comp);
}
template <class Comparator, class DataType, class... Properties,
          class = std::enable_if_t<!is_execution_space<Comparator>::value>>
void sort(const View<DataType, Properties...>& view, const Comparator& comp) {
  Kokkos::fence("Kokkos::sort: before");
  Impl::host_sort(view, 0, view.extent(0), comp);
}
template <class ExecutionSpace, class ViewType, class = std::enable_if_t<is_execution_space<ExecutionSpace>::value>>
void sort(const ExecutionSpace& exec, ViewType view, size_t begin, size_t end) {
  using T = typename ViewType::non_const_value_type;
  if (end <= begin + 1) return;
  if constexpr (Impl::radix_sortable<ViewType>() && std::is_same<ExecutionSpace, MI355X>::value) {
    if (view.span_is_contiguous()) {
      Impl::radix_sort<T>(exec, view.data() + begin, int64_t(end - begin));
      return;
    }
    // strided keys: the same gather / sort / scatter over [begin, end)
    View<T*, MI355XSpace> packed(view_alloc(WithoutInitializing, std::string("Kokkos::sort::packed")), end - begin);
    auto window = subview(view, std::make_pair(begin, end));
    deep_copy(exec, packed, window);
    Impl::radix_sort<T>(exec, packed.data(), int64_t(packed.extent(0)));
    deep_copy(exec, window, packed);
    exec.fence("Kokkos::sort: strided keys written back");
    return;
  }
  exec.fence();
  Impl::host_sort(view, begin, end, std::less<T>());
}
template <class ViewType>
void sort(ViewType view, size_t begin, size_t end) {
  Kokkos::fence("Kokkos::sort: before");
  sort(MI355X(), view, begin, end);
  Kokkos::fence("Kokkos::sort: after");
}

// ------------------------------------------------------------------ sort_by_key
// algorithms/src/sorting/Kokkos_SortByKeyPublicAPI.hpp:27-108 and
// impl/Kokkos_SortByKeyImpl.hpp:98-109 (admissible Views), 262-324
// (permutation fallback).  Keys that the radix sort takes are sorted on the
// device with the values riding along as radix payload (contiguous values of
// 4 or 8 bytes) or as (key, position) pairs whose permutation then gathers
// the values; either way equal keys keep their input order, as with the
// rocThrust path (impl/Kokkos_SortByKeyImpl.hpp:145-158).  A comparator, or keys
// the radix sort does not take, sort a permutation on the host with
// std::stable_sort.
namespace Impl {

template <class V>
constexpr bool sort_by_key_admissible() {
  using L = typename V::array_layout;
  return is_view<V>::value && V::rank == 1 &&
         (std::is_same<L, LayoutLeft>::value || std::is_same<L, LayoutRight>::value ||
          std::is_same<L, LayoutStride>::value);
}

template <class Dst, class Src>
struct SortByKeyCopyF {
  Dst dst;
  Src src;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { dst(i) = src(i); }
};
template <class Dst, class Src, class P>
struct SortByKeyGatherF {
  Dst dst;
  Src src;
  P perm;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t i) const { dst(i) = src(perm(i)); }
};

// view(i) = view(perm(i)) for every i, through a device copy of view.
template <class VV, class P>
void apply_permutation(const MI355X& exec, const P& perm, const VV& view) {
  using T         = typename VV::non_const_value_type;
  const int64_t n = int64_t(view.extent(0));
  View<T*, MI355XSpace> copy(view_alloc(WithoutInitializing, std::string("Kokkos::sort_by_key::permute_copy")),
                             size_t(n));
  parallel_for("Kokkos::sort_by_key::copy", RangePolicy<MI355X, IndexType<int64_t>>(exec, 0, n),
               SortByKeyCopyF<decltype(copy), VV>{copy, view});
  parallel_for("Kokkos::sort_by_key::gather", RangePolicy<MI355X, IndexType<int64_t>>(exec, 0, n),
               SortByKeyGatherF<VV, decltype(copy), P>{view, copy, perm});
  exec.fence("Kokkos::sort_by_key: permuted");
}

template <class KV, class VV, class Comp>
void sort_by_key_host(const KV& keys, const VV& values, Comp comp) {
  const size_t n = keys.extent(0);
  auto hk        = create_mirror_view_and_copy(HostSpace(), keys);
  auto hv        = create_mirror_view_and_copy(HostSpace(), values);
  std::vector<uint32_t> idx(n);
  for (size_t i = 0; i < n; ++i) idx[i] = uint32_t(i);
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return comp(hk(a), hk(b)); });
  std::vector<typename KV::non_const_value_type> k(n);
  std::vector<typename VV::non_const_value_type> v(n);
  for (size_t i = 0; i < n; ++i) {
    k[i] = hk(idx[i]);
    v[i] = hv(idx[i]);
  }
  for (size_t i = 0; i < n; ++i) {
    hk(i) = k[i];
    hv(i) = v[i];
  }
  deep_copy(keys, hk);
  deep_copy(values, hv);
}

template <class ExecutionSpace, class KV, class VV>
void check_sort_by_key(const KV& keys, const VV& values) {
  static_assert(sort_by_key_admissible<KV>() && sort_by_key_admissible<VV>(),
                "Kokkos::sort_by_key only accepts 1D values View with LayoutRight, LayoutLeft or LayoutStride.");
  static_assert(SpaceAccessibility<ExecutionSpace, typename KV::memory_space>::accessible,
                "Kokkos::sort: execution space instance is not able to access the memory space of the keys View "
                "argument!");
  static_assert(SpaceAccessibility<ExecutionSpace, typename VV::memory_space>::accessible,
                "Kokkos::sort: execution space instance is not able to access the memory space of the values View "
                "argument!");
  if (values.size() != keys.size())
    Kokkos::abort((std::string("values and keys extents must be the same. The values extent is ") +
                   std::to_string(values.size()) + ", and the keys extent is " + std::to_string(keys.size()) + ".")
                      .c_str());
}

}  // namespace Impl

namespace Experimental {

template <class ExecutionSpace, class KeysDataType, class... KeysProperties, class ValuesDataType,
          class... ValuesProperties>
void sort_by_key(const ExecutionSpace& exec, const View<KeysDataType, KeysProperties...>& keys,
                 const View<ValuesDataType, ValuesProperties...>& values) {
  using KV = View<KeysDataType, KeysProperties...>;
  using VV = View<ValuesDataType, ValuesProperties...>;
  using T  = typename KV::non_const_value_type;
  Kokkos::Impl::check_sort_by_key<ExecutionSpace>(keys, values);
  if (keys.extent(0) <= 1) return;
  if constexpr (Kokkos::Impl::radix_sortable<KV>() && std::is_same<ExecutionSpace, MI355X>::value) {
    if (keys.span_is_contiguous()) {
      using W = typename VV::non_const_value_type;
      if constexpr ((sizeof(W) == 4 || sizeof(W) == 8) && std::is_trivially_copyable<W>::value) {
        if (values.span_is_contiguous()) {
          // the values ride through the radix passes with their keys
          using PT = std::conditional_t<sizeof(W) == 4, uint32_t, uint64_t>;
          Kokkos::Impl::radix_sort<T, PT>(exec, keys.data(), int64_t(keys.extent(0)),
                                          reinterpret_cast<PT*>(values.data()));
          return;
        }
      }
      View<uint32_t*, MI355XSpace> perm(view_alloc(WithoutInitializing, std::string("Kokkos::sort_by_key::permute")),
                                        keys.extent(0));
      Kokkos::Impl::radix_sort<T, uint32_t, true>(exec, keys.data(), int64_t(keys.extent(0)), perm.data());
      Kokkos::Impl::apply_permutation(exec, perm, values);
      return;
    }
  }
  exec.fence("Kokkos::sort_by_key: before host sort");
  Kokkos::Impl::sort_by_key_host(keys, values, std::less<T>());
}

template <class ExecutionSpace, class ComparatorType, class KeysDataType, class... KeysProperties,
          class ValuesDataType, class... ValuesProperties>
void sort_by_key(const ExecutionSpace& exec, const View<KeysDataType, KeysProperties...>& keys,
                 const View<ValuesDataType, ValuesProperties...>& values, const ComparatorType& comparator) {
  Kokkos::Impl::check_sort_by_key<ExecutionSpace>(keys, values);
  if (keys.extent(0) <= 1) return;
  exec.fence("Kokkos::sort_by_key: before host sort");
  Kokkos::Impl::sort_by_key_host(keys, values, comparator);
}

}  // namespace Experimental

}  // namespace Kokkos

#endif
