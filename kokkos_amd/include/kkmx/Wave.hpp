// Wave64 primitives for gfx950: ordered inclusive/exclusive scans and
// reductions across the 64 lanes of a wavefront built from DPP row shifts and
// row broadcasts, plus the LDS combine across the wavefronts of a workgroup.
//
// Every combine keeps index order (the value from the lower lane is the left
// operand of join), so the primitives are correct for any associative join,
// commutative or not, and give a fixed, run-to-run reproducible combination
// tree for floating point.
//
// DPP controls used (GFX9 encoding, valid on CDNA3/CDNA4):
//   row_shr:n   0x110+n  lane i reads lane i-n of its 16-lane row
//   row_bcast15 0x142    row r reads lane 15 of row r-1 (rows masked by 0xa)
//   row_bcast31 0x143    rows 2,3 read lane 31 (rows masked by 0xc)
//   wave_shr:1  0x138    lane i reads lane i-1 of the wavefront
// Lanes that are masked off or whose source is out of range keep the `old`
// operand, which is always the identity here.
#ifndef KKMX_WAVE_HPP
#define KKMX_WAVE_HPP

#include <kkmx/Core_fwd.hpp>

namespace Kokkos {
namespace Impl {
namespace wave {

constexpr int kWave = 64;

__device__ __forceinline__ int lane() { return static_cast<int>(__lane_id()); }

template <class T>
struct words {
  static constexpr int n = (sizeof(T) + 3) / 4;
};

template <int CTRL, int ROW_MASK, int BANK_MASK, class T>
__device__ __forceinline__ T dpp(const T& old, const T& src) {
  constexpr int W = words<T>::n;
  uint32_t o[W], s[W], r[W];
  o[W - 1] = 0;
  s[W - 1] = 0;
  __builtin_memcpy(o, &old, sizeof(T));
  __builtin_memcpy(s, &src, sizeof(T));
#pragma unroll
  for (int w = 0; w < W; ++w)
    r[w] = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
        static_cast<int>(o[w]), static_cast<int>(s[w]), CTRL, ROW_MASK, BANK_MASK, false));
  T out;
  __builtin_memcpy(&out, r, sizeof(T));
  return out;
}

template <class T>
__device__ __forceinline__ T readlane(const T& v, int l) {
  constexpr int W = words<T>::n;
  uint32_t s[W], r[W];
  s[W - 1] = 0;
  __builtin_memcpy(s, &v, sizeof(T));
#pragma unroll
  for (int w = 0; w < W; ++w)
    r[w] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s[w]), l));
  T out;
  __builtin_memcpy(&out, r, sizeof(T));
  return out;
}

// Value of lane `src_lane` (per lane, any lane) through ds_bpermute.
template <class T>
__device__ __forceinline__ T shuffle(const T& v, int src_lane) {
  constexpr int W = words<T>::n;
  uint32_t s[W], r[W];
  s[W - 1] = 0;
  __builtin_memcpy(s, &v, sizeof(T));
#pragma unroll
  for (int w = 0; w < W; ++w)
    r[w] = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_lane << 2, static_cast<int>(s[w])));
  T out;
  __builtin_memcpy(&out, r, sizeof(T));
  return out;
}

// new = join(from_lower_lane, mine), identity where there is no source.
template <int CTRL, int RM, class T, class Join>
__device__ __forceinline__ void step(T& v, const T& identity, const Join& join) {
  T t = dpp<CTRL, RM, 0xf>(identity, v);
  join(t, v);
  v = t;
}

// Inclusive scan over the 64 lanes; all lanes must be active.
template <class T, class Join>
__device__ __forceinline__ T inclusive_scan(T v, const T& identity, const Join& join) {
  step<0x111, 0xf>(v, identity, join);
  step<0x112, 0xf>(v, identity, join);
  step<0x114, 0xf>(v, identity, join);
  step<0x118, 0xf>(v, identity, join);
  step<0x142, 0xa>(v, identity, join);
  step<0x143, 0xc>(v, identity, join);
  return v;
}

// Shift by one lane: lane i gets lane i-1's value, lane 0 gets identity.
template <class T>
__device__ __forceinline__ T shift_up(const T& v, const T& identity) {
  return dpp<0x138, 0xf, 0xf>(identity, v);
}

// Ordered reduction of all 64 lanes; result uniform (read from lane 63).
template <class T, class Join>
__device__ __forceinline__ T reduce(T v, const T& identity, const Join& join) {
  return readlane(inclusive_scan(v, identity, join), 63);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

}  // namespace wave

// Ordered reduction over a workgroup of nthreads (multiple of 64, <= 1024).
// `lds` must hold 16 values.  The result is returned to every thread.
template <class T, class Join>
__device__ __forceinline__ T block_reduce(T v, const T& identity, const Join& join, T* lds,
                                          int nthreads) {
  const int w   = threadIdx.x >> 6;
  const int nw  = nthreads >> 6;
  T r           = wave::reduce(v, identity, join);
  if (nw == 1) return r;
  if (wave::lane() == 0) lds[w] = r;
  __syncthreads();
  T x = (wave::lane() < nw) ? lds[wave::lane()] : identity;
  r   = wave::reduce(x, identity, join);
  __syncthreads();  // lds may be reused by the caller
  return r;
}

}  // namespace Impl
}  // namespace Kokkos

#endif
