// Kokkos atomics for the MI355X execution space.
//
// Semantics follow the reference wrapper core/src/Kokkos_Atomics_Desul_Wrapper.hpp:
// relaxed memory order, device ("agent") scope on the GPU (:46-50), the value
// argument never participates in type deduction (:86-146), fetch_op returns
// the old value, op_fetch the new one, atomic_op returns nothing.
//
// Mapping onto gfx950:
//  * 4- and 8-byte integers: native add/sub/min/max/and/or/xor (global_atomic_*
//    at the L2, ds_* when the pointer is in LDS; the compiler picks the
//    no-return form when the result is unused).
//  * float/double add: native global_atomic_add_f32/f64 (and min/max through the
//    compiler's atomicrmw, which is native for f64 on CDNA).
//  * every other operator and every other 4/8-byte type: a compare-and-swap
//    loop on the bit pattern.
//  * 1- and 2-byte types: CAS on the aligned 4-byte word that contains them.
//  * larger types: a hashed lock array in device memory; the lanes of a
//    wavefront take turns so no two of them ever spin for a lock together
//    (the reference uses a 2^17-entry lock hash, desul Lock_Array_HIP.hpp:76-116).
// Host: GCC/Clang __atomic builtins with the same rules.
#ifndef KKMX_ATOMIC_HPP
#define KKMX_ATOMIC_HPP

#include <kkmx/Core_fwd.hpp>

#include <cstring>

namespace Kokkos {
namespace Impl {

template <class T>
struct nodeduce {
  using type = T;
};
template <class T>
using nodeduce_t = typename nodeduce<T>::type;

template <class T>
KOKKOS_INLINE_FUNCTION constexpr bool atomic_native_int() {
  return std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8);
}
template <class T>
KOKKOS_INLINE_FUNCTION constexpr bool atomic_native_fp() {
  return std::is_same<T, float>::value || std::is_same<T, double>::value;
}

template <int N>
struct uint_of_size;
template <>
struct uint_of_size<4> {
  using type = uint32_t;
};
template <>
struct uint_of_size<8> {
  using type = uint64_t;
};

template <class To, class From>
KOKKOS_FORCEINLINE_FUNCTION To bit_cast_(const From& f) {
  static_assert(sizeof(To) == sizeof(From), "size mismatch");
  To t;
  __builtin_memcpy(&t, &f, sizeof(To));
  return t;
}

// Lock arrays for types wider than 8 bytes: one for the device, one for the
// host, each shared by every operation and type, so any two atomic
// operations on one address (a fetch_add and an exchange, say) take the same
// lock.
constexpr int kLockCount = 1 << 16;
#if defined(__HIP__)
__device__ inline int g_atomic_locks[kLockCount];
#endif
inline int g_host_atomic_locks[kLockCount];

KOKKOS_INLINE_FUNCTION unsigned lock_index(const void* p) {
  auto a = reinterpret_cast<uintptr_t>(p);
  return static_cast<unsigned>((a >> 4) ^ (a >> 20)) & (kLockCount - 1);
}

// Memory scopes of the scoped operations below: the HIP scope constants,
// plus kScopeCaller for operations that need no atomicity (desul's
// MemoryScopeCaller).  Kokkos:: atomics use kScopeAgent (the whole GPU).
constexpr int kScopeCaller    = -1;
constexpr int kScopeWorkgroup = __HIP_MEMORY_SCOPE_WORKGROUP;
constexpr int kScopeAgent     = __HIP_MEMORY_SCOPE_AGENT;
constexpr int kScopeSystem    = __HIP_MEMORY_SCOPE_SYSTEM;

// ---------------------------------------------------------- raw CAS (returns old)
template <int S, class T>
KOKKOS_FORCEINLINE_FUNCTION T cas_word(T* p, T expected, T desired) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "word CAS");
  using U = typename uint_of_size<sizeof(T)>::type;
  U e     = bit_cast_<U>(expected);
  U d     = bit_cast_<U>(desired);
  if constexpr (S == kScopeCaller) {
    U* q = reinterpret_cast<U*>(p);
    if (*q == e) *q = d;
    else e = *q;
    return bit_cast_<T>(e);
  }
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_compare_exchange_strong(reinterpret_cast<U*>(p), &e, d, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       S == kScopeCaller ? kScopeAgent : S);
#else
  __atomic_compare_exchange_n(reinterpret_cast<U*>(p), &e, d, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
#endif
  return bit_cast_<T>(e);
}

template <int S, class T>
KOKKOS_FORCEINLINE_FUNCTION T load_word(T* p) {
  using U = typename uint_of_size<sizeof(T)>::type;
  if constexpr (S == kScopeCaller) return *p;
#if defined(__HIP_DEVICE_COMPILE__)
  U v = __hip_atomic_load(reinterpret_cast<U*>(p), __ATOMIC_RELAXED, S == kScopeCaller ? kScopeAgent : S);
#else
  U v = __atomic_load_n(reinterpret_cast<U*>(p), __ATOMIC_RELAXED);
#endif
  return bit_cast_<T>(v);
}

// Generic fetch-op: returns the old value; new value = op(old, val).
template <int S, class T, class Op>
KOKKOS_INLINE_FUNCTION T fetch_oper(T* p, const T& val, Op op) {
  if constexpr (S == kScopeCaller) {
    T old = *p;
    *p    = op(old, val);
    return old;
  } else if constexpr (sizeof(T) == 4 || sizeof(T) == 8) {
    T old = load_word<S>(p);
    while (true) {
      T desired = op(old, val);
      T seen    = cas_word<S>(p, old, desired);
      if (bit_cast_<typename uint_of_size<sizeof(T)>::type>(seen) ==
          bit_cast_<typename uint_of_size<sizeof(T)>::type>(old))
        return old;
      old = seen;
    }
  } else if constexpr (sizeof(T) == 1 || sizeof(T) == 2) {
    // CAS on the containing aligned 32-bit word.
    uintptr_t a     = reinterpret_cast<uintptr_t>(p);
    uint32_t* w     = reinterpret_cast<uint32_t*>(a & ~uintptr_t(3));
    unsigned shift  = static_cast<unsigned>(a & 3) * 8;
    uint32_t mask   = (sizeof(T) == 1 ? 0xffu : 0xffffu) << shift;
    uint32_t oldw   = load_word<S>(w);
    while (true) {
      T oldv;
      uint32_t bits = (oldw & mask) >> shift;
      __builtin_memcpy(&oldv, &bits, sizeof(T));
      T newv       = op(oldv, val);
      uint32_t nb  = 0;
      __builtin_memcpy(&nb, &newv, sizeof(T));
      uint32_t neww = (oldw & ~mask) | ((nb << shift) & mask);
      uint32_t seen = cas_word<S>(w, oldw, neww);
      if (seen == oldw) return oldv;
      oldw = seen;
    }
  } else {
#if defined(__HIP_DEVICE_COMPILE__)
    // The lanes of a wavefront take turns, lowest lane first: only the lane
    // whose turn it is spins for the lock, updates and releases it.  A spin
    // shared by several lanes would let the compiler move the update and
    // the release after the spin loop, where a lane holding the lock cannot
    // release it while another lane of its wavefront still spins for it.
    int* lock           = &g_atomic_locks[lock_index(p)];
    T result{};
    const unsigned lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    uint64_t pending    = __builtin_amdgcn_ballot_w64(true);
    while (pending) {
      if (lane == unsigned(__builtin_ctzll(pending))) {
        // The lock array is shared by every workgroup, so the lock itself is
        // taken at agent scope whatever scope the caller asked for.
        while (__hip_atomic_exchange(lock, 1, __ATOMIC_ACQUIRE, kScopeAgent) != 0) {
        }
        result = *p;
        *p     = op(result, val);
        __hip_atomic_store(lock, 0, __ATOMIC_RELEASE, kScopeAgent);
      }
      pending &= pending - 1;
    }
    return result;
#else
    int* lock = &g_host_atomic_locks[lock_index(p)];
    while (__atomic_exchange_n(lock, 1, __ATOMIC_ACQUIRE) != 0) {
    }
    T result = *p;
    *p       = op(result, val);
    __atomic_store_n(lock, 0, __ATOMIC_RELEASE);
    return result;
#endif
  }
}

struct OpAdd { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a + b; } };
struct OpSub { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a - b; } };
struct OpMul { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a * b; } };
struct OpDiv { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a / b; } };
struct OpMod { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a % b; } };
struct OpMin { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return b < a ? b : a; } };
struct OpMax { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a < b ? b : a; } };
struct OpAnd { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a & b; } };
struct OpOr  { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a | b; } };
struct OpXor { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a ^ b; } };
struct OpNand { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return ~(a & b); } };
struct OpLshift { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return T(a << b); } };
struct OpRshift { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return T(a >> b); } };
struct OpIncMod { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return a >= b ? T(0) : T(a + 1); } };
struct OpDecMod { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T& b) const { return (a == 0 || a > b) ? b : T(a - 1); } };
struct OpSecond { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T&, const T& b) const { return b; } };
struct OpFirst { template <class T> KOKKOS_FORCEINLINE_FUNCTION T operator()(const T& a, const T&) const { return a; } };

// ---------------------------------------------------------- scoped fetch_op
// Native read-modify-writes where gfx950 has them (4/8-byte integers: add,
// sub, min, max, and, or, xor; float/double add), CAS or locks otherwise.
template <int S, class T>
KOKKOS_INLINE_FUNCTION T s_fetch_add(T* p, const T v) {
  if constexpr (S != kScopeCaller && (atomic_native_int<T>() || atomic_native_fp<T>())) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, S);
#else
    if constexpr (atomic_native_int<T>()) return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
    else return fetch_oper<S>(p, v, OpAdd{});
#endif
  } else {
    return fetch_oper<S>(p, v, OpAdd{});
  }
}
template <int S, class T>
KOKKOS_INLINE_FUNCTION T s_fetch_sub(T* p, const T v) {
  if constexpr (S != kScopeCaller && atomic_native_int<T>()) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __hip_atomic_fetch_add(p, T(0) - v, __ATOMIC_RELAXED, S);
#else
    return __atomic_fetch_sub(p, v, __ATOMIC_RELAXED);
#endif
  } else if constexpr (S != kScopeCaller && atomic_native_fp<T>()) {
    return s_fetch_add<S>(p, T(-v));
  } else {
    return fetch_oper<S>(p, v, OpSub{});
  }
}
template <int S, class T>
KOKKOS_INLINE_FUNCTION T s_fetch_min(T* p, const T v) {
  if constexpr (S != kScopeCaller && atomic_native_int<T>()) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, S);
#else
    return fetch_oper<S>(p, v, OpMin{});
#endif
  } else {
    return fetch_oper<S>(p, v, OpMin{});
  }
}
template <int S, class T>
KOKKOS_INLINE_FUNCTION T s_fetch_max(T* p, const T v) {
  if constexpr (S != kScopeCaller && atomic_native_int<T>()) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, S);
#else
    return fetch_oper<S>(p, v, OpMax{});
#endif
  } else {
    return fetch_oper<S>(p, v, OpMax{});
  }
}
#define KKMX_ATOMIC_BITOP(NAME, HIPB, GCCB, OPS)                     \
  template <int S, class T>                                          \
  KOKKOS_INLINE_FUNCTION T s_fetch_##NAME(T* p, const T v) {         \
    if constexpr (S != kScopeCaller && atomic_native_int<T>()) {     \
      KOKKOS_IF_ON_DEVICE((return HIPB(p, v, __ATOMIC_RELAXED, S);)) \
      KOKKOS_IF_ON_HOST((return GCCB(p, v, __ATOMIC_RELAXED);))      \
    } else {                                                         \
      return fetch_oper<S>(p, v, OPS{});                             \
    }                                                                \
  }
KKMX_ATOMIC_BITOP(and, __hip_atomic_fetch_and, __atomic_fetch_and, OpAnd)
KKMX_ATOMIC_BITOP(or, __hip_atomic_fetch_or, __atomic_fetch_or, OpOr)
KKMX_ATOMIC_BITOP(xor, __hip_atomic_fetch_xor, __atomic_fetch_xor, OpXor)
#undef KKMX_ATOMIC_BITOP

template <int S, class T>
KOKKOS_INLINE_FUNCTION T s_load(T* p) {
  if constexpr (sizeof(T) == 4 || sizeof(T) == 8) return load_word<S>(p);
  else return fetch_oper<S>(p, T{}, OpFirst{});
}
template <int S, class T>
KOKKOS_INLINE_FUNCTION void s_store(T* p, const T v) {
  if constexpr (S == kScopeCaller) {
    *p = v;
  } else if constexpr (sizeof(T) == 4 || sizeof(T) == 8) {
    using U = typename uint_of_size<sizeof(T)>::type;
    KOKKOS_IF_ON_DEVICE((__hip_atomic_store(reinterpret_cast<U*>(p), bit_cast_<U>(v), __ATOMIC_RELAXED, S);))
    KOKKOS_IF_ON_HOST((__atomic_store_n(reinterpret_cast<U*>(p), bit_cast_<U>(v), __ATOMIC_RELAXED);))
  } else {
    (void)fetch_oper<S>(p, v, OpSecond{});
  }
}
// Compare-exchange: returns the old value; stores `desired` when the old
// value equals `expected` bit for bit.
template <int S, class T>
KOKKOS_INLINE_FUNCTION T s_compare_exchange(T* p, const T expected, const T desired) {
  if constexpr (S != kScopeCaller && (sizeof(T) == 4 || sizeof(T) == 8)) {
    return cas_word<S>(p, expected, desired);
  } else {
    T old{};
    (void)fetch_oper<S>(p, desired, [&old, expected] KOKKOS_FUNCTION(const T& a, const T& d) {
      old = a;
      return (__builtin_memcmp(&a, &expected, sizeof(T)) == 0) ? d : a;
    });
    return old;
  }
}

}  // namespace Impl

// ------------------------------------------------------------------ fetch_op
// Relaxed order, device (agent) scope, as core/src/Kokkos_Atomics_Desul_Wrapper.hpp:46-50.
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_add(T* const p, Impl::nodeduce_t<const T> v) { return Impl::s_fetch_add<Impl::kScopeAgent>(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_sub(T* const p, Impl::nodeduce_t<const T> v) { return Impl::s_fetch_sub<Impl::kScopeAgent>(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_min(T* const p, Impl::nodeduce_t<const T> v) { return Impl::s_fetch_min<Impl::kScopeAgent>(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_max(T* const p, Impl::nodeduce_t<const T> v) { return Impl::s_fetch_max<Impl::kScopeAgent>(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_and(T* const p, Impl::nodeduce_t<const T> v) { return Impl::s_fetch_and<Impl::kScopeAgent>(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_or(T* const p, Impl::nodeduce_t<const T> v) { return Impl::s_fetch_or<Impl::kScopeAgent>(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_xor(T* const p, Impl::nodeduce_t<const T> v) { return Impl::s_fetch_xor<Impl::kScopeAgent>(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_mul(T* const p, Impl::nodeduce_t<const T> v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, v, Impl::OpMul{}); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_div(T* const p, Impl::nodeduce_t<const T> v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, v, Impl::OpDiv{}); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_mod(T* const p, Impl::nodeduce_t<const T> v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, v, Impl::OpMod{}); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_nand(T* const p, Impl::nodeduce_t<const T> v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, v, Impl::OpNand{}); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_lshift(T* const p, const unsigned v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, T(v), Impl::OpLshift{}); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_rshift(T* const p, const unsigned v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, T(v), Impl::OpRshift{}); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_inc(T* const p) { return atomic_fetch_add(p, T(1)); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_dec(T* const p) { return atomic_fetch_sub(p, T(1)); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_inc_mod(T* const p, Impl::nodeduce_t<const T> v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, v, Impl::OpIncMod{}); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_fetch_dec_mod(T* const p, Impl::nodeduce_t<const T> v) { return Impl::fetch_oper<Impl::kScopeAgent>(p, v, Impl::OpDecMod{}); }

// ------------------------------------------------------------------ op_fetch
template <class T> KOKKOS_INLINE_FUNCTION T atomic_add_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_add(p, v) + v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_sub_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_sub(p, v) - v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_min_fetch(T* const p, Impl::nodeduce_t<const T> v) { T o = atomic_fetch_min(p, v); return v < o ? v : o; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_max_fetch(T* const p, Impl::nodeduce_t<const T> v) { T o = atomic_fetch_max(p, v); return o < v ? v : o; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_mul_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_mul(p, v) * v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_div_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_div(p, v) / v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_mod_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_mod(p, v) % v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_and_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_and(p, v) & v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_or_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_or(p, v) | v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_xor_fetch(T* const p, Impl::nodeduce_t<const T> v) { return atomic_fetch_xor(p, v) ^ v; }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_nand_fetch(T* const p, Impl::nodeduce_t<const T> v) { return ~(atomic_fetch_nand(p, v) & v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_lshift_fetch(T* const p, const unsigned v) { return T(atomic_fetch_lshift(p, v) << v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_rshift_fetch(T* const p, const unsigned v) { return T(atomic_fetch_rshift(p, v) >> v); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_inc_fetch(T* const p) { return atomic_fetch_inc(p) + T(1); }
template <class T> KOKKOS_INLINE_FUNCTION T atomic_dec_fetch(T* const p) { return atomic_fetch_dec(p) - T(1); }

// ------------------------------------------------------------------ op (no result)
template <class T> KOKKOS_INLINE_FUNCTION void atomic_add(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_add(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_sub(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_sub(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_min(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_min(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_max(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_max(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_mul(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_mul(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_div(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_div(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_mod(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_mod(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_and(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_and(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_or(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_or(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_xor(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_xor(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_nand(T* const p, Impl::nodeduce_t<const T> v) { (void)atomic_fetch_nand(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_lshift(T* const p, const unsigned v) { (void)atomic_fetch_lshift(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_rshift(T* const p, const unsigned v) { (void)atomic_fetch_rshift(p, v); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_inc(T* const p) { (void)atomic_fetch_inc(p); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_dec(T* const p) { (void)atomic_fetch_dec(p); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_increment(T* const p) { (void)atomic_fetch_inc(p); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_decrement(T* const p) { (void)atomic_fetch_dec(p); }

// ------------------------------------------------------------------ load/store/exchange/CAS
template <class T> KOKKOS_INLINE_FUNCTION T atomic_load(T* const p) { return Impl::s_load<Impl::kScopeAgent>(p); }
template <class T> KOKKOS_INLINE_FUNCTION void atomic_store(T* const p, Impl::nodeduce_t<const T> v) { Impl::s_store<Impl::kScopeAgent>(p, v); }
template <class T>
KOKKOS_INLINE_FUNCTION T atomic_exchange(T* const p, Impl::nodeduce_t<const T> v) {
  return Impl::fetch_oper<Impl::kScopeAgent>(p, v, Impl::OpSecond{});
}
template <class T>
KOKKOS_INLINE_FUNCTION void atomic_assign(T* const p, Impl::nodeduce_t<const T> v) {
  (void)atomic_exchange(p, v);
}
template <class T>
KOKKOS_INLINE_FUNCTION T atomic_compare_exchange(T* const p, Impl::nodeduce_t<const T> expected,
                                                 Impl::nodeduce_t<const T> desired) {
  return Impl::s_compare_exchange<Impl::kScopeAgent>(p, expected, desired);
}
template <class T>
KOKKOS_INLINE_FUNCTION bool atomic_compare_exchange_strong(T* const p, Impl::nodeduce_t<const T> expected,
                                                           Impl::nodeduce_t<const T> desired) {
  T old = atomic_compare_exchange(p, expected, desired);
  return __builtin_memcmp(&old, &expected, sizeof(T)) == 0;
}

KOKKOS_INLINE_FUNCTION void memory_fence() {
  KOKKOS_IF_ON_DEVICE((__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");))
  KOKKOS_IF_ON_HOST((__atomic_thread_fence(__ATOMIC_SEQ_CST);))
}
KOKKOS_INLINE_FUNCTION void load_fence() { memory_fence(); }
KOKKOS_INLINE_FUNCTION void store_fence() { memory_fence(); }

namespace Impl {
// Proxy returned by Views with MemoryTraits<Atomic>
// (reference core/src/View/Kokkos_ViewAtomic.hpp).
template <class T>
class AtomicDataElement {
 public:
  using value_type = std::remove_const_t<T>;
  KOKKOS_INLINE_FUNCTION explicit AtomicDataElement(T* p) : m_ptr(const_cast<value_type*>(p)) {}
  KOKKOS_INLINE_FUNCTION operator value_type() const { return atomic_load(m_ptr); }
  KOKKOS_INLINE_FUNCTION value_type operator=(const value_type& v) const { atomic_store(m_ptr, v); return v; }
  KOKKOS_INLINE_FUNCTION value_type operator+=(const value_type& v) const { return atomic_add_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator-=(const value_type& v) const { return atomic_sub_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator*=(const value_type& v) const { return atomic_mul_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator/=(const value_type& v) const { return atomic_div_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator%=(const value_type& v) const { return atomic_mod_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator&=(const value_type& v) const { return atomic_and_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator|=(const value_type& v) const { return atomic_or_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator^=(const value_type& v) const { return atomic_xor_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator<<=(const unsigned v) const { return atomic_lshift_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator>>=(const unsigned v) const { return atomic_rshift_fetch(m_ptr, v); }
  KOKKOS_INLINE_FUNCTION value_type operator++() const { return atomic_inc_fetch(m_ptr); }
  KOKKOS_INLINE_FUNCTION value_type operator--() const { return atomic_dec_fetch(m_ptr); }
  KOKKOS_INLINE_FUNCTION value_type operator++(int) const { return atomic_fetch_inc(m_ptr); }
  KOKKOS_INLINE_FUNCTION value_type operator--(int) const { return atomic_fetch_dec(m_ptr); }

 private:
  value_type* m_ptr;
};
}  // namespace Impl

}  // namespace Kokkos

#endif
