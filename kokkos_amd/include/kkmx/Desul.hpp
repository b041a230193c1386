// The desul atomic entry points (namespace desul) that Kokkos code calls
// directly: operations taking a memory-order tag and a memory-scope tag
// (tpls/desul/include/desul/atomics/Generic.hpp:20-453, tag types
// tpls/desul/include/desul/atomics/Common.hpp:32-55).  The reference's
// Kokkos::atomic_* wrapper (core/src/Kokkos_Atomics_Desul_Wrapper.hpp:52-176)
// is these with MemoryOrderRelaxed and MemoryScopeDevice.
//
// On MI355X every call lands on the scoped operations of kkmx/Atomic.hpp:
//   MemoryScopeCaller        plain loads and stores (no other thread races)
//   MemoryScopeCore          workgroup scope, as the reference maps it
//                            (tpls/desul/include/desul/atomics/Adapt_HIP.hpp:55):
//                            only threads of one workgroup (one CU) touch the
//                            address, so the atomic resolves in the XCD's L2
//   MemoryScopeDevice        agent scope: native L2 atomics across the 8 XCDs
//   MemoryScopeNode/System   system scope (host-coherent memory, peer GPUs)
// Memory orders other than relaxed add fences of the same scope around the
// relaxed operation: release (and acq_rel, seq_cst) before it, acquire (and
// acq_rel, seq_cst) after it.
#ifndef KKMX_DESUL_HPP
#define KKMX_DESUL_HPP

#include <kkmx/Atomic.hpp>

namespace desul {

struct MemoryOrderSeqCst {};
struct MemoryOrderAcqRel {};
struct MemoryOrderAcquire {};
struct MemoryOrderRelease {};
struct MemoryOrderRelaxed {};

struct MemoryScopeSystem {};
struct MemoryScopeNode {};
struct MemoryScopeDevice {};
struct MemoryScopeCore {};
struct MemoryScopeCaller {};

namespace Impl {

template <class T>
struct dont_deduce_this_parameter {
  using type = T;
};
template <class T>
using dont_deduce_this_parameter_t = typename dont_deduce_this_parameter<T>::type;

template <class MS>
struct hip_scope;
template <>
struct hip_scope<MemoryScopeCaller> {
  static constexpr int value = Kokkos::Impl::kScopeCaller;
};
template <>
struct hip_scope<MemoryScopeCore> {
  static constexpr int value = Kokkos::Impl::kScopeWorkgroup;
};
template <>
struct hip_scope<MemoryScopeDevice> {
  static constexpr int value = Kokkos::Impl::kScopeAgent;
};
template <>
struct hip_scope<MemoryScopeNode> {
  static constexpr int value = Kokkos::Impl::kScopeSystem;
};
template <>
struct hip_scope<MemoryScopeSystem> {
  static constexpr int value = Kokkos::Impl::kScopeSystem;
};

template <class MO>
constexpr bool releases() {
  return std::is_same<MO, MemoryOrderRelease>::value || std::is_same<MO, MemoryOrderAcqRel>::value ||
         std::is_same<MO, MemoryOrderSeqCst>::value;
}
template <class MO>
constexpr bool acquires() {
  return std::is_same<MO, MemoryOrderAcquire>::value || std::is_same<MO, MemoryOrderAcqRel>::value ||
         std::is_same<MO, MemoryOrderSeqCst>::value;
}

// The fence builtin takes its order as a literal, so each order has a branch.
template <int Order, int S>
KOKKOS_FORCEINLINE_FUNCTION void fence() {
  if constexpr (S != Kokkos::Impl::kScopeCaller) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (S == Kokkos::Impl::kScopeSystem) {
      if constexpr (Order == __ATOMIC_SEQ_CST) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "");
      else if constexpr (Order == __ATOMIC_ACQ_REL) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "");
      else if constexpr (Order == __ATOMIC_ACQUIRE) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
      else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    } else {
      if constexpr (Order == __ATOMIC_SEQ_CST) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
      else if constexpr (Order == __ATOMIC_ACQ_REL) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      else if constexpr (Order == __ATOMIC_ACQUIRE) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    }
#else
    __atomic_thread_fence(Order);
#endif
  }
}

template <class MO, class MS>
KOKKOS_FORCEINLINE_FUNCTION void before() {
  if constexpr (std::is_same<MO, MemoryOrderSeqCst>::value) fence<__ATOMIC_SEQ_CST, hip_scope<MS>::value>();
  else if constexpr (releases<MO>()) fence<__ATOMIC_RELEASE, hip_scope<MS>::value>();
}
template <class MO, class MS>
KOKKOS_FORCEINLINE_FUNCTION void after() {
  if constexpr (std::is_same<MO, MemoryOrderSeqCst>::value) fence<__ATOMIC_SEQ_CST, hip_scope<MS>::value>();
  else if constexpr (acquires<MO>()) fence<__ATOMIC_ACQUIRE, hip_scope<MS>::value>();
}

// Runs `f` between the fences that `MO` asks for and returns its result.
template <class MO, class MS, class F>
KOKKOS_FORCEINLINE_FUNCTION auto ordered(const F& f) {
  before<MO, MS>();
  if constexpr (std::is_void<decltype(f())>::value) {
    f();
    after<MO, MS>();
  } else {
    auto r = f();
    after<MO, MS>();
    return r;
  }
}

}  // namespace Impl

template <class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_thread_fence(MO, MS) {
  if constexpr (std::is_same<MO, MemoryOrderSeqCst>::value)
    Impl::fence<__ATOMIC_SEQ_CST, Impl::hip_scope<MS>::value>();
  else if constexpr (std::is_same<MO, MemoryOrderAcqRel>::value)
    Impl::fence<__ATOMIC_ACQ_REL, Impl::hip_scope<MS>::value>();
  else if constexpr (std::is_same<MO, MemoryOrderAcquire>::value)
    Impl::fence<__ATOMIC_ACQUIRE, Impl::hip_scope<MS>::value>();
  else if constexpr (std::is_same<MO, MemoryOrderRelease>::value)
    Impl::fence<__ATOMIC_RELEASE, Impl::hip_scope<MS>::value>();
}

#define KKMX_DESUL_SCOPED(NAME, CALL)                                                                        \
  template <class T, class MO, class MS>                                                                      \
  KOKKOS_INLINE_FUNCTION T NAME(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO, MS) { \
    constexpr int S = Impl::hip_scope<MS>::value;                                                             \
    return Impl::ordered<MO, MS>([&]() { return CALL; });                                                    \
  }

// fetch_op: the value before the operation
KKMX_DESUL_SCOPED(atomic_fetch_add, (Kokkos::Impl::s_fetch_add<S>(dest, val)))
KKMX_DESUL_SCOPED(atomic_fetch_sub, (Kokkos::Impl::s_fetch_sub<S>(dest, val)))
KKMX_DESUL_SCOPED(atomic_fetch_min, (Kokkos::Impl::s_fetch_min<S>(dest, val)))
KKMX_DESUL_SCOPED(atomic_fetch_max, (Kokkos::Impl::s_fetch_max<S>(dest, val)))
KKMX_DESUL_SCOPED(atomic_fetch_and, (Kokkos::Impl::s_fetch_and<S>(dest, val)))
KKMX_DESUL_SCOPED(atomic_fetch_or, (Kokkos::Impl::s_fetch_or<S>(dest, val)))
KKMX_DESUL_SCOPED(atomic_fetch_xor, (Kokkos::Impl::s_fetch_xor<S>(dest, val)))
KKMX_DESUL_SCOPED(atomic_fetch_mul, (Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpMul{})))
KKMX_DESUL_SCOPED(atomic_fetch_div, (Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpDiv{})))
KKMX_DESUL_SCOPED(atomic_fetch_mod, (Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpMod{})))
KKMX_DESUL_SCOPED(atomic_fetch_nand, (Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpNand{})))
KKMX_DESUL_SCOPED(atomic_fetch_inc_mod, (Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpIncMod{})))
KKMX_DESUL_SCOPED(atomic_fetch_dec_mod, (Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpDecMod{})))
KKMX_DESUL_SCOPED(atomic_exchange, (Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpSecond{})))

// op_fetch: the value after the operation
KKMX_DESUL_SCOPED(atomic_add_fetch, (T(Kokkos::Impl::s_fetch_add<S>(dest, val) + val)))
KKMX_DESUL_SCOPED(atomic_sub_fetch, (T(Kokkos::Impl::s_fetch_sub<S>(dest, val) - val)))
KKMX_DESUL_SCOPED(atomic_min_fetch, (Kokkos::Impl::OpMin{}(Kokkos::Impl::s_fetch_min<S>(dest, val), T(val))))
KKMX_DESUL_SCOPED(atomic_max_fetch, (Kokkos::Impl::OpMax{}(Kokkos::Impl::s_fetch_max<S>(dest, val), T(val))))
KKMX_DESUL_SCOPED(atomic_and_fetch, (T(Kokkos::Impl::s_fetch_and<S>(dest, val) & val)))
KKMX_DESUL_SCOPED(atomic_or_fetch, (T(Kokkos::Impl::s_fetch_or<S>(dest, val) | val)))
KKMX_DESUL_SCOPED(atomic_xor_fetch, (T(Kokkos::Impl::s_fetch_xor<S>(dest, val) ^ val)))
KKMX_DESUL_SCOPED(atomic_mul_fetch, (T(Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpMul{}) * val)))
KKMX_DESUL_SCOPED(atomic_div_fetch, (T(Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpDiv{}) / val)))
KKMX_DESUL_SCOPED(atomic_mod_fetch, (T(Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpMod{}) % val)))
KKMX_DESUL_SCOPED(atomic_nand_fetch,
                  (T(~(Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpNand{}) & val))))
#undef KKMX_DESUL_SCOPED

template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_fetch_lshift(T* const dest, const unsigned val, MO, MS) {
  constexpr int S = Impl::hip_scope<MS>::value;
  return Impl::ordered<MO, MS>([&]() { return Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpLshift{}); });
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_fetch_rshift(T* const dest, const unsigned val, MO, MS) {
  constexpr int S = Impl::hip_scope<MS>::value;
  return Impl::ordered<MO, MS>([&]() { return Kokkos::Impl::fetch_oper<S>(dest, T(val), Kokkos::Impl::OpRshift{}); });
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_lshift_fetch(T* const dest, const unsigned val, MO mo, MS ms) {
  return T(atomic_fetch_lshift(dest, val, mo, ms) << val);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_rshift_fetch(T* const dest, const unsigned val, MO mo, MS ms) {
  return T(atomic_fetch_rshift(dest, val, mo, ms) >> val);
}

template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_fetch_inc(T* const dest, MO mo, MS ms) {
  return atomic_fetch_add(dest, T(1), mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_fetch_dec(T* const dest, MO mo, MS ms) {
  return atomic_fetch_sub(dest, T(1), mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_inc_fetch(T* const dest, MO mo, MS ms) {
  return atomic_add_fetch(dest, T(1), mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_dec_fetch(T* const dest, MO mo, MS ms) {
  return atomic_sub_fetch(dest, T(1), mo, ms);
}

// op: no result
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_add(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO mo, MS ms) {
  (void)atomic_fetch_add(dest, val, mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_sub(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO mo, MS ms) {
  (void)atomic_fetch_sub(dest, val, mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_mul(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO mo, MS ms) {
  (void)atomic_fetch_mul(dest, val, mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_div(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO mo, MS ms) {
  (void)atomic_fetch_div(dest, val, mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_min(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO mo, MS ms) {
  (void)atomic_fetch_min(dest, val, mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_max(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO mo, MS ms) {
  (void)atomic_fetch_max(dest, val, mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_inc(T* const dest, MO mo, MS ms) {
  (void)atomic_fetch_add(dest, T(1), mo, ms);
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_dec(T* const dest, MO mo, MS ms) {
  (void)atomic_fetch_sub(dest, T(1), mo, ms);
}

template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_load(const T* const dest, MO, MS) {
  constexpr int S = Impl::hip_scope<MS>::value;
  return Impl::ordered<MO, MS>([&]() { return Kokkos::Impl::s_load<S>(const_cast<T*>(dest)); });
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION void atomic_store(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> val, MO, MS) {
  constexpr int S = Impl::hip_scope<MS>::value;
  Impl::ordered<MO, MS>([&]() { Kokkos::Impl::s_store<S>(dest, T(val)); });
}
template <class T, class MO, class MS>
KOKKOS_INLINE_FUNCTION T atomic_compare_exchange(T* const dest, const Impl::dont_deduce_this_parameter_t<const T> cmp,
                                                 const Impl::dont_deduce_this_parameter_t<const T> val, MO, MS) {
  constexpr int S = Impl::hip_scope<MS>::value;
  return Impl::ordered<MO, MS>([&]() { return Kokkos::Impl::s_compare_exchange<S>(dest, T(cmp), T(val)); });
}
template <class T, class SuccessMO, class FailureMO, class MS>
KOKKOS_INLINE_FUNCTION bool atomic_compare_exchange_strong(T* const dest, T& expected, T desired, SuccessMO success,
                                                           FailureMO, MS scope) {
  const T old = atomic_compare_exchange(dest, expected, desired, success, scope);
  if (__builtin_memcmp(&old, &expected, sizeof(T)) != 0) {
    expected = old;
    return false;
  }
  return true;
}
template <class T, class SuccessMO, class FailureMO, class MS>
KOKKOS_INLINE_FUNCTION bool atomic_compare_exchange_weak(T* const dest, T& expected, T desired, SuccessMO success,
                                                         FailureMO failure, MS scope) {
  return atomic_compare_exchange_strong(dest, expected, desired, success, failure, scope);
}

}  // namespace desul

#endif
