// Forward declarations, error handling and small utilities shared by the
// whole Kokkos-API layer of the MI355X execution space.
//
// Reference: core/src/Kokkos_Core_fwd.hpp (primary templates of the dispatch
// closures, :290-328), core/src/impl/Kokkos_Error.hpp (throw_runtime_exception,
// RawMemoryAllocationFailure) and core/src/Kokkos_Abort.hpp (Kokkos::abort).
#ifndef KKMX_CORE_FWD_HPP
#define KKMX_CORE_FWD_HPP

#include <Kokkos_Macros.hpp>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "kkmx_runtime.h"

namespace Kokkos {

// ---------------------------------------------------------------- printf
// Kokkos::printf (reference core/src/Kokkos_Printf.hpp): C printf on the
// host and from kernels, returning nothing; a lone format string is printed
// through "%s" so that a stray '%' in it is harmless.
template <class... Args>
KOKKOS_FORCEINLINE_FUNCTION void printf(const char* format, Args... args) {
  if constexpr (sizeof...(Args) == 0) {
    ::printf("%s", format);
  } else {
    ::printf(format, args...);
  }
}

// ---------------------------------------------------------------- abort
// Host: print and abort the process (reference core/src/impl/Kokkos_Error.cpp
// host_abort).  Device: trap the wavefront; the fault surfaces at the next
// fence as a HIP error.
KOKKOS_INLINE_FUNCTION void abort(const char* msg) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)msg;
  __builtin_trap();
#else
  std::fprintf(stderr, "%s\n", msg);
  std::fflush(stderr);
  std::abort();
#endif
}

namespace Impl {

[[noreturn]] inline void throw_runtime_exception(const std::string& msg) {
  throw std::runtime_error(msg);
}

// HIP runtime status -> exception, like KOKKOS_IMPL_HIP_SAFE_CALL
// (reference core/src/HIP/Kokkos_HIP_Error.hpp, Kokkos_HIP_Instance.cpp:456-465).
inline void mx_check(int status, const char* what, const char* file, int line) {
  if (status != 0) {
    std::string msg = std::string(what) + " error( " +
                      mx_error_string(status) + "): " + file + ":" +
                      std::to_string(line);
    throw std::runtime_error(msg);
  }
}

}  // namespace Impl

namespace Experimental {
// Allocation failure type of the reference (core/src/impl/Kokkos_Error.hpp).
class RawMemoryAllocationFailure : public std::bad_alloc {
 public:
  RawMemoryAllocationFailure(size_t bytes, std::string msg)
      : m_bytes(bytes), m_msg(std::move(msg)) {}
  const char* what() const noexcept override { return m_msg.c_str(); }
  size_t attempted_size() const noexcept { return m_bytes; }

 private:
  size_t m_bytes;
  std::string m_msg;
};
}  // namespace Experimental

// ---------------------------------------------------------------- spaces
class HostSpace;
class MI355XSpace;
class MI355XHostPinnedSpace;
class Serial;
class MI355X;

using DefaultExecutionSpace     = MI355X;
using DefaultHostExecutionSpace = Serial;

template <class ExecutionSpace, class MemorySpace>
struct Device {
  using execution_space = ExecutionSpace;
  using memory_space    = MemorySpace;
  using device_type     = Device<execution_space, memory_space>;
};

struct AUTO_t {
  KOKKOS_INLINE_FUNCTION constexpr const AUTO_t& operator()() const { return *this; }
};
inline constexpr AUTO_t AUTO{};

struct InvalidType {};

// Dispatch closures; the execution spaces specialise them
// (reference core/src/Kokkos_Core_fwd.hpp:290-328).
namespace Impl {
template <class FunctorType, class ExecPolicy, class ExecutionSpace = void>
class ParallelFor;
template <class CombinedFunctorReducerType, class PolicyType,
          class ExecutionSpaceType = void>
class ParallelReduce;
template <class FunctorType, class ExecPolicy, class ExecutionSpace = void>
class ParallelScan;
template <class FunctorType, class ExecPolicy, class ReturnType = InvalidType,
          class ExecutionSpace = void>
class ParallelScanWithTotal;
}  // namespace Impl

// ---------------------------------------------------------------- traits
template <class T, class = void>
struct is_execution_space : std::false_type {};
template <class T>
struct is_execution_space<T, std::void_t<typename T::execution_space,
                                         typename T::memory_space,
                                         decltype(&T::concurrency)>>
    : std::is_same<T, typename T::execution_space> {};
template <class T>
inline constexpr bool is_execution_space_v = is_execution_space<T>::value;

template <class T, class = void>
struct is_memory_space : std::false_type {};
template <class T>
struct is_memory_space<T, std::void_t<typename T::memory_space>>
    : std::is_same<T, typename T::memory_space> {};
template <class T>
inline constexpr bool is_memory_space_v = is_memory_space<T>::value;

namespace Impl {
template <class T, class = void>
struct has_execution_space : std::false_type {};
template <class T>
struct has_execution_space<T, std::void_t<typename T::execution_space>>
    : std::true_type {};

template <class T>
struct remove_cvref {
  using type = std::remove_cv_t<std::remove_reference_t<T>>;
};
template <class T>
using remove_cvref_t = typename remove_cvref<T>::type;

// Non-allocating integer helpers.
KOKKOS_INLINE_FUNCTION constexpr int64_t ceil_div(int64_t a, int64_t b) {
  return (a + b - 1) / b;
}
KOKKOS_INLINE_FUNCTION constexpr bool is_pow2(uint64_t v) {
  return v && !(v & (v - 1));
}
}  // namespace Impl

}  // namespace Kokkos

#define KKMX_SAFE_CALL(call) \
  ::Kokkos::Impl::mx_check((call), #call, __FILE__, __LINE__)

#endif
