// Function-qualifier and lambda macros of the Kokkos programming model for the
// MI355X execution space.  Names follow core/src/Kokkos_Macros.hpp and
// core/src/setup/Kokkos_Setup_HIP.hpp:33-43 of the reference (same spelling,
// same meaning: functions callable on host and device, lambdas captured by
// value and compiled for both).
#ifndef KOKKOS_MACROS_HPP
#define KOKKOS_MACROS_HPP

#include <hip/hip_runtime.h>

#define KOKKOS_VERSION 40699
#define KOKKOS_VERSION_MAJOR 4
#define KOKKOS_VERSION_MINOR 6
#define KOKKOS_VERSION_PATCH 99

#define KOKKOS_ENABLE_MI355X 1
#define KOKKOS_ENABLE_SERIAL 1
#define KOKKOS_ENABLE_TUNING 1

#define KOKKOS_FUNCTION __host__ __device__
#define KOKKOS_INLINE_FUNCTION __host__ __device__ inline
#define KOKKOS_FORCEINLINE_FUNCTION __host__ __device__ __forceinline__
#define KOKKOS_INLINE_FUNCTION_DELETED __host__ __device__ inline
#define KOKKOS_DEFAULTED_FUNCTION
#define KOKKOS_DEDUCTION_GUIDE
#define KOKKOS_IMPL_FORCEINLINE __forceinline__
#define KOKKOS_IMPL_HOST_FUNCTION __host__
#define KOKKOS_IMPL_DEVICE_FUNCTION __device__

#define KOKKOS_LAMBDA [=] __host__ __device__
#define KOKKOS_CLASS_LAMBDA [ =, *this ] __host__ __device__

#if defined(__HIP_DEVICE_COMPILE__)
#define KOKKOS_IF_ON_DEVICE(CODE) \
  { KOKKOS_IMPL_STRIP_PARENS CODE }
#define KOKKOS_IF_ON_HOST(CODE) \
  {}
#else
#define KOKKOS_IF_ON_DEVICE(CODE) \
  {}
#define KOKKOS_IF_ON_HOST(CODE) \
  { KOKKOS_IMPL_STRIP_PARENS CODE }
#endif
#define KOKKOS_IMPL_STRIP_PARENS(...) __VA_ARGS__

#define KOKKOS_IMPL_CONCAT_(a, b) a##b
#define KOKKOS_IMPL_CONCAT(a, b) KOKKOS_IMPL_CONCAT_(a, b)

#define KOKKOS_EXPECTS(...)                                              \
  do {                                                                   \
    if (!(__VA_ARGS__)) ::Kokkos::abort("Kokkos contract violation: " #__VA_ARGS__); \
  } while (0)
#define KOKKOS_ENSURES(...) KOKKOS_EXPECTS(__VA_ARGS__)
#define KOKKOS_ASSERT(...)                                                        \
  do {                                                                            \
    if (!(__VA_ARGS__)) ::Kokkos::abort("Kokkos assertion failed: " #__VA_ARGS__); \
  } while (0)

// A wavefront is 64 lanes on CDNA.
#define KOKKOS_IMPL_MI355X_WAVE 64

#endif
