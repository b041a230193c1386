// Kokkos math functions, min / max / clamp / minmax, kokkos_swap and the
// Kokkos::numbers constants, callable on the host and in kernels.
//
// Reference interfaces: core/src/Kokkos_MathematicalFunctions.hpp (the
// <cmath> set, abs overloads, rsqrt), core/src/Kokkos_MinMax.hpp:27-180,
// core/src/Kokkos_Swap.hpp, core/src/Kokkos_MathematicalConstants.hpp.
//
// Each function forwards to the C library function of the same name on the
// host and to the device math library on gfx950 (the global-namespace
// overloads HIP declares for both sides); float arguments take the float
// function, integer arguments are computed in double as <cmath> does.
#ifndef KKMX_MATH_HPP
#define KKMX_MATH_HPP

#include <Kokkos_Macros.hpp>

#include <cmath>
#include <cstdlib>
#include <initializer_list>
#include <type_traits>
#include <utility>

namespace Kokkos {

template <class T1, class T2>
struct pair;

// ------------------------------------------------------------------ min/max
template <class T>
KOKKOS_INLINE_FUNCTION constexpr const T& min(const T& a, const T& b) {
  return (b < a) ? b : a;
}
template <class T, class Compare>
KOKKOS_INLINE_FUNCTION constexpr const T& min(const T& a, const T& b, Compare comp) {
  return comp(b, a) ? b : a;
}
template <class T>
KOKKOS_INLINE_FUNCTION constexpr T min(std::initializer_list<T> l) {
  const T* p = l.begin();
  T m        = *p;
  for (++p; p != l.end(); ++p)
    if (*p < m) m = *p;
  return m;
}
template <class T>
KOKKOS_INLINE_FUNCTION constexpr const T& max(const T& a, const T& b) {
  return (a < b) ? b : a;
}
template <class T, class Compare>
KOKKOS_INLINE_FUNCTION constexpr const T& max(const T& a, const T& b, Compare comp) {
  return comp(a, b) ? b : a;
}
template <class T>
KOKKOS_INLINE_FUNCTION constexpr T max(std::initializer_list<T> l) {
  const T* p = l.begin();
  T m        = *p;
  for (++p; p != l.end(); ++p)
    if (m < *p) m = *p;
  return m;
}
template <class T>
KOKKOS_INLINE_FUNCTION constexpr const T& clamp(const T& v, const T& lo, const T& hi) {
  return (v < lo) ? lo : (hi < v) ? hi : v;
}
template <class T, class Compare>
KOKKOS_INLINE_FUNCTION constexpr const T& clamp(const T& v, const T& lo, const T& hi, Compare comp) {
  return comp(v, lo) ? lo : comp(hi, v) ? hi : v;
}

// ------------------------------------------------------------------ swap
template <class T>
KOKKOS_INLINE_FUNCTION constexpr void kokkos_swap(T& a, T& b) noexcept {
  T t = std::move(a);
  a   = std::move(b);
  b   = std::move(t);
}
template <class T, std::size_t N>
KOKKOS_INLINE_FUNCTION constexpr void kokkos_swap(T (&a)[N], T (&b)[N]) noexcept {
  for (std::size_t i = 0; i < N; ++i) kokkos_swap(a[i], b[i]);
}

// ------------------------------------------------------------------ abs
KOKKOS_INLINE_FUNCTION int abs(int n) { return n < 0 ? -n : n; }
KOKKOS_INLINE_FUNCTION long abs(long n) { return n < 0 ? -n : n; }
KOKKOS_INLINE_FUNCTION long long abs(long long n) { return n < 0 ? -n : n; }
KOKKOS_INLINE_FUNCTION float abs(float x) { return ::fabsf(x); }
KOKKOS_INLINE_FUNCTION double abs(double x) { return ::fabs(x); }

// ------------------------------------------------------------------ <cmath>
#define KKMX_MATH_UNARY(NAME)                                                             \
  KOKKOS_INLINE_FUNCTION float NAME(float x) { return ::NAME##f(x); }                     \
  KOKKOS_INLINE_FUNCTION double NAME(double x) { return ::NAME(x); }                      \
  template <class T, class = std::enable_if_t<std::is_integral<T>::value>>                \
  KOKKOS_INLINE_FUNCTION double NAME(T x) {                                               \
    return ::NAME(static_cast<double>(x));                                                \
  }                                                                                       \
  KOKKOS_INLINE_FUNCTION float NAME##f(float x) { return ::NAME##f(x); }

#define KKMX_MATH_BINARY(NAME)                                                            \
  KOKKOS_INLINE_FUNCTION float NAME(float x, float y) { return ::NAME##f(x, y); }         \
  KOKKOS_INLINE_FUNCTION double NAME(double x, double y) { return ::NAME(x, y); }         \
  template <class T, class U,                                                             \
            class = std::enable_if_t<std::is_arithmetic<T>::value && std::is_arithmetic<U>::value && \
                                     !(std::is_same<T, float>::value && std::is_same<U, float>::value) && \
                                     !(std::is_same<T, double>::value && std::is_same<U, double>::value)>> \
  KOKKOS_INLINE_FUNCTION double NAME(T x, U y) {                                          \
    return ::NAME(static_cast<double>(x), static_cast<double>(y));                        \
  }                                                                                       \
  KOKKOS_INLINE_FUNCTION float NAME##f(float x, float y) { return ::NAME##f(x, y); }

KKMX_MATH_UNARY(fabs)
KKMX_MATH_BINARY(fmod)
KKMX_MATH_BINARY(remainder)
KKMX_MATH_BINARY(fmax)
KKMX_MATH_BINARY(fmin)
KKMX_MATH_BINARY(fdim)
KKMX_MATH_UNARY(exp)
KKMX_MATH_UNARY(exp2)
KKMX_MATH_UNARY(expm1)
KKMX_MATH_UNARY(log)
KKMX_MATH_UNARY(log10)
KKMX_MATH_UNARY(log2)
KKMX_MATH_UNARY(log1p)
KKMX_MATH_BINARY(pow)
KKMX_MATH_UNARY(sqrt)
KKMX_MATH_UNARY(cbrt)
KKMX_MATH_BINARY(hypot)
KKMX_MATH_UNARY(sin)
KKMX_MATH_UNARY(cos)
KKMX_MATH_UNARY(tan)
KKMX_MATH_UNARY(asin)
KKMX_MATH_UNARY(acos)
KKMX_MATH_UNARY(atan)
KKMX_MATH_BINARY(atan2)
KKMX_MATH_UNARY(sinh)
KKMX_MATH_UNARY(cosh)
KKMX_MATH_UNARY(tanh)
KKMX_MATH_UNARY(asinh)
KKMX_MATH_UNARY(acosh)
KKMX_MATH_UNARY(atanh)
KKMX_MATH_UNARY(erf)
KKMX_MATH_UNARY(erfc)
KKMX_MATH_UNARY(tgamma)
KKMX_MATH_UNARY(lgamma)
KKMX_MATH_UNARY(ceil)
KKMX_MATH_UNARY(floor)
KKMX_MATH_UNARY(trunc)
KKMX_MATH_UNARY(round)
KKMX_MATH_UNARY(nearbyint)
KKMX_MATH_UNARY(logb)
KKMX_MATH_BINARY(nextafter)
KKMX_MATH_BINARY(copysign)

#undef KKMX_MATH_UNARY
#undef KKMX_MATH_BINARY

KOKKOS_INLINE_FUNCTION float fma(float x, float y, float z) { return ::fmaf(x, y, z); }
KOKKOS_INLINE_FUNCTION double fma(double x, double y, double z) { return ::fma(x, y, z); }
KOKKOS_INLINE_FUNCTION float fmaf(float x, float y, float z) { return ::fmaf(x, y, z); }
// 1 / sqrt(x), with the square root correctly rounded before the division
KOKKOS_INLINE_FUNCTION float rsqrt(float x) { return 1.0f / ::sqrtf(x); }
KOKKOS_INLINE_FUNCTION double rsqrt(double x) { return 1.0 / ::sqrt(x); }

KOKKOS_INLINE_FUNCTION bool isfinite(float x) { return ::isfinite(x); }
KOKKOS_INLINE_FUNCTION bool isfinite(double x) { return ::isfinite(x); }
KOKKOS_INLINE_FUNCTION bool isinf(float x) { return ::isinf(x); }
KOKKOS_INLINE_FUNCTION bool isinf(double x) { return ::isinf(x); }
KOKKOS_INLINE_FUNCTION bool isnan(float x) { return ::isnan(x); }
KOKKOS_INLINE_FUNCTION bool isnan(double x) { return ::isnan(x); }
KOKKOS_INLINE_FUNCTION bool signbit(float x) { return ::signbit(x); }
KOKKOS_INLINE_FUNCTION bool signbit(double x) { return ::signbit(x); }

// ------------------------------------------------------------------ numbers
namespace numbers {
template <class T>
inline constexpr T e_v = T(2.718281828459045235360287471352662498L);
template <class T>
inline constexpr T log2e_v = T(1.442695040888963407359924681001892137L);
template <class T>
inline constexpr T log10e_v = T(0.434294481903251827651128918916605082L);
template <class T>
inline constexpr T pi_v = T(3.141592653589793238462643383279502884L);
template <class T>
inline constexpr T inv_pi_v = T(0.318309886183790671537767526745028724L);
template <class T>
inline constexpr T inv_sqrtpi_v = T(0.564189583547756286948079451560772586L);
template <class T>
inline constexpr T ln2_v = T(0.693147180559945309417232121458176568L);
template <class T>
inline constexpr T ln10_v = T(2.302585092994045684017991454684364208L);
template <class T>
inline constexpr T sqrt2_v = T(1.414213562373095048801688724209698079L);
template <class T>
inline constexpr T sqrt3_v = T(1.732050807568877293527446341505872367L);
template <class T>
inline constexpr T inv_sqrt3_v = T(0.577350269189625764509148780501957456L);
template <class T>
inline constexpr T egamma_v = T(0.577215664901532860606512090082402431L);
template <class T>
inline constexpr T phi_v = T(1.618033988749894848204586834365638118L);

inline constexpr double e          = e_v<double>;
inline constexpr double log2e      = log2e_v<double>;
inline constexpr double log10e     = log10e_v<double>;
inline constexpr double pi         = pi_v<double>;
inline constexpr double inv_pi     = inv_pi_v<double>;
inline constexpr double inv_sqrtpi = inv_sqrtpi_v<double>;
inline constexpr double ln2        = ln2_v<double>;
inline constexpr double ln10       = ln10_v<double>;
inline constexpr double sqrt2      = sqrt2_v<double>;
inline constexpr double sqrt3      = sqrt3_v<double>;
inline constexpr double inv_sqrt3  = inv_sqrt3_v<double>;
inline constexpr double egamma     = egamma_v<double>;
inline constexpr double phi        = phi_v<double>;
}  // namespace numbers

}  // namespace Kokkos

#endif
