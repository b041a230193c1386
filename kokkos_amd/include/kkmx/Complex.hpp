// Kokkos::complex<RealType>, usable on the host and in kernels.
//
// Reference interface: core/src/Kokkos_Complex.hpp (class complex, its
// arithmetic with complex and real operands, comparisons, real / imag /
// abs / arg / norm / conj / polar, exp / log / log10 / sqrt / pow, the
// trigonometric and hyperbolic functions, stream insertion and
// extraction, conversion to and from std::complex).
//
// Layout is two RealType values, real then imaginary, aligned to twice the
// size of RealType like std::complex, so Views of complex are plain arrays
// of pairs and atomics on them take the wide-type path.  Division uses
// Smith's scaled algorithm (R. L. Smith, "Algorithm 116: Complex division",
// CACM 5(8), 1962), which avoids overflow when |denominator| is large; the
// functions follow their textbook definitions in terms of the real math
// functions (exp, log, sqrt, sin, cos, sinh, cosh, atan2, hypot).
#ifndef KKMX_COMPLEX_HPP
#define KKMX_COMPLEX_HPP

#include <kkmx/Math.hpp>

#include <complex>
#include <istream>
#include <ostream>
#include <type_traits>

namespace Kokkos {

template <class RealType>
class alignas(2 * sizeof(RealType)) complex {
  static_assert(std::is_floating_point<RealType>::value, "Kokkos::complex is defined for floating-point types");
  RealType re_{};
  RealType im_{};

 public:
  using value_type = RealType;

  KOKKOS_DEFAULTED_FUNCTION complex() = default;
  KOKKOS_INLINE_FUNCTION constexpr complex(const RealType& re) noexcept : re_(re), im_(RealType(0)) {}
  KOKKOS_INLINE_FUNCTION constexpr complex(const RealType& re, const RealType& im) noexcept : re_(re), im_(im) {}
  template <class U, class = std::enable_if_t<std::is_convertible<U, RealType>::value>>
  KOKKOS_INLINE_FUNCTION constexpr complex(const complex<U>& z) noexcept
      : re_(RealType(z.real())), im_(RealType(z.imag())) {}
  complex(const std::complex<RealType>& z) noexcept : re_(z.real()), im_(z.imag()) {}
  operator std::complex<RealType>() const noexcept { return std::complex<RealType>(re_, im_); }

  KOKKOS_INLINE_FUNCTION complex& operator=(const RealType& re) noexcept {
    re_ = re;
    im_ = RealType(0);
    return *this;
  }

  KOKKOS_INLINE_FUNCTION constexpr RealType real() const noexcept { return re_; }
  KOKKOS_INLINE_FUNCTION constexpr RealType imag() const noexcept { return im_; }
  KOKKOS_INLINE_FUNCTION RealType& real() noexcept { return re_; }
  KOKKOS_INLINE_FUNCTION RealType& imag() noexcept { return im_; }
  KOKKOS_INLINE_FUNCTION void real(RealType v) noexcept { re_ = v; }
  KOKKOS_INLINE_FUNCTION void imag(RealType v) noexcept { im_ = v; }

  KOKKOS_INLINE_FUNCTION complex& operator+=(const complex& z) noexcept {
    re_ += z.re_;
    im_ += z.im_;
    return *this;
  }
  KOKKOS_INLINE_FUNCTION complex& operator+=(const RealType& v) noexcept {
    re_ += v;
    return *this;
  }
  KOKKOS_INLINE_FUNCTION complex& operator-=(const complex& z) noexcept {
    re_ -= z.re_;
    im_ -= z.im_;
    return *this;
  }
  KOKKOS_INLINE_FUNCTION complex& operator-=(const RealType& v) noexcept {
    re_ -= v;
    return *this;
  }
  KOKKOS_INLINE_FUNCTION complex& operator*=(const complex& z) noexcept {
    const RealType r = re_ * z.re_ - im_ * z.im_;
    im_              = re_ * z.im_ + im_ * z.re_;
    re_              = r;
    return *this;
  }
  KOKKOS_INLINE_FUNCTION complex& operator*=(const RealType& v) noexcept {
    re_ *= v;
    im_ *= v;
    return *this;
  }
  KOKKOS_INLINE_FUNCTION complex& operator/=(const complex& z) noexcept {
    // Smith: divide by the larger of |c|, |d| first
    const RealType c = z.re_, d = z.im_;
    if (Kokkos::fabs(c) >= Kokkos::fabs(d)) {
      const RealType r = d / c, den = c + d * r;
      const RealType x = (re_ + im_ * r) / den, y = (im_ - re_ * r) / den;
      re_ = x;
      im_ = y;
    } else {
      const RealType r = c / d, den = c * r + d;
      const RealType x = (re_ * r + im_) / den, y = (im_ * r - re_) / den;
      re_ = x;
      im_ = y;
    }
    return *this;
  }
  KOKKOS_INLINE_FUNCTION complex& operator/=(const RealType& v) noexcept {
    re_ /= v;
    im_ /= v;
    return *this;
  }
};

// ------------------------------------------------------------------ arithmetic
#define KKMX_COMPLEX_BINARY(OP, CAT)                                                              \
  template <class T>                                                                              \
  KOKKOS_INLINE_FUNCTION complex<T> operator OP(const complex<T>& a, const complex<T>& b) {       \
    complex<T> r = a;                                                                             \
    r CAT b;                                                                                      \
    return r;                                                                                     \
  }                                                                                               \
  template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>             \
  KOKKOS_INLINE_FUNCTION complex<T> operator OP(const complex<T>& a, const S& b) {                \
    complex<T> r = a;                                                                             \
    r CAT T(b);                                                                                   \
    return r;                                                                                     \
  }                                                                                               \
  template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>             \
  KOKKOS_INLINE_FUNCTION complex<T> operator OP(const S& a, const complex<T>& b) {                \
    complex<T> r{static_cast<T>(a), T(0)};                                                        \
    r CAT b;                                                                                      \
    return r;                                                                                     \
  }
KKMX_COMPLEX_BINARY(+, +=)
KKMX_COMPLEX_BINARY(-, -=)
KKMX_COMPLEX_BINARY(*, *=)
KKMX_COMPLEX_BINARY(/, /=)
#undef KKMX_COMPLEX_BINARY

template <class T>
KOKKOS_INLINE_FUNCTION complex<T> operator+(const complex<T>& a) {
  return a;
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> operator-(const complex<T>& a) {
  return complex<T>(-a.real(), -a.imag());
}
template <class T>
KOKKOS_INLINE_FUNCTION bool operator==(const complex<T>& a, const complex<T>& b) {
  return a.real() == b.real() && a.imag() == b.imag();
}
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
KOKKOS_INLINE_FUNCTION bool operator==(const complex<T>& a, const S& b) {
  return a.real() == T(b) && a.imag() == T(0);
}
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
KOKKOS_INLINE_FUNCTION bool operator==(const S& a, const complex<T>& b) {
  return b == a;
}
template <class T>
KOKKOS_INLINE_FUNCTION bool operator!=(const complex<T>& a, const complex<T>& b) {
  return !(a == b);
}
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
KOKKOS_INLINE_FUNCTION bool operator!=(const complex<T>& a, const S& b) {
  return !(a == b);
}
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
KOKKOS_INLINE_FUNCTION bool operator!=(const S& a, const complex<T>& b) {
  return !(b == a);
}

// ------------------------------------------------------------------ functions
template <class T>
KOKKOS_INLINE_FUNCTION T real(const complex<T>& z) {
  return z.real();
}
template <class T>
KOKKOS_INLINE_FUNCTION T imag(const complex<T>& z) {
  return z.imag();
}
template <class T>
KOKKOS_INLINE_FUNCTION T abs(const complex<T>& z) {
  return Kokkos::hypot(z.real(), z.imag());
}
template <class T>
KOKKOS_INLINE_FUNCTION T arg(const complex<T>& z) {
  return Kokkos::atan2(z.imag(), z.real());
}
template <class T>
KOKKOS_INLINE_FUNCTION T norm(const complex<T>& z) {
  return z.real() * z.real() + z.imag() * z.imag();
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> conj(const complex<T>& z) {
  return complex<T>(z.real(), -z.imag());
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> polar(const T& r, const T& theta = T()) {
  return complex<T>(r * Kokkos::cos(theta), r * Kokkos::sin(theta));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> exp(const complex<T>& z) {
  return Kokkos::exp(z.real()) * complex<T>(Kokkos::cos(z.imag()), Kokkos::sin(z.imag()));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> log(const complex<T>& z) {
  return complex<T>(Kokkos::log(Kokkos::abs(z)), Kokkos::arg(z));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> log10(const complex<T>& z) {
  return Kokkos::log(z) / Kokkos::log(T(10));
}
// principal root: sqrt(r) * (cos(theta/2), sin(theta/2)) in half-angle form
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> sqrt(const complex<T>& z) {
  const T x = z.real(), y = z.imag();
  if (x == T(0) && y == T(0)) return complex<T>(T(0), y);
  const T r = Kokkos::abs(z);
  T a       = Kokkos::sqrt((r + Kokkos::fabs(x)) / T(2));
  if (x >= T(0)) return complex<T>(a, y / (T(2) * a));
  return complex<T>(Kokkos::fabs(y) / (T(2) * a), Kokkos::copysign(a, y));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> pow(const complex<T>& z, const complex<T>& w) {
  if (z.real() == T(0) && z.imag() == T(0)) return complex<T>(T(0), T(0));
  return Kokkos::exp(w * Kokkos::log(z));
}
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
KOKKOS_INLINE_FUNCTION complex<T> pow(const complex<T>& z, const S& e) {
  return Kokkos::pow(z, complex<T>(T(e)));
}
template <class T, class S, class = std::enable_if_t<std::is_arithmetic<S>::value>>
KOKKOS_INLINE_FUNCTION complex<T> pow(const S& b, const complex<T>& w) {
  return Kokkos::pow(complex<T>(T(b)), w);
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> sin(const complex<T>& z) {
  return complex<T>(Kokkos::sin(z.real()) * Kokkos::cosh(z.imag()), Kokkos::cos(z.real()) * Kokkos::sinh(z.imag()));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> cos(const complex<T>& z) {
  return complex<T>(Kokkos::cos(z.real()) * Kokkos::cosh(z.imag()), -Kokkos::sin(z.real()) * Kokkos::sinh(z.imag()));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> tan(const complex<T>& z) {
  return Kokkos::sin(z) / Kokkos::cos(z);
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> sinh(const complex<T>& z) {
  return complex<T>(Kokkos::sinh(z.real()) * Kokkos::cos(z.imag()), Kokkos::cosh(z.real()) * Kokkos::sin(z.imag()));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> cosh(const complex<T>& z) {
  return complex<T>(Kokkos::cosh(z.real()) * Kokkos::cos(z.imag()), Kokkos::sinh(z.real()) * Kokkos::sin(z.imag()));
}
template <class T>
KOKKOS_INLINE_FUNCTION complex<T> tanh(const complex<T>& z) {
  return Kokkos::sinh(z) / Kokkos::cosh(z);
}

template <class T>
std::ostream& operator<<(std::ostream& os, const complex<T>& z) {
  return os << "(" << z.real() << "," << z.imag() << ")";
}
template <class T>
std::istream& operator>>(std::istream& is, complex<T>& z) {
  std::complex<T> s;
  is >> s;
  z = complex<T>(s);
  return is;
}

}  // namespace Kokkos

#endif
