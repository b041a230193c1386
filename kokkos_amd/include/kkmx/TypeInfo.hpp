// Kokkos::Impl::TypeInfo<T>::name(): the spelling of a type as a
// compile-time string_view, used in test messages and tool labels.
//
// Reference interface: core/src/Kokkos_TypeInfo.hpp:33-80.
//
// The name is cut out of __PRETTY_FUNCTION__ for a function template
// instantiated with T: clang's text ends "[T = <type>]", gcc's
// "[with T = <type>; ...]".  The characters are copied into a std::array at
// compile time so the string_view outlives the function.
#ifndef KKMX_TYPE_INFO_HPP
#define KKMX_TYPE_INFO_HPP

#include <array>
#include <cstddef>
#include <string_view>

#define KOKKOS_ENABLE_IMPL_TYPEINFO

namespace Kokkos {
namespace Impl {

template <class T>
constexpr std::string_view type_info_signature() {
  return __PRETTY_FUNCTION__;
}

template <class T>
constexpr std::string_view type_info_slice() {
  constexpr std::string_view sig = type_info_signature<T>();
  constexpr std::string_view key = "T = ";
  constexpr size_t at            = sig.find(key);
  static_assert(at != std::string_view::npos, "unexpected __PRETTY_FUNCTION__ form");
  constexpr size_t first = at + key.size();
  constexpr size_t semi  = sig.find(';', first);
  constexpr size_t last  = semi != std::string_view::npos ? semi : sig.rfind(']');
  static_assert(last != std::string_view::npos && last > first, "unexpected __PRETTY_FUNCTION__ form");
  return sig.substr(first, last - first);
}

template <class T>
struct TypeInfoStorage {
  static constexpr size_t size = type_info_slice<T>().size();
  static constexpr std::array<char, size> chars = [] {
    std::array<char, size> a{};
    constexpr std::string_view s = type_info_slice<T>();
    for (size_t i = 0; i < size; ++i) a[i] = s[i];
    return a;
  }();
};

template <class T>
class TypeInfo {
 public:
  static constexpr std::string_view name() noexcept {
    return {TypeInfoStorage<T>::chars.data(), TypeInfoStorage<T>::size};
  }
};

}  // namespace Impl
}  // namespace Kokkos

#endif
