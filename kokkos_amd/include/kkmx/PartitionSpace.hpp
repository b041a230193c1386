// Kokkos::Experimental::partition_space (reference core/src/Kokkos_Core.hpp
// :290-346, HIP backend HIP/Kokkos_HIP_Instance.hpp:355-375): one new
// instance per weight, each on a stream of its own that the instance owns
// (default priority), so work submitted to different partitions may run
// concurrently; the weights are accepted and, as in the reference's GPU
// backends, do not limit a partition's share of the device.
#ifndef KKMX_PARTITION_SPACE_HPP
#define KKMX_PARTITION_SPACE_HPP

#include <kkmx/Spaces.hpp>

#include <array>
#include <type_traits>
#include <vector>

namespace Kokkos {
namespace Impl {

template <class ExecSpace, class T>
std::vector<ExecSpace> impl_partition_space(const ExecSpace& base, const std::vector<T>& weights) {
  std::vector<ExecSpace> out;
  out.reserve(weights.size());
  for (size_t i = 0; i < weights.size(); ++i) {
    if constexpr (std::is_same<ExecSpace, MI355X>::value) {
      mx_stream_t s = nullptr;
      KKMX_SAFE_CALL(mx_stream_create(&s));
      out.push_back(MI355X(s, true));
    } else {
      out.push_back(base);
    }
  }
  return out;
}

}  // namespace Impl

namespace Experimental {

template <class ExecSpace, class... Args>
std::array<ExecSpace, sizeof...(Args)> partition_space(const ExecSpace& base, Args... args) {
  static_assert(is_execution_space<ExecSpace>::value,
                "Kokkos Error: partition_space expects an Execution Space as first argument");
  static_assert((... && std::is_arithmetic<Args>::value),
                "Kokkos Error: partitioning arguments must be integers or floats");
  std::vector<std::common_type_t<Args...>> weights = {args...};
  auto v = Kokkos::Impl::impl_partition_space(base, weights);
  std::array<ExecSpace, sizeof...(Args)> out;
  for (size_t i = 0; i < out.size(); ++i) out[i] = v[i];
  return out;
}

template <class ExecSpace, class T>
std::vector<ExecSpace> partition_space(const ExecSpace& base, const std::vector<T>& weights) {
  static_assert(is_execution_space<ExecSpace>::value,
                "Kokkos Error: partition_space expects an Execution Space as first argument");
  static_assert(std::is_arithmetic<T>::value, "Kokkos Error: partitioning arguments must be integers or floats");
  return Kokkos::Impl::impl_partition_space(base, weights);
}

}  // namespace Experimental
}  // namespace Kokkos

#endif
