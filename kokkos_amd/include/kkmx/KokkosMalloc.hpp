// Kokkos::kokkos_malloc / kokkos_free / kokkos_realloc (reference
// core/src/Kokkos_Core.hpp:174-207): raw allocations from a memory space,
// remembered with their size and label so that free and realloc need only
// the pointer.  Allocation and release go through the space's allocate /
// deallocate, so profiling tools see them like View allocations.
#ifndef KKMX_KOKKOS_MALLOC_HPP
#define KKMX_KOKKOS_MALLOC_HPP

#include <kkmx/CopyViews.hpp>

#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>

namespace Kokkos {
namespace Impl {

struct RawAllocations {
  std::mutex mtx;
  std::unordered_map<void*, std::pair<size_t, std::string>> live;
};
inline RawAllocations& raw_allocations() {
  static RawAllocations r;
  return r;
}
inline void check_raw_alloc_state(const char* what) {
  if (!is_initialized())
    Kokkos::abort((std::string(what) + " called outside of Kokkos::initialize / Kokkos::finalize").c_str());
}

}  // namespace Impl

template <class Space = DefaultExecutionSpace::memory_space>
inline void* kokkos_malloc(const std::string& label, const size_t bytes) {
  Impl::check_raw_alloc_state("kokkos_malloc");
  using Mem = typename Space::memory_space;
  void* p   = Mem().allocate(label.c_str(), bytes);
  if (p) {
    auto& r = Impl::raw_allocations();
    std::lock_guard<std::mutex> lock(r.mtx);
    r.live[p] = {bytes, label};
  }
  return p;
}
template <class Space = DefaultExecutionSpace::memory_space>
inline void* kokkos_malloc(const size_t bytes) {
  return kokkos_malloc<Space>("no-label", bytes);
}

template <class Space = DefaultExecutionSpace::memory_space>
inline void kokkos_free(void* p) {
  Impl::check_raw_alloc_state("kokkos_free");
  if (!p) return;
  using Mem = typename Space::memory_space;
  std::pair<size_t, std::string> rec;
  {
    auto& r = Impl::raw_allocations();
    std::lock_guard<std::mutex> lock(r.mtx);
    auto it = r.live.find(p);
    if (it == r.live.end()) Kokkos::abort("kokkos_free: the pointer was not allocated by kokkos_malloc");
    rec = it->second;
    r.live.erase(it);
  }
  Mem().deallocate(rec.second.c_str(), p, rec.first);
}

template <class Space = DefaultExecutionSpace::memory_space>
inline void* kokkos_realloc(void* p, const size_t bytes) {
  Impl::check_raw_alloc_state("kokkos_realloc");
  if (!p) return kokkos_malloc<Space>(bytes);
  std::pair<size_t, std::string> rec;
  {
    auto& r = Impl::raw_allocations();
    std::lock_guard<std::mutex> lock(r.mtx);
    auto it = r.live.find(p);
    if (it == r.live.end()) Kokkos::abort("kokkos_realloc: the pointer was not allocated by kokkos_malloc");
    rec = it->second;
  }
  void* q            = kokkos_malloc<Space>(rec.second, bytes);
  const size_t keep  = rec.first < bytes ? rec.first : bytes;
  using Mem          = typename Space::memory_space;
  if (keep) {
    if constexpr (Impl::host_accessible<Mem>()) {
      Kokkos::fence("Kokkos::kokkos_realloc: before the host copy");
      std::memcpy(q, p, keep);
    } else {
      KKMX_SAFE_CALL(mx_memcpy(q, p, keep));
    }
  }
  kokkos_free<Space>(p);
  return q;
}

}  // namespace Kokkos

#endif
