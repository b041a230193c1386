// Out-of-line parts of the spaces: allocation, initialisation, finalisation,
// fences and configuration printing for Kokkos::MI355X.
//
// Reference: core/src/HIP/Kokkos_HIP.cpp:47-116 (impl_initialize),
// core/src/HIP/Kokkos_HIP_Space.cpp:92-131 (allocation errors become
// RawMemoryAllocationFailure), core/src/impl/Kokkos_Core.cpp:1029-1053
// (initialize: spaces, then tools), core/src/HIP/Kokkos_HIP_Instance.cpp:87-133
// (print_configuration).
#ifndef KKMX_SPACES_IMPL_HPP
#define KKMX_SPACES_IMPL_HPP

#include <kkmx/ExecSpaceManager.hpp>
#include <kkmx/Tools.hpp>

namespace Kokkos {

inline void* MI355XSpace::allocate(const char* label, size_t n, size_t) const {
  if (n == 0) return nullptr;
  void* p = nullptr;
  int rc  = mx_malloc(&p, n, nullptr);
  if (rc != 0 || p == nullptr) {
    (void)mx_get_last_error();
    throw Experimental::RawMemoryAllocationFailure(
        n, std::string("Kokkos failed to allocate memory for label \"") + label +
               "\".  Allocation using MemorySpace named \"MI355XSpace\" failed with the following error: " +
               mx_error_string(rc));
  }
  if (Tools::profileLibraryLoaded()) Tools::allocateData(Tools::make_space_handle(name()), label, p, n);
  return p;
}
inline void MI355XSpace::deallocate(const char* label, void* p, size_t n, size_t) const {
  if (!p) return;
  if (Tools::profileLibraryLoaded()) Tools::deallocateData(Tools::make_space_handle(name()), label, p, n);
  KKMX_SAFE_CALL(mx_free(p, nullptr));
}
inline void* MI355XHostPinnedSpace::allocate(const char* label, size_t n, size_t) const {
  if (n == 0) return nullptr;
  void* p = nullptr;
  int rc  = mx_host_alloc(&p, n);
  if (rc != 0 || p == nullptr)
    throw Experimental::RawMemoryAllocationFailure(
        n, std::string("Kokkos failed to allocate pinned memory for label \"") + label + "\": " +
               mx_error_string(rc));
  return p;
}
inline void MI355XHostPinnedSpace::deallocate(const char*, void* p, size_t, size_t) const {
  if (p) KKMX_SAFE_CALL(mx_host_free(p));
}

inline void MI355X::fence(const std::string& name) const {
  if (Tools::profileLibraryLoaded()) {
    uint64_t h = 0;
    Tools::beginFence(name, Tools::Experimental::device_id(*this), &h);
    KKMX_SAFE_CALL(mx_stream_sync(m_inst->stream));
    Tools::endFence(h);
    return;
  }
  KKMX_SAFE_CALL(mx_stream_sync(m_inst->stream));
}
inline void MI355X::impl_static_fence(const std::string& name) {
  if (!Impl::mi355x_globals().default_instance) return;
  if (Tools::profileLibraryLoaded()) {
    uint64_t h = 0;
    Tools::beginFence(name, 0, &h);
    KKMX_SAFE_CALL(mx_device_sync());
    Tools::endFence(h);
    return;
  }
  KKMX_SAFE_CALL(mx_device_sync());
}
inline void MI355X::impl_initialize(const InitializationSettings& s) {
  auto& g = Impl::mi355x_globals();
  if (g.default_instance) return;
  const int dev = s.has_device_id() ? s.get_device_id() : 0;
  KKMX_SAFE_CALL(mx_init(dev));
  KKMX_SAFE_CALL(mx_device_props_get(&g.props));
  mx_stream_t st = nullptr;
  KKMX_SAFE_CALL(mx_default_stream(&st));
  g.default_instance = std::make_shared<Impl::MI355XInstance>(st, false, 1u);
}
inline void MI355X::impl_finalize() {
  auto& g = Impl::mi355x_globals();
  if (!g.default_instance) return;
  (void)mx_device_sync();
  g.default_instance.reset();
}
inline void MI355X::print_configuration(std::ostream& os, bool verbose) const {
  const auto& p = Impl::mi355x_globals().props;
  os << "Device Execution Space:\n  KOKKOS_ENABLE_MI355X: yes\n";
  if (!Impl::mi355x_globals().default_instance) {
    os << "  (not initialized)\n";
    return;
  }
  os << "  Device " << p.device << ": " << p.name << " (" << p.arch << ")\n";
  os << "  CUs: " << p.cu_count << ", wavefront: " << p.wave_size << " lanes, XCDs: " << p.xcc_count << "\n";
  os << "  LDS per workgroup: " << p.lds_per_block << " B, LDS per CU: " << p.lds_per_cu << " B\n";
  os << "  L2: " << p.l2_bytes << " B, HBM: " << p.global_mem << " B, bus width " << p.mem_bus_width
     << " bit, memory clock " << p.mem_clock_khz << " kHz\n";
  if (verbose) os << "  max threads per block: " << p.max_threads_per_block << "\n";
}

namespace Impl {
inline ExecSpaceManager::ExecSpaceManager() {
  m_spaces["100_Serial"] = make_exec_space_hooks<Serial>();
  m_spaces["150_MI355X"] = make_exec_space_hooks<MI355X>();
}
}  // namespace Impl

inline void fence(const std::string& name) { Impl::ExecSpaceManager::get_instance().static_fence(name); }

inline void initialize(const InitializationSettings& settings_in) {
  auto& g = Impl::global_state();
  if (g.state.load() == 1) return;
  if (g.state.load() == 2) Kokkos::abort("Error: Kokkos::initialize() has already been called. Kokkos can be initialized at most once.\n");
  InitializationSettings settings = settings_in;
  Impl::parse_kokkos_env(settings);
  g.tune_internals   = settings.has_tune_internals() && settings.get_tune_internals();
  g.disable_warnings = settings.has_disable_warnings() && settings.get_disable_warnings();
  Impl::ExecSpaceManager::get_instance().initialize_spaces(settings);
  g.state.store(1);
  Tools::Impl::tools_start(settings.has_tools_libs() ? settings.get_tools_libs() : std::string(),
                           settings.has_tools_args() ? settings.get_tools_args() : std::string());
  if (settings.has_print_configuration() && settings.get_print_configuration())
    print_configuration(std::cout, false);
}

inline void finalize() {
  auto& g = Impl::global_state();
  if (g.state.load() != 1) Kokkos::abort("Kokkos::finalize: Kokkos was not initialized or was already finalized");
  for (auto it = g.finalize_hooks.rbegin(); it != g.finalize_hooks.rend(); ++it) (*it)();
  g.finalize_hooks.clear();
  Tools::Impl::tools_stop();
  Impl::ExecSpaceManager::get_instance().finalize_spaces();
  g.state.store(2);
}

inline void print_configuration(std::ostream& os, bool verbose) {
  os << "Kokkos Version: 4.6.99 (MI355X execution space)\n";
  Impl::ExecSpaceManager::get_instance().print_configuration(os, verbose);
}

}  // namespace Kokkos

#endif
