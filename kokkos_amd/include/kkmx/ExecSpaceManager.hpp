// Registry of execution spaces that Kokkos::initialize, finalize, fence and
// print_configuration walk, and the registration call every backend makes.
//
// Reference: core/src/impl/Kokkos_ExecSpaceManager.hpp:28-107 (what a
// registered space must provide), :157-163 (initialize_space_factory), and
// core/src/impl/Kokkos_Core.cpp:216-252 (spaces are handled in the order of
// their keys, e.g. "100_Serial" before "150_HIP").
//
// A space is stored as four plain function pointers bound to its static
// members; keys sort the spaces, and registering a key again replaces the
// entry.  Serial ("100_Serial") and MI355X ("150_MI355X") are entered when the
// registry is first used, so a registration from another translation unit's
// static initialiser never races the built-in spaces.
#ifndef KKMX_EXEC_SPACE_MANAGER_HPP
#define KKMX_EXEC_SPACE_MANAGER_HPP

#include <map>
#include <mutex>
#include <ostream>
#include <string>
#include <type_traits>

namespace Kokkos {
namespace Impl {

struct ExecSpaceHooks {
  void (*initialize)(const InitializationSettings&) = nullptr;
  void (*finalize)()                                 = nullptr;
  void (*static_fence)(const std::string&)           = nullptr;
  void (*print_configuration)(std::ostream&, bool)   = nullptr;
};

template <class Space>
ExecSpaceHooks make_exec_space_hooks() {
  ExecSpaceHooks h;
  h.initialize          = [](const InitializationSettings& s) { Space::impl_initialize(s); };
  h.finalize            = [] { Space::impl_finalize(); };
  h.static_fence        = [](const std::string& label) { Space::impl_static_fence(label); };
  h.print_configuration = [](std::ostream& os, bool verbose) { Space().print_configuration(os, verbose); };
  return h;
}

// The compile-time part of the contract a registered space satisfies.
template <class Space>
constexpr bool check_exec_space_contract() {
  static_assert(std::is_default_constructible<Space>::value, "an execution space is default-constructible");
  static_assert(std::is_copy_constructible<Space>::value && std::is_copy_assignable<Space>::value,
                "an execution space is copyable");
  static_assert(std::is_same<typename Space::execution_space, Space>::value,
                "Space::execution_space names the space itself");
  static_assert(sizeof(typename Space::memory_space) > 0 && sizeof(typename Space::device_type) > 0 &&
                    sizeof(typename Space::array_layout) > 0 && sizeof(typename Space::size_type) > 0 &&
                    sizeof(typename Space::scratch_memory_space) > 0,
                "an execution space names its memory space, device type, layout, size type and scratch space");
  static_assert(std::is_convertible<decltype(std::declval<const Space&>() == std::declval<const Space&>()),
                                    bool>::value,
                "execution space instances compare with ==");
  static_assert(std::is_convertible<decltype(std::declval<const Space&>().concurrency()), int>::value,
                "an execution space reports its concurrency");
  static_assert(sizeof(Space) <= 2 * sizeof(void*), "an execution space instance is a handle");
  return true;
}

class ExecSpaceManager {
 public:
  static ExecSpaceManager& get_instance() {
    static ExecSpaceManager m;
    return m;
  }
  void register_space_factory(const std::string& key, const ExecSpaceHooks& h) {
    std::lock_guard<std::mutex> lk(m_mtx);
    m_spaces[key] = h;
  }
  void initialize_spaces(const InitializationSettings& settings) {
    for (auto& e : snapshot()) e.second.initialize(settings);
  }
  void finalize_spaces() {
    for (auto& e : snapshot()) e.second.finalize();
  }
  void static_fence(const std::string& label) {
    std::lock_guard<std::mutex> lk(m_mtx);
    for (auto& e : m_spaces) e.second.static_fence(label);
  }
  void print_configuration(std::ostream& os, bool verbose) {
    for (auto& e : snapshot()) e.second.print_configuration(os, verbose);
  }
  bool is_registered(const std::string& key) {
    std::lock_guard<std::mutex> lk(m_mtx);
    return m_spaces.count(key) != 0;
  }

 private:
  ExecSpaceManager();
  // A copy, so a space's hook may register another space without deadlock.
  std::map<std::string, ExecSpaceHooks> snapshot() {
    std::lock_guard<std::mutex> lk(m_mtx);
    return m_spaces;
  }
  std::mutex m_mtx;
  std::map<std::string, ExecSpaceHooks> m_spaces;
};

// Registers Space under `key` (its position among the spaces) and returns 1,
// so a backend can register from a namespace-scope initialiser:
//   int registered = Kokkos::Impl::initialize_space_factory<MySpace>("160_MySpace");
template <class Space>
int initialize_space_factory(const std::string& key) {
  static_assert(check_exec_space_contract<Space>());
  ExecSpaceManager::get_instance().register_space_factory(key, make_exec_space_hooks<Space>());
  return 1;
}

}  // namespace Impl
}  // namespace Kokkos

#endif
