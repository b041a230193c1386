// Memory spaces, execution spaces, initialisation and fences.
//
// Kokkos::MI355X is the new execution space.  It satisfies the execution-space
// concept the reference checks in core/src/impl/Kokkos_ExecSpaceManager.hpp:28-107
// (regular type, sizeof <= 2 pointers, execution_space / memory_space /
// device_type / array_layout / size_type / scratch_memory_space typedefs,
// print_configuration, impl_initialize, impl_finalize, fence,
// impl_static_fence, concurrency, name).  Each instance owns a HIP stream
// (through the C-ABI runtime) plus grow-only device scratch for reductions and
// scans, like HIPInternal (core/src/HIP/Kokkos_HIP_Instance.hpp:158-353), but
// the scratch here is self-cleaning: every kernel leaves its counters at zero,
// so no memset runs between dispatches and dispatches can be graph-captured.
//
// Kokkos::Serial is a minimal host execution space (the reference's
// core/src/Serial/), used for host mirrors and host-side parallel loops.
#ifndef KKMX_SPACES_HPP
#define KKMX_SPACES_HPP

#include <kkmx/Core_fwd.hpp>

#include <atomic>
#include <cstring>
#include <iostream>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

namespace Kokkos {

struct LayoutLeft;
struct LayoutRight;
template <class ExecSpace>
class ScratchMemorySpace;

// ------------------------------------------------------------------ settings
// Same option names as core/src/impl/Kokkos_InitializationSettings.hpp:66-76.
class InitializationSettings {
 public:
  InitializationSettings& set_num_threads(int v) { m_num_threads = v; m_has_num_threads = true; return *this; }
  bool has_num_threads() const { return m_has_num_threads; }
  int get_num_threads() const { return m_num_threads; }
  InitializationSettings& set_device_id(int v) { m_device_id = v; m_has_device_id = true; return *this; }
  bool has_device_id() const { return m_has_device_id; }
  int get_device_id() const { return m_device_id; }
  InitializationSettings& set_disable_warnings(bool v) { m_disable_warnings = v; m_has_disable_warnings = true; return *this; }
  bool has_disable_warnings() const { return m_has_disable_warnings; }
  bool get_disable_warnings() const { return m_disable_warnings; }
  InitializationSettings& set_print_configuration(bool v) { m_print_configuration = v; m_has_print_configuration = true; return *this; }
  bool has_print_configuration() const { return m_has_print_configuration; }
  bool get_print_configuration() const { return m_print_configuration; }
  InitializationSettings& set_tune_internals(bool v) { m_tune_internals = v; m_has_tune_internals = true; return *this; }
  bool has_tune_internals() const { return m_has_tune_internals; }
  bool get_tune_internals() const { return m_tune_internals; }
  InitializationSettings& set_tools_libs(std::string v) { m_tools_libs = std::move(v); m_has_tools_libs = true; return *this; }
  bool has_tools_libs() const { return m_has_tools_libs; }
  std::string get_tools_libs() const { return m_tools_libs; }
  InitializationSettings& set_tools_args(std::string v) { m_tools_args = std::move(v); m_has_tools_args = true; return *this; }
  bool has_tools_args() const { return m_has_tools_args; }
  std::string get_tools_args() const { return m_tools_args; }

 private:
  int m_num_threads = 1, m_device_id = 0;
  bool m_disable_warnings = false, m_print_configuration = false, m_tune_internals = false;
  std::string m_tools_libs, m_tools_args;
  bool m_has_num_threads = false, m_has_device_id = false, m_has_disable_warnings = false,
       m_has_print_configuration = false, m_has_tune_internals = false,
       m_has_tools_libs = false, m_has_tools_args = false;
};

namespace Impl {
struct GlobalState {
  std::atomic<int> state{0};  // 0 = never, 1 = initialised, 2 = finalised
  bool tune_internals = false;
  bool disable_warnings = false;
  int device_id         = 0;
  std::vector<void (*)()> finalize_hooks;
};
inline GlobalState& global_state() {
  static GlobalState g;
  return g;
}
}  // namespace Impl

inline bool is_initialized() noexcept { return Impl::global_state().state.load() == 1; }
inline bool is_finalized() noexcept { return Impl::global_state().state.load() == 2; }
inline void push_finalize_hook(void (*f)()) { Impl::global_state().finalize_hooks.push_back(f); }
namespace Impl {
inline bool& tune_internals() noexcept { return global_state().tune_internals; }
}  // namespace Impl

// ------------------------------------------------------------------ HostSpace
class HostSpace {
 public:
  using memory_space    = HostSpace;
  using execution_space = Serial;
  using size_type       = size_t;
  using device_type     = Device<execution_space, memory_space>;

  HostSpace() = default;
  void* allocate(size_t n) const { return allocate("[unlabeled]", n); }
  void* allocate(const char* label, size_t n, size_t = 0) const {
    if (n == 0) return nullptr;
    void* p = ::operator new(n, std::align_val_t(64), std::nothrow);
    if (!p)
      throw Experimental::RawMemoryAllocationFailure(
          n, std::string("Kokkos failed to allocate memory for label \"") +
                 label + "\".  Allocation using MemorySpace named \"Host\" failed with the following error: Allocation of size " +
                 std::to_string(n) + " B failed: out of memory");
    return p;
  }
  void deallocate(void* p, size_t n) const { deallocate("[unlabeled]", p, n); }
  void deallocate(const char*, void* p, size_t, size_t = 0) const {
    if (p) ::operator delete(p, std::align_val_t(64));
  }
  static constexpr const char* name() { return "Host"; }
};

// ------------------------------------------------------------------ device spaces
class MI355XSpace {
 public:
  using memory_space    = MI355XSpace;
  using execution_space = MI355X;
  using size_type       = unsigned int;
  using device_type     = Device<execution_space, memory_space>;

  MI355XSpace() = default;
  void* allocate(size_t n) const { return allocate("[unlabeled]", n); }
  void* allocate(const char* label, size_t n, size_t = 0) const;
  void deallocate(void* p, size_t n) const { deallocate("[unlabeled]", p, n); }
  void deallocate(const char*, void* p, size_t, size_t = 0) const;
  static constexpr const char* name() { return "MI355XSpace"; }
};

class MI355XHostPinnedSpace {
 public:
  using memory_space    = MI355XHostPinnedSpace;
  using execution_space = Serial;
  using size_type       = size_t;
  using device_type     = Device<execution_space, memory_space>;

  void* allocate(size_t n) const { return allocate("[unlabeled]", n); }
  void* allocate(const char* label, size_t n, size_t = 0) const;
  void deallocate(void* p, size_t n) const { deallocate("[unlabeled]", p, n); }
  void deallocate(const char*, void* p, size_t, size_t = 0) const;
  static constexpr const char* name() { return "MI355XHostPinned"; }
};

// ------------------------------------------------------------------ Serial
// Tag asking for an execution space instance of its own
// (core/src/Serial/Kokkos_Serial.hpp NewInstance).
struct NewInstance {
  explicit NewInstance() = default;
};

class Serial {
 public:
  using execution_space      = Serial;
  using memory_space         = HostSpace;
  using device_type          = Device<execution_space, memory_space>;
  using array_layout         = LayoutRight;
  using size_type            = memory_space::size_type;
  using scratch_memory_space = ScratchMemorySpace<Serial>;

  Serial() = default;
  // Serial work runs inline on the calling thread, so a new instance
  // behaves as the default one.
  explicit Serial(NewInstance) {}
  static constexpr const char* name() { return "Serial"; }
  static int concurrency() { return 1; }
  KOKKOS_INLINE_FUNCTION static unsigned impl_hardware_thread_id() { return 0; }
  KOKKOS_INLINE_FUNCTION static unsigned impl_max_hardware_threads() { return 1; }
  void fence(const std::string& = "") const {}
  static void impl_static_fence(const std::string& = "") {}
  void print_configuration(std::ostream& os, bool = false) const {
    os << "Host Serial Execution Space:\n  KOKKOS_ENABLE_SERIAL: yes\n";
  }
  static void impl_initialize(const InitializationSettings&) {}
  static void impl_finalize() {}
  uint32_t impl_instance_id() const noexcept { return 1; }
  bool operator==(const Serial&) const { return true; }
  bool operator!=(const Serial&) const { return false; }
  bool in_parallel() const { return false; }
};

// ------------------------------------------------------------------ MI355X
namespace Impl {

// Per-instance state: stream plus grow-only scratch.  The device control
// block holds, per instance: [0] last-block counter of reductions,
// [2..3] one 64-bit scan word {tile counter (low half), epoch (high half)},
// the rest spare.  Kernels return the counters to zero and advance the epoch
// themselves.
struct MI355XInstance {
  mx_stream_t stream = nullptr;
  bool owns_stream   = false;
  uint32_t id        = 0;
  std::mutex mtx;

  unsigned* control       = nullptr;  // device, 64 x uint32
  void* partials          = nullptr;  // device, reduction partials
  size_t partials_bytes   = 0;
  void* scan_status       = nullptr;  // device, scan tile status words
  size_t scan_status_bytes = 0;
  void* scan_values       = nullptr;  // device, scan tile values (generic path)
  void* scan_flags        = nullptr;  // device, scan tile flags (generic path)
  size_t scan_flags_bytes = 0;
  size_t scan_values_bytes = 0;
  void* pinned            = nullptr;  // host-mapped result staging
  size_t pinned_bytes     = 0;
  void* kernarg_stage     = nullptr;  // device copy of large closures
  size_t kernarg_stage_bytes = 0;
  void* team_l1           = nullptr;  // level-1 team scratch pool
  size_t team_l1_bytes    = 0;
  // Scan launches made against scan_status / scan_flags since they were last
  // zeroed (tile status words keep only 30 bits of the launch epoch).
  uint64_t scan_since_clear = 0;

  explicit MI355XInstance(mx_stream_t s, bool owns, uint32_t ident)
      : stream(s), owns_stream(owns), id(ident) {
    void* p = nullptr;
    KKMX_SAFE_CALL(mx_malloc(&p, 64 * sizeof(unsigned), nullptr));
    control = static_cast<unsigned*>(p);
    unsigned init[64] = {};
    init[3]           = 1;  // scan epochs start at 1 so zeroed status words are invalid
    KKMX_SAFE_CALL(mx_memcpy(control, init, sizeof(init)));
  }
  ~MI355XInstance() {
    if (stream) (void)mx_stream_sync(stream);
    (void)mx_free(control, nullptr);
    (void)mx_free(partials, nullptr);
    (void)mx_free(scan_status, nullptr);
    (void)mx_free(scan_values, nullptr);
    (void)mx_free(scan_flags, nullptr);
    (void)mx_free(kernarg_stage, nullptr);
    (void)mx_free(team_l1, nullptr);
    (void)mx_host_free(pinned);
    if (owns_stream) (void)mx_stream_destroy(stream);
  }
  MI355XInstance(const MI355XInstance&) = delete;
  MI355XInstance& operator=(const MI355XInstance&) = delete;

  // Grow-only device buffers; the old buffer is released in stream order so
  // kernels still queued on this stream finish with it first.
  void* grow(void*& buf, size_t& cur, size_t need, bool zero) {
    if (need <= cur) return buf;
    size_t n = cur ? cur : 4096;
    while (n < need) n *= 2;
    if (buf) KKMX_SAFE_CALL(mx_free(buf, stream));
    void* p = nullptr;
    KKMX_SAFE_CALL(mx_malloc(&p, n, stream));
    if (zero) KKMX_SAFE_CALL(mx_memset_async(p, 0, n, stream));
    buf = p;
    cur = n;
    return buf;
  }
  void* partials_space(size_t need) { return grow(partials, partials_bytes, need, false); }
  void* scan_status_space(size_t need) { return grow(scan_status, scan_status_bytes, need, true); }
  void* scan_values_space(size_t need) { return grow(scan_values, scan_values_bytes, need, false); }
  // The generic scan state keeps 32-bit flags apart from the 64-bit
  // {flag, value} words of the 4- and 8-byte states: a flag word read where
  // another layout left a value could otherwise pass for a current flag.
  void* scan_flags_space(size_t need) { return grow(scan_flags, scan_flags_bytes, need, true); }
  void* kernarg_space(size_t need) { return grow(kernarg_stage, kernarg_stage_bytes, need, false); }
  void* team_l1_space(size_t need) { return grow(team_l1, team_l1_bytes, need, false); }
  void* pinned_space(size_t need) {
    if (need <= pinned_bytes) return pinned;
    size_t n = pinned_bytes ? pinned_bytes : 4096;
    while (n < need) n *= 2;
    if (pinned) {
      KKMX_SAFE_CALL(mx_stream_sync(stream));
      KKMX_SAFE_CALL(mx_host_free(pinned));
    }
    KKMX_SAFE_CALL(mx_host_alloc(&pinned, n));
    pinned_bytes = n;
    return pinned;
  }
};

struct MI355XGlobals {
  std::shared_ptr<MI355XInstance> default_instance;
  mx_device_props props{};
  std::atomic<uint32_t> next_id{2};
};
inline MI355XGlobals& mi355x_globals() {
  static MI355XGlobals g;
  return g;
}

}  // namespace Impl

class MI355X {
 public:
  using execution_space      = MI355X;
  using memory_space         = MI355XSpace;
  using device_type          = Device<execution_space, memory_space>;
  using array_layout         = LayoutLeft;
  using size_type            = MI355XSpace::size_type;
  using scratch_memory_space = ScratchMemorySpace<MI355X>;

  // Default instance: the stream created at initialisation.
  MI355X() : m_inst(Impl::mi355x_globals().default_instance) {
    if (!m_inst)
      Kokkos::abort(
          "Kokkos::MI355X: the execution space was used before "
          "Kokkos::initialize or after Kokkos::finalize");
  }
  // Instance on a caller-provided stream (not owned), like HIP(hipStream_t).
  explicit MI355X(mx_stream_t stream, bool manage_stream = false)
      : m_inst(std::make_shared<Impl::MI355XInstance>(
            stream, manage_stream, Impl::mi355x_globals().next_id++)) {}

  // Instance wrapping existing per-instance state (graph nodes own one each).
  explicit MI355X(std::shared_ptr<Impl::MI355XInstance> inst) : m_inst(std::move(inst)) {}

  static constexpr const char* name() { return "MI355X"; }
  int concurrency() const {
    // Resident lanes: 8 wavefronts of 64 lanes per SIMD, 4 SIMDs per CU
    // (the reference uses 32 waves x wave size x CUs, Kokkos_HIP_Instance.cpp:80-85).
    return 32 * 64 * Impl::mi355x_globals().props.cu_count;
  }
  void fence(const std::string& name = "Kokkos::MI355X::fence: Unnamed Instance Fence") const;
  static void impl_static_fence(const std::string& name = "Kokkos::MI355X::fence: Unnamed Static Fence");
  void print_configuration(std::ostream& os, bool verbose = false) const;
  static void impl_initialize(const InitializationSettings& settings);
  static void impl_finalize();
  bool operator==(const MI355X& o) const { return m_inst == o.m_inst; }
  bool operator!=(const MI355X& o) const { return m_inst != o.m_inst; }
  mx_stream_t stream() const { return m_inst->stream; }
  Impl::MI355XInstance* impl_internal_space_instance() const { return m_inst.get(); }
  uint32_t impl_instance_id() const noexcept { return m_inst->id; }
  static int impl_cu_count() { return Impl::mi355x_globals().props.cu_count; }
  static const mx_device_props& impl_props() { return Impl::mi355x_globals().props; }
  // Slot of the calling lane among all lanes the GPU can hold at once: the
  // wavefront's hardware position (XCD from the XCC_ID register; shader
  // engine, shader array, CU, SIMD and wave slot from HW_ID) times 64 plus
  // the lane.  Unique among lanes that run at the same time and below
  // impl_max_hardware_threads(); 0 on the host.
  KOKKOS_INLINE_FUNCTION static unsigned impl_hardware_thread_id() {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned hw   = __builtin_amdgcn_s_getreg((31 << 11) | 4);         // HW_ID, bits 31:0
    const unsigned xcc  = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u;  // XCC_ID, bits 3:0
    const unsigned wave = hw & 15u, simd = (hw >> 4) & 3u, cu = (hw >> 8) & 15u, sh = (hw >> 12) & 1u,
                   se = (hw >> 13) & 7u;
    const unsigned slot = ((((xcc * 8u + se) * 2u + sh) * 16u + cu) * 4u + simd) * 16u + wave;
    return slot * 64u + static_cast<unsigned>(__lane_id());
#else
    return 0;
#endif
  }
  KOKKOS_INLINE_FUNCTION static unsigned impl_max_hardware_threads() { return 16u * 8u * 2u * 16u * 4u * 16u * 64u; }
  bool in_parallel() const { return false; }

 private:
  std::shared_ptr<Impl::MI355XInstance> m_inst;
};
static_assert(sizeof(MI355X) <= 2 * sizeof(void*), "execution space must stay small");

// ------------------------------------------------------------------ accessibility
template <class DstMem, class SrcMem>
struct MemorySpaceAccess {
  enum : bool { assignable = std::is_same<DstMem, SrcMem>::value };
  enum : bool { accessible = std::is_same<DstMem, SrcMem>::value };
  enum : bool { deepcopy = true };
};
template <>
struct MemorySpaceAccess<HostSpace, MI355XHostPinnedSpace> {
  enum : bool { assignable = false, accessible = true, deepcopy = true };
};
template <>
struct MemorySpaceAccess<MI355XSpace, MI355XHostPinnedSpace> {
  enum : bool { assignable = false, accessible = true, deepcopy = true };
};

namespace Impl {
template <class T>
struct as_memory_space {
  using type = typename T::memory_space;
};
// Whether code running in execution space Exec may dereference Mem.
template <class Exec, class Mem>
struct exec_can_access : std::false_type {};
template <>
struct exec_can_access<MI355X, MI355XSpace> : std::true_type {};
template <>
struct exec_can_access<MI355X, MI355XHostPinnedSpace> : std::true_type {};
template <>
struct exec_can_access<Serial, HostSpace> : std::true_type {};
template <>
struct exec_can_access<Serial, MI355XHostPinnedSpace> : std::true_type {};
template <class Mem>
struct exec_can_access<HostSpace, Mem> : exec_can_access<Serial, Mem> {};
template <class Mem>
struct exec_can_access<MI355XSpace, Mem> : exec_can_access<MI355X, Mem> {};
}  // namespace Impl

template <class AccessSpace, class MemorySpace>
struct SpaceAccessibility {
  using exec = std::conditional_t<is_execution_space<AccessSpace>::value,
                                  AccessSpace,
                                  typename AccessSpace::execution_space>;
  using mem  = typename MemorySpace::memory_space;
  enum : bool {
    accessible = Impl::exec_can_access<exec, mem>::value ||
                 (std::is_same<typename AccessSpace::memory_space, mem>::value)
  };
  enum : bool { assignable = std::is_same<typename AccessSpace::memory_space, mem>::value };
  using space = std::conditional_t<accessible, typename AccessSpace::memory_space, HostSpace>;
};

// ------------------------------------------------------------------ init / fence
namespace Tools {
namespace Impl {
void tools_initialize(const InitializationSettings&);
void tools_finalize();
}  // namespace Impl
}  // namespace Tools

// Fences every registered execution space (defined with the registry).
inline void fence(const std::string& name = "Kokkos::fence: Unnamed Global Fence");

namespace Impl {
inline void parse_kokkos_args(int& argc, char* argv[], InitializationSettings& s) {
  auto starts = [](const char* a, const char* p) { return std::strncmp(a, p, std::strlen(p)) == 0; };
  int w = 1;
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    bool consumed = true;
    if (starts(a, "--kokkos-device-id=")) s.set_device_id(std::atoi(a + 19));
    else if (starts(a, "--kokkos-num-threads=")) s.set_num_threads(std::atoi(a + 21));
    else if (!std::strcmp(a, "--kokkos-tune-internals")) s.set_tune_internals(true);
    else if (starts(a, "--kokkos-tools-libs=")) s.set_tools_libs(a + 20);
    else if (starts(a, "--kokkos-tools-args=")) s.set_tools_args(a + 20);
    else if (!std::strcmp(a, "--kokkos-print-configuration")) s.set_print_configuration(true);
    else if (!std::strcmp(a, "--kokkos-disable-warnings")) s.set_disable_warnings(true);
    else consumed = false;
    if (!consumed) argv[w++] = argv[i];
  }
  argc = w;
}
inline void parse_kokkos_env(InitializationSettings& s) {
  if (const char* v = std::getenv("KOKKOS_DEVICE_ID"); v && !s.has_device_id()) s.set_device_id(std::atoi(v));
  if (const char* v = std::getenv("KOKKOS_TUNE_INTERNALS"); v && !s.has_tune_internals())
    s.set_tune_internals(std::atoi(v) != 0 || !std::strcmp(v, "true") || !std::strcmp(v, "ON"));
  if (const char* v = std::getenv("KOKKOS_TOOLS_LIBS"); v && !s.has_tools_libs()) s.set_tools_libs(v);
  if (const char* v = std::getenv("KOKKOS_TOOLS_ARGS"); v && !s.has_tools_args()) s.set_tools_args(v);
  if (const char* v = std::getenv("KOKKOS_PRINT_CONFIGURATION"); v && !s.has_print_configuration())
    s.set_print_configuration(std::atoi(v) != 0);
}
}  // namespace Impl

void initialize(const InitializationSettings& settings = InitializationSettings());
inline void initialize(int& argc, char* argv[]) {
  InitializationSettings s;
  Impl::parse_kokkos_args(argc, argv, s);
  initialize(s);
}
void finalize();
void print_configuration(std::ostream& os, bool verbose = false);

class ScopeGuard {
 public:
  ScopeGuard(int& argc, char* argv[]) { initialize(argc, argv); }
  explicit ScopeGuard(const InitializationSettings& s = InitializationSettings()) { initialize(s); }
  ~ScopeGuard() { finalize(); }
  ScopeGuard(const ScopeGuard&) = delete;
  ScopeGuard& operator=(const ScopeGuard&) = delete;
};

}  // namespace Kokkos

#endif
