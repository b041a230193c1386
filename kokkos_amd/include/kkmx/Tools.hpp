// Kokkos Tools: profiling callbacks and the tuning interface.
//
// Behaviour follows core/src/impl/Kokkos_Profiling.cpp of the reference:
//  * callbacks are either found with dlsym("kokkosp_*") in the libraries named
//    by KOKKOS_TOOLS_LIBS / --kokkos-tools-libs (:596-718), or installed in
//    process with set_*_callback (:969-996);
//  * kernel begin/end callbacks may fence the device first unless the tool
//    declares requires_global_fencing = false through request_tool_settings
//    (:277-295);
//  * with tune_internals on, every begin_parallel_* opens a tuning context
//    carrying kernel_name and kernel_type input values and every end closes it
//    (:308-470);
//  * declare_{input,output}_type hand out increasing variable ids,
//    set_input_values / request_output_values / begin_context / end_context
//    keep the per-context feature bookkeeping of :1032-1132.
// The MI355X dispatch adds its own output variables (Range block size and
// items per lane, Team size and vector length, MDRange tile) and asks the
// tool for them inside that context; see KernelRange.hpp / KernelMDRange.hpp /
// KernelTeam.hpp.
#ifndef KKMX_TOOLS_HPP
#define KKMX_TOOLS_HPP

#include <kkmx/Spaces.hpp>

#include <dlfcn.h>

#include <cstring>
#include <algorithm>
#include <deque>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "kkmx_tools.h"

namespace Kokkos {
namespace Tools {

using SpaceHandle = Kokkos_Profiling_SpaceHandle;
inline SpaceHandle make_space_handle(const char* name) {
  SpaceHandle h;
  std::memset(h.name, 0, sizeof(h.name));
  std::strncpy(h.name, name, sizeof(h.name) - 1);
  return h;
}

using initFunction                  = Kokkos_Profiling_initFunction;
using finalizeFunction              = Kokkos_Profiling_finalizeFunction;
using parseArgsFunction             = Kokkos_Profiling_parseArgsFunction;
using printHelpFunction             = Kokkos_Profiling_printHelpFunction;
using beginFunction                 = Kokkos_Profiling_beginFunction;
using endFunction                   = Kokkos_Profiling_endFunction;
using pushFunction                  = Kokkos_Profiling_pushFunction;
using popFunction                   = Kokkos_Profiling_popFunction;
using allocateDataFunction          = Kokkos_Profiling_allocateDataFunction;
using deallocateDataFunction        = Kokkos_Profiling_deallocateDataFunction;
using createProfileSectionFunction  = Kokkos_Profiling_createProfileSectionFunction;
using startProfileSectionFunction   = Kokkos_Profiling_startProfileSectionFunction;
using stopProfileSectionFunction    = Kokkos_Profiling_stopProfileSectionFunction;
using destroyProfileSectionFunction = Kokkos_Profiling_destroyProfileSectionFunction;
using profileEventFunction          = Kokkos_Profiling_profileEventFunction;
using beginDeepCopyFunction         = Kokkos_Profiling_beginDeepCopyFunction;
using endDeepCopyFunction           = Kokkos_Profiling_endDeepCopyFunction;
using beginFenceFunction            = Kokkos_Profiling_beginFenceFunction;
using endFenceFunction              = Kokkos_Profiling_endFenceFunction;
using declareMetadataFunction       = Kokkos_Profiling_declareMetadataFunction;
using dualViewSyncFunction          = Kokkos_Profiling_dualViewSyncFunction;
using dualViewModifyFunction        = Kokkos_Profiling_dualViewModifyFunction;
using requestToolSettingsFunction   = Kokkos_Tools_requestToolSettingsFunction;
using provideToolProgrammingInterfaceFunction = Kokkos_Tools_provideToolProgrammingInterfaceFunction;
using ToolSettings             = Kokkos_Tools_ToolSettings;
using ToolProgrammingInterface = Kokkos_Tools_ToolProgrammingInterface;

namespace Experimental {
// The tool-settings and programming-interface types live in
// Tools::Experimental in the reference (impl/Kokkos_Profiling_Interface.hpp:132-154).
using ToolSettings                            = Kokkos_Tools_ToolSettings;
using ToolProgrammingInterface                = Kokkos_Tools_ToolProgrammingInterface;
using toolInvokedFenceFunction                = Kokkos_Tools_toolInvokedFenceFunction;
using provideToolProgrammingInterfaceFunction = Kokkos_Tools_provideToolProgrammingInterfaceFunction;
using requestToolSettingsFunction             = Kokkos_Tools_requestToolSettingsFunction;
using VariableInfo          = Kokkos_Tools_VariableInfo;
using VariableValue         = Kokkos_Tools_VariableValue;
using SetOrRange            = Kokkos_Tools_VariableInfo_SetOrRange;
using ValueSet              = Kokkos_Tools_ValueSet;
using ValueRange            = Kokkos_Tools_ValueRange;
using StatisticalCategory   = Kokkos_Tools_VariableInfo_StatisticalCategory;
using ValueType             = Kokkos_Tools_VariableInfo_ValueType;
using CandidateValueType    = Kokkos_Tools_VariableInfo_CandidateValueType;
using OptimizationGoal      = Kokkos_Tools_OptimizationGoal;
using OptimizationType      = Kokkos_Tools_OptimizationType;
using TuningString          = Kokkos_Tools_Tuning_String;
using outputTypeDeclarationFunction       = Kokkos_Tools_outputTypeDeclarationFunction;
using inputTypeDeclarationFunction        = Kokkos_Tools_inputTypeDeclarationFunction;
using requestValueFunction                = Kokkos_Tools_requestValueFunction;
using contextBeginFunction                = Kokkos_Tools_contextBeginFunction;
using contextEndFunction                  = Kokkos_Tools_contextEndFunction;
using optimizationGoalDeclarationFunction = Kokkos_Tools_optimizationGoalDeclarationFunction;

// One slot per callback (the reference keeps them in an EventSet struct).
struct EventSet {
  initFunction init                                   = nullptr;
  finalizeFunction finalize                           = nullptr;
  parseArgsFunction parse_args                        = nullptr;
  printHelpFunction print_help                        = nullptr;
  beginFunction begin_parallel_for                    = nullptr;
  endFunction end_parallel_for                        = nullptr;
  beginFunction begin_parallel_reduce                 = nullptr;
  endFunction end_parallel_reduce                     = nullptr;
  beginFunction begin_parallel_scan                   = nullptr;
  endFunction end_parallel_scan                       = nullptr;
  pushFunction push_region                            = nullptr;
  popFunction pop_region                              = nullptr;
  allocateDataFunction allocate_data                  = nullptr;
  deallocateDataFunction deallocate_data              = nullptr;
  createProfileSectionFunction create_profile_section = nullptr;
  startProfileSectionFunction start_profile_section   = nullptr;
  stopProfileSectionFunction stop_profile_section     = nullptr;
  destroyProfileSectionFunction destroy_profile_section = nullptr;
  profileEventFunction profile_event                  = nullptr;
  beginDeepCopyFunction begin_deep_copy               = nullptr;
  endDeepCopyFunction end_deep_copy                   = nullptr;
  beginFenceFunction begin_fence                      = nullptr;
  endFenceFunction end_fence                          = nullptr;
  declareMetadataFunction declare_metadata            = nullptr;
  dualViewSyncFunction sync_dual_view                 = nullptr;
  dualViewModifyFunction modify_dual_view             = nullptr;
  provideToolProgrammingInterfaceFunction provide_tool_programming_interface = nullptr;
  requestToolSettingsFunction request_tool_settings   = nullptr;
  outputTypeDeclarationFunction declare_output_type   = nullptr;
  inputTypeDeclarationFunction declare_input_type     = nullptr;
  requestValueFunction request_output_values          = nullptr;
  contextBeginFunction begin_tuning_context           = nullptr;
  contextEndFunction end_tuning_context               = nullptr;
  optimizationGoalDeclarationFunction declare_optimization_goal = nullptr;
};
}  // namespace Experimental

namespace Impl {
struct ToolsState {
  Experimental::EventSet cb;
  Experimental::EventSet paused;
  bool paused_flag          = false;
  bool requires_fence       = true;
  bool loaded               = false;
  std::vector<void*> handles;
  // tuning bookkeeping
  size_t context_counter  = 0;
  size_t variable_counter = 0;
  // Variable ids and context ids are small consecutive integers, so the
  // per-request bookkeeping is kept in id-indexed arrays (a request on every
  // dispatch must cost well under a microsecond).  metadata is a deque:
  // tools keep pointers into it.
  std::deque<Experimental::VariableInfo> metadata;             // [id]
  std::vector<std::vector<size_t>> features_per_context;       // [ctx]
  std::vector<size_t> active_features;                         // ids, insertion order
  std::vector<Experimental::VariableValue> feature_values;     // [id]
  std::unordered_map<size_t, size_t> goals;
  std::vector<Experimental::VariableValue> request_scratch;
  size_t kernel_name_id = 0, kernel_type_id = 0;
  bool ids_declared     = false;
  std::unordered_map<std::string, uint32_t> sections;
  uint32_t next_section = 0;
};
inline ToolsState& tools_state() {
  static ToolsState s;
  return s;
}
inline bool is_loaded() {
  const auto& c = tools_state().cb;
  const char* b = reinterpret_cast<const char*>(&c);
  for (size_t i = 0; i < sizeof(c); ++i)
    if (b[i]) return true;
  return false;
}
inline void update_loaded() { tools_state().loaded = is_loaded(); }
enum class MayFence : bool { No, Yes };
template <class Cb, class... Args>
inline void invoke(MayFence mf, Cb cb, Args&&... args) {
  if (cb) {
    if (mf == MayFence::Yes && tools_state().requires_fence)
      Kokkos::fence("Kokkos::Tools::invoke_kokkosp_callback: Kokkos Profile Tool Fence");
    cb(std::forward<Args>(args)...);
  }
}
}  // namespace Impl

inline bool profileLibraryLoaded() { return Impl::tools_state().loaded; }

namespace Experimental {
// ------------------------------------------------------------ tuning API
inline size_t get_new_context_id() { return ++Impl::tools_state().context_counter; }
inline size_t get_current_context_id() { return Impl::tools_state().context_counter; }
inline size_t get_new_variable_id() { return ++Impl::tools_state().variable_counter; }
inline bool have_tuning_tool() { return Impl::tools_state().cb.request_output_values != nullptr; }

inline void impl_register_variable(size_t id, const VariableInfo& info) {
  auto& s = Tools::Impl::tools_state();
  if (s.metadata.size() <= id) s.metadata.resize(id + 1);
  s.metadata[id] = info;
  if (s.feature_values.size() <= id) s.feature_values.resize(id + 1);
}
inline VariableInfo* impl_variable_info(size_t id) {
  auto& s = Tools::Impl::tools_state();
  return id < s.metadata.size() ? &s.metadata[id] : nullptr;
}

inline size_t declare_output_type(const std::string& name, VariableInfo info) {
  auto& s   = Impl::tools_state();
  size_t id = get_new_variable_id();
  Impl::invoke(Impl::MayFence::No, s.cb.declare_output_type, name.c_str(), id, &info);
  impl_register_variable(id, info);
  return id;
}
inline size_t declare_input_type(const std::string& name, VariableInfo info) {
  auto& s   = Impl::tools_state();
  size_t id = get_new_variable_id();
  Impl::invoke(Impl::MayFence::No, s.cb.declare_input_type, name.c_str(), id, &info);
  impl_register_variable(id, info);
  return id;
}
inline void set_input_values(size_t ctx, size_t count, VariableValue* values) {
  auto& s = Impl::tools_state();
  if (s.features_per_context.size() <= ctx) s.features_per_context.resize(ctx + 1);
  auto& f = s.features_per_context[ctx];
  for (size_t i = 0; i < count; ++i) {
    const size_t id    = values[i].type_id;
    values[i].metadata = impl_variable_info(id);
    if (std::find(f.begin(), f.end(), id) == f.end()) f.push_back(id);
    if (std::find(s.active_features.begin(), s.active_features.end(), id) == s.active_features.end())
      s.active_features.push_back(id);
    if (s.feature_values.size() <= id) s.feature_values.resize(id + 1);
    s.feature_values[id] = values[i];
  }
}
inline void request_output_values(size_t ctx, size_t count, VariableValue* values) {
  auto& s = Impl::tools_state();
  if (!s.cb.request_output_values) return;
  auto& in = s.request_scratch;
  in.clear();
  for (size_t id : s.active_features) in.push_back(s.feature_values[id]);
  for (size_t i = 0; i < count; ++i) values[i].metadata = impl_variable_info(values[i].type_id);
  Impl::invoke(Impl::MayFence::No, s.cb.request_output_values, ctx, in.size(), in.data(), count,
               values);
}
inline void begin_context(size_t ctx) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.begin_tuning_context, ctx);
}
inline void end_context(size_t ctx) {
  auto& s = Impl::tools_state();
  if (ctx < s.features_per_context.size()) {
    auto& f = s.features_per_context[ctx];
    for (size_t id : f) {
      auto it = std::find(s.active_features.begin(), s.active_features.end(), id);
      if (it != s.active_features.end()) s.active_features.erase(it);
    }
    f.clear();
  }
  VariableValue goal{};
  auto g = s.goals.find(ctx);
  if (g != s.goals.end()) {
    if (g->second < s.feature_values.size()) goal = s.feature_values[g->second];
    s.goals.erase(g);
  }
  Impl::invoke(Impl::MayFence::No, s.cb.end_tuning_context, ctx, goal);
  --s.context_counter;
}
inline void declare_optimization_goal(const size_t ctx, const OptimizationGoal& goal) {
  auto& s = Impl::tools_state();
  Impl::invoke(Impl::MayFence::No, s.cb.declare_optimization_goal, ctx, goal);
  s.goals[ctx] = goal.type_id;
}

inline VariableValue make_variable_value(size_t id, int64_t v) {
  VariableValue r{};
  r.type_id         = id;
  r.value.int_value = v;
  return r;
}
inline VariableValue make_variable_value(size_t id, int v) { return make_variable_value(id, int64_t(v)); }
inline VariableValue make_variable_value(size_t id, double v) {
  VariableValue r{};
  r.type_id            = id;
  r.value.double_value = v;
  return r;
}
inline VariableValue make_variable_value(size_t id, const std::string& v) {
  VariableValue r{};
  r.type_id = id;
  std::strncpy(r.value.string_value, v.c_str(), KOKKOS_TOOLS_TUNING_STRING_LENGTH - 1);
  return r;
}
inline VariableValue make_variable_value(size_t id, const char* v) {
  return make_variable_value(id, std::string(v));
}
inline SetOrRange make_candidate_set(size_t size, int64_t* data) {
  SetOrRange r{};
  r.set.size             = size;
  r.set.values.int_value = data;
  return r;
}
inline SetOrRange make_candidate_set(size_t size, double* data) {
  SetOrRange r{};
  r.set.size                = size;
  r.set.values.double_value = data;
  return r;
}
inline SetOrRange make_candidate_set(size_t size, std::string* data) {
  SetOrRange r{};
  r.set.size                = size;
  r.set.values.string_value = new TuningString[size];
  for (size_t i = 0; i < size; ++i)
    std::strncpy(r.set.values.string_value[i], data[i].c_str(), KOKKOS_TOOLS_TUNING_STRING_LENGTH - 1);
  return r;
}
inline SetOrRange make_candidate_range(int64_t lo, int64_t hi, int64_t step, bool open_lo = false,
                                       bool open_hi = false) {
  SetOrRange r{};
  r.range.lower.int_value = lo;
  r.range.upper.int_value = hi;
  r.range.step.int_value  = step;
  r.range.openLower       = open_lo;
  r.range.openUpper       = open_hi;
  return r;
}
inline SetOrRange make_candidate_range(double lo, double hi, double step, bool open_lo = false,
                                       bool open_hi = false) {
  SetOrRange r{};
  r.range.lower.double_value = lo;
  r.range.upper.double_value = hi;
  r.range.step.double_value  = step;
  r.range.openLower          = open_lo;
  r.range.openUpper          = open_hi;
  return r;
}

// ------------------------------------------------------------ callback setters
#define KKMX_SETTER(NAME, FIELD, TYPE)                     \
  inline void NAME(TYPE cb) {                              \
    Impl::tools_state().cb.FIELD = cb;                     \
    Impl::update_loaded();                                 \
  }
KKMX_SETTER(set_init_callback, init, initFunction)
KKMX_SETTER(set_finalize_callback, finalize, finalizeFunction)
KKMX_SETTER(set_parse_args_callback, parse_args, parseArgsFunction)
KKMX_SETTER(set_print_help_callback, print_help, printHelpFunction)
KKMX_SETTER(set_begin_parallel_for_callback, begin_parallel_for, beginFunction)
KKMX_SETTER(set_end_parallel_for_callback, end_parallel_for, endFunction)
KKMX_SETTER(set_begin_parallel_reduce_callback, begin_parallel_reduce, beginFunction)
KKMX_SETTER(set_end_parallel_reduce_callback, end_parallel_reduce, endFunction)
KKMX_SETTER(set_begin_parallel_scan_callback, begin_parallel_scan, beginFunction)
KKMX_SETTER(set_end_parallel_scan_callback, end_parallel_scan, endFunction)
KKMX_SETTER(set_push_region_callback, push_region, pushFunction)
KKMX_SETTER(set_pop_region_callback, pop_region, popFunction)
KKMX_SETTER(set_allocate_data_callback, allocate_data, allocateDataFunction)
KKMX_SETTER(set_deallocate_data_callback, deallocate_data, deallocateDataFunction)
KKMX_SETTER(set_create_profile_section_callback, create_profile_section, createProfileSectionFunction)
KKMX_SETTER(set_start_profile_section_callback, start_profile_section, startProfileSectionFunction)
KKMX_SETTER(set_stop_profile_section_callback, stop_profile_section, stopProfileSectionFunction)
KKMX_SETTER(set_destroy_profile_section_callback, destroy_profile_section, destroyProfileSectionFunction)
KKMX_SETTER(set_profile_event_callback, profile_event, profileEventFunction)
KKMX_SETTER(set_begin_deep_copy_callback, begin_deep_copy, beginDeepCopyFunction)
KKMX_SETTER(set_end_deep_copy_callback, end_deep_copy, endDeepCopyFunction)
KKMX_SETTER(set_begin_fence_callback, begin_fence, beginFenceFunction)
KKMX_SETTER(set_end_fence_callback, end_fence, endFenceFunction)
KKMX_SETTER(set_declare_metadata_callback, declare_metadata, declareMetadataFunction)
KKMX_SETTER(set_dual_view_sync_callback, sync_dual_view, dualViewSyncFunction)
KKMX_SETTER(set_dual_view_modify_callback, modify_dual_view, dualViewModifyFunction)
KKMX_SETTER(set_request_tool_settings_callback, request_tool_settings, requestToolSettingsFunction)
KKMX_SETTER(set_provide_tool_programming_interface_callback, provide_tool_programming_interface,
            provideToolProgrammingInterfaceFunction)
KKMX_SETTER(set_declare_output_type_callback, declare_output_type, outputTypeDeclarationFunction)
KKMX_SETTER(set_declare_input_type_callback, declare_input_type, inputTypeDeclarationFunction)
KKMX_SETTER(set_request_output_values_callback, request_output_values, requestValueFunction)
KKMX_SETTER(set_begin_context_callback, begin_tuning_context, contextBeginFunction)
KKMX_SETTER(set_end_context_callback, end_tuning_context, contextEndFunction)
KKMX_SETTER(set_declare_optimization_goal_callback, declare_optimization_goal,
            optimizationGoalDeclarationFunction)
#undef KKMX_SETTER

inline EventSet get_callbacks() { return Impl::tools_state().cb; }
inline void set_callbacks(EventSet e) {
  Impl::tools_state().cb = e;
  Impl::update_loaded();
}
inline void pause_tools() {
  auto& s = Impl::tools_state();
  if (s.paused_flag) return;
  s.paused      = s.cb;
  s.cb          = EventSet{};
  s.paused_flag = true;
  Impl::update_loaded();
}
inline void resume_tools() {
  auto& s = Impl::tools_state();
  if (!s.paused_flag) return;
  s.cb          = s.paused;
  s.paused_flag = false;
  Impl::update_loaded();
}

// Device ids handed to begin_* callbacks: the top bits name the device type
// (reference core/src/impl/Kokkos_Profiling_Interface.hpp:40-100).
enum struct DeviceType { Serial, OpenMP, Cuda, HIP, OpenMPTarget, HPX, Threads, SYCL, OpenACC, Unknown };
template <class Space>
struct DeviceTypeTraits {
  static constexpr DeviceType id = DeviceType::Unknown;
  static int device_id(const Space&) { return 0; }
};
template <>
struct DeviceTypeTraits<Serial> {
  static constexpr DeviceType id = DeviceType::Serial;
  static int device_id(const Serial&) { return 0; }
};
template <>
struct DeviceTypeTraits<MI355X> {
  // Tools see the MI355X space as an AMD GPU device.
  static constexpr DeviceType id = DeviceType::HIP;
  static int device_id(const MI355X&) { return MI355X::impl_props().device; }
};
template <class Space>
inline uint32_t device_id(const Space& s) {
  return (static_cast<uint32_t>(DeviceTypeTraits<Space>::id) << 24) |
         (static_cast<uint32_t>(DeviceTypeTraits<Space>::device_id(s)) << 17) |
         (s.impl_instance_id() & ((1u << 17) - 1));
}
}  // namespace Experimental

namespace Impl {
inline void declare_context_ids() {
  auto& s = tools_state();
  if (s.ids_declared) return;
  s.ids_declared = true;
  Experimental::VariableInfo info{};
  info.type          = kokkos_value_string;
  info.category      = kokkos_value_categorical;
  info.valueQuantity = kokkos_value_unbounded;
  s.kernel_name_id   = Experimental::declare_input_type("kokkos.kernel_name", info);
  s.kernel_type_id   = Experimental::declare_input_type("kokkos.kernel_type", info);
}
inline void open_kernel_context(const std::string& name, const char* type) {
  if (!Kokkos::Impl::tune_internals()) return;
  declare_context_ids();
  auto& s  = tools_state();
  size_t c = Experimental::get_new_context_id();
  Experimental::begin_context(c);
  Experimental::VariableValue v[2] = {Experimental::make_variable_value(s.kernel_name_id, name),
                                      Experimental::make_variable_value(s.kernel_type_id, std::string(type))};
  Experimental::set_input_values(c, 2, v);
}
inline void close_kernel_context() {
  if (!Kokkos::Impl::tune_internals()) return;
  Experimental::end_context(Experimental::get_current_context_id());
}
}  // namespace Impl

// ------------------------------------------------------------ profiling entry points
inline void beginParallelFor(const std::string& name, const uint32_t dev, uint64_t* kid) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.begin_parallel_for, name.c_str(), dev, kid);
  Impl::open_kernel_context(name, "parallel_for");
}
inline void endParallelFor(const uint64_t kid) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.end_parallel_for, kid);
  Impl::close_kernel_context();
}
inline void beginParallelReduce(const std::string& name, const uint32_t dev, uint64_t* kid) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.begin_parallel_reduce, name.c_str(), dev, kid);
  Impl::open_kernel_context(name, "parallel_reduce");
}
inline void endParallelReduce(const uint64_t kid) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.end_parallel_reduce, kid);
  Impl::close_kernel_context();
}
inline void beginParallelScan(const std::string& name, const uint32_t dev, uint64_t* kid) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.begin_parallel_scan, name.c_str(), dev, kid);
  Impl::open_kernel_context(name, "parallel_scan");
}
inline void endParallelScan(const uint64_t kid) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.end_parallel_scan, kid);
  Impl::close_kernel_context();
}
inline void pushRegion(const std::string& n) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.push_region, n.c_str());
}
inline void popRegion() { Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.pop_region); }
inline void markEvent(const std::string& n) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.profile_event, n.c_str());
}
inline void allocateData(const SpaceHandle h, const std::string& l, const void* p, const uint64_t n) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.allocate_data, h, l.c_str(), p, n);
}
inline void deallocateData(const SpaceHandle h, const std::string& l, const void* p, const uint64_t n) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.deallocate_data, h, l.c_str(), p, n);
}
inline void beginDeepCopy(const SpaceHandle d, const std::string& dl, const void* dp, const SpaceHandle s,
                          const std::string& sl, const void* sp, const uint64_t n) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.begin_deep_copy, d, dl.c_str(), dp, s,
               sl.c_str(), sp, n);
}
inline void endDeepCopy() { Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.end_deep_copy); }
inline void beginFence(const std::string& n, const uint32_t dev, uint64_t* h) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.begin_fence, n.c_str(), dev, h);
}
inline void endFence(const uint64_t h) { Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.end_fence, h); }
inline void declareMetadata(const std::string& k, const std::string& v) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.declare_metadata, k.c_str(), v.c_str());
}
inline void createProfileSection(const std::string& n, uint32_t* id) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.create_profile_section, n.c_str(), id);
}
inline void startSection(const uint32_t id) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.start_profile_section, id);
}
inline void stopSection(const uint32_t id) {
  Impl::invoke(Impl::MayFence::Yes, Impl::tools_state().cb.stop_profile_section, id);
}
inline void destroyProfileSection(const uint32_t id) {
  Impl::invoke(Impl::MayFence::No, Impl::tools_state().cb.destroy_profile_section, id);
}

class ProfilingSection {
 public:
  explicit ProfilingSection(const std::string& name) { createProfileSection(name, &m_id); }
  ~ProfilingSection() { destroyProfileSection(m_id); }
  void start() { startSection(m_id); }
  void stop() { stopSection(m_id); }

 private:
  uint32_t m_id = 0;
};
class ScopedRegion {
 public:
  explicit ScopedRegion(const std::string& name) { pushRegion(name); }
  ~ScopedRegion() { popRegion(); }
};

namespace Impl {
// Load the kokkosp_* symbols of each library in a ';'-separated list.
inline void load_tool_libraries(const std::string& libs) {
  auto& s    = tools_state();
  size_t pos = 0;
  while (pos <= libs.size()) {
    size_t end      = libs.find_first_of(";:", pos);
    std::string lib = libs.substr(pos, end == std::string::npos ? std::string::npos : end - pos);
    pos             = (end == std::string::npos) ? libs.size() + 1 : end + 1;
    if (lib.empty()) continue;
    void* h = dlopen(lib.c_str(), RTLD_NOW | RTLD_GLOBAL);
    if (!h) {
      std::fprintf(stderr, "KokkosP: Error: could not load tool library %s: %s\n", lib.c_str(), dlerror());
      continue;
    }
    s.handles.push_back(h);
    auto& c = s.cb;
#define KKMX_SYM(FIELD, NAME) \
  if (auto* p = dlsym(h, NAME)) c.FIELD = reinterpret_cast<decltype(c.FIELD)>(p);
    KKMX_SYM(init, "kokkosp_init_library")
    KKMX_SYM(finalize, "kokkosp_finalize_library")
    KKMX_SYM(parse_args, "kokkosp_parse_args")
    KKMX_SYM(print_help, "kokkosp_print_help")
    KKMX_SYM(begin_parallel_for, "kokkosp_begin_parallel_for")
    KKMX_SYM(end_parallel_for, "kokkosp_end_parallel_for")
    KKMX_SYM(begin_parallel_reduce, "kokkosp_begin_parallel_reduce")
    KKMX_SYM(end_parallel_reduce, "kokkosp_end_parallel_reduce")
    KKMX_SYM(begin_parallel_scan, "kokkosp_begin_parallel_scan")
    KKMX_SYM(end_parallel_scan, "kokkosp_end_parallel_scan")
    KKMX_SYM(push_region, "kokkosp_push_profile_region")
    KKMX_SYM(pop_region, "kokkosp_pop_profile_region")
    KKMX_SYM(allocate_data, "kokkosp_allocate_data")
    KKMX_SYM(deallocate_data, "kokkosp_deallocate_data")
    KKMX_SYM(create_profile_section, "kokkosp_create_profile_section")
    KKMX_SYM(start_profile_section, "kokkosp_start_profile_section")
    KKMX_SYM(stop_profile_section, "kokkosp_stop_profile_section")
    KKMX_SYM(destroy_profile_section, "kokkosp_destroy_profile_section")
    KKMX_SYM(profile_event, "kokkosp_profile_event")
    KKMX_SYM(begin_deep_copy, "kokkosp_begin_deep_copy")
    KKMX_SYM(end_deep_copy, "kokkosp_end_deep_copy")
    KKMX_SYM(begin_fence, "kokkosp_begin_fence")
    KKMX_SYM(end_fence, "kokkosp_end_fence")
    KKMX_SYM(declare_metadata, "kokkosp_declare_metadata")
    KKMX_SYM(provide_tool_programming_interface, "kokkosp_provide_tool_programming_interface")
    KKMX_SYM(request_tool_settings, "kokkosp_request_tool_settings")
    KKMX_SYM(declare_output_type, "kokkosp_declare_output_type")
    KKMX_SYM(declare_input_type, "kokkosp_declare_input_type")
    KKMX_SYM(request_output_values, "kokkosp_request_values")
    KKMX_SYM(begin_tuning_context, "kokkosp_begin_context")
    KKMX_SYM(end_tuning_context, "kokkosp_end_context")
    KKMX_SYM(declare_optimization_goal, "kokkosp_declare_optimization_goal")
#undef KKMX_SYM
  }
  update_loaded();
}

inline void tool_invoked_fence(const uint32_t) { Kokkos::fence("Kokkos::Tools::Fence"); }

inline void tools_start(const std::string& libs, const std::string& args) {
  if (!libs.empty()) load_tool_libraries(libs);
  auto& s = tools_state();
  ToolSettings settings{};
  settings.requires_global_fencing = true;
  invoke(MayFence::No, s.cb.request_tool_settings, 1u, &settings);
  s.requires_fence = settings.requires_global_fencing;
  if (s.cb.provide_tool_programming_interface) {
    ToolProgrammingInterface tpi{};
    tpi.fence = &tool_invoked_fence;
    s.cb.provide_tool_programming_interface(1u, tpi);
  }
  Kokkos_Profiling_KokkosPDeviceInfo info{0};
  invoke(MayFence::No, s.cb.init, 0, uint64_t(KOKKOS_VERSION), uint32_t(KOKKOSP_INTERFACE_VERSION), &info);
  if (s.cb.parse_args && !args.empty()) {
    std::string a = args;
    std::vector<char*> argv;
    char* tok = std::strtok(a.data(), " ");
    while (tok) {
      argv.push_back(tok);
      tok = std::strtok(nullptr, " ");
    }
    s.cb.parse_args(static_cast<int>(argv.size()), argv.data());
  }
}
inline void tools_stop() {
  auto& s = tools_state();
  invoke(MayFence::No, s.cb.finalize);
  s.cb = Experimental::EventSet{};
  update_loaded();
}
}  // namespace Impl

// Settings the dispatch layer re-reads after a tool changes them at run time.
inline void impl_request_tool_settings_refresh() {
  auto& s = Impl::tools_state();
  ToolSettings settings{};
  settings.requires_global_fencing = true;
  Impl::invoke(Impl::MayFence::No, s.cb.request_tool_settings, 1u, &settings);
  if (s.cb.request_tool_settings) s.requires_fence = settings.requires_global_fencing;
}

}  // namespace Tools

namespace Profiling {
// impl/Kokkos_Profiling_DeviceInfo.hpp:24 and the Profiling spellings of the
// Tools types.
using KokkosPDeviceInfo = Kokkos_Profiling_KokkosPDeviceInfo;
using Tools::beginParallelFor;
using Tools::beginParallelReduce;
using Tools::beginParallelScan;
using Tools::endParallelFor;
using Tools::endParallelReduce;
using Tools::endParallelScan;
using Tools::make_space_handle;
using Tools::markEvent;
using Tools::popRegion;
using Tools::profileLibraryLoaded;
using Tools::ProfilingSection;
using Tools::pushRegion;
using Tools::ScopedRegion;
using Tools::SpaceHandle;
}  // namespace Profiling

}  // namespace Kokkos

#endif
