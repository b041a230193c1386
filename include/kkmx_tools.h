/*
 * kkmx_tools.h -- the Kokkos Tools C ABI as the MI355X execution space
 * exposes it to profiling and tuning tools.
 *
 * A tool library (loaded with dlopen from KOKKOS_TOOLS_LIBS, or installed in
 * process with the Kokkos::Tools::Experimental::set_*_callback functions)
 * exports kokkosp_* functions with exactly these signatures.  Struct layouts
 * and enumerator values are kept identical to the reference interface so that
 * tools built against it load unchanged:
 *   core/src/impl/Kokkos_Profiling_C_Interface.h:33-276 (types, callbacks,
 *   EventSet), looked up by name in core/src/impl/Kokkos_Profiling.cpp:646-718.
 *
 * Tuning entry points a tool provides (reference :215-227):
 *   kokkosp_declare_output_type, kokkosp_declare_input_type,
 *   kokkosp_request_values, kokkosp_begin_context, kokkosp_end_context,
 *   kokkosp_declare_optimization_goal.
 */
#ifndef KKMX_TOOLS_H
#define KKMX_TOOLS_H

#ifdef __cplusplus
#include <cstddef>
#include <cstdint>
extern "C" {
#else
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#endif

#define KOKKOSP_INTERFACE_VERSION 20240906

struct Kokkos_Profiling_KokkosPDeviceInfo {
  size_t deviceID;
};
struct Kokkos_Profiling_SpaceHandle {
  char name[64];
};

typedef void (*Kokkos_Profiling_initFunction)(const int, const uint64_t, const uint32_t,
                                              struct Kokkos_Profiling_KokkosPDeviceInfo*);
typedef void (*Kokkos_Profiling_finalizeFunction)();
typedef void (*Kokkos_Profiling_parseArgsFunction)(int, char**);
typedef void (*Kokkos_Profiling_printHelpFunction)(char*);
typedef void (*Kokkos_Profiling_beginFunction)(const char*, const uint32_t, uint64_t*);
typedef void (*Kokkos_Profiling_endFunction)(uint64_t);
typedef void (*Kokkos_Profiling_pushFunction)(const char*);
typedef void (*Kokkos_Profiling_popFunction)();
typedef void (*Kokkos_Profiling_allocateDataFunction)(const struct Kokkos_Profiling_SpaceHandle,
                                                      const char*, const void*, const uint64_t);
typedef void (*Kokkos_Profiling_deallocateDataFunction)(const struct Kokkos_Profiling_SpaceHandle,
                                                        const char*, const void*, const uint64_t);
typedef void (*Kokkos_Profiling_createProfileSectionFunction)(const char*, uint32_t*);
typedef void (*Kokkos_Profiling_startProfileSectionFunction)(const uint32_t);
typedef void (*Kokkos_Profiling_stopProfileSectionFunction)(const uint32_t);
typedef void (*Kokkos_Profiling_destroyProfileSectionFunction)(const uint32_t);
typedef void (*Kokkos_Profiling_profileEventFunction)(const char*);
typedef void (*Kokkos_Profiling_beginDeepCopyFunction)(struct Kokkos_Profiling_SpaceHandle, const char*,
                                                       const void*, struct Kokkos_Profiling_SpaceHandle,
                                                       const char*, const void*, uint64_t);
typedef void (*Kokkos_Profiling_endDeepCopyFunction)();
typedef void (*Kokkos_Profiling_beginFenceFunction)(const char*, const uint32_t, uint64_t*);
typedef void (*Kokkos_Profiling_endFenceFunction)(uint64_t);
typedef void (*Kokkos_Profiling_dualViewSyncFunction)(const char*, const void* const, bool);
typedef void (*Kokkos_Profiling_dualViewModifyFunction)(const char*, const void* const, bool);
typedef void (*Kokkos_Profiling_declareMetadataFunction)(const char*, const char*);
typedef void (*Kokkos_Tools_toolInvokedFenceFunction)(const uint32_t);
typedef void (*Kokkos_Tools_functionPointer)();

struct Kokkos_Tools_ToolProgrammingInterface {
  Kokkos_Tools_toolInvokedFenceFunction fence;
  Kokkos_Tools_functionPointer padding[31];
};
struct Kokkos_Tools_ToolSettings {
  bool requires_global_fencing;
  bool padding[255];
};
typedef void (*Kokkos_Tools_provideToolProgrammingInterfaceFunction)(
    const uint32_t, struct Kokkos_Tools_ToolProgrammingInterface);
typedef void (*Kokkos_Tools_requestToolSettingsFunction)(const uint32_t,
                                                          struct Kokkos_Tools_ToolSettings*);

#define KOKKOS_TOOLS_TUNING_STRING_LENGTH 64
typedef char Kokkos_Tools_Tuning_String[KOKKOS_TOOLS_TUNING_STRING_LENGTH];
union Kokkos_Tools_VariableValue_ValueUnion {
  int64_t int_value;
  double double_value;
  Kokkos_Tools_Tuning_String string_value;
};
union Kokkos_Tools_VariableValue_ValueUnionSet {
  int64_t* int_value;
  double* double_value;
  Kokkos_Tools_Tuning_String* string_value;
};
struct Kokkos_Tools_ValueSet {
  size_t size;
  union Kokkos_Tools_VariableValue_ValueUnionSet values;
};
enum Kokkos_Tools_OptimizationType { Kokkos_Tools_Minimize, Kokkos_Tools_Maximize };
struct Kokkos_Tools_OptimizationGoal {
  size_t type_id;
  enum Kokkos_Tools_OptimizationType goal;
};
struct Kokkos_Tools_ValueRange {
  union Kokkos_Tools_VariableValue_ValueUnion lower;
  union Kokkos_Tools_VariableValue_ValueUnion upper;
  union Kokkos_Tools_VariableValue_ValueUnion step;
  bool openLower;
  bool openUpper;
};
enum Kokkos_Tools_VariableInfo_ValueType {
  kokkos_value_double,
  kokkos_value_int64,
  kokkos_value_string,
};
enum Kokkos_Tools_VariableInfo_StatisticalCategory {
  kokkos_value_categorical,
  kokkos_value_ordinal,
  kokkos_value_interval,
  kokkos_value_ratio
};
enum Kokkos_Tools_VariableInfo_CandidateValueType {
  kokkos_value_set,
  kokkos_value_range,
  kokkos_value_unbounded
};
union Kokkos_Tools_VariableInfo_SetOrRange {
  struct Kokkos_Tools_ValueSet set;
  struct Kokkos_Tools_ValueRange range;
};
struct Kokkos_Tools_VariableInfo {
  enum Kokkos_Tools_VariableInfo_ValueType type;
  enum Kokkos_Tools_VariableInfo_StatisticalCategory category;
  enum Kokkos_Tools_VariableInfo_CandidateValueType valueQuantity;
  union Kokkos_Tools_VariableInfo_SetOrRange candidates;
  void* toolProvidedInfo;
};
struct Kokkos_Tools_VariableValue {
  size_t type_id;
  union Kokkos_Tools_VariableValue_ValueUnion value;
  struct Kokkos_Tools_VariableInfo* metadata;
};

typedef void (*Kokkos_Tools_outputTypeDeclarationFunction)(const char*, const size_t,
                                                           struct Kokkos_Tools_VariableInfo* info);
typedef void (*Kokkos_Tools_inputTypeDeclarationFunction)(const char*, const size_t,
                                                          struct Kokkos_Tools_VariableInfo* info);
typedef void (*Kokkos_Tools_requestValueFunction)(const size_t, const size_t,
                                                  const struct Kokkos_Tools_VariableValue*,
                                                  const size_t count,
                                                  struct Kokkos_Tools_VariableValue*);
typedef void (*Kokkos_Tools_contextBeginFunction)(const size_t);
typedef void (*Kokkos_Tools_contextEndFunction)(const size_t, struct Kokkos_Tools_VariableValue);
typedef void (*Kokkos_Tools_optimizationGoalDeclarationFunction)(
    const size_t, const struct Kokkos_Tools_OptimizationGoal goal);

#ifdef __cplusplus
}
#endif

#endif /* KKMX_TOOLS_H */
