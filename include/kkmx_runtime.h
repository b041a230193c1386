/*
 * kkmx_runtime.h -- the thin C-ABI runtime between the Kokkos-API host layer
 * (kokkos_amd/include/Kokkos_Core.hpp) and the HIP runtime.
 *
 * Every host-side HIP call the MI355X execution space makes goes through one
 * of these functions: plain pointers, sizes and opaque handles, no C++ types.
 * Each function returns an int status: 0 is success, anything else is a HIP
 * error code that mx_error_string() turns into text.  The C++ layer turns a
 * non-zero status into the exception types the reference throws.
 *
 * Reference interfaces each entry replaces (paths relative to the kokkos
 * source tree):
 *   mx_init / mx_finalize        core/src/HIP/Kokkos_HIP.cpp:47-116 (HIP::impl_initialize),
 *                                core/src/HIP/Kokkos_HIP_Instance.cpp:183-245,372-421
 *   mx_device_props              core/src/HIP/Kokkos_HIP_Instance.cpp:87-133 (print_configuration)
 *   mx_malloc / mx_free          core/src/HIP/Kokkos_HIP_Space.cpp:70-131,238-260 (HIPSpace::allocate/deallocate)
 *   mx_host_alloc / mx_host_free core/src/HIP/Kokkos_HIP_Space.cpp:133-163 (HIPHostPinnedSpace::allocate)
 *   mx_memcpy_async / mx_memcpy  core/src/HIP/Kokkos_HIP_DeepCopy.cpp:23-51 (DeepCopyHIP / DeepCopyAsyncHIP)
 *   mx_memset_async              core/src/HIP/Kokkos_HIP_ZeroMemset.cpp
 *   mx_stream_*                  core/src/HIP/Kokkos_HIP_Instance.cpp:160-181 (fence), :183-245 (stream)
 *   mx_event_*                   core/src/HIP/Kokkos_HIP_Instance.cpp (fence by event), benchmarks timers
 *   mx_launch                    core/src/HIP/Kokkos_HIP_KernelLaunch.hpp:382-389,594-641 (hip_parallel_launch)
 *   mx_func_attributes           core/src/HIP/Kokkos_HIP_KernelLaunch.hpp:194-206 (cached hipFuncGetAttributes)
 *   mx_occupancy                 core/src/HIP/Kokkos_HIP_BlockSize_Deduction.hpp:107-150
 *   mx_graph_*                   core/src/HIP/Kokkos_HIP_Graph_Impl.hpp:82-89,118-221
 *   mx_error_string              core/src/HIP/Kokkos_HIP_Instance.cpp:456-465 (hip_internal_error_throw)
 */
#ifndef KKMX_RUNTIME_H
#define KKMX_RUNTIME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* mx_stream_t;     /* hipStream_t */
typedef void* mx_event_t;      /* hipEvent_t */
typedef void* mx_graph_t;      /* hipGraph_t */
typedef void* mx_graph_exec_t; /* hipGraphExec_t */

typedef struct mx_device_props {
  char name[256];
  char arch[64];             /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int device;
  int cu_count;              /* multiProcessorCount */
  int wave_size;             /* warpSize: 64 on CDNA */
  int max_threads_per_block;
  int max_threads_per_cu;
  int regs_per_block;
  int clock_khz;
  int mem_clock_khz;
  int mem_bus_width;
  int l2_bytes;
  int xcc_count;             /* derived: cu_count / CUs per XCD when known */
  size_t lds_per_block;      /* sharedMemPerBlock */
  size_t lds_per_cu;         /* maxSharedMemoryPerMultiProcessor */
  size_t global_mem;
  int max_grid_x;
  int concurrent_kernels;
} mx_device_props;

/* Device selection and set-up.  Safe to call more than once. */
int mx_init(int device);
int mx_finalize(void);
int mx_is_initialized(void);
int mx_device_count(int* count);
int mx_device_props_get(mx_device_props* out);
int mx_current_device(int* device);

/* Device memory, stream ordered when a stream is given (NULL = synchronous). */
int mx_malloc(void** ptr, size_t bytes, mx_stream_t stream);
int mx_free(void* ptr, mx_stream_t stream);
/* Pinned, device-mapped host memory. */
int mx_host_alloc(void** ptr, size_t bytes);
int mx_host_free(void* ptr);
/* 1 if ptr is device memory, 2 if pinned host memory, 0 if plain host memory. */
int mx_pointer_kind(const void* ptr, int* kind);

int mx_memcpy_async(void* dst, const void* src, size_t bytes, mx_stream_t stream);
int mx_memcpy(void* dst, const void* src, size_t bytes);
int mx_memset_async(void* dst, int value, size_t bytes, mx_stream_t stream);

/* Streams.  All streams keep the default priority. */
int mx_stream_create(mx_stream_t* stream);
int mx_stream_destroy(mx_stream_t stream);
int mx_stream_sync(mx_stream_t stream);
int mx_stream_query(mx_stream_t stream); /* 0 = idle, 600 (hipErrorNotReady) = busy */
int mx_device_sync(void);

/* Events for fences between streams and for timing. */
int mx_event_create(mx_event_t* ev);
int mx_event_destroy(mx_event_t ev);
int mx_event_record(mx_event_t ev, mx_stream_t stream);
int mx_event_sync(mx_event_t ev);
int mx_event_query(mx_event_t ev); /* 0 = complete, 600 (hipErrorNotReady) = pending */
int mx_event_elapsed_ms(float* ms, mx_event_t start, mx_event_t stop);
/* *capturing = 1 while the stream records into a graph (stream capture). */
int mx_stream_capturing(mx_stream_t stream, int* capturing);
int mx_stream_wait_event(mx_stream_t stream, mx_event_t ev);

/* Kernel launch: fn is the host-side handle of a __global__ function,
 * args an array of pointers to each kernel argument. */
int mx_launch(const void* fn, unsigned gx, unsigned gy, unsigned gz,
              unsigned bx, unsigned by, unsigned bz, size_t lds_bytes,
              void** args, mx_stream_t stream);
int mx_func_attributes(const void* fn, int* max_threads_per_block,
                       int* num_regs, size_t* static_lds_bytes,
                       size_t* private_bytes);
/* Resident workgroups per CU for a block size and dynamic LDS size. */
int mx_occupancy(const void* fn, int block_size, size_t dyn_lds_bytes,
                 int* blocks_per_cu);
int mx_get_last_error(void);

/* Graphs: stream capture, instantiation and replay. */
int mx_graph_capture_begin(mx_stream_t stream);
int mx_graph_capture_end(mx_stream_t stream, mx_graph_t* graph);
int mx_graph_instantiate(mx_graph_exec_t* exec, mx_graph_t graph);
int mx_graph_launch(mx_graph_exec_t exec, mx_stream_t stream);
int mx_graph_exec_destroy(mx_graph_exec_t exec);
int mx_graph_destroy(mx_graph_t graph);
int mx_graph_node_count(mx_graph_t graph, size_t* count);

const char* mx_error_string(int status);

/* The stream the default MI355X instance dispatches on (created by mx_init,
 * like the reference's default HIP instance stream). */
int mx_default_stream(mx_stream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* KKMX_RUNTIME_H */
