// C-ABI runtime: the only translation unit of the MI355X execution space that
// calls the HIP host API.  Everything above it (the Kokkos-API headers, the
// dispatch templates, the graph and batch layers) calls these functions.
//
// Design notes
//  * One device per process; mx_init selects it and creates the default
//    instance stream (reference: core/src/HIP/Kokkos_HIP.cpp:47-116 creates a
//    stream and the HIPInternal singleton, Kokkos_HIP_Instance.cpp:183-245).
//  * Allocation is stream ordered (hipMallocAsync) when a stream is given, like
//    the reference's HIPSpace::impl_allocate (Kokkos_HIP_Space.cpp:92-131).
//  * Streams keep the default priority.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "kkmx_runtime.h"

namespace {

struct RuntimeState {
  std::mutex mtx;
  int device           = -1;
  bool initialized     = false;
  hipStream_t stream   = nullptr;
  mx_device_props props{};
};

RuntimeState& state() {
  static RuntimeState s;
  return s;
}

inline int rc(hipError_t e) { return static_cast<int>(e); }

}  // namespace

extern "C" {

int mx_device_count(int* count) {
  int n     = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = (e == hipSuccess) ? n : 0;
  return rc(e);
}

int mx_init(int device) {
  RuntimeState& s = state();
  std::lock_guard<std::mutex> lk(s.mtx);
  if (s.initialized) return 0;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return rc(e);
  if (n <= 0) return rc(hipErrorNoDevice);
  if (device < 0) device = 0;
  if (device >= n) return rc(hipErrorInvalidDevice);
  e = hipSetDevice(device);
  if (e != hipSuccess) return rc(e);
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, device);
  if (e != hipSuccess) return rc(e);
  mx_device_props& d = s.props;
  std::memset(&d, 0, sizeof(d));
  std::snprintf(d.name, sizeof(d.name), "%s", p.name);
  std::snprintf(d.arch, sizeof(d.arch), "%s", p.gcnArchName);
  d.device                = device;
  d.cu_count              = p.multiProcessorCount;
  d.wave_size             = p.warpSize;
  d.max_threads_per_block = p.maxThreadsPerBlock;
  d.max_threads_per_cu    = p.maxThreadsPerMultiProcessor;
  d.regs_per_block        = p.regsPerBlock;
  d.clock_khz             = p.clockRate;
  d.mem_clock_khz         = p.memoryClockRate;
  d.mem_bus_width         = p.memoryBusWidth;
  d.l2_bytes              = p.l2CacheSize;
  d.lds_per_block         = p.sharedMemPerBlock;
  d.lds_per_cu            = p.maxSharedMemoryPerMultiProcessor;
  d.global_mem            = p.totalGlobalMem;
  d.max_grid_x            = p.maxGridSize[0];
  d.concurrent_kernels    = p.concurrentKernels;
  // CDNA3/CDNA4 parts group 32 (38 physical, 32 or 36 enabled) CUs per XCD.
  d.xcc_count = (d.cu_count % 32 == 0) ? d.cu_count / 32 : 1;
  e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
  if (e != hipSuccess) return rc(e);
  s.device      = device;
  s.initialized = true;
  return 0;
}

int mx_finalize(void) {
  RuntimeState& s = state();
  std::lock_guard<std::mutex> lk(s.mtx);
  if (!s.initialized) return 0;
  hipError_t e = hipStreamSynchronize(s.stream);
  hipError_t e2 = hipStreamDestroy(s.stream);
  s.stream      = nullptr;
  s.initialized = false;
  return rc(e != hipSuccess ? e : e2);
}

int mx_is_initialized(void) { return state().initialized ? 1 : 0; }

int mx_device_props_get(mx_device_props* out) {
  RuntimeState& s = state();
  if (!s.initialized) return rc(hipErrorNotInitialized);
  *out = s.props;
  return 0;
}

int mx_current_device(int* device) { return rc(hipGetDevice(device)); }

int mx_default_stream(mx_stream_t* stream) {
  RuntimeState& s = state();
  if (!s.initialized) return rc(hipErrorNotInitialized);
  *stream = s.stream;
  return 0;
}

int mx_malloc(void** ptr, size_t bytes, mx_stream_t stream) {
  *ptr = nullptr;
  if (bytes == 0) return 0;
  if (stream)
    return rc(hipMallocAsync(ptr, bytes, static_cast<hipStream_t>(stream)));
  return rc(hipMalloc(ptr, bytes));
}

int mx_free(void* ptr, mx_stream_t stream) {
  if (!ptr) return 0;
  if (stream) return rc(hipFreeAsync(ptr, static_cast<hipStream_t>(stream)));
  return rc(hipFree(ptr));
}

int mx_host_alloc(void** ptr, size_t bytes) {
  *ptr = nullptr;
  if (bytes == 0) return 0;
  return rc(hipHostMalloc(ptr, bytes, hipHostMallocMapped));
}

int mx_host_free(void* ptr) {
  if (!ptr) return 0;
  return rc(hipHostFree(ptr));
}

int mx_pointer_kind(const void* ptr, int* kind) {
  *kind = 0;
  if (!ptr) return 0;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // plain host memory reports an error: clear it
    return 0;
  }
  if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged)
    *kind = 1;
  else if (a.type == hipMemoryTypeHost)
    *kind = 2;
  return 0;
}

int mx_memcpy_async(void* dst, const void* src, size_t bytes,
                    mx_stream_t stream) {
  if (bytes == 0) return 0;
  return rc(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault,
                           static_cast<hipStream_t>(stream)));
}

int mx_memcpy(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  return rc(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
}

int mx_memset_async(void* dst, int value, size_t bytes, mx_stream_t stream) {
  if (bytes == 0) return 0;
  return rc(hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream)));
}

int mx_stream_create(mx_stream_t* stream) {
  hipStream_t s = nullptr;
  hipError_t e  = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  *stream       = s;
  return rc(e);
}

int mx_stream_destroy(mx_stream_t stream) {
  if (!stream) return 0;
  return rc(hipStreamDestroy(static_cast<hipStream_t>(stream)));
}

int mx_stream_sync(mx_stream_t stream) {
  return rc(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

int mx_stream_query(mx_stream_t stream) {
  return rc(hipStreamQuery(static_cast<hipStream_t>(stream)));
}

int mx_device_sync(void) { return rc(hipDeviceSynchronize()); }

int mx_event_create(mx_event_t* ev) {
  hipEvent_t e = nullptr;
  hipError_t r = hipEventCreate(&e);
  *ev          = e;
  return rc(r);
}

int mx_event_destroy(mx_event_t ev) {
  if (!ev) return 0;
  return rc(hipEventDestroy(static_cast<hipEvent_t>(ev)));
}

int mx_event_record(mx_event_t ev, mx_stream_t stream) {
  return rc(hipEventRecord(static_cast<hipEvent_t>(ev),
                           static_cast<hipStream_t>(stream)));
}

int mx_event_sync(mx_event_t ev) {
  return rc(hipEventSynchronize(static_cast<hipEvent_t>(ev)));
}

int mx_event_query(mx_event_t ev) {
  const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(ev));
  if (e == hipErrorNotReady) (void)hipGetLastError();  // pending is not an error
  return rc(e);
}

int mx_stream_capturing(mx_stream_t stream, int* capturing) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  const int r = rc(hipStreamIsCapturing(static_cast<hipStream_t>(stream), &st));
  if (capturing) *capturing = (st != hipStreamCaptureStatusNone);
  return r;
}

int mx_event_elapsed_ms(float* ms, mx_event_t start, mx_event_t stop) {
  return rc(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                static_cast<hipEvent_t>(stop)));
}

int mx_stream_wait_event(mx_stream_t stream, mx_event_t ev) {
  return rc(hipStreamWaitEvent(static_cast<hipStream_t>(stream),
                               static_cast<hipEvent_t>(ev), 0));
}

int mx_launch(const void* fn, unsigned gx, unsigned gy, unsigned gz,
              unsigned bx, unsigned by, unsigned bz, size_t lds_bytes,
              void** args, mx_stream_t stream) {
  if (gx == 0 || gy == 0 || gz == 0) return 0;  // empty dispatch
  return rc(hipLaunchKernel(fn, dim3(gx, gy, gz), dim3(bx, by, bz), args,
                            lds_bytes, static_cast<hipStream_t>(stream)));
}

int mx_func_attributes(const void* fn, int* max_threads_per_block,
                       int* num_regs, size_t* static_lds_bytes,
                       size_t* private_bytes) {
  hipFuncAttributes a;
  hipError_t e = hipFuncGetAttributes(&a, fn);
  if (e != hipSuccess) return rc(e);
  if (max_threads_per_block) *max_threads_per_block = a.maxThreadsPerBlock;
  if (num_regs) *num_regs = a.numRegs;
  if (static_lds_bytes) *static_lds_bytes = a.sharedSizeBytes;
  if (private_bytes) *private_bytes = a.localSizeBytes;
  return 0;
}

int mx_occupancy(const void* fn, int block_size, size_t dyn_lds_bytes,
                 int* blocks_per_cu) {
  return rc(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks_per_cu, fn, block_size, dyn_lds_bytes));
}

int mx_get_last_error(void) { return rc(hipGetLastError()); }

int mx_graph_capture_begin(mx_stream_t stream) {
  return rc(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                  hipStreamCaptureModeThreadLocal));
}

int mx_graph_capture_end(mx_stream_t stream, mx_graph_t* graph) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream), &g);
  *graph       = g;
  return rc(e);
}

int mx_graph_instantiate(mx_graph_exec_t* exec, mx_graph_t graph) {
  hipGraphExec_t x = nullptr;
  hipError_t e =
      hipGraphInstantiate(&x, static_cast<hipGraph_t>(graph), nullptr, nullptr, 0);
  *exec = x;
  return rc(e);
}

int mx_graph_launch(mx_graph_exec_t exec, mx_stream_t stream) {
  return rc(hipGraphLaunch(static_cast<hipGraphExec_t>(exec),
                           static_cast<hipStream_t>(stream)));
}

int mx_graph_exec_destroy(mx_graph_exec_t exec) {
  if (!exec) return 0;
  return rc(hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec)));
}

int mx_graph_destroy(mx_graph_t graph) {
  if (!graph) return 0;
  return rc(hipGraphDestroy(static_cast<hipGraph_t>(graph)));
}

int mx_graph_node_count(mx_graph_t graph, size_t* count) {
  return rc(hipGraphGetNodes(static_cast<hipGraph_t>(graph), nullptr, count));
}

const char* mx_error_string(int status) {
  return hipGetErrorString(static_cast<hipError_t>(status));
}

}  // extern "C"
