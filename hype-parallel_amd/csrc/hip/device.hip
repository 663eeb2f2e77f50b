// device.hip -- thin C ABI over the HIP runtime used by the C host layer:
// allocation, copies, streams, graph capture and kernel launches.  Every
// function returns 0 on success or HYPE_E_DEVICE and records the HIP error
// string for hyped_last_error().
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "../../../include/hype2d.h"
#include "../hype_internal.h"
#include "physics.hpp"

namespace hype {
hipError_t launch_stage(const hype_launch& L, hipStream_t s);
int blocks_per_cu(const hype_variant& v);
int pair_modes(int system);
hipError_t launch_cons2prim(int system, const double* U, double* W, long n, Eos e, hipStream_t s);
}  // namespace hype
namespace hype_fc {  // the same kernels compiled with FMA contraction (sweep_fc.hip)
hipError_t launch_stage(const hype_launch& L, hipStream_t s);
int blocks_per_cu(const hype_variant& v);
}  // namespace hype_fc

static thread_local char g_err[256];

static int fail(hipError_t e, const char* what) {
  snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
  return HYPE_E_DEVICE;
}
#define HCHECK(call, what)                  \
  do {                                      \
    hipError_t e_ = (call);                 \
    if (e_ != hipSuccess) return fail(e_, what); \
  } while (0)

extern "C" {

const char* hyped_last_error(void) { return g_err; }

int hyped_init(void) {
  int n = 0;
  HCHECK(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (n < 1) {
    snprintf(g_err, sizeof(g_err), "no HIP device");
    return HYPE_E_DEVICE;
  }
  return HYPE_OK;
}

int hyped_malloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    snprintf(g_err, sizeof(g_err), "hipMalloc(%zu): out of memory", bytes);
    return HYPE_E_NOMEM;
  }
  HCHECK(e, "hipMalloc");
  return HYPE_OK;
}

int hyped_free(void* p) {
  if (p) HCHECK(hipFree(p), "hipFree");
  return HYPE_OK;
}

// pinned host memory (staging of asynchronous copies)
int hyped_host_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    snprintf(g_err, sizeof(g_err), "hipHostMalloc(%zu): out of memory", bytes);
    return HYPE_E_NOMEM;
  }
  HCHECK(e, "hipHostMalloc");
  return HYPE_OK;
}

int hyped_host_free(void* p) {
  if (p) HCHECK(hipHostFree(p), "hipHostFree");
  return HYPE_OK;
}

int hyped_memcpy_h2d(void* dst, const void* src, size_t bytes, void* s) {
  HCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s), "hipMemcpyAsync H2D");
  return HYPE_OK;
}

int hyped_memcpy_d2h(void* dst, const void* src, size_t bytes, void* s) {
  HCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s), "hipMemcpyAsync D2H");
  return HYPE_OK;
}

int hyped_memcpy_d2d(void* dst, const void* src, size_t bytes, void* s) {
  HCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s), "hipMemcpyAsync D2D");
  return HYPE_OK;
}

__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// a fill kernel rather than hipMemsetAsync, so the fill may sit in a graph
int hyped_fill_u64(void* dst, unsigned long long value, size_t count, void* s) {
  if (count == 0) return HYPE_OK;
  const unsigned bs = 256;
  const unsigned grid = (unsigned)((count + bs - 1) / bs);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid), dim3(bs), 0, (hipStream_t)s, (unsigned long long*)dst, value,
                     count);
  HCHECK(hipGetLastError(), "fill_u64_kernel");
  return HYPE_OK;
}

__global__ void copy_kernel(double* dst, const double* src, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int hyped_copy(double* dst, const double* src, long n, void* s) {
  if (n <= 0) return HYPE_OK;
  const unsigned bs = 256;
  const unsigned grid = (unsigned)((n + bs - 1) / bs);
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(bs), 0, (hipStream_t)s, dst, src, n);
  HCHECK(hipGetLastError(), "copy_kernel");
  return HYPE_OK;
}

int hyped_stream_create(void** s) {
  hipStream_t st;
  HCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  *s = (void*)st;
  return HYPE_OK;
}

int hyped_stream_destroy(void* s) {
  if (s) HCHECK(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy");
  return HYPE_OK;
}

int hyped_stream_sync(void* s) {
  HCHECK(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize");
  return HYPE_OK;
}

int hyped_launch(const hype_launch* L, void* s) {
  HCHECK(L->v.fma ? hype_fc::launch_stage(*L, (hipStream_t)s) : hype::launch_stage(*L, (hipStream_t)s),
         "stage kernel launch");
  return HYPE_OK;
}

int hyped_cons2prim(int system, const double* U, double* W, long ncells, double g1, double g2, double p1,
                    double p2, void* s) {
  hype::Eos e{g1, g2, p1, p2, false};
  HCHECK(hype::launch_cons2prim(system, U, W, ncells, e, (hipStream_t)s), "cons2prim launch");
  return HYPE_OK;
}

int hyped_mem_info(size_t* free_bytes, size_t* total_bytes) {
  HCHECK(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
  return HYPE_OK;
}

int hyped_num_cus(void) {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return n;
}

int hyped_blocks_per_cu(const hype_variant* v) { return v->fma ? hype_fc::blocks_per_cu(*v) : hype::blocks_per_cu(*v); }

int hyped_block_waves(const hype_variant* v) { return v->pair ? 2 : 1; }

int hyped_pair_modes(int system) { return hype::pair_modes(system); }  // the same in both builds

int hyped_graph_begin(void* s) {
  HCHECK(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  return HYPE_OK;
}

int hyped_graph_end(void* s, void** exec) {
  hipGraph_t g;
  HCHECK(hipStreamEndCapture((hipStream_t)s, &g), "hipStreamEndCapture");
  hipGraphExec_t ge;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HCHECK(e, "hipGraphInstantiate");
  *exec = (void*)ge;
  return HYPE_OK;
}

int hyped_graph_launch(void* exec, void* s) {
  HCHECK(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)s), "hipGraphLaunch");
  return HYPE_OK;
}

int hyped_graph_destroy(void* exec) {
  if (exec) HCHECK(hipGraphExecDestroy((hipGraphExec_t)exec), "hipGraphExecDestroy");
  return HYPE_OK;
}

int hyped_stream_wait(void* s, void* ev) {
  HCHECK(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)ev, 0), "hipStreamWaitEvent");
  return HYPE_OK;
}

int hyped_event_create_sync(void** ev) {
  hipEvent_t e;
  HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
  *ev = (void*)e;
  return HYPE_OK;
}

int hyped_event_create(void** ev) {
  hipEvent_t e;
  HCHECK(hipEventCreate(&e), "hipEventCreate");
  *ev = (void*)e;
  return HYPE_OK;
}

int hyped_event_destroy(void* ev) {
  if (ev) HCHECK(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy");
  return HYPE_OK;
}

int hyped_event_record(void* ev, void* s) {
  HCHECK(hipEventRecord((hipEvent_t)ev, (hipStream_t)s), "hipEventRecord");
  return HYPE_OK;
}

int hyped_event_sync(void* ev) {
  HCHECK(hipEventSynchronize((hipEvent_t)ev), "hipEventSynchronize");
  return HYPE_OK;
}

int hyped_event_elapsed_ms(void* a, void* b, float* ms) {
  HCHECK(hipEventSynchronize((hipEvent_t)b), "hipEventSynchronize");
  HCHECK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
  return HYPE_OK;
}

int hyped_num_stages(int order) { return order == 0 ? 1 : (order == 1 ? 2 : (order == 2 ? 4 : 10)); }

}  // extern "C"
