// xg_hip.hip -- HIP runtime side of the thin C-ABI layer declared in xg_hip.h.
//
// Replaces the CUDA runtime calls the xGPU host code makes directly
// (reference src/cuda_xengine.cu:208-269 device/stream/event setup,
// :245-252 allocation, :584-672 copies and synchronisation).  Host memory is
// only ever allocated pinned; existing caller memory is never registered.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "xg_hip.h"

static thread_local char g_err[256];

static int check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
  return (int)e;
}

extern "C" {

const char *xgh_last_error(void) { return g_err; }

int xgh_device_count(int *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = (e == hipSuccess) ? n : 0;
  return check(e, "hipGetDeviceCount");
}

int xgh_set_device(int device) { return check(hipSetDevice(device), "hipSetDevice"); }
int xgh_get_device(int *device) { return check(hipGetDevice(device), "hipGetDevice"); }

int xgh_device_name(int device, char *buf, size_t len) {
  hipDeviceProp_t p;
  int rc = check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
  if (rc == 0 && len) {
    strncpy(buf, p.name, len - 1);
    buf[len - 1] = 0;
  }
  return rc;
}

int xgh_device_cu_count(int device, int *cus) {
  return check(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, device),
               "hipDeviceGetAttribute");
}

int xgh_malloc(void **ptr, size_t bytes) { return check(hipMalloc(ptr, bytes), "hipMalloc"); }
int xgh_free(void *ptr) { return ptr ? check(hipFree(ptr), "hipFree") : 0; }
int xgh_host_alloc(void **ptr, size_t bytes) {
  return check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
}
int xgh_host_free(void *ptr) { return ptr ? check(hipHostFree(ptr), "hipHostFree") : 0; }

int xgh_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream) {
  return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
               "hipMemcpyAsync(H2D)");
}
int xgh_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream) {
  return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
               "hipMemcpyAsync(D2H)");
}
int xgh_memcpy_d2d_async(void *dst, const void *src, size_t bytes, void *stream) {
  return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream),
               "hipMemcpyAsync(D2D)");
}
int xgh_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  return check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy(D2H)");
}
int xgh_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  return check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy(H2D)");
}

int xgh_stream_create(void **stream) {
  hipStream_t s = nullptr;
  int rc = check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  *stream = (void *)s;
  return rc;
}
int xgh_stream_destroy(void *stream) {
  return stream ? check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy") : 0;
}
int xgh_stream_sync(void *stream) {
  return check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}
int xgh_stream_wait_event(void *stream, void *event) {
  return check(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0),
               "hipStreamWaitEvent");
}

int xgh_event_create(void **event, int timing) {
  hipEvent_t e = nullptr;
  int rc = check(hipEventCreateWithFlags(&e, timing ? hipEventDefault : hipEventDisableTiming),
                 "hipEventCreate");
  *event = (void *)e;
  return rc;
}
int xgh_event_destroy(void *event) {
  return event ? check(hipEventDestroy((hipEvent_t)event), "hipEventDestroy") : 0;
}
int xgh_event_record(void *event, void *stream) {
  return check(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}
int xgh_event_sync(void *event) {
  return check(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}
int xgh_event_query(void *event) {
  hipError_t e = hipEventQuery((hipEvent_t)event);
  if (e == hipSuccess) return 0;
  if (e == hipErrorNotReady) return 1;
  check(e, "hipEventQuery");
  return -1;
}
int xgh_event_elapsed_ms(float *ms, void *start, void *stop) {
  return check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop),
               "hipEventElapsedTime");
}
int xgh_device_sync(void) { return check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int xgh_graph_begin(void *stream) {
  return check(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
               "hipStreamBeginCapture");
}
int xgh_graph_end(void *stream, void **exec) {
  hipGraph_t g = nullptr;
  int rc = check(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
  if (rc) return rc;
  hipGraphExec_t x = nullptr;
  rc = check(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  *exec = (void *)x;
  return rc;
}
int xgh_graph_launch(void *exec, void *stream) {
  return check(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream), "hipGraphLaunch");
}
int xgh_graph_destroy(void *exec) {
  return exec ? check(hipGraphExecDestroy((hipGraphExec_t)exec), "hipGraphExecDestroy") : 0;
}

}  // extern "C"

// Shared with xg_kernels.hip for launch error reporting.
int xgh_report(hipError_t e, const char *what) { return check(e, what); }
