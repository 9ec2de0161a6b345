// oit_device_api.hip - device plumbing of the C ABI: allocation, copies,
// streams and events on the library's HIP runtime, so that hosts (and the
// Python tests) need no other HIP binding in the process.
#include <hip/hip_runtime.h>

#include "oit.h"

namespace {
int st(hipError_t e) {
  if (e == hipSuccess) return OIT_OK;
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory) return OIT_ERR_OOM;
  if (e == hipErrorInvalidValue) return OIT_ERR_INVALID_ARG;
  if (e == hipErrorNotReady) return OIT_ERR_NOT_READY;
  return OIT_ERR_DEVICE;
}
}  // namespace

extern "C" {

int oit_device_count(int* out) {
  if (!out) return OIT_ERR_INVALID_ARG;
  *out = 0;
  return st(hipGetDeviceCount(out));
}

int oit_set_device(int device) { return st(hipSetDevice(device)); }

int oit_device_alloc(void** out, uint64_t bytes) {
  if (!out) return OIT_ERR_INVALID_ARG;
  *out = nullptr;
  if (bytes == 0) return OIT_OK;
  return st(hipMalloc(out, bytes));
}

int oit_device_free(void* ptr) { return ptr ? st(hipFree(ptr)) : OIT_OK; }

int oit_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes == 0) return OIT_OK;
  if (!dst || !src) return OIT_ERR_INVALID_ARG;
  return st(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
}

int oit_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes == 0) return OIT_OK;
  if (!dst || !src) return OIT_ERR_INVALID_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = st(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  if (rc) return rc;
  return st(hipStreamSynchronize(s));
}

int oit_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes == 0) return OIT_OK;
  if (!dst || !src) return OIT_ERR_INVALID_ARG;
  return st(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
}

int oit_stream_create(void** out) {
  if (!out) return OIT_ERR_INVALID_ARG;
  hipStream_t s = nullptr;
  int rc = st(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = s;
  return rc;
}

int oit_stream_destroy(void* stream) { return stream ? st(hipStreamDestroy(static_cast<hipStream_t>(stream))) : OIT_OK; }

int oit_stream_synchronize(void* stream) { return st(hipStreamSynchronize(static_cast<hipStream_t>(stream))); }

int oit_event_create(void** out) {
  if (!out) return OIT_ERR_INVALID_ARG;
  hipEvent_t e = nullptr;
  int rc = st(hipEventCreate(&e));
  *out = e;
  return rc;
}

int oit_event_destroy(void* event) { return event ? st(hipEventDestroy(static_cast<hipEvent_t>(event))) : OIT_OK; }

int oit_event_record(void* event, void* stream) {
  if (!event) return OIT_ERR_INVALID_ARG;
  return st(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
}

int oit_event_synchronize(void* event) {
  if (!event) return OIT_ERR_INVALID_ARG;
  return st(hipEventSynchronize(static_cast<hipEvent_t>(event)));
}

int oit_event_elapsed_ms(float* out, void* start, void* stop) {
  if (!out || !start || !stop) return OIT_ERR_INVALID_ARG;
  return st(hipEventElapsedTime(out, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
}

}  // extern "C"
