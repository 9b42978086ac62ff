// capi_misc.cpp — error reporting, device queries and raw device-memory helpers of the C ABI.
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <new>
#include <stdexcept>

#include "common.h"

namespace gary {

namespace {
thread_local char g_err[1024] = "";
}

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = '\0'; }

const char *last_error() { return g_err; }

}  // namespace gary

namespace {
int use_device(int device) {
  int cur = -1;
  GARY_HIP(hipGetDevice(&cur));
  if (cur != device) GARY_HIP(hipSetDevice(device));
  return GARY_OK;
}
}  // namespace

extern "C" {

const char *gary_last_error(void) { return gary::last_error(); }

int gary_version(void) { return 100; }

int gary_device_count(int *count) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(count, "gary_device_count: null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    c = 0;
  }
  *count = c;
  return GARY_OK;
  GARY_API_END
}

int gary_device_synchronize(int device) {
  GARY_API_BEGIN
  GARY_TRY(use_device(device));
  GARY_HIP(hipDeviceSynchronize());
  return GARY_OK;
  GARY_API_END
}

int gary_malloc(int device, size_t bytes, void **ptr) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ptr, "gary_malloc: null out pointer");
  GARY_TRY(use_device(device));
  *ptr = nullptr;
  if (bytes == 0) return GARY_OK;
  GARY_HIP(hipMalloc(ptr, bytes));
  return GARY_OK;
  GARY_API_END
}

int gary_free(int device, void *ptr) {
  GARY_API_BEGIN
  if (!ptr) return GARY_OK;
  GARY_TRY(use_device(device));
  GARY_HIP(hipFree(ptr));
  return GARY_OK;
  GARY_API_END
}

int gary_memcpy_h2d(int device, void *dst, const void *src, size_t bytes) {
  GARY_API_BEGIN
  if (bytes == 0) return GARY_OK;
  GARY_CHECK_ARG(dst && src, "gary_memcpy_h2d: null pointer");
  GARY_TRY(use_device(device));
  GARY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GARY_OK;
  GARY_API_END
}

int gary_memcpy_d2h(int device, void *dst, const void *src, size_t bytes) {
  GARY_API_BEGIN
  if (bytes == 0) return GARY_OK;
  GARY_CHECK_ARG(dst && src, "gary_memcpy_d2h: null pointer");
  GARY_TRY(use_device(device));
  GARY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

int gary_memcpy_d2d(int device, void *dst, const void *src, size_t bytes) {
  GARY_API_BEGIN
  if (bytes == 0) return GARY_OK;
  GARY_CHECK_ARG(dst && src, "gary_memcpy_d2d: null pointer");
  GARY_TRY(use_device(device));
  GARY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return GARY_OK;
  GARY_API_END
}

int gary_memset(int device, void *dst, int value, size_t bytes) {
  GARY_API_BEGIN
  if (bytes == 0) return GARY_OK;
  GARY_CHECK_ARG(dst, "gary_memset: null pointer");
  GARY_TRY(use_device(device));
  GARY_HIP(hipMemset(dst, value, bytes));
  return GARY_OK;
  GARY_API_END
}

int gary_stream_create(int device, void **stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(stream, "gary_stream_create: null out pointer");
  GARY_TRY(use_device(device));
  hipStream_t s = nullptr;
  GARY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return GARY_OK;
  GARY_API_END
}

int gary_stream_destroy(int device, void *stream) {
  GARY_API_BEGIN
  if (!stream) return GARY_OK;
  GARY_TRY(use_device(device));
  GARY_HIP(hipStreamDestroy((hipStream_t)stream));
  return GARY_OK;
  GARY_API_END
}

int gary_stream_synchronize(int device, void *stream) {
  GARY_API_BEGIN
  GARY_TRY(use_device(device));
  GARY_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"

// ---- workspace cache ------------------------------------------------------------------------
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "workspace.h"

namespace gary {

namespace {
struct Buf {
  void *ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<std::pair<int, void *>, Buf> *g_ws = new std::map<std::pair<int, void *>, Buf>();
}  // namespace

int workspace_get(int device, void *stream, size_t bytes, void **ptr) {
  std::lock_guard<std::mutex> lock(g_ws_mu);
  Buf &b = (*g_ws)[{device, stream}];
  if (b.bytes < bytes) {
    if (b.ptr) {
      // the previous buffer may still be in use by queued work (keys are not always streams)
      GARY_HIP(hipDeviceSynchronize());
      GARY_HIP(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    GARY_HIP(hipMalloc(&b.ptr, want));
    b.bytes = want;
  }
  *ptr = b.ptr;
  return GARY_OK;
}

int current_device(int *device) {
  GARY_HIP(hipGetDevice(device));
  return GARY_OK;
}

namespace {
std::mutex g_tm_mu;
bool g_tm_on = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> *g_tm_pairs =
    new std::vector<std::pair<hipEvent_t, hipEvent_t>>();
std::vector<hipEvent_t> *g_tm_pool = new std::vector<hipEvent_t>();
hipEvent_t g_tm_open = nullptr;

int take_event(hipEvent_t *e) {
  if (!g_tm_pool->empty()) {
    *e = g_tm_pool->back();
    g_tm_pool->pop_back();
    return GARY_OK;
  }
  GARY_HIP(hipEventCreate(e));
  return GARY_OK;
}
}  // namespace

int timing_begin(void *stream) {
  std::lock_guard<std::mutex> lock(g_tm_mu);
  if (!g_tm_on) return GARY_OK;
  GARY_TRY(take_event(&g_tm_open));
  GARY_HIP(hipEventRecord(g_tm_open, (hipStream_t)stream));
  return GARY_OK;
}

int timing_end(void *stream) {
  std::lock_guard<std::mutex> lock(g_tm_mu);
  if (!g_tm_on || !g_tm_open) return GARY_OK;
  hipEvent_t e = nullptr;
  GARY_TRY(take_event(&e));
  GARY_HIP(hipEventRecord(e, (hipStream_t)stream));
  g_tm_pairs->push_back({g_tm_open, e});
  g_tm_open = nullptr;
  return GARY_OK;
}

}  // namespace gary

extern "C" {

int gary_kernel_timing(int enable) {
  GARY_API_BEGIN
  std::lock_guard<std::mutex> lock(gary::g_tm_mu);
  for (auto &p : *gary::g_tm_pairs) {
    gary::g_tm_pool->push_back(p.first);
    gary::g_tm_pool->push_back(p.second);
  }
  gary::g_tm_pairs->clear();
  gary::g_tm_on = enable != 0;
  return GARY_OK;
  GARY_API_END
}

int gary_kernel_timing_read(int *launches, float *total_ms) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(launches && total_ms, "gary_kernel_timing_read: null pointer");
  std::lock_guard<std::mutex> lock(gary::g_tm_mu);
  float tot = 0;
  for (auto &p : *gary::g_tm_pairs) {
    GARY_HIP(hipEventSynchronize(p.second));
    float ms = 0;
    GARY_HIP(hipEventElapsedTime(&ms, p.first, p.second));
    tot += ms;
    gary::g_tm_pool->push_back(p.first);
    gary::g_tm_pool->push_back(p.second);
  }
  *launches = (int)gary::g_tm_pairs->size();
  *total_ms = tot;
  gary::g_tm_pairs->clear();
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"

// ---- event timer ------------------------------------------------------------------------------
struct gary_timer {
  hipEvent_t start = nullptr, stop = nullptr;
};

extern "C" {

int gary_timer_create(gary_timer **out) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(out, "gary_timer_create: null out");
  auto *t = new (std::nothrow) gary_timer();
  if (!t) return gary::set_error(GARY_ERR_OOM, "gary_timer_create: out of memory");
  hipError_t e1 = hipEventCreate(&t->start), e2 = hipEventCreate(&t->stop);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    if (t->start) (void)hipEventDestroy(t->start);
    if (t->stop) (void)hipEventDestroy(t->stop);
    delete t;
    (void)hipGetLastError();
    return gary::set_error(GARY_ERR_HIP, "gary_timer_create: hipEventCreate failed");
  }
  *out = t;
  return GARY_OK;
  GARY_API_END
}

int gary_timer_destroy(gary_timer *t) {
  GARY_API_BEGIN
  if (!t) return GARY_OK;
  (void)hipEventDestroy(t->start);
  (void)hipEventDestroy(t->stop);
  delete t;
  return GARY_OK;
  GARY_API_END
}

int gary_timer_start(gary_timer *t, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(t, "gary_timer_start: null timer");
  GARY_HIP(hipEventRecord(t->start, (hipStream_t)stream));
  return GARY_OK;
  GARY_API_END
}

int gary_timer_stop(gary_timer *t, void *stream, float *ms) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(t && ms, "gary_timer_stop: null pointer");
  GARY_HIP(hipEventRecord(t->stop, (hipStream_t)stream));
  GARY_HIP(hipEventSynchronize(t->stop));
  GARY_HIP(hipEventElapsedTime(ms, t->start, t->stop));
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
