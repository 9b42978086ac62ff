// common.h — host-side status / error plumbing shared by every translation unit of libgary.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "gary.h"

namespace gary {

// Sets the thread-local message returned by gary_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

}  // namespace gary

// Evaluate a HIP call; on failure record the message and return GARY_ERR_HIP (or OOM).
#define GARY_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      (void)hipGetLastError();                                                               \
      return ::gary::set_error(e_ == hipErrorOutOfMemory ? GARY_ERR_OOM : GARY_ERR_HIP,       \
                               "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),         \
                               __FILE__, __LINE__);                                          \
    }                                                                                        \
  } while (0)

// Propagate a non-OK status from a helper.
#define GARY_TRY(expr)          \
  do {                          \
    int s_ = (expr);            \
    if (s_ != GARY_OK) return s_; \
  } while (0)

#define GARY_CHECK_ARG(cond, ...)                                   \
  do {                                                              \
    if (!(cond)) return ::gary::set_error(GARY_ERR_INVALID, __VA_ARGS__); \
  } while (0)

// Convert a C++ exception escaping a helper into a status (no exception crosses the ABI).
#define GARY_API_BEGIN try {
#define GARY_API_END                                                          \
  }                                                                           \
  catch (const std::bad_alloc &) {                                            \
    return ::gary::set_error(GARY_ERR_OOM, "host allocation failed");         \
  }                                                                           \
  catch (const std::exception &ex_) {                                         \
    return ::gary::set_error(GARY_ERR_INVALID, "exception: %s", ex_.what());  \
  }                                                                           \
  catch (...) {                                                               \
    return ::gary::set_error(GARY_ERR_INVALID, "unknown exception");          \
  }
