// Host helpers shared by the C-ABI translation units: the thread-local error
// message, HIP status checks and device allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <new>

#include "fieldspice_hip.h"

namespace fs {

// Record a printf-style message for fs_last_error() and return `code`.
int fail(int code, const char* fmt, ...);
const char* last_error();

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(FS_ENOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                hipGetErrorString(e));
  }
  return FS_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace fs

#define FS_HIP(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fs::fail(FS_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                      __FILE__, __LINE__);                                           \
  } while (0)

#define FS_GUARD_BEGIN try {
#define FS_GUARD_END                                                      \
  }                                                                       \
  catch (const std::bad_alloc&) {                                         \
    return fs::fail(FS_ENOMEM, "host allocation failed");                 \
  }                                                                       \
  catch (...) {                                                           \
    return fs::fail(FS_EINVAL, "unexpected internal exception");          \
  }
