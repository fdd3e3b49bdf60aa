// Host-side helpers shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/fpie_hip.h"
#include "kernels.h"

// Records the message returned by pie_last_error() on this thread and
// returns code (defined in solver.hip).
int pie_fail(int code, const std::string& msg);

#define PIE_HIP(expr)                                                       \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return pie::host::fail(PIE_EHIP, std::string(#expr) + ": " +          \
                                           hipGetErrorString(e_));          \
  } while (0)

#define PIE_CHECK(expr)          \
  do {                           \
    int r_ = (expr);             \
    if (r_ != PIE_OK) return r_; \
  } while (0)

namespace pie {
namespace host {

inline int fail(int code, const std::string& msg) { return pie_fail(code, msg); }

// Owning device allocation.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  hipError_t alloc(size_t count) {
    release();
    n = count;
    return hipMalloc(reinterpret_cast<void**>(&p),
                     std::max<size_t>(count, 1) * sizeof(T));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Row pitch: left padding, the W values, right padding (see kernels.h).
inline int pitch_of(int W) {
  return round_up(kPadLeft + W + kPadRight, 64);
}

inline int upload(const void* host, size_t bytes, void* dev, hipStream_t st) {
  PIE_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
  return PIE_OK;
}

}  // namespace host
}  // namespace pie
