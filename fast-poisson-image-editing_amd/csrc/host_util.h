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
  // n records the capacity only once the allocation succeeded, so a failed
  // alloc leaves an empty buffer (p null, n 0) that the next grow-if-needed
  // test allocates again.
  hipError_t alloc(size_t count) {
    release();
    T* q = nullptr;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&q),
                                   std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    p = q;
    n = count;
    return hipSuccess;
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

// Layout of a solve's three float channels and mask on an H x W grid.
// Row-interleaved (default): row i holds channel 0, 1, 2 segments of
// pitch_of(W) floats each, so the channel stride ("plane") is one segment
// and the row pitch P three; the loads of one row of all channels then fall
// in one stretch of memory.  Planar: channel planes of H x pitch_of(W).
// The mask uses the same row pitch (only its first segment is used).
#ifndef PIE_ROW_INTERLEAVED
#define PIE_ROW_INTERLEAVED 1
#endif
struct GridLayout {
  int P;             // row pitch (elements)
  long long plane;   // channel stride (elements)
  long long floats;  // elements of one three-channel buffer
  long long mask;    // bytes of the mask buffer
};

inline GridLayout grid_layout(int H, int W) {
  const int p = pitch_of(W);
  if (PIE_ROW_INTERLEAVED)
    return {3 * p, p, 3LL * H * p, 3LL * H * p};
  return {p, (long long)H * p, 3LL * H * p, (long long)H * p};
}

// Sweep tile below 2^22 pixels.  Short row ranges keep enough wave tasks in
// flight: measured on MI355X, Jacobi sweeps of a 1024^2 composite take
// 3.1 us with (8, 1, 16) one channel a task, 3.5 us with all three, 4.9 us
// with (8, 2, 32) and 6.5 us with (8, 1, 64); a 512^2 one 1.8 us against
// 2.4 us; a 2048^2 one 6.3 us with (8, 2, 32) against 8.4 us with
// (8, 1, 16) and 8.3 us with (8, 1, 64).
inline TileCfg small_grid_tile(long long pixels) {
  if (pixels <= (1LL << 21)) return TileCfg{8, 1, 16, 1};
  return TileCfg{8, 2, 32};
}

inline int upload(const void* host, size_t bytes, void* dev, hipStream_t st) {
  PIE_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
  return PIE_OK;
}

}  // namespace host

// unet.hip: the warm-start network's forward pass on device buffers,
// stream-ordered against `st` (no host wait)
int unet_forward_on_stream(pie_unet* u, int h, int w, const float* x,
                           float* out, hipStream_t st);

}  // namespace pie
