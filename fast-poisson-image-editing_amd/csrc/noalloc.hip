// pie_grid_solve_noalloc: a GridSolver solve that allocates nothing.
//
// For callers that may not allocate device memory (tight serving loops,
// harnesses that forbid hipMalloc): reset + step + read of a GridSolver in
// one call, with the caller's device buffers as inputs and outputs and every
// intermediate (the two state buffers, the guidance and mask planes, the
// residual partial sums and counters) in `scratch`, a file-scope device
// array of kScratchBytes.  Each GPU the library runs on holds its own copy
// of the array until the process ends.  Host code never uses the array's
// name as an address (in host code it names a host-side stand-in): the
// device address is looked up once per device with hipGetSymbolAddress and
// kept in g_scratch[].
//
// Every call on one GPU shares that GPU's array and finds the previous
// call's contents there, so the call clears what must start cleared (the
// padding of the planes and the counters).  Calls that overlap on one GPU
// (from two host threads, or on two streams without ordering) race on the
// array: callers must serialise them.  The call returns when its outputs
// are written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "host_util.h"

using pie::MarchArgs;
using pie::TileCfg;
using pie::host::fail;
using pie::host::pitch_of;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on first use
// with the device current).
int scratch_of(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(PIE_EUNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PIE_EHIP, std::string("scratch address: ") +
                                hipGetErrorString(e));
    g_scratch[device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[device]);
  return PIE_OK;
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Byte offsets of the intermediates inside the scratch array.
struct Carve {
  size_t xa, xb, g, m, counts, partial, total;
};

Carve carve(long long floats, long long mask_bytes, int nblocks) {
  Carve c{};
  size_t o = 0;
  c.xa = o;
  o += align256(floats * sizeof(float));
  c.xb = o;
  o += align256(floats * sizeof(float));
  c.g = o;
  o += align256(floats * sizeof(float));
  c.m = o;
  o += align256((size_t)mask_bytes);
  c.counts = o;
  o += align256(3 * sizeof(unsigned long long));
  c.partial = o;
  o += align256((size_t)nblocks * 3 * sizeof(double));
  c.total = o;
  return c;
}

}  // namespace

extern "C" {

int pie_grid_solve_noalloc(int h, int w, const int32_t* mask,
                           const float* tgt, const float* grad,
                           int iterations, int schedule, int order,
                           uint8_t* img, float* err, void* stream) {
  if (!mask || !tgt || !grad || !img || !err)
    return fail(PIE_EINVAL, "NULL argument");
  if (h < 3 || w < 3) return fail(PIE_EINVAL, "grid smaller than 3 x 3");
  if (iterations < 0) return fail(PIE_EINVAL, "negative iteration count");
  if (schedule != PIE_JACOBI && schedule != PIE_GS_LEX &&
      schedule != PIE_REDBLACK)
    return fail(PIE_EINVAL, "unknown schedule");
  if (order != PIE_ORDER_UDLR && order != PIE_ORDER_ULRD)
    return fail(PIE_EINVAL, "unknown neighbour order");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const pie::host::GridLayout L = pie::host::grid_layout(h, w);
  const int nb = pie::finalize_blocks(h, w);
  const Carve c = carve(L.floats, L.mask, nb);
  if ((long long)h * L.P > (1LL << 31) - 1 || c.total > kScratchBytes)
    return fail(PIE_EINVAL,
                "grid of " + std::to_string(h) + " x " + std::to_string(w) +
                    " needs " + std::to_string(c.total >> 20) +
                    " MiB of scratch; pie_grid_solve_noalloc holds " +
                    std::to_string(kScratchBytes >> 20) +
                    " MiB (about a 1000 x 1000 crop)");
  int device = 0;
  PIE_HIP(hipGetDevice(&device));
  unsigned char* base = nullptr;
  PIE_CHECK(scratch_of(device, &base));
  float* xa = reinterpret_cast<float*>(base + c.xa);
  float* xb = reinterpret_cast<float*>(base + c.xb);
  float* g = reinterpret_cast<float*>(base + c.g);
  uint8_t* m = base + c.m;
  auto* counts = reinterpret_cast<unsigned long long*>(base + c.counts);
  double* partial = reinterpret_cast<double*>(base + c.partial);
  const size_t fbytes = (size_t)L.floats * sizeof(float);

  // the previous call's contents: state padding zero, mask unsolved,
  // counters 0 (every guidance value is written below: the crop by the
  // deinterleave, the rest by the box kernel)
  PIE_HIP(hipMemsetAsync(xa, 0, fbytes, st));
  PIE_HIP(hipMemsetAsync(m, 0, (size_t)L.mask, st));
  PIE_HIP(hipMemsetAsync(counts, 0, 3 * sizeof(unsigned long long), st));
  // mask plane and its checks (solved count, solved pixels on the border)
  PIE_HIP(pie::launch_mask(mask, m + pie::kPadLeft, h, w, L.P, counts, st));
  // state = target, guidance = grad (both deinterleaved into the planes)
  PIE_HIP(pie::launch_deinterleave(tgt, xa + pie::kPadLeft, L.plane, h, w,
                                   L.P, st));
  PIE_HIP(pie::launch_deinterleave(grad, g + pie::kPadLeft, L.plane, h, w,
                                   L.P, st));
  PIE_HIP(pie::launch_box_guidance(g, m, L.plane, h, L.P, pitch_of(w),
                                   counts + 2, st));
  // counts: solved pixels, solved pixels on the border, solved pixels whose
  // guidance is NaN
  unsigned long long cnt[3] = {0, 0, 0};
  PIE_HIP(hipMemcpyAsync(cnt, counts, sizeof(cnt), hipMemcpyDeviceToHost, st));
  PIE_HIP(hipStreamSynchronize(st));
  if (cnt[1])
    return fail(PIE_EINVAL,
                "mask has solved pixels on the outermost rows/columns");
  if (cnt[2])
    return fail(PIE_EINVAL, std::to_string(cnt[2]) +
                                " guidance values of solved pixels are NaN");
  PIE_HIP(hipMemcpyAsync(xb, xa, fbytes, hipMemcpyDeviceToDevice, st));

  // sweeps: the GridSolver's schedule and its default tile for the size
  int steps = iterations;
  if (schedule == PIE_REDBLACK)
    steps = 2 * iterations;
  else if (schedule == PIE_GS_LEX)
    steps = iterations > 0 ? 2 * (iterations - 1) + (h - 1) + (w - 1) + 1 : 0;
  const TileCfg cfg = pie::host::small_grid_tile((long long)h * w);
  float* cur = xa;
  float* nxt = xb;
  for (int t0 = 0; t0 < steps; t0 += cfg.S) {
    MarchArgs a{};
    a.xin = cur + pie::kPadLeft;
    a.xout = nxt + pie::kPadLeft;
    a.g = g + pie::kPadLeft;
    a.m = m + pie::kPadLeft;
    a.plane = L.plane;
    a.H = h;
    a.W = w;
    a.P = L.P;
    a.t0 = t0;
    a.t_lo = t0;
    a.t_end = steps;
    a.band = 2 * iterations;
    PIE_HIP(pie::launch_march(cfg, schedule, order, a, st));
    std::swap(cur, nxt);
  }
  // outputs: clamped uint8 image and the residual sums (pie_step's pass)
  PIE_HIP(pie::launch_finalize(order, cur + pie::kPadLeft, g + pie::kPadLeft,
                               m + pie::kPadLeft, L.plane, h, w, L.P, img,
                               partial, nb, st));
  PIE_HIP(pie::launch_reduce(partial, nb, err, st));
  PIE_HIP(hipStreamSynchronize(st));
  return PIE_OK;
}

}  // extern "C"
