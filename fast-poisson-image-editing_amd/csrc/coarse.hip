// Coarse-grid algebra of the learned warm start (pie_warm_net_start): the
// coarse residual, the network's input and output, the damped smoothing of
// the network directions, the coarse operator and the per-channel inner
// products of the flexible conjugate-gradient steps.
//
// Fields are double [hc][wc][3] (channels interleaved), the coarse mask mc
// uint8 [hc][wc]; its outer ring is unsolved, so a solved cell's four
// neighbours always lie inside the grid.  Unsolved cells hold 0 in every
// field.  Sums follow the order of the host restatement of the algorithm
// (neighbours up + down + left + right, left to right).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace pie {
namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

int blocks_for(long long n, int cap) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return (int)(b < cap ? b : cap);
}

__global__ void __launch_bounds__(kBlock)
    coarse_load_kernel(const float* __restrict__ rc,
                       const uint8_t* __restrict__ mc, int n,
                       double* __restrict__ r, double* __restrict__ c) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n;
       k += gridDim.x * kBlock) {
    const bool on = mc[k] != 0;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      r[k * 3 + ch] = on ? (double)rc[k * 3 + ch] : 0.0;
      c[k * 3 + ch] = 0.0;
    }
  }
}

// Network input [hp][wp][8]: channels 0..2 = r / sig (float), channel 6 =
// the mask, the rest 0; cells past (hc, wc) are 0.
__global__ void __launch_bounds__(kBlock)
    net_input_kernel(const double* __restrict__ r,
                     const uint8_t* __restrict__ mc, int hc, int wc, int hp,
                     int wp, const double* __restrict__ sig,
                     float* __restrict__ x) {
  const double s0 = sig[0], s1 = sig[1], s2 = sig[2];
  for (int id = blockIdx.x * kBlock + threadIdx.x; id < hp * wp;
       id += gridDim.x * kBlock) {
    const int i = id / wp, j = id % wp;
    const bool in = i < hc && j < wc;
    const int k = in ? i * wc + j : 0;
    float* o = x + (long long)id * 8;
    o[0] = in ? (float)(r[k * 3 + 0] / s0) : 0.0f;
    o[1] = in ? (float)(r[k * 3 + 1] / s1) : 0.0f;
    o[2] = in ? (float)(r[k * 3 + 2] / s2) : 0.0f;
    o[3] = o[4] = o[5] = o[7] = 0.0f;
    o[6] = (in && mc[k]) ? 1.0f : 0.0f;
  }
}

// z = the network output [hp][wp][3] on the solved cells; 0 everywhere
// when *any is 0 (the residual vanished: no direction)
__global__ void __launch_bounds__(kBlock)
    net_output_kernel(const float* __restrict__ out,
                      const uint8_t* __restrict__ mc, int hc, int wc, int wp,
                      const double* __restrict__ any,
                      double* __restrict__ z) {
  const bool use = *any != 0.0;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < hc * wc;
       k += gridDim.x * kBlock) {
    const int i = k / wc, j = k % wc;
    const bool on = use && mc[k] != 0;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
      z[k * 3 + ch] =
          on ? (double)out[((long long)i * wp + j) * 3 + ch] : 0.0;
  }
}

// One damped sweep (weight 1/2) of the homogeneous coarse problem:
// q' = (q + (up + down + left + right) / 4) / 2 on solved cells.
__global__ void __launch_bounds__(kBlock)
    lowpass_kernel(const double* __restrict__ q,
                   const uint8_t* __restrict__ mc, int hc, int wc,
                   double* __restrict__ out) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < hc * wc;
       k += gridDim.x * kBlock) {
    if (!mc[k]) {
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) out[k * 3 + ch] = 0.0;
      continue;
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const double s = q[(k - wc) * 3 + ch] + q[(k + wc) * 3 + ch] +
                       q[(k - 1) * 3 + ch] + q[(k + 1) * 3 + ch];
      const double avg = 0.25 * s;
      out[k * 3 + ch] = 0.5 * (q[k * 3 + ch] + avg);
    }
  }
}

// A_c p = 4 p - (up + down + left + right) on solved cells
__global__ void __launch_bounds__(kBlock)
    coarse_apply_kernel(const double* __restrict__ p,
                        const uint8_t* __restrict__ mc, int hc, int wc,
                        double* __restrict__ out) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < hc * wc;
       k += gridDim.x * kBlock) {
    if (!mc[k]) {
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) out[k * 3 + ch] = 0.0;
      continue;
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const double s = p[(k - wc) * 3 + ch] + p[(k + wc) * 3 + ch] +
                       p[(k - 1) * 3 + ch] + p[(k + 1) * 3 + ch];
      out[k * 3 + ch] = 4.0 * p[k * 3 + ch] - s;
    }
  }
}

// Per-channel inner products of up to kMaxDots field pairs: block partial
// sums partial[block][pair][3] (sum_partials_kernel adds the blocks in
// order).
__global__ void __launch_bounds__(kBlock)
    coarse_dots_kernel(CoarseDots d, int n, double* __restrict__ partial) {
  double acc[kMaxDots][3];
#pragma unroll
  for (int q = 0; q < kMaxDots; ++q)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) acc[q][ch] = 0.0;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n;
       k += gridDim.x * kBlock) {
#pragma unroll
    for (int q = 0; q < kMaxDots; ++q) {
      if (q >= d.count) break;
#pragma unroll
      for (int ch = 0; ch < 3; ++ch)
        acc[q][ch] += d.a[q][k * 3 + ch] * d.b[q][k * 3 + ch];
    }
  }
  __shared__ double red[kBlock / kWave][kMaxDots][3];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < kMaxDots; ++q)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      double v = acc[q][ch];
#pragma unroll
      for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) red[wave][q][ch] = v;
    }
  __syncthreads();
  if (threadIdx.x < kMaxDots * 3) {
    const int q = threadIdx.x / 3, ch = threadIdx.x % 3;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) v += red[w][q][ch];
    partial[((long long)blockIdx.x * kMaxDots + q) * 3 + ch] = v;
  }
}

// y += a_c x (a: three doubles in device memory)
__global__ void __launch_bounds__(kBlock)
    coarse_axpy_kernel(double* __restrict__ y, const double* __restrict__ x,
                       const double* __restrict__ a, int n) {
  const double a0 = a[0], a1 = a[1], a2 = a[2];
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n;
       k += gridDim.x * kBlock) {
    y[k * 3 + 0] = y[k * 3 + 0] + a0 * x[k * 3 + 0];
    y[k * 3 + 1] = y[k * 3 + 1] + a1 * x[k * 3 + 1];
    y[k * 3 + 2] = y[k * 3 + 2] + a2 * x[k * 3 + 2];
  }
}

__global__ void __launch_bounds__(kBlock)
    coarse_to_float_kernel(const double* __restrict__ c, int n3,
                           float* __restrict__ out) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n3;
       k += gridDim.x * kBlock)
    out[k] = (float)c[k];
}

// v = float(double(v) * a_c); *maxbits = max |v| as float bits (atomicMax
// on the bits of non-negative floats orders them as the floats)
__global__ void __launch_bounds__(kBlock)
    coarse_scale_kernel(float* __restrict__ v,
                        const double* __restrict__ alpha, int n,
                        unsigned* __restrict__ maxbits) {
  const double a0 = alpha[0], a1 = alpha[1], a2 = alpha[2];
  unsigned m = 0;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n;
       k += gridDim.x * kBlock) {
    const double a[3] = {a0, a1, a2};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float s = (float)((double)v[k * 3 + ch] * a[ch]);
      v[k * 3 + ch] = s;
      m = max(m, __float_as_uint(fabsf(s)));
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
    m = max(m, (unsigned)__shfl_xor((int)m, o));
  if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicMax(maxbits, m);
}

// out[k] = sum over blocks b of partial[b * width + k], k < width <= 12:
// thread t adds blocks t, t + 256, ... in order, then a fixed tree over the
// threads (deterministic; the grouping differs from a sequential sum only
// in the last bits)
constexpr int kSumWidth = kMaxDots * 3;
__global__ void __launch_bounds__(kBlock)
    sum_partials_kernel(const double* __restrict__ partial, int nb, int width,
                        double* __restrict__ out) {
  __shared__ double red[kSumWidth][kBlock];
  double acc[kSumWidth];
#pragma unroll
  for (int k = 0; k < kSumWidth; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock)
#pragma unroll
    for (int k = 0; k < kSumWidth; ++k)
      if (k < width) acc[k] += partial[(long long)b * width + k];
#pragma unroll
  for (int k = 0; k < kSumWidth; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < kSumWidth; ++k)
        red[k][threadIdx.x] = red[k][threadIdx.x] + red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < width) out[threadIdx.x] = red[threadIdx.x][0];
}

// The scalar steps of the flexible CG on the device (one thread), on the
// scalars of WarmScalars (see kernels.h)
__global__ void warm_scalar_kernel(double* __restrict__ sc, int op, int k,
                                   double cells) {
  if (threadIdx.x != 0) return;
  const double* d = sc + kWsDots;
  if (op == kWarmSig) {
    // per-channel rms of r over the solved cells (1 where it vanishes);
    // any = whether some channel is non-zero
    bool any = false;
    for (int ch = 0; ch < 3; ++ch) {
      double v = cells > 0.0 ? sqrt(d[ch] / cells) : 0.0;
      any = any || v > 0.0;
      sc[kWsSig + ch] = v > 0.0 ? v : 1.0;
    }
    sc[kWsAny] = any ? 1.0 : 0.0;
  } else if (op == kWarmOrth) {
    // coefficient removing the A_c-component along direction k
    for (int ch = 0; ch < 3; ++ch)
      sc[kWsCoef + ch] = -(d[ch] / sc[kWsPap + 3 * k + ch]);
  } else if (op == kWarmStep) {
    // <z, A z>, <z, z>, <r, z> -> the energy-optimal step along z
    for (int ch = 0; ch < 3; ++ch) {
      const double zaz = d[ch], zz = d[3 + ch], rz = d[6 + ch];
      const bool ok = zaz > 1e-30 * fmax(zz, 1e-300);
      const double a = ok ? rz / zaz : 0.0;
      sc[kWsStep + ch] = a;
      sc[kWsNegStep + ch] = -a;
      sc[kWsPap + 3 * k + ch] = ok ? zaz : 1.0;
    }
  } else if (op == kWarmAlpha) {
    // fine line search: alpha = max(0, <r, d> / <d, A d>)
    for (int ch = 0; ch < 3; ++ch) {
      const double den = d[3 + ch];
      const double a = den > 0.0 ? d[ch] / den : 0.0;
      sc[kWsAlpha + ch] = fmax(a, 0.0);
    }
  }
}

}  // namespace

hipError_t launch_sum_partials(const double* partial, int nb, int width,
                               double* out, hipStream_t st) {
  if (width > kSumWidth) return hipErrorInvalidValue;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kBlock), 0, st,
                     partial, nb, width, out);
  return hipGetLastError();
}

hipError_t launch_warm_scalar(double* sc, int op, int k, double cells,
                              hipStream_t st) {
  hipLaunchKernelGGL(warm_scalar_kernel, dim3(1), dim3(kWave), 0, st, sc, op,
                     k, cells);
  return hipGetLastError();
}

int coarse_dots_blocks(int n) { return blocks_for(n, 256); }

hipError_t launch_coarse_load(const float* rc, const uint8_t* mc, int n,
                              double* r, double* c, hipStream_t st) {
  hipLaunchKernelGGL(coarse_load_kernel, dim3(blocks_for(n, 4096)),
                     dim3(kBlock), 0, st, rc, mc, n, r, c);
  return hipGetLastError();
}

hipError_t launch_net_input(const double* r, const uint8_t* mc, int hc,
                            int wc, int hp, int wp, const double* sig,
                            float* x, hipStream_t st) {
  hipLaunchKernelGGL(net_input_kernel, dim3(blocks_for((long long)hp * wp, 4096)),
                     dim3(kBlock), 0, st, r, mc, hc, wc, hp, wp, sig, x);
  return hipGetLastError();
}

hipError_t launch_net_output(const float* out, const uint8_t* mc, int hc,
                             int wc, int wp, const double* any, double* z,
                             hipStream_t st) {
  hipLaunchKernelGGL(net_output_kernel, dim3(blocks_for((long long)hc * wc, 4096)),
                     dim3(kBlock), 0, st, out, mc, hc, wc, wp, any, z);
  return hipGetLastError();
}

hipError_t launch_lowpass(const double* q, const uint8_t* mc, int hc, int wc,
                          double* out, hipStream_t st) {
  hipLaunchKernelGGL(lowpass_kernel, dim3(blocks_for((long long)hc * wc, 4096)),
                     dim3(kBlock), 0, st, q, mc, hc, wc, out);
  return hipGetLastError();
}

hipError_t launch_coarse_apply(const double* p, const uint8_t* mc, int hc,
                               int wc, double* out, hipStream_t st) {
  hipLaunchKernelGGL(coarse_apply_kernel,
                     dim3(blocks_for((long long)hc * wc, 4096)), dim3(kBlock),
                     0, st, p, mc, hc, wc, out);
  return hipGetLastError();
}

hipError_t launch_coarse_dots(const CoarseDots& d, int n, double* partial,
                              hipStream_t st) {
  hipLaunchKernelGGL(coarse_dots_kernel, dim3(coarse_dots_blocks(n)),
                     dim3(kBlock), 0, st, d, n, partial);
  return hipGetLastError();
}

hipError_t launch_coarse_axpy(double* y, const double* x, const double* a,
                              int n, hipStream_t st) {
  hipLaunchKernelGGL(coarse_axpy_kernel, dim3(blocks_for(n, 4096)),
                     dim3(kBlock), 0, st, y, x, a, n);
  return hipGetLastError();
}

hipError_t launch_coarse_to_float(const double* c, int n3, float* out,
                                  hipStream_t st) {
  hipLaunchKernelGGL(coarse_to_float_kernel, dim3(blocks_for(n3, 4096)),
                     dim3(kBlock), 0, st, c, n3, out);
  return hipGetLastError();
}

hipError_t launch_coarse_scale(float* v, const double* alpha, int n,
                               unsigned* maxbits, hipStream_t st) {
  hipLaunchKernelGGL(coarse_scale_kernel, dim3(blocks_for(n, 4096)),
                     dim3(kBlock), 0, st, v, alpha, n, maxbits);
  return hipGetLastError();
}

}  // namespace pie
