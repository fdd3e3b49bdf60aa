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
                     int wp, double s0, double s1, double s2,
                     float* __restrict__ x) {
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

// z = the network output [hp][wp][3] on the solved cells
__global__ void __launch_bounds__(kBlock)
    net_output_kernel(const float* __restrict__ out,
                      const uint8_t* __restrict__ mc, int hc, int wc, int wp,
                      double* __restrict__ z) {
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < hc * wc;
       k += gridDim.x * kBlock) {
    const int i = k / wc, j = k % wc;
    const bool on = mc[k] != 0;
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
// sums partial[block][pair][3] (the host adds the blocks in order).
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

// y += a_c x
__global__ void __launch_bounds__(kBlock)
    coarse_axpy_kernel(double* __restrict__ y, const double* __restrict__ x,
                       double a0, double a1, double a2, int n) {
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
    coarse_scale_kernel(float* __restrict__ v, double a0, double a1,
                        double a2, int n, unsigned* __restrict__ maxbits) {
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

}  // namespace

int coarse_dots_blocks(int n) { return blocks_for(n, 256); }

hipError_t launch_coarse_load(const float* rc, const uint8_t* mc, int n,
                              double* r, double* c, hipStream_t st) {
  hipLaunchKernelGGL(coarse_load_kernel, dim3(blocks_for(n, 4096)),
                     dim3(kBlock), 0, st, rc, mc, n, r, c);
  return hipGetLastError();
}

hipError_t launch_net_input(const double* r, const uint8_t* mc, int hc,
                            int wc, int hp, int wp, const double sig[3],
                            float* x, hipStream_t st) {
  hipLaunchKernelGGL(net_input_kernel, dim3(blocks_for((long long)hp * wp, 4096)),
                     dim3(kBlock), 0, st, r, mc, hc, wc, hp, wp, sig[0],
                     sig[1], sig[2], x);
  return hipGetLastError();
}

hipError_t launch_net_output(const float* out, const uint8_t* mc, int hc,
                             int wc, int wp, double* z, hipStream_t st) {
  hipLaunchKernelGGL(net_output_kernel, dim3(blocks_for((long long)hc * wc, 4096)),
                     dim3(kBlock), 0, st, out, mc, hc, wc, wp, z);
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

hipError_t launch_coarse_axpy(double* y, const double* x, const double a[3],
                              int n, hipStream_t st) {
  hipLaunchKernelGGL(coarse_axpy_kernel, dim3(blocks_for(n, 4096)),
                     dim3(kBlock), 0, st, y, x, a[0], a[1], a[2], n);
  return hipGetLastError();
}

hipError_t launch_coarse_to_float(const double* c, int n3, float* out,
                                  hipStream_t st) {
  hipLaunchKernelGGL(coarse_to_float_kernel, dim3(blocks_for(n3, 4096)),
                     dim3(kBlock), 0, st, c, n3, out);
  return hipGetLastError();
}

hipError_t launch_coarse_scale(float* v, const double a[3], int n,
                               unsigned* maxbits, hipStream_t st) {
  hipLaunchKernelGGL(coarse_scale_kernel, dim3(blocks_for(n, 4096)),
                     dim3(kBlock), 0, st, v, a[0], a[1], a[2], n, maxbits);
  return hipGetLastError();
}

}  // namespace pie
