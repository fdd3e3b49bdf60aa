// gfx950 kernels of the Poisson solver other than the sweep kernel: layout
// conversion, the residual/output pass, the EquSolver gather path, and the
// dispatch of the sweep kernel over the compiled tile variants.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "march.h"

namespace pie {

namespace {

using namespace detail;

// ------------------------------------------------------------ layout ----

__global__ void __launch_bounds__(kBlock)
    deinterleave_kernel(const float* __restrict__ src, float* __restrict__ dst,
                        long long plane, int H, int W, int P) {
  const long long n = (long long)H * W;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const long long o = (long long)i * P + j;
#pragma unroll
    for (int c = 0; c < 3; ++c) dst[c * plane + o] = src[id * 3 + c];
  }
}

__global__ void __launch_bounds__(kBlock)
    interleave_kernel(const float* __restrict__ src, float* __restrict__ dst,
                      long long plane, int H, int W, int P) {
  const long long n = (long long)H * W;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const long long o = (long long)i * P + j;
#pragma unroll
    for (int c = 0; c < 3; ++c) dst[id * 3 + c] = src[c * plane + o];
  }
}

// fpie/process.py:113-122: the source difference ('src'), the mean of both
// ('avg'), or the larger in magnitude, ties keeping the source ('max').
__device__ __forceinline__ float mixgrad(float a, float b, int mode) {
  if (mode == 0) return a;
  if (mode == 1) return (a + b) / 2.0f;
  return fabsf(a) < fabsf(b) ? b : a;
}

__global__ void __launch_bounds__(kBlock)
    divergence_kernel(const uint8_t* __restrict__ src,
                      const uint8_t* __restrict__ tgt,
                      const uint8_t* __restrict__ m, float* __restrict__ x,
                      float* __restrict__ g, long long plane, int H, int W,
                      int P, int mode, int equ) {
  const long long n = (long long)H * W;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const long long o = (long long)i * P + j;
    const bool solved = m[o] != 0;  // never on the outermost rows/columns
    const long long nb[4] = {id - W, id + W, id - 1, id + 1};  // U, D, L, R
    const long long nm[4] = {o - P, o + P, o - 1, o + 1};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float tc = (float)tgt[id * 3 + c];
      // equation form: unsolved pixels hold X[0] = 0 (id 0)
      x[c * plane + o] = (equ && !solved) ? 0.0f : tc;
      float acc = 0.0f;
      if (solved) {
        const float sc = (float)src[id * 3 + c];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          acc = acc + mixgrad(sc - (float)src[nb[k] * 3 + c],
                              tc - (float)tgt[nb[k] * 3 + c], mode);
        // equation form: B adds the target of every unsolved neighbour
        if (equ) {
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (!m[nm[k]]) acc = acc + (float)tgt[nb[k] * 3 + c];
        }
      }
      g[c * plane + o] = acc;
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    keep_unsolved_kernel(uint8_t* __restrict__ img,
                         const uint8_t* __restrict__ tgt,
                         const uint8_t* __restrict__ m, int H, int W, int P) {
  const long long n = (long long)H * W;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    if (!m[(long long)i * P + j]) {
#pragma unroll
      for (int c = 0; c < 3; ++c) img[id * 3 + c] = tgt[id * 3 + c];
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    mask_kernel(const int32_t* __restrict__ src, uint8_t* __restrict__ dst,
                int H, int W, int P, unsigned long long* counts) {
  const long long n = (long long)H * W;
  unsigned long long solved = 0, border = 0;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const uint8_t v = src[id] != 0;
    dst[(long long)i * P + j] = v;
    solved += v;
    if (v && (i == 0 || j == 0 || i == H - 1 || j == W - 1)) ++border;
  }
  // wave sums through shuffles, then one atomic per wave
  for (int off = kWave / 2; off > 0; off >>= 1) {
    solved += __shfl_down(solved, off, kWave);
    border += __shfl_down(border, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    if (solved) atomicAdd(&counts[0], solved);
    if (border) atomicAdd(&counts[1], border);
  }
}

__global__ void __launch_bounds__(kBlock)
    fill3_kernel(float* __restrict__ dst, long long n, int P, long long plane,
                 float v0, float v1, float v2) {
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    // channel of element id: planar (plane >= P) or row-interleaved layout
    const long long c = plane >= P ? id / plane : (id % P) / plane;
    dst[id] = c == 0 ? v0 : (c == 1 ? v1 : v2);
  }
}

__global__ void __launch_bounds__(kBlock)
    equ_scatter_kernel(const float* __restrict__ X, const float* __restrict__ B,
                       const int32_t* __restrict__ lin, int N,
                       float* __restrict__ x, float* __restrict__ g,
                       long long plane) {
  for (int id = 1 + blockIdx.x * kBlock + threadIdx.x; id < N;
       id += gridDim.x * kBlock) {
    const long long o = lin[id];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c * plane + o] = X[(long long)id * 3 + c];
      g[c * plane + o] = B[(long long)id * 3 + c];
    }
  }
}

__device__ __forceinline__ uint8_t clamp_u8(float v) {
  // fpie/core/gcc/equ.cc:82-84: clamp to [0, 255], then truncate
  return (uint8_t)(v < 0.0f ? 0.0f : v > 255.0f ? 255.0f : v);
}

__global__ void __launch_bounds__(kBlock)
    equ_gather_kernel(const float* __restrict__ x,
                      const int32_t* __restrict__ lin, int N, long long plane,
                      float x00, float x01, float x02, uint8_t* out_u8,
                      float* out_f32) {
  for (int id = blockIdx.x * kBlock + threadIdx.x; id < N;
       id += gridDim.x * kBlock) {
    float v[3];
    if (id == 0) {
      v[0] = x00;
      v[1] = x01;
      v[2] = x02;
    } else {
      const long long o = lin[id];
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = x[c * plane + o];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (out_u8) out_u8[(long long)id * 3 + c] = clamp_u8(v[c]);
      if (out_f32) out_f32[(long long)id * 3 + c] = v[c];
    }
  }
}

// Block-wide fixed-order sum of three doubles; thread 0 gets the totals.
__device__ __forceinline__ void block_sum3(double (&v)[3]) {
  __shared__ double red[3][kBlock];
#pragma unroll
  for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        red[c][threadIdx.x] = red[c][threadIdx.x] + red[c][threadIdx.x + s];
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = red[c][0];
}

// Residual |(b + n1 + n2 + n3 + n4) - 4x| of the reference's calc_error
// (fpie/core/gcc/equ.cc:56-72, gcc/grid.cc:38-55): the neighbour sum in
// float in the solver's order, the subtraction and absolute value in
// double; the terms are summed in double in a fixed order.
template <int ORDER>
__global__ void __launch_bounds__(kBlock)
    finalize_kernel(const float* __restrict__ x, const float* __restrict__ g,
                    const uint8_t* __restrict__ m, long long plane, int H,
                    int W, int P, uint8_t* __restrict__ img,
                    double* __restrict__ partial) {
  double acc[3] = {0.0, 0.0, 0.0};
  const long long n = (long long)H * W;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const long long o = (long long)i * P + j;
    if (m[o]) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float* xc = x + c * plane;
        float s = g[c * plane + o];
        if (ORDER == ORDER_UDLR) {
          s = s + xc[o - P];
          s = s + xc[o + P];
          s = s + xc[o - 1];
          s = s + xc[o + 1];
        } else {
          s = s + xc[o - P];
          s = s + xc[o - 1];
          s = s + xc[o + 1];
          s = s + xc[o + P];
        }
        acc[c] += fabs((double)s - (double)xc[o] * 4.0);
      }
    }
    if (img) {
#pragma unroll
      for (int c = 0; c < 3; ++c) img[id * 3 + c] = clamp_u8(x[c * plane + o]);
    }
  }
  block_sum3(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) partial[blockIdx.x * 3 + c] = acc[c];
}

__global__ void __launch_bounds__(kBlock)
    reduce_kernel(const double* __restrict__ partial, int n, float* err) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += kBlock)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] += partial[i * 3 + c];
  block_sum3(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) err[c] = (float)acc[c];
}

// ------------------------------------------- EquSolver on arbitrary A ---

__global__ void __launch_bounds__(kBlock)
    equ_pack_kernel(const float* __restrict__ X, float4* __restrict__ out,
                    int N) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < N;
       i += gridDim.x * kBlock)
    out[i] = make_float4(X[3LL * i], X[3LL * i + 1], X[3LL * i + 2], 0.0f);
}

// Jacobi on the id system: x'[i] = (B[i] + x[A0] + x[A1] + x[A2] + x[A3])
// / 4 for ids 1..N-1 (the order of fpie/core/gcc/equ.cc:48-53 and
// np_solver.py:33-41); id 0 keeps its value.
__global__ void __launch_bounds__(kBlock)
    equ_jacobi_kernel(const int4* __restrict__ A, const float4* __restrict__ B,
                      const float4* __restrict__ xin,
                      float4* __restrict__ xout, int N) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  if (i == 0) {
    xout[0] = xin[0];
    return;
  }
  const int4 a = A[i];
  const float4 b = B[i], u = xin[a.x], d = xin[a.y], l = xin[a.z],
               r = xin[a.w];
  float4 o;
  o.x = ((((b.x + u.x) + d.x) + l.x) + r.x) * 0.25f;
  o.y = ((((b.y + u.y) + d.y) + l.y) + r.y) * 0.25f;
  o.z = ((((b.z + u.z) + d.z) + l.z) + r.z) * 0.25f;
  o.w = 0.0f;
  xout[i] = o;
}

__global__ void __launch_bounds__(kBlock)
    equ_finalize_kernel(const int4* __restrict__ A,
                        const float4* __restrict__ B,
                        const float4* __restrict__ x, int N,
                        uint8_t* __restrict__ img, float* __restrict__ state,
                        double* __restrict__ partial) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < N;
       i += gridDim.x * kBlock) {
    const float4 xi = x[i];
    const float v[3] = {xi.x, xi.y, xi.z};
    if (i > 0) {
      const int4 a = A[i];
      const float4 b = B[i], u = x[a.x], d = x[a.y], l = x[a.z], r = x[a.w];
      const float bs[3] = {b.x, b.y, b.z}, us[3] = {u.x, u.y, u.z},
                  ds[3] = {d.x, d.y, d.z}, ls[3] = {l.x, l.y, l.z},
                  rs[3] = {r.x, r.y, r.z};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float s = bs[c];
        s = s + us[c];
        s = s + ds[c];
        s = s + ls[c];
        s = s + rs[c];
        acc[c] += fabs((double)s - (double)v[c] * 4.0);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (img) img[3LL * i + c] = clamp_u8(v[c]);
      if (state) state[3LL * i + c] = v[c];
    }
  }
  block_sum3(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) partial[blockIdx.x * 3 + c] = acc[c];
}

// After each batched launch: advance every live item by S sweeps, stop it
// when its last sweep changed no value by tol or more, or when its sweep
// budget is used up; count the live items and reset the work counter.
__global__ void __launch_bounds__(kBlock)
    batch_epilogue_kernel(BatchItem* items, int n, int S, unsigned* queue,
                          unsigned* solo_queue, unsigned* live,
                          unsigned* shared_live) {
  __shared__ unsigned cnt[2][kBlock];
  const unsigned long long now = wall_clock64();
  unsigned c = 0, cs = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) {
    BatchItem& it = items[i];
    if (it.done) continue;
    // solo items were advanced by their work-group inside the round
    if (!it.solo) batch_advance(it, S, now);
    c += !it.done;
    cs += !it.done && !it.solo;
  }
  cnt[0][threadIdx.x] = c;
  cnt[1][threadIdx.x] = cs;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      cnt[0][threadIdx.x] += cnt[0][threadIdx.x + s];
      cnt[1][threadIdx.x] += cnt[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *live = cnt[0][0];
    *shared_live = cnt[1][0];
    *queue = 0u;
    *solo_queue = 0u;
  }
}

__global__ void clock_kernel(unsigned long long* out) {
  if (threadIdx.x == 0) *out = wall_clock64();
}

// ------------------------------------------- coarse-grid correction ----

// float sum g + four neighbours in the solver's order (no rounding other
// than the float adds, as the sweep kernel)
template <int ORDER>
__device__ __forceinline__ float neighbour_sum(const float* __restrict__ xc,
                                               float gv, long long o, int P) {
  float s = gv;
  if (ORDER == ORDER_UDLR) {
    s = s + xc[o - P];
    s = s + xc[o + P];
    s = s + xc[o - 1];
    s = s + xc[o + 1];
  } else {
    s = s + xc[o - P];
    s = s + xc[o - 1];
    s = s + xc[o + 1];
    s = s + xc[o + P];
  }
  return s;
}

// out[I][J][c] = sum over the solved pixels of the f x f block (I, J) of the
// signed residual g + four neighbours - 4 x, each term and the sums in double
// (exact for float inputs: a float32 sum would round to the state's own
// fixed point and hide smooth errors below half an ulp per pixel), in
// row-major order inside the block.  One thread per coarse cell.
__global__ void __launch_bounds__(kBlock)
    restrict_kernel(const float* __restrict__ x, const float* __restrict__ g,
                    const uint8_t* __restrict__ m, long long plane, int P,
                    int f, int hc, int wc, float* __restrict__ out) {
  const int id = blockIdx.x * kBlock + threadIdx.x;
  if (id >= hc * wc) return;
  const int I = id / wc, J = id % wc;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int di = 0; di < f; ++di) {
    const long long row = (long long)(I * f + di) * P + (long long)J * f;
    for (int dj = 0; dj < f; ++dj) {
      const long long o = row + dj;
      if (!m[o]) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float* xc = x + c * plane;
        const double s = (double)g[c * plane + o] + (double)xc[o - P] +
                         (double)xc[o - 1] + (double)xc[o + 1] +
                         (double)xc[o + P];
        acc[c] += s - (double)xc[o] * 4.0;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[id * 3 + c] = (float)acc[c];
}

// Bilinear sample of corr[hc][wc][3] at fine pixel (i, j), cell-centred:
// the pixel sits at coarse coordinate ((i + 0.5) / f - 0.5,
// (j + 0.5) / f - 0.5), clamped to the coarse grid.
__device__ __forceinline__ void bilinear3(const float* __restrict__ corr,
                                          int i, int j, float inv, int hc,
                                          int wc, float v[3]) {
  const float ys = ((float)i + 0.5f) * inv - 0.5f;
  const float xs = ((float)j + 0.5f) * inv - 0.5f;
  const int y0 = min(max((int)floorf(ys), 0), hc - 1);
  const int x0 = min(max((int)floorf(xs), 0), wc - 1);
  const int y1 = min(y0 + 1, hc - 1), x1 = min(x0 + 1, wc - 1);
  const float fy = fminf(fmaxf(ys - (float)y0, 0.0f), 1.0f);
  const float fx = fminf(fmaxf(xs - (float)x0, 0.0f), 1.0f);
  const float* c00 = corr + ((long long)y0 * wc + x0) * 3;
  const float* c01 = corr + ((long long)y0 * wc + x1) * 3;
  const float* c10 = corr + ((long long)y1 * wc + x0) * 3;
  const float* c11 = corr + ((long long)y1 * wc + x1) * 3;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float top = c00[c] * (1.0f - fx) + c01[c] * fx;
    const float bot = c10[c] * (1.0f - fx) + c11[c] * fx;
    v[c] = top * (1.0f - fy) + bot * fy;
  }
}

// x[i][j] += bilinear sample of corr at the solved pixels.
__global__ void __launch_bounds__(kBlock)
    prolong_kernel(float* __restrict__ x, const uint8_t* __restrict__ m,
                   long long plane, int H, int W, int P, int f, int hc, int wc,
                   const float* __restrict__ corr) {
  const long long n = (long long)H * W;
  const float inv = 1.0f / (float)f;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const long long o = (long long)i * P + j;
    if (!m[o]) continue;
    float v[3];
    bilinear3(corr, i, j, inv, hc, wc, v);
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c * plane + o] += v[c];
  }
}

// Line-search sums of a prolonged correction d (d = bilinear sample of corr
// at solved pixels, 0 elsewhere), per channel, in double:
// partial[block][0..2] += r . d with r the exact residual g + neighbours
// - 4 x of the state, partial[block][3..5] += d . (A d) with
// (A d)_p = 4 d_p - sum of the neighbours' d.
__global__ void __launch_bounds__(kBlock)
    prolong_dots_kernel(const float* __restrict__ x,
                        const float* __restrict__ g,
                        const uint8_t* __restrict__ m, long long plane, int H,
                        int W, int P, int f, int hc, int wc,
                        const float* __restrict__ corr,
                        double* __restrict__ partial) {
  double rd[3] = {0.0, 0.0, 0.0}, dad[3] = {0.0, 0.0, 0.0};
  const long long n = (long long)H * W;
  const float inv = 1.0f / (float)f;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const long long o = (long long)i * P + j;
    if (!m[o]) continue;  // solved pixels are never on the crop border
    float d[3], dn[3];
    double ad[3];
    bilinear3(corr, i, j, inv, hc, wc, d);
#pragma unroll
    for (int c = 0; c < 3; ++c) ad[c] = 4.0 * (double)d[c];
    const int ni[4] = {i - 1, i, i, i + 1}, nj[4] = {j, j - 1, j + 1, j};
    const long long no[4] = {o - P, o - 1, o + 1, o + P};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!m[no[k]]) continue;
      bilinear3(corr, ni[k], nj[k], inv, hc, wc, dn);
#pragma unroll
      for (int c = 0; c < 3; ++c) ad[c] -= (double)dn[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float* xc = x + c * plane;
      const double r = (double)g[c * plane + o] + (double)xc[o - P] +
                       (double)xc[o - 1] + (double)xc[o + 1] +
                       (double)xc[o + P] - 4.0 * (double)xc[o];
      rd[c] += r * (double)d[c];
      dad[c] += (double)d[c] * ad[c];
    }
  }
  block_sum3(rd);
  __syncthreads();  // the shared buffer is reused by the second sum
  block_sum3(dad);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      partial[blockIdx.x * 6 + c] = rd[c];
      partial[blockIdx.x * 6 + 3 + c] = dad[c];
    }
}

// One weighted Jacobi sweep: xout = xin + omega * ((sum) * 0.25 - xin) at
// the solved pixels (exactly (sum) * 0.25 when omega is 1, the plain sweep),
// xout = xin elsewhere (every pixel of the H x W crop).
template <int ORDER>
__global__ void __launch_bounds__(kBlock)
    weighted_kernel(const float* __restrict__ xin, float* __restrict__ xout,
                    const float* __restrict__ g, const uint8_t* __restrict__ m,
                    long long plane, int H, int W, int P, float omega) {
  const long long n = (long long)H * W;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    const long long o = (long long)i * P + j;
    const bool solved = m[o] != 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float* xc = xin + c * plane;
      const float v = xc[o];
      float r = v;
      if (solved) {
        const float jv = neighbour_sum<ORDER>(xc, g[c * plane + o], o, P) *
                         0.25f;
        r = omega == 1.0f ? jv : v + omega * (jv - v);
      }
      xout[c * plane + o] = r;
    }
  }
}

// The sweep kernel tells solved pixels by their guidance: every guidance
// value of an unsolved pixel or of the padding becomes NaN (a quiet NaN; a
// solved pixel's guidance keeps its value).  n = 3 * H * seg elements, seg
// = the padded segment length; element (c, i, jp) sits at c * plane + i * P
// + jp of the allocation and its mask byte at i * P + jp.  *nan_solved
// counts solved pixels whose guidance already was NaN (rejected by reset).
__global__ void __launch_bounds__(kBlock)
    box_kernel(float* __restrict__ g, const uint8_t* __restrict__ m,
               long long plane, int H, int P, int seg,
               unsigned long long* __restrict__ nan_solved) {
  const long long n = 3LL * H * seg;
  unsigned long long bad = 0;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const long long per = (long long)H * seg;
    const int c = (int)(id / per);
    const long long rem = id % per;
    const int i = (int)(rem / seg), jp = (int)(rem % seg);
    float& v = g[c * plane + (long long)i * P + jp];
    if (m[(long long)i * P + jp])
      bad += v != v;
    else
      v = __builtin_nanf("");
  }
  if (bad) atomicAdd(nan_solved, bad);
}

// occ[stripe * H + r] = 1 if row r of the columns stripe `stripe` writes
// ([stripe * U, min(stripe * U + U, W))) holds a solved pixel, else 0; one
// thread per (stripe, row).  Unsolved pixels never change, so a task whose
// rows are all 0 has nothing to do (the host turns this into task lists).
__global__ void __launch_bounds__(kBlock)
    stripe_rows_kernel(const uint8_t* __restrict__ m, int H, int W, int P,
                       int U, int stripes, uint8_t* __restrict__ occ) {
  const long long id = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (id >= (long long)stripes * H) return;
  const int stripe = (int)(id % stripes), r = (int)(id / stripes);
  const int c0 = stripe * U, c1 = min(c0 + U, W);
  const uint8_t* row = m + (long long)r * P;
  uint8_t any = 0;
  for (int c = c0; c < c1; ++c) any |= row[c];
  occ[(long long)stripe * H + r] = any != 0;
}

// Mask preparation of the processors (fpie/process.py:209-217, 338-345):
// a pixel is solved when the mean of its channels is >= 128 (for uint8
// channels: their sum >= 128 * channels, the same test without rounding),
// and the outermost rows and columns are cleared.  out[i][j] = 0/1;
// rows[i] / cols[j] |= 1 for every solved pixel (the bounding box).
__global__ void __launch_bounds__(kBlock)
    prep_mask_kernel(const uint8_t* __restrict__ mask, int H, int W,
                     int channels, uint8_t* __restrict__ out,
                     unsigned* __restrict__ rows, unsigned* __restrict__ cols) {
  const long long n = (long long)H * W;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int i = (int)(id / W), j = (int)(id % W);
    int sum = 0;
    for (int c = 0; c < channels; ++c) sum += mask[id * channels + c];
    const bool edge = i == 0 || j == 0 || i == H - 1 || j == W - 1;
    const uint8_t v = !edge && sum >= 128 * channels;
    out[id] = v;
    if (v) {
      // one store per row / column suffices; racing stores write the same 1
      if (!rows[i]) rows[i] = 1u;
      if (!cols[j]) cols[j] = 1u;
    }
  }
}

// crop[r][c] = (int32) m[x0 + r][y0 + c] for the h x w window
__global__ void __launch_bounds__(kBlock)
    crop_mask_kernel(const uint8_t* __restrict__ m, int W, int x0, int y0,
                     int h, int w, int32_t* __restrict__ crop) {
  const long long n = (long long)h * w;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const int r = (int)(id / w), c = (int)(id % w);
    crop[id] = m[(long long)(x0 + r) * W + y0 + c];
  }
}

int grid_blocks(long long n, int cap) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return (int)(b < cap ? b : cap);
}

}  // namespace

// ------------------------------------------------------------ launchers --

bool tile_supported(const TileCfg& cfg) {
  const int N = tile_nch(cfg);
#define PIE_T(s, c, n) \
  if (cfg.S == s && cfg.C == c && N == n) return true;
  PIE_TILES(PIE_T)
#undef PIE_T
  return false;
}

int tile_variants(int (*out)[3], int max) {
  int n = 0;
#define PIE_T(s, c, ch) \
  if (n < max) {        \
    out[n][0] = s;      \
    out[n][1] = c;      \
    out[n][2] = ch;     \
  }                     \
  ++n;
  PIE_TILES(PIE_T)
#undef PIE_T
  return n;
}

void tile_tasks(const TileCfg& cfg, int H, int W, int* stripes, int* segs) {
  const int U = kWave * cfg.C - 2 * cfg.S;
  *stripes = (W + U - 1) / U;
  *segs = (H + cfg.seg_rows - 1) / cfg.seg_rows;
}

hipError_t launch_march(const TileCfg& cfg, int sched, int order, MarchArgs a,
                        hipStream_t st) {
  tile_tasks(cfg, a.H, a.W, &a.stripes, &a.segs);
  a.seg_rows = cfg.seg_rows;
  if (!a.tasks) a.ntask = a.stripes * a.segs * tile_groups(cfg);
  if (a.ntask < 1) return hipSuccess;  // nothing to run
  const int N = tile_nch(cfg);
#define PIE_T(s, c, n)                              \
  if (cfg.S == s && cfg.C == c && N == n)           \
    return launch_march_##s##_##c##_##n(sched, order, a, st);
  PIE_TILES(PIE_T)
#undef PIE_T
  return hipErrorInvalidValue;
}

hipError_t launch_batch(const TileCfg& cfg, int sched, int order,
                        const BatchArgs& b, int blocks, hipStream_t st) {
  const int N = tile_nch(cfg);
#define PIE_T(s, c, n)                              \
  if (cfg.S == s && cfg.C == c && N == n)           \
    return launch_batch_##s##_##c##_##n(sched, order, b, blocks, st);
  PIE_TILES(PIE_T)
#undef PIE_T
  return hipErrorInvalidValue;
}

hipError_t launch_stripe_rows(const uint8_t* m, int H, int W, int P, int U,
                              int stripes, uint8_t* occ, hipStream_t st) {
  const long long n = (long long)stripes * H;
  hipLaunchKernelGGL(stripe_rows_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, st, m, H, W, P, U, stripes, occ);
  return hipGetLastError();
}

int march_blocks_per_cu(const TileCfg& cfg, int sched, int order) {
  const int N = tile_nch(cfg);
#define PIE_T(s, c, n)                              \
  if (cfg.S == s && cfg.C == c && N == n)           \
    return march_blocks_per_cu_##s##_##c##_##n(sched, order);
  PIE_TILES(PIE_T)
#undef PIE_T
  return 1;
}

int batch_blocks_per_cu(const TileCfg& cfg, int sched, int order) {
  const int N = tile_nch(cfg);
#define PIE_T(s, c, n)                              \
  if (cfg.S == s && cfg.C == c && N == n)           \
    return batch_blocks_per_cu_##s##_##c##_##n(sched, order);
  PIE_TILES(PIE_T)
#undef PIE_T
  return 1;
}

hipError_t launch_batch_epilogue(BatchItem* items, int n, int S,
                                 unsigned* queue, unsigned* solo_queue,
                                 unsigned* live, unsigned* shared_live,
                                 hipStream_t st) {
  hipLaunchKernelGGL(batch_epilogue_kernel, dim3(1), dim3(kBlock), 0, st, items,
                     n, S, queue, solo_queue, live, shared_live);
  return hipGetLastError();
}

hipError_t launch_clock(unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(clock_kernel, dim3(1), dim3(kWave), 0, st, out);
  return hipGetLastError();
}

hipError_t launch_deinterleave(const float* src, float* dst, long long plane,
                               int H, int W, int P, hipStream_t st) {
  hipLaunchKernelGGL(deinterleave_kernel, dim3(grid_blocks((long long)H * W, 8192)),
                     dim3(kBlock), 0, st, src, dst, plane, H, W, P);
  return hipGetLastError();
}

hipError_t launch_interleave(const float* src, float* dst, long long plane,
                             int H, int W, int P, hipStream_t st) {
  hipLaunchKernelGGL(interleave_kernel, dim3(grid_blocks((long long)H * W, 8192)),
                     dim3(kBlock), 0, st, src, dst, plane, H, W, P);
  return hipGetLastError();
}

hipError_t launch_mask(const int32_t* src, uint8_t* dst, int H, int W, int P,
                       unsigned long long* counts, hipStream_t st) {
  hipLaunchKernelGGL(mask_kernel, dim3(grid_blocks((long long)H * W, 8192)),
                     dim3(kBlock), 0, st, src, dst, H, W, P, counts);
  return hipGetLastError();
}

hipError_t launch_divergence(const uint8_t* src, const uint8_t* tgt,
                             const uint8_t* m, float* x, float* g,
                             long long plane, int H, int W, int P, int mode,
                             int equ, hipStream_t st) {
  hipLaunchKernelGGL(divergence_kernel,
                     dim3(grid_blocks((long long)H * W, 8192)), dim3(kBlock),
                     0, st, src, tgt, m, x, g, plane, H, W, P, mode, equ);
  return hipGetLastError();
}

hipError_t launch_keep_unsolved(uint8_t* img, const uint8_t* tgt,
                                const uint8_t* m, int H, int W, int P,
                                hipStream_t st) {
  hipLaunchKernelGGL(keep_unsolved_kernel,
                     dim3(grid_blocks((long long)H * W, 8192)), dim3(kBlock),
                     0, st, img, tgt, m, H, W, P);
  return hipGetLastError();
}

hipError_t launch_box_guidance(float* g, const uint8_t* m, long long plane,
                               int H, int P, int seg,
                               unsigned long long* nan_solved,
                               hipStream_t st) {
  hipLaunchKernelGGL(box_kernel, dim3(grid_blocks(3LL * H * seg, 8192)),
                     dim3(kBlock), 0, st, g, m, plane, H, P, seg, nan_solved);
  return hipGetLastError();
}

hipError_t launch_prep_mask(const uint8_t* mask, int H, int W, int channels,
                            uint8_t* out, unsigned* rows, unsigned* cols,
                            hipStream_t st) {
  hipLaunchKernelGGL(prep_mask_kernel,
                     dim3(grid_blocks((long long)H * W, 8192)), dim3(kBlock),
                     0, st, mask, H, W, channels, out, rows, cols);
  return hipGetLastError();
}

hipError_t launch_crop_mask(const uint8_t* m, int W, int x0, int y0, int h,
                            int w, int32_t* crop, hipStream_t st) {
  hipLaunchKernelGGL(crop_mask_kernel,
                     dim3(grid_blocks((long long)h * w, 8192)), dim3(kBlock),
                     0, st, m, W, x0, y0, h, w, crop);
  return hipGetLastError();
}

hipError_t launch_fill3(float* dst, long long n, int P, long long plane,
                        float v0, float v1, float v2, hipStream_t st) {
  hipLaunchKernelGGL(fill3_kernel, dim3(grid_blocks(n, 8192)), dim3(kBlock),
                     0, st, dst, n, P, plane, v0, v1, v2);
  return hipGetLastError();
}

hipError_t launch_equ_scatter(const float* X, const float* B,
                              const int32_t* lin, int N, float* x, float* g,
                              long long plane, hipStream_t st) {
  hipLaunchKernelGGL(equ_scatter_kernel, dim3(grid_blocks(N, 8192)), dim3(kBlock),
                     0, st, X, B, lin, N, x, g, plane);
  return hipGetLastError();
}

hipError_t launch_equ_gather(const float* x, const int32_t* lin, int N,
                             long long plane, float x00, float x01, float x02,
                             uint8_t* out_u8, float* out_f32,
                             hipStream_t st) {
  hipLaunchKernelGGL(equ_gather_kernel, dim3(grid_blocks(N, 8192)), dim3(kBlock),
                     0, st, x, lin, N, plane, x00, x01, x02, out_u8, out_f32);
  return hipGetLastError();
}

int finalize_blocks(int H, int W) { return grid_blocks((long long)H * W, 2048); }

hipError_t launch_finalize(int order, const float* x, const float* g,
                           const uint8_t* m, long long plane, int H, int W,
                           int P, uint8_t* img, double* partial, int nblocks,
                           hipStream_t st) {
  if (order == ORDER_UDLR)
    hipLaunchKernelGGL(finalize_kernel<ORDER_UDLR>, dim3(nblocks), dim3(kBlock),
                       0, st, x, g, m, plane, H, W, P, img, partial);
  else
    hipLaunchKernelGGL(finalize_kernel<ORDER_ULRD>, dim3(nblocks), dim3(kBlock),
                       0, st, x, g, m, plane, H, W, P, img, partial);
  return hipGetLastError();
}

hipError_t launch_restrict(const float* x, const float* g, const uint8_t* m,
                           long long plane, int P, int f, int hc, int wc,
                           float* out, hipStream_t st) {
  const int blocks = (int)(((long long)hc * wc + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(restrict_kernel, dim3(blocks), dim3(kBlock), 0, st, x, g,
                     m, plane, P, f, hc, wc, out);
  return hipGetLastError();
}

hipError_t launch_prolong(float* x, const uint8_t* m, long long plane, int H,
                          int W, int P, int f, int hc, int wc,
                          const float* corr, hipStream_t st) {
  hipLaunchKernelGGL(prolong_kernel,
                     dim3(grid_blocks((long long)H * W, 8192)), dim3(kBlock),
                     0, st, x, m, plane, H, W, P, f, hc, wc, corr);
  return hipGetLastError();
}

int prolong_dots_blocks(int H, int W) {
  return grid_blocks((long long)H * W, 2048);
}

hipError_t launch_prolong_dots(const float* x, const float* g,
                               const uint8_t* m, long long plane, int H,
                               int W, int P, int f, int hc, int wc,
                               const float* corr, double* partial,
                               int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(prolong_dots_kernel, dim3(nblocks), dim3(kBlock), 0, st,
                     x, g, m, plane, H, W, P, f, hc, wc, corr, partial);
  return hipGetLastError();
}

hipError_t launch_weighted(int order, const float* xin, float* xout,
                           const float* g, const uint8_t* m, long long plane,
                           int H, int W, int P, float omega, hipStream_t st) {
  const dim3 grid(grid_blocks((long long)H * W, 8192));
  if (order == ORDER_UDLR)
    hipLaunchKernelGGL(weighted_kernel<ORDER_UDLR>, grid, dim3(kBlock), 0, st,
                       xin, xout, g, m, plane, H, W, P, omega);
  else
    hipLaunchKernelGGL(weighted_kernel<ORDER_ULRD>, grid, dim3(kBlock), 0, st,
                       xin, xout, g, m, plane, H, W, P, omega);
  return hipGetLastError();
}

hipError_t launch_reduce(const double* partial, int n, float* err,
                         hipStream_t st) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kBlock), 0, st, partial, n,
                     err);
  return hipGetLastError();
}

hipError_t launch_equ_pack(const float* X, float4* out, int N,
                           hipStream_t st) {
  hipLaunchKernelGGL(equ_pack_kernel, dim3(grid_blocks(N, 8192)), dim3(kBlock),
                     0, st, X, out, N);
  return hipGetLastError();
}

hipError_t launch_equ_jacobi(const int4* A, const float4* B,
                             const float4* xin, float4* xout, int N,
                             hipStream_t st) {
  hipLaunchKernelGGL(equ_jacobi_kernel, dim3((N + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, st, A, B, xin, xout, N);
  return hipGetLastError();
}

int equ_generic_blocks(int N) { return grid_blocks(N, 2048); }

hipError_t launch_equ_finalize(const int4* A, const float4* B,
                               const float4* x, int N, uint8_t* img,
                               float* state, double* partial, int nblocks,
                               hipStream_t st) {
  hipLaunchKernelGGL(equ_finalize_kernel, dim3(nblocks), dim3(kBlock), 0, st, A,
                     B, x, N, img, state, partial);
  return hipGetLastError();
}

}  // namespace pie
