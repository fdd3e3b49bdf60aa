// gfx950 kernels of the warm-start U-Net.  Convolutions are implicit GEMMs
// on v_mfma_f32_16x16x32_bf16 (bf16 operands, fp32 accumulation); the rest
// is elementwise or a small reduction.
//
// MFMA 16x16x32 operand layout (per lane l, group g = l / 16, i = l % 16):
//   A (16 x 32): row i, k = 8g .. 8g+7   (8 bf16 in four VGPRs)
//   B (32 x 16): col i, k = 8g .. 8g+7
//   D (16 x 16): col i, rows 4g .. 4g+3  (4 fp32)
#include <hip/hip_runtime.h>

#include "unet_kernels.h"

namespace pie {
namespace unet {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16_t v) {
  return __uint_as_float((uint32_t)v << 16);
}

// float -> bfloat16, round to nearest even
__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const bf16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ bf16x8 zero8() {
  return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
}

struct ConvArgs {
  const bf16_t* in;
  int in_stride, in_off, cin;
  int n, h, w;
  const bf16_t* wk;
  int cout;
  const float* bias;
  int relu;
  bf16_t* out_bf;
  int out_stride, out_off;
  float* out_f32;
  int out_f32_stride, accumulate;
};

// Wave tile: MT x 16 pixels by NT x 16 output channels; a block is four
// waves along the pixels.
template <int MT, int NT>
__global__ void __launch_bounds__(kBlock) conv3x3_kernel(ConvArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int g = lane >> 4, i = lane & 15;
  const int hw = a.h * a.w;
  const long long P = (long long)a.n * hw;
  const long long p0 = ((long long)blockIdx.x * 4 + wave) * (MT * 16);
  if (p0 >= P) return;
  const int co0 = blockIdx.y * (NT * 16);
  const int CB = a.cin / 8;
  const int KB = 9 * CB;

  int py[MT], px[MT];
  long long pimg[MT];
  bool pok[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const long long p = p0 + mt * 16 + i;
    pok[mt] = p < P;
    const long long pc = pok[mt] ? p : 0;
    const int img = (int)(pc / hw), rem = (int)(pc % hw);
    py[mt] = rem / a.w;
    px[mt] = rem % a.w;
    pimg[mt] = (long long)img * hw;
  }

  f32x4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int s = 0; s * 4 < KB; ++s) {
    const int kb = 4 * s + g;
    const bool kok = kb < KB;
    const int tap = kok ? kb / CB : 0, cb = kok ? kb % CB : 0;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    bf16x8 A[MT], B[NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int y = py[mt] + dy, x = px[mt] + dx;
      const bool ok = kok && pok[mt] && y >= 0 && y < a.h && x >= 0 && x < a.w;
      A[mt] = ok ? load8(a.in + (pimg[mt] + (long long)y * a.w + x) *
                                    a.in_stride + a.in_off + cb * 8)
                 : zero8();
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      B[nt] = kok ? load8(a.wk + ((long long)kb * a.cout + co0 + nt * 16 + i) * 8)
                  : zero8();
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            A[mt], B[nt], acc[mt][nt], 0, 0, 0);
  }

#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int co = co0 + nt * 16 + i;
      const float b = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long p = p0 + mt * 16 + 4 * g + r;
        if (p >= P) continue;
        float v = acc[mt][nt][r] + b;
        if (a.relu) v = fmaxf(v, 0.0f);
        if (a.out_bf) a.out_bf[p * a.out_stride + a.out_off + co] = f2bf(v);
        if (a.out_f32) {
          float* o = a.out_f32 + p * a.out_f32_stride + co;
          *o = a.accumulate ? *o + v : v;
        }
      }
    }
}

// Weight gradient: a block owns 32 input x 32 output channels for all nine
// taps and a run of 32-pixel chunks.  Each chunk's shifted inputs
// (tap x ci x pixel) and output gradients (co x pixel) are staged in LDS
// with the pixel index contiguous, so each lane reads its eight K values
// (pixels) of an MFMA operand as one 16-byte LDS load.
constexpr int kWgPix = 32;

__global__ void __launch_bounds__(kBlock)
    wgrad_kernel(const bf16_t* __restrict__ x, int x_stride, int x_off,
                 int cin, int n, int h, int w, const bf16_t* __restrict__ dy,
                 int cout, float* __restrict__ dw, float* __restrict__ db,
                 int chunks_per_block) {
  __shared__ __attribute__((aligned(16))) bf16_t xs[9][32][kWgPix];
  __shared__ __attribute__((aligned(16))) bf16_t ds[32][kWgPix];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int g = lane >> 4, i = lane & 15;
  const int hw = h * w;
  const long long P = (long long)n * hw;
  const long long nchunk = (P + kWgPix - 1) / kWgPix;
  const int nci = (cin + 31) / 32;
  const int ci0 = (blockIdx.y % nci) * 32, co0 = (blockIdx.y / nci) * 32;
  // taps of this wave: wave, wave + 4, wave + 8
  f32x4 acc[3][2][2];
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[t][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  // bias gradient: the blocks of the first input-channel tile also sum
  // their staged output-gradient columns (thread t < 32: channel co0 + t)
  const bool bias = db && ci0 == 0 && threadIdx.x < 32;
  float bsum = 0.0f;
  const long long c_begin = (long long)blockIdx.x * chunks_per_block;
  const long long c_end = min(nchunk, c_begin + chunks_per_block);
  for (long long c = c_begin; c < c_end; ++c) {
    const long long pbase = c * kWgPix;
    // stage: 9 taps x 32 pixels x 4 channel groups of 8 = 1152 loads of 16 B
    for (int job = threadIdx.x; job < 9 * kWgPix * 4; job += kBlock) {
      const int tap = job / (kWgPix * 4);
      const int rem = job % (kWgPix * 4);
      const int pix = rem / 4, cg = rem % 4;
      const long long p = pbase + pix;
      const int ci = ci0 + cg * 8;
      bool ok = p < P && ci < cin;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (ok) {
        const int img = (int)(p / hw), r = (int)(p % hw);
        const int yy = r / w + tap / 3 - 1, xx = r % w + tap % 3 - 1;
        if (yy >= 0 && yy < h && xx >= 0 && xx < w)
          v = *reinterpret_cast<const uint4*>(
              x + ((long long)img * hw + (long long)yy * w + xx) * x_stride +
              x_off + ci);
      }
      const bf16_t* e = reinterpret_cast<const bf16_t*>(&v);
#pragma unroll
      for (int k = 0; k < 8; ++k) xs[tap][cg * 8 + k][pix] = e[k];
    }
    for (int job = threadIdx.x; job < kWgPix * 4; job += kBlock) {
      const int pix = job / 4, cg = job % 4;
      const long long p = pbase + pix;
      const int co = co0 + cg * 8;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (p < P && co < cout)
        v = *reinterpret_cast<const uint4*>(dy + p * cout + co);
      const bf16_t* e = reinterpret_cast<const bf16_t*>(&v);
#pragma unroll
      for (int k = 0; k < 8; ++k) ds[cg * 8 + k][pix] = e[k];
    }
    __syncthreads();
    if (bias)
#pragma unroll 8
      for (int pix = 0; pix < kWgPix; ++pix) bsum += bf2f(ds[threadIdx.x][pix]);
    bf16x8 B[2];
#pragma unroll
    for (int b = 0; b < 2; ++b)
      B[b] = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(&ds[b * 16 + i][8 * g]));
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int tap = wave + 4 * t;
      if (tap < 9) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const bf16x8 A = __builtin_bit_cast(
              bf16x8,
              *reinterpret_cast<const uint4*>(&xs[tap][a * 16 + i][8 * g]));
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[t][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                A, B[b], acc[t][a][b], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  if (bias && co0 + (int)threadIdx.x < cout)
    atomicAdd(&db[co0 + threadIdx.x], bsum);
  // D[ci][co]: lane holds col co = co0 + 16b + i, rows ci = ci0 + 16a + 4g + r
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int tap = wave + 4 * t;
    if (tap >= 9) continue;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int co = co0 + b * 16 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ci = ci0 + a * 16 + 4 * g + r;
          if (ci < cin && co < cout)
            atomicAdd(&dw[((long long)tap * cin + ci) * cout + co],
                      acc[t][a][b][r]);
        }
      }
  }
}

__global__ void __launch_bounds__(kBlock)
    pack_kernel(const float* __restrict__ w, int cin, int cout,
                bf16_t* __restrict__ wk, bf16_t* __restrict__ wkt) {
  const long long n = 9LL * cin * cout;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    // forward operand: id = (kb * cout + co) * 8 + e, kb = tap * (cin/8) + cb
    {
      const int e = (int)(id % 8);
      const long long t = id / 8;
      const int co = (int)(t % cout);
      const int kb = (int)(t / cout);
      const int tap = kb / (cin / 8), ci = (kb % (cin / 8)) * 8 + e;
      wk[id] = f2bf(w[((long long)tap * cin + ci) * cout + co]);
    }
    if (wkt && cout % 8 == 0) {
      // data-gradient operand: id = (kb' * cin + ci) * 8 + e,
      // kb' = tap' * (cout/8) + ob, value w[8 - tap'][ci][8 ob + e]
      const int e = (int)(id % 8);
      const long long t = id / 8;
      const int ci = (int)(t % cin);
      const int kb = (int)(t / cin);
      const int tp = kb / (cout / 8), co = (kb % (cout / 8)) * 8 + e;
      wkt[id] = f2bf(w[((long long)(8 - tp) * cin + ci) * cout + co]);
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    to_bf16_kernel(const float* __restrict__ x, long long p, int ch,
                   bf16_t* __restrict__ out, int stride, int off) {
  const long long n = p * ch;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock)
    out[(id / ch) * stride + off + id % ch] = f2bf(x[id]);
}

__global__ void __launch_bounds__(kBlock)
    pool_fwd_kernel(const bf16_t* __restrict__ in, int in_stride, int in_off,
                    int n, int h, int w, int ch, bf16_t* __restrict__ out,
                    int out_stride, int out_off) {
  const int h2 = h / 2, w2 = w / 2;
  const long long total = (long long)n * h2 * w2 * ch;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < total;
       id += (long long)gridDim.x * kBlock) {
    const int c = (int)(id % ch);
    const long long q = id / ch;
    const int x2 = (int)(q % w2), y2 = (int)((q / w2) % h2);
    const long long img = q / ((long long)w2 * h2);
    const long long b = (img * h + 2 * y2) * w + 2 * x2;
    const float s = bf2f(in[b * in_stride + in_off + c]) +
                    bf2f(in[(b + 1) * in_stride + in_off + c]) +
                    bf2f(in[(b + w) * in_stride + in_off + c]) +
                    bf2f(in[(b + w + 1) * in_stride + in_off + c]);
    out[q * out_stride + out_off + c] = f2bf(s * 0.25f);
  }
}

__global__ void __launch_bounds__(kBlock)
    pool_bwd_kernel(const float* __restrict__ dy, int n, int h2, int w2,
                    int ch, float* __restrict__ dx, int dx_stride) {
  const int h = 2 * h2, w = 2 * w2;
  const long long total = (long long)n * h * w * ch;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < total;
       id += (long long)gridDim.x * kBlock) {
    const int c = (int)(id % ch);
    const long long q = id / ch;
    const int x = (int)(q % w), y = (int)((q / w) % h);
    const long long img = q / ((long long)w * h);
    const long long s = (img * h2 + y / 2) * w2 + x / 2;
    dx[q * dx_stride + c] += dy[s * ch + c] * 0.25f;
  }
}

__global__ void __launch_bounds__(kBlock)
    up_fwd_kernel(const bf16_t* __restrict__ in, int in_stride, int in_off,
                  int n, int h, int w, int ch, bf16_t* __restrict__ out,
                  int out_stride, int out_off) {
  const int H = 2 * h, W = 2 * w;
  const long long total = (long long)n * H * W * ch;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < total;
       id += (long long)gridDim.x * kBlock) {
    const int c = (int)(id % ch);
    const long long q = id / ch;
    const int x = (int)(q % W), y = (int)((q / W) % H);
    const long long img = q / ((long long)W * H);
    const long long s = (img * h + y / 2) * w + x / 2;
    out[q * out_stride + out_off + c] = in[s * in_stride + in_off + c];
  }
}

// up_fwd_kernel with eight channels (16 bytes) per thread: channel counts,
// strides and offsets are multiples of 8.
__global__ void __launch_bounds__(kBlock)
    up_fwd8_kernel(const bf16_t* __restrict__ in, int in_stride, int in_off,
                   int n, int h, int w, int ch, bf16_t* __restrict__ out,
                   int out_stride, int out_off) {
  const int H = 2 * h, W = 2 * w, cg = ch / 8;
  const long long total = (long long)n * H * W * cg;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < total;
       id += (long long)gridDim.x * kBlock) {
    const int c = (int)(id % cg) * 8;
    const long long q = id / cg;
    const int x = (int)(q % W), y = (int)((q / W) % H);
    const long long img = q / ((long long)W * H);
    const long long sp = (img * h + y / 2) * w + x / 2;
    *reinterpret_cast<uint4*>(out + q * out_stride + out_off + c) =
        *reinterpret_cast<const uint4*>(in + sp * in_stride + in_off + c);
  }
}

__global__ void __launch_bounds__(kBlock)
    up_bwd_kernel(const float* __restrict__ dy, int dy_stride, int n, int h,
                  int w, int ch, float* __restrict__ dx) {
  const int W = 2 * w;
  const long long total = (long long)n * h * w * ch;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < total;
       id += (long long)gridDim.x * kBlock) {
    const int c = (int)(id % ch);
    const long long q = id / ch;
    const int x = (int)(q % w), y = (int)((q / w) % h);
    const long long img = q / ((long long)w * h);
    const long long b = (img * 2 * h + 2 * y) * W + 2 * x;
    dx[q * ch + c] = dy[b * dy_stride + c] + dy[(b + 1) * dy_stride + c] +
                     dy[(b + W) * dy_stride + c] +
                     dy[(b + W + 1) * dy_stride + c];
  }
}

__global__ void __launch_bounds__(kBlock)
    relu_grad_kernel(const float* __restrict__ g, int g_stride, int g_off,
                     const bf16_t* __restrict__ y, int y_stride, int y_off,
                     long long p, int ch, bf16_t* __restrict__ out) {
  const long long total = p * ch;
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < total;
       id += (long long)gridDim.x * kBlock) {
    const int c = (int)(id % ch);
    const long long q = id / ch;
    const float yv = bf2f(y[q * y_stride + y_off + c]);
    out[id] = f2bf(yv > 0.0f ? g[q * g_stride + g_off + c] : 0.0f);
  }
}

__global__ void __launch_bounds__(kBlock)
    head_fwd_kernel(const bf16_t* __restrict__ x, int x_stride, int x_off,
                    long long p, const float* __restrict__ w,
                    const float* __restrict__ b, float* __restrict__ out) {
  __shared__ float ws[32 * 3];
  for (int k = threadIdx.x; k < 96; k += kBlock) ws[k] = bf2f(f2bf(w[k]));
  __syncthreads();
  for (long long q = blockIdx.x * (long long)kBlock + threadIdx.x; q < p;
       q += (long long)gridDim.x * kBlock) {
    float o0 = 0.f, o1 = 0.f, o2 = 0.f;
    for (int c = 0; c < 32; ++c) {
      const float v = bf2f(x[q * x_stride + x_off + c]);
      o0 += v * ws[c * 3 + 0];
      o1 += v * ws[c * 3 + 1];
      o2 += v * ws[c * 3 + 2];
    }
    out[q * 3 + 0] = o0 + b[0];
    out[q * 3 + 1] = o1 + b[1];
    out[q * 3 + 2] = o2 + b[2];
  }
}

__global__ void __launch_bounds__(kBlock)
    head_bwd_kernel(const bf16_t* __restrict__ x, int x_stride, int x_off,
                    long long p, const float* __restrict__ w,
                    const float* __restrict__ dout, float* dw, float* db,
                    float* __restrict__ dx) {
  __shared__ float ws[32 * 3];
  __shared__ float red[kBlock];
  for (int k = threadIdx.x; k < 96; k += kBlock) ws[k] = bf2f(f2bf(w[k]));
  __syncthreads();
  float gw[96];
#pragma unroll
  for (int k = 0; k < 96; ++k) gw[k] = 0.0f;
  float gb[3] = {0.f, 0.f, 0.f};
  for (long long q = blockIdx.x * (long long)kBlock + threadIdx.x; q < p;
       q += (long long)gridDim.x * kBlock) {
    const float d0 = dout[q * 3 + 0], d1 = dout[q * 3 + 1],
                d2 = dout[q * 3 + 2];
    gb[0] += d0;
    gb[1] += d1;
    gb[2] += d2;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
      const float v = bf2f(x[q * x_stride + x_off + c]);
      gw[c * 3 + 0] += v * d0;
      gw[c * 3 + 1] += v * d1;
      gw[c * 3 + 2] += v * d2;
      dx[q * 32 + c] = d0 * ws[c * 3 + 0] + d1 * ws[c * 3 + 1] +
                       d2 * ws[c * 3 + 2];
    }
  }
  // block sums of the 99 parameter gradients, one atomic each
  for (int k = 0; k < 99; ++k) {
    red[threadIdx.x] = k < 96 ? gw[k] : gb[k - 96];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(k < 96 ? &dw[k] : &db[k - 96], red[0]);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock)
    mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt,
               const uint8_t* __restrict__ mask, long long p,
               float* __restrict__ dout, double* acc) {
  __shared__ double red[2][kBlock];
  double s = 0.0, c = 0.0;
  for (long long q = blockIdx.x * (long long)kBlock + threadIdx.x; q < p;
       q += (long long)gridDim.x * kBlock) {
    const bool m = mask[q] != 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float d = m ? out[q * 3 + k] - tgt[q * 3 + k] : 0.0f;
      dout[q * 3 + k] = 2.0f * d;
      s += (double)d * d;
    }
    c += m ? 3.0 : 0.0;
  }
  red[0][threadIdx.x] = s;
  red[1][threadIdx.x] = c;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      red[0][threadIdx.x] += red[0][threadIdx.x + k];
      red[1][threadIdx.x] += red[1][threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&acc[0], red[0][0]);
    atomicAdd(&acc[1], red[1][0]);
  }
}

__global__ void __launch_bounds__(kBlock)
    scale_kernel(float* x, long long n, const double* acc) {
  const float inv = (float)(1.0 / fmax(1.0, acc[1]));
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock)
    x[id] *= inv;
}

__global__ void __launch_bounds__(kBlock)
    adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                float* __restrict__ m, float* __restrict__ v, long long n,
                float lr, float b1, float b2, float eps, float c1, float c2) {
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < n;
       id += (long long)gridDim.x * kBlock) {
    const float gi = g[id];
    const float mi = b1 * m[id] + (1.0f - b1) * gi;
    const float vi = b2 * v[id] + (1.0f - b2) * gi * gi;
    m[id] = mi;
    v[id] = vi;
    p[id] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

int blocks_for(long long n, int cap = 8192) {
  long long b = (n + kBlock - 1) / kBlock;
  return (int)(b < 1 ? 1 : (b < cap ? b : cap));
}

}  // namespace

hipError_t conv3x3(const Slice& in, int n, int h, int w, const bf16_t* wk,
                   int cout, const float* bias, bool relu, const Slice* out_bf,
                   float* out_f32, int out_f32_stride, bool accumulate,
                   hipStream_t st) {
  if (in.ch % 8 || cout % 16 || in.stride % 8 || in.off % 8)
    return hipErrorInvalidValue;
  ConvArgs a{};
  a.in = static_cast<const bf16_t*>(in.base);
  a.in_stride = in.stride;
  a.in_off = in.off;
  a.cin = in.ch;
  a.n = n;
  a.h = h;
  a.w = w;
  a.wk = wk;
  a.cout = cout;
  a.bias = bias;
  a.relu = relu;
  a.out_bf = out_bf ? static_cast<bf16_t*>(out_bf->base) : nullptr;
  a.out_stride = out_bf ? out_bf->stride : 0;
  a.out_off = out_bf ? out_bf->off : 0;
  a.out_f32 = out_f32;
  a.out_f32_stride = out_f32_stride;
  a.accumulate = accumulate;
  const long long P = (long long)n * h * w;
  constexpr int MT = 2;
  if (cout % 32 == 0) {
    constexpr int NT = 2;
    dim3 grid((unsigned)((P + 4 * MT * 16 - 1) / (4 * MT * 16)), cout / 32);
    hipLaunchKernelGGL((conv3x3_kernel<MT, NT>), grid, dim3(kBlock), 0, st, a);
  } else {
    dim3 grid((unsigned)((P + 4 * MT * 16 - 1) / (4 * MT * 16)), cout / 16);
    hipLaunchKernelGGL((conv3x3_kernel<MT, 1>), grid, dim3(kBlock), 0, st, a);
  }
  return hipGetLastError();
}

hipError_t conv3x3_wgrad(const Slice& x, int n, int h, int w,
                         const bf16_t* dy, int cout, float* dw, float* db,
                         hipStream_t st) {
  if (x.ch % 8 || cout % 8 || x.stride % 8 || x.off % 8)
    return hipErrorInvalidValue;
  const long long P = (long long)n * h * w;
  const long long nchunk = (P + kWgPix - 1) / kWgPix;
  const int tiles = ((x.ch + 31) / 32) * ((cout + 31) / 32);
  // aim at about 2048 blocks in all
  long long per = (nchunk * tiles + 2047) / 2048;
  if (per < 1) per = 1;
  const int xb = (int)((nchunk + per - 1) / per);
  hipLaunchKernelGGL(wgrad_kernel, dim3(xb, tiles), dim3(kBlock), 0, st,
                     static_cast<const bf16_t*>(x.base), x.stride, x.off,
                     x.ch, n, h, w, dy, cout, dw, db, (int)per);
  return hipGetLastError();
}

hipError_t pack_weights(const float* w, int cin, int cout, bf16_t* wk,
                        bf16_t* wkt, hipStream_t st) {
  hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(9LL * cin * cout)),
                     dim3(kBlock), 0, st, w, cin, cout, wk, wkt);
  return hipGetLastError();
}

hipError_t to_bf16(const float* x, int p, int ch, const Slice& out,
                   hipStream_t st) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for((long long)p * ch)),
                     dim3(kBlock), 0, st, x, (long long)p, ch,
                     static_cast<bf16_t*>(out.base), out.stride, out.off);
  return hipGetLastError();
}

hipError_t pool_fwd(const Slice& in, int n, int h, int w, const Slice& out,
                    hipStream_t st) {
  hipLaunchKernelGGL(pool_fwd_kernel,
                     dim3(blocks_for((long long)n * h * w / 4 * in.ch)),
                     dim3(kBlock), 0, st, static_cast<const bf16_t*>(in.base),
                     in.stride, in.off, n, h, w, in.ch,
                     static_cast<bf16_t*>(out.base), out.stride, out.off);
  return hipGetLastError();
}

hipError_t pool_bwd(const float* dy, int n, int h2, int w2, int ch, float* dx,
                    int dx_stride, hipStream_t st) {
  hipLaunchKernelGGL(pool_bwd_kernel,
                     dim3(blocks_for((long long)n * h2 * w2 * 4 * ch)),
                     dim3(kBlock), 0, st, dy, n, h2, w2, ch, dx, dx_stride);
  return hipGetLastError();
}

hipError_t up_fwd(const Slice& in, int n, int h, int w, const Slice& out,
                  hipStream_t st) {
  if (in.ch % 8 == 0 && in.stride % 8 == 0 && in.off % 8 == 0 &&
      out.stride % 8 == 0 && out.off % 8 == 0) {
    hipLaunchKernelGGL(up_fwd8_kernel,
                       dim3(blocks_for((long long)n * h * w * 4 * in.ch / 8)),
                       dim3(kBlock), 0, st,
                       static_cast<const bf16_t*>(in.base), in.stride, in.off,
                       n, h, w, in.ch, static_cast<bf16_t*>(out.base),
                       out.stride, out.off);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(up_fwd_kernel,
                     dim3(blocks_for((long long)n * h * w * 4 * in.ch)),
                     dim3(kBlock), 0, st, static_cast<const bf16_t*>(in.base),
                     in.stride, in.off, n, h, w, in.ch,
                     static_cast<bf16_t*>(out.base), out.stride, out.off);
  return hipGetLastError();
}

hipError_t up_bwd(const float* dy, int dy_stride, int n, int h, int w, int ch,
                  float* dx, hipStream_t st) {
  hipLaunchKernelGGL(up_bwd_kernel,
                     dim3(blocks_for((long long)n * h * w * ch)),
                     dim3(kBlock), 0, st, dy, dy_stride, n, h, w, ch, dx);
  return hipGetLastError();
}

hipError_t relu_grad(const float* g, int g_stride, int g_off, const Slice& y,
                     int p, bf16_t* out, hipStream_t st) {
  hipLaunchKernelGGL(relu_grad_kernel, dim3(blocks_for((long long)p * y.ch)),
                     dim3(kBlock), 0, st, g, g_stride, g_off,
                     static_cast<const bf16_t*>(y.base), y.stride, y.off,
                     (long long)p, y.ch, out);
  return hipGetLastError();
}

hipError_t head_fwd(const Slice& x, int p, const float* w, const float* b,
                    float* out, hipStream_t st) {
  hipLaunchKernelGGL(head_fwd_kernel, dim3(blocks_for(p)), dim3(kBlock), 0, st,
                     static_cast<const bf16_t*>(x.base), x.stride, x.off,
                     (long long)p, w, b, out);
  return hipGetLastError();
}

hipError_t head_bwd(const Slice& x, int p, const float* w, const float* dout,
                    float* dw, float* db, float* dx, hipStream_t st) {
  hipLaunchKernelGGL(head_bwd_kernel, dim3(blocks_for(p, 1024)), dim3(kBlock),
                     0, st, static_cast<const bf16_t*>(x.base), x.stride,
                     x.off, (long long)p, w, dout, dw, db, dx);
  return hipGetLastError();
}

hipError_t mse(const float* out, const float* target, const uint8_t* mask,
               int p, float* dout, double* acc, hipStream_t st) {
  hipLaunchKernelGGL(mse_kernel, dim3(blocks_for(p, 1024)), dim3(kBlock), 0,
                     st, out, target, mask, (long long)p, dout, acc);
  return hipGetLastError();
}

hipError_t scale(float* x, long long n, const double* acc, hipStream_t st) {
  hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, st, x,
                     n, acc);
  return hipGetLastError();
}

hipError_t adam(float* p, const float* g, float* m, float* v, long long n,
                float lr, float b1, float b2, float eps, int t,
                hipStream_t st) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, st, p,
                     g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace unet
}  // namespace pie
