// gfx950 kernels of the warm-start U-Net.  Convolutions are implicit GEMMs
// on v_mfma_f32_16x16x32_bf16 (bf16 operands, fp32 accumulation); the rest
// is elementwise or a small reduction.
//
// MFMA 16x16x32 operand layout (per lane l, group g = l / 16, i = l % 16):
//   A (16 x 32): row i, k = 8g .. 8g+7   (8 bf16 in four VGPRs)
//   B (32 x 16): col i, k = 8g .. 8g+7
//   D (16 x 16): col i, rows 4g .. 4g+3  (4 fp32)
#include <atomic>
#include <hip/hip_runtime.h>

#include "unet_kernels.h"

namespace pie {
namespace unet {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// staging registers: a clang vector, not HIP's uint4 struct, so that
// fixed-index arrays of them stay in VGPRs
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

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

// Blocks that read the same data (the output-channel blocks of one pixel
// tile, the channel blocks of one tile range) are numbered so that they run
// on one XCD, one after another, and share its L2: the dispatcher hands
// block L to XCD L % 8, so group g's nm members are blocks
// 8 (nm (g / 8) + m) + g % 8, m = 0 .. nm - 1.  The grid holds
// ceil(groups / 8) * 8 * nm blocks; groups past the last are empty.
__device__ __forceinline__ void xcd_group(int L, int nm, int& grp, int& mem) {
  const int k = L >> 3;
  mem = k % nm;
  grp = (k / nm) * 8 + (L & 7);
}

inline unsigned xcd_grid(long long groups, int nm) {
  return (unsigned)(((groups + 7) / 8) * 8 * nm);
}

// The backward of the ReLU on a level's output folded into the producer of
// its gradient: out[p][c] = bf16(y[p][c] > 0 ? g : 0), with g the fp32
// gradient the producer sums (eight channels per thread; channel counts,
// strides and offsets multiples of 8).
__device__ __forceinline__ u32x4 relu_mask_bf16(const u32x4 y, f32x4 g0,
                                                f32x4 g1) {
  u32x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float ga = j < 2 ? g0[2 * j] : g1[2 * j - 4];
    const float gb = j < 2 ? g0[2 * j + 1] : g1[2 * j - 3];
    const float ya = __uint_as_float(y[j] << 16);
    const float yb = __uint_as_float(y[j] & 0xffff0000u);
    o[j] = (uint32_t)f2bf(ya > 0.0f ? ga : 0.0f) |
           ((uint32_t)f2bf(yb > 0.0f ? gb : 0.0f) << 16);
  }
  return o;
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
  const bf16_t* mask;  // bf16 output masked by mask > 0 (LDS-tiled kernel)
  int mask_stride, mask_off;
  // upsampling backward folded in (LDS-tiled kernel): output channels below
  // up_ch are summed over 2 x 2 pixel blocks, masked by up_mask > 0 of the
  // half-resolution level and written as bf16 [n][h/2][w/2][up_ch] to up_out;
  // the other channels go to out_f32 or out_bf as usual
  bf16_t* up_out;
  int up_ch;
  const bf16_t* up_mask;
  int up_mask_stride, up_mask_off;
  // forward fold (LDS-tiled kernel, with out_bf): the bf16 output's 2 x 2
  // average pool into pool_out
  bf16_t* pool_out;
  int pool_stride, pool_off;
  // pooling backward folded in (data gradient; no fp32 output): for the
  // four double-resolution pixels q under output pixel p, pb_out[q][co] =
  // bf16(pb_y[q][co] > 0 ? pb_gs[q][co] + v[p][co] / 4 : 0)
  bf16_t* pb_out;
  const bf16_t* pb_gs;
  int pb_gs_stride, pb_gs_off;
  const bf16_t* pb_y;
  int pb_y_stride, pb_y_off;
  // first layer (conv3x3_in8_kernel): fp32 input [P][8] instead of in, and
  // its bf16 rounding written to in_copy [P][8] when in_copy is not null
  const float* in_f32;
  bf16_t* in_copy;
  // input channels [0, inu_ch) read at half resolution from inu (the
  // nearest 2x upsample of a coarser level, not materialised): pixel (y, x)
  // of image n takes inu's pixel (n, y / 2, x / 2) (LDS-tiled kernel)
  const bf16_t* inu;
  int inu_stride, inu_off, inu_ch;
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

// LDS-tiled variant (input channels a multiple of 32): a block computes a
// (4 MT) x 16-pixel spatial tile of one image for 16 NT output channels.
// For every 32 input channels it stages the input window (tile plus a
// one-pixel halo) and the weights of all nine taps in LDS once; each wave
// then runs its MT x NT MFMA tiles for the nine taps from shifted views of
// the window (lane group g holds channels 8g .. 8g+7 of the 32). The global
// loads of the next 32 channels are issued before the MFMAs of the current
// ones (addresses clamped into the image, halo values selected to zero, so
// every load is unconditional and they all overlap).
// Global loads of one 32-channel stage of conv3x3_lds_kernel: the input
// window (WH x WW pixels x 4 groups of 8 channels) and the weights of the
// nine taps for CT output channels, into registers. Addresses are clamped
// into the image and every load is issued; bit k of the returned mask says
// whether window word k lies inside the image (the halo outside is zero,
// applied when the registers are stored to LDS, so that the compiler keeps
// the loads unconditional and in flight together).
// The window part of conv_stage_load.
template <int WH, int WW, bool UPIN>
__device__ __forceinline__ uint32_t conv_win_load(
    const ConvArgs& a, long long img_base, int row0, int col0, int c32,
    u32x4 (&wv)[(WH * WW * 4 + kBlock - 1) / kBlock]) {
  constexpr int kWinJobs = WH * WW * 4, kWinIt = (kWinJobs + kBlock - 1) / kBlock;
  uint32_t ok = 0;
  // the stage's source (wave-uniform): the half-resolution one for
  // channels below inu_ch; the image's base folded into the pointer, so
  // that a load needs only a 32-bit pixel offset
  const bool half = UPIN && c32 * 32 < a.inu_ch;
  const int stride = half ? a.inu_stride : a.in_stride;
  const int sh = half ? 1 : 0;
  const int wr = a.w >> sh;  // row pitch of the source in pixels
  const bf16_t* src =
      (half ? a.inu + a.inu_off + (img_base >> 2) * stride
            : a.in + a.in_off + img_base * stride) +
      c32 * 32;
#pragma unroll
  for (int k = 0; k < kWinIt; ++k) {
    const int job = threadIdx.x + k * kBlock;
    const int jj = min(job, kWinJobs - 1);
    const int pix = jj >> 2, gg = jj & 3;
    const int y = row0 - 1 + pix / WW, x = col0 - 1 + pix % WW;
    if (y >= 0 && y < a.h && x >= 0 && x < a.w) ok |= 1u << k;
    const int yc = min(max(y, 0), a.h - 1), xc = min(max(x, 0), a.w - 1);
    const int p = (yc >> sh) * wr + (xc >> sh);
    wv[k] = *reinterpret_cast<const u32x4*>(src + (long long)p * stride +
                                            gg * 8);
  }
  return ok;
}

// The window part of conv_stage_store (win: [WH][WW][4]).
template <int WH, int WW>
__device__ __forceinline__ void conv_win_store(
    const u32x4 (&wv)[(WH * WW * 4 + kBlock - 1) / kBlock], uint32_t ok,
    u32x4* win) {
  constexpr int kWinJobs = WH * WW * 4, kWinIt = (kWinJobs + kBlock - 1) / kBlock;
#pragma unroll
  for (int k = 0; k < kWinIt; ++k) {
    const int job = threadIdx.x + k * kBlock;
    if (job < kWinJobs)
      win[job] = ((ok >> k) & 1u) ? wv[k] : u32x4{0u, 0u, 0u, 0u};
  }
}

template <int WH, int WW, int CT, bool UPIN>
__device__ __forceinline__ uint32_t conv_stage_load(
    const ConvArgs& a, long long img_base, int row0, int col0, int co0,
    int c32, u32x4 (&wv)[(WH * WW * 4 + kBlock - 1) / kBlock],
    u32x4 (&wt)[(9 * 4 * CT + kBlock - 1) / kBlock]) {
  constexpr int kWtJobs = 9 * 4 * CT, kWtIt = (kWtJobs + kBlock - 1) / kBlock;
  const int CB = a.cin / 8;
  const uint32_t ok =
      conv_win_load<WH, WW, UPIN>(a, img_base, row0, col0, c32, wv);
#pragma unroll
  for (int k = 0; k < kWtIt; ++k) {
    const int jj = min((int)threadIdx.x + k * kBlock, kWtJobs - 1);
    const int co = jj % CT, rest = jj / CT;
    const int kb = (rest >> 2) * CB + c32 * 4 + (rest & 3);
    wt[k] = *reinterpret_cast<const u32x4*>(
        a.wk + ((long long)kb * a.cout + co0 + co) * 8);
  }
  return ok;
}

template <int WH, int WW, int CT>
__device__ __forceinline__ void conv_stage_store(
    const u32x4 (&wv)[(WH * WW * 4 + kBlock - 1) / kBlock],
    const u32x4 (&wt)[(9 * 4 * CT + kBlock - 1) / kBlock], uint32_t ok,
    u32x4 (&win)[WH][WW][4], u32x4 (&ws)[9][4][CT]) {
  constexpr int kWtJobs = 9 * 4 * CT, kWtIt = (kWtJobs + kBlock - 1) / kBlock;
  conv_win_store<WH, WW>(wv, ok, &win[0][0][0]);
#pragma unroll
  for (int k = 0; k < kWtIt; ++k) {
    const int job = threadIdx.x + k * kBlock;
    if (job < kWtJobs) {
      const int co = job % CT, rest = job / CT;
      ws[rest >> 2][rest & 3][co] = wt[k];
    }
  }
}

// The MFMAs of one 32-channel stage: win = the staged input window
// [WH][WW][4] (lane group g holds channels 8g .. 8g+7), ws = the stage's
// weights [9][4][CT].  Wave `wave` owns output rows wave MT .. wave MT +
// MT - 1; every tap reads a shifted view of the window, and every output
// tile accumulates the taps in the order 0 .. 8 (dy outer), whatever MT.
// (Taps column by column, each loaded window row serving the up to three
// output rows it neighbours, read fewer LDS words but hold three taps'
// weights at once: 192 instead of 160 VGPRs for MT = 2, NT = 4, one wave a
// SIMD less.)
template <int MT, int NT>
__device__ __forceinline__ void conv_taps(f32x4 (&acc)[MT][NT],
                                          const u32x4* win, const u32x4* ws,
                                          int wave, int g, int i) {
  constexpr int WW = 18, CT = 16 * NT;
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const int dy = tap / 3, dx = tap % 3;
    bf16x8 A[MT], B[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      B[nt] = __builtin_bit_cast(bf16x8, ws[(tap * 4 + g) * CT + nt * 16 + i]);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
      A[mt] = __builtin_bit_cast(
          bf16x8, win[((wave * MT + mt + dy) * WW + i + dx) * 4 + g]);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            A[mt], B[nt], acc[mt][nt], 0, 0, 0);
  }
}

// Epilogue of the LDS-tiled convolutions, through the fp32 output tile T
// ([pixel][channel], rows padded to CT + 4 floats; it may alias the
// staging arrays once every wave is past their last read): bias and ReLU
// per lane into T, then 16-byte stores of 8 bf16 or 4 fp32 channels per
// thread, with the folded pooling / upsampling / ReLU-mask outputs.
template <int MT, int NT>
__device__ __forceinline__ void conv_epilogue(
    const ConvArgs& a, const f32x4 (&acc)[MT][NT], float* T, int wave, int g,
    int i, int img, int row0, int col0, int co0, long long img_base) {
  constexpr int TH = 4 * MT, TW = 16, CT = 16 * NT, TS = CT + 4;
  const bool up = a.up_out && co0 < a.up_ch;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int co = nt * 16 + i;
    const float b = a.bias ? a.bias[co0 + co] : 0.0f;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = acc[mt][nt][r] + b;
        if (a.relu) v = fmaxf(v, 0.0f);
        T[((wave * MT + mt) * TW + 4 * g + r) * TS + co] = v;
      }
  }
  __syncthreads();
  if (a.out_bf && !up) {
    constexpr int kJobs = TH * TW * (CT / 8);
#pragma unroll
    for (int k = 0; k < (kJobs + kBlock - 1) / kBlock; ++k) {
      const int job = threadIdx.x + k * kBlock;
      const int pix = job / (CT / 8), c8 = job % (CT / 8);
      const int y = row0 + pix / TW, x = col0 + pix % TW;
      if (job < kJobs && y < a.h && x < a.w) {
        f32x4 v0 = *reinterpret_cast<const f32x4*>(&T[pix * TS + c8 * 8]);
        f32x4 v1 = *reinterpret_cast<const f32x4*>(&T[pix * TS + c8 * 8 + 4]);
        const long long p = img_base + (long long)y * a.w + x;
        if (a.mask) {
          const u32x4 m = *reinterpret_cast<const u32x4*>(
              a.mask + p * a.mask_stride + a.mask_off + co0 + c8 * 8);
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            if (!(__uint_as_float(m[j] << 16) > 0.0f)) v0[2 * j] = 0.0f;
            if (!(__uint_as_float(m[j] & 0xffff0000u) > 0.0f)) v0[2 * j + 1] = 0.0f;
            if (!(__uint_as_float(m[j + 2] << 16) > 0.0f)) v1[2 * j] = 0.0f;
            if (!(__uint_as_float(m[j + 2] & 0xffff0000u) > 0.0f))
              v1[2 * j + 1] = 0.0f;
          }
        }
        u32x4 o;
        o[0] = (uint32_t)f2bf(v0[0]) | ((uint32_t)f2bf(v0[1]) << 16);
        o[1] = (uint32_t)f2bf(v0[2]) | ((uint32_t)f2bf(v0[3]) << 16);
        o[2] = (uint32_t)f2bf(v1[0]) | ((uint32_t)f2bf(v1[1]) << 16);
        o[3] = (uint32_t)f2bf(v1[2]) | ((uint32_t)f2bf(v1[3]) << 16);
        *reinterpret_cast<u32x4*>(a.out_bf + p * a.out_stride + a.out_off +
                                  co0 + c8 * 8) = o;
      }
    }
  }
  if (a.pool_out) {
    // average of the four bf16 outputs of each 2 x 2 block, in the order
    // (0, 0), (0, 1), (1, 0), (1, 1)
    constexpr int kJobs = (TH / 2) * (TW / 2) * (CT / 8);
    const int h2 = a.h / 2, w2 = a.w / 2;
#pragma unroll
    for (int k = 0; k < (kJobs + kBlock - 1) / kBlock; ++k) {
      const int job = threadIdx.x + k * kBlock;
      const int cp = job / (CT / 8), c8 = job % (CT / 8);
      const int cy = cp / (TW / 2), cx = cp % (TW / 2);
      const int y2 = row0 / 2 + cy, x2 = col0 / 2 + cx;
      if (job < kJobs && y2 < h2 && x2 < w2) {
        const float* t0 = &T[((2 * cy) * TW + 2 * cx) * TS + c8 * 8];
        const float* tk[4] = {t0, t0 + TS, t0 + TW * TS, t0 + TW * TS + TS};
        float sum[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) sum[e] = bf2f(f2bf(tk[0][e]));
#pragma unroll
        for (int k2 = 1; k2 < 4; ++k2)
#pragma unroll
          for (int e = 0; e < 8; ++e) sum[e] += bf2f(f2bf(tk[k2][e]));
        u32x4 o;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          o[j] = (uint32_t)f2bf(sum[2 * j] * 0.25f) |
                 ((uint32_t)f2bf(sum[2 * j + 1] * 0.25f) << 16);
        const long long q = ((long long)img * h2 + y2) * w2 + x2;
        *reinterpret_cast<u32x4*>(a.pool_out + q * a.pool_stride +
                                  a.pool_off + co0 + c8 * 8) = o;
      }
    }
  }
  if (up) {
    // coarse pixel (row0 / 2 + cy, col0 / 2 + cx): the fine pixels in the
    // order (0, 0), (0, 1), (1, 0), (1, 1)
    constexpr int kJobs = (TH / 2) * (TW / 2) * (CT / 8);
    const int h2 = a.h / 2, w2 = a.w / 2;
#pragma unroll
    for (int k = 0; k < (kJobs + kBlock - 1) / kBlock; ++k) {
      const int job = threadIdx.x + k * kBlock;
      const int cp = job / (CT / 8), c8 = job % (CT / 8);
      const int cy = cp / (TW / 2), cx = cp % (TW / 2);
      const int y2 = row0 / 2 + cy, x2 = col0 / 2 + cx;
      if (job < kJobs && y2 < h2 && x2 < w2) {
        const int f = (2 * cy) * TW + 2 * cx;
        const float* t0 = &T[f * TS + c8 * 8];
        const float* t1 = t0 + TS;
        const float* t2 = t0 + TW * TS;
        const float* t3 = t2 + TS;
        f32x4 s0 = *reinterpret_cast<const f32x4*>(t0);
        f32x4 s1 = *reinterpret_cast<const f32x4*>(t0 + 4);
        s0 = s0 + *reinterpret_cast<const f32x4*>(t1);
        s1 = s1 + *reinterpret_cast<const f32x4*>(t1 + 4);
        s0 = s0 + *reinterpret_cast<const f32x4*>(t2);
        s1 = s1 + *reinterpret_cast<const f32x4*>(t2 + 4);
        s0 = s0 + *reinterpret_cast<const f32x4*>(t3);
        s1 = s1 + *reinterpret_cast<const f32x4*>(t3 + 4);
        const long long q = ((long long)img * h2 + y2) * w2 + x2;
        const u32x4 m = *reinterpret_cast<const u32x4*>(
            a.up_mask + q * a.up_mask_stride + a.up_mask_off + co0 + c8 * 8);
        *reinterpret_cast<u32x4*>(a.up_out + q * a.up_ch + co0 + c8 * 8) =
            relu_mask_bf16(m, s0, s1);
      }
    }
  }
  if (a.pb_out) {
    constexpr int kJobs = TH * TW * (CT / 8);
    const int W2 = 2 * a.w;
#pragma unroll
    for (int k = 0; k < (kJobs + kBlock - 1) / kBlock; ++k) {
      const int job = threadIdx.x + k * kBlock;
      const int pix = job / (CT / 8), c8 = job % (CT / 8);
      const int y = row0 + pix / TW, x = col0 + pix % TW;
      if (job < kJobs && y < a.h && x < a.w) {
        const f32x4 v0 =
            *reinterpret_cast<const f32x4*>(&T[pix * TS + c8 * 8]) * 0.25f;
        const f32x4 v1 =
            *reinterpret_cast<const f32x4*>(&T[pix * TS + c8 * 8 + 4]) * 0.25f;
        const long long f = ((long long)img * 2 * a.h + 2 * y) * W2 + 2 * x;
        const long long q[4] = {f, f + 1, f + W2, f + W2 + 1};
        const int c = co0 + c8 * 8;
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
          const u32x4 gs = *reinterpret_cast<const u32x4*>(
              a.pb_gs + q[k2] * a.pb_gs_stride + a.pb_gs_off + c);
          const u32x4 yv = *reinterpret_cast<const u32x4*>(
              a.pb_y + q[k2] * a.pb_y_stride + a.pb_y_off + c);
          f32x4 g0, g1;
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            g0[2 * j] = __uint_as_float(gs[j] << 16);
            g0[2 * j + 1] = __uint_as_float(gs[j] & 0xffff0000u);
            g1[2 * j] = __uint_as_float(gs[j + 2] << 16);
            g1[2 * j + 1] = __uint_as_float(gs[j + 2] & 0xffff0000u);
          }
          *reinterpret_cast<u32x4*>(a.pb_out + q[k2] * a.cout + c) =
              relu_mask_bf16(yv, g0 + v0, g1 + v1);
        }
      }
    }
  }
  if (a.out_f32 && !up) {
    constexpr int kJobs = TH * TW * (CT / 4);
#pragma unroll
    for (int k = 0; k < (kJobs + kBlock - 1) / kBlock; ++k) {
      const int job = threadIdx.x + k * kBlock;
      const int pix = job / (CT / 4), c4 = job % (CT / 4);
      const int y = row0 + pix / TW, x = col0 + pix % TW;
      if (job < kJobs && y < a.h && x < a.w) {
        f32x4 v = *reinterpret_cast<const f32x4*>(&T[pix * TS + c4 * 4]);
        const long long p = img_base + (long long)y * a.w + x;
        f32x4* o = reinterpret_cast<f32x4*>(a.out_f32 + p * a.out_f32_stride +
                                            co0 + c4 * 4);
        if (a.accumulate) v += *o;
        *o = v;
      }
    }
  }
}

// UPIN: the input's first inu_ch channels are read at half resolution
template <int MT, int NT, bool UPIN>
__global__ void __launch_bounds__(kBlock) conv3x3_lds_kernel(ConvArgs a) {
  constexpr int TH = 4 * MT, TW = 16, WH = TH + 2, WW = TW + 2;
  constexpr int CT = 16 * NT;
  constexpr int kWinJobs = WH * WW * 4, kWinIt = (kWinJobs + kBlock - 1) / kBlock;
  constexpr int kWtJobs = 9 * 4 * CT, kWtIt = (kWtJobs + kBlock - 1) / kBlock;
  // staging arrays, reused for the output tile after the last barrier
  struct Stage {
    u32x4 win[WH][WW][4];
    u32x4 ws[9][4][CT];
  };
  constexpr int TS = CT + 4;  // padded output row: lane groups on distinct banks
  constexpr int kOutBytes = TH * TW * TS * 4;
  __shared__ __attribute__((aligned(16)))
  unsigned char smem[sizeof(Stage) > kOutBytes ? sizeof(Stage) : kOutBytes];
  auto& win = reinterpret_cast<Stage*>(smem)->win;
  auto& ws = reinterpret_cast<Stage*>(smem)->ws;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int g = lane >> 4, i = lane & 15;
  const int tiles_x = (a.w + TW - 1) / TW, tiles_y = (a.h + TH - 1) / TH;
  int tile, cob;
  xcd_group(blockIdx.x, a.cout / CT, tile, cob);
  if (tile >= a.n * tiles_x * tiles_y) return;  // whole block, before any barrier
  const int img = tile / (tiles_x * tiles_y);
  const int trem = tile % (tiles_x * tiles_y);
  const int row0 = (trem / tiles_x) * TH, col0 = (trem % tiles_x) * TW;
  const int co0 = cob * CT;
  const long long img_base = (long long)img * a.h * a.w;
  u32x4 wv[kWinIt], wt[kWtIt];
  f32x4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int nc32 = a.cin / 32;
  uint32_t ok = conv_stage_load<WH, WW, CT, UPIN>(a, img_base, row0, col0, co0, 0,
                                            wv, wt);
  for (int c32 = 0; c32 < nc32; ++c32) {
    conv_stage_store<WH, WW, CT>(wv, wt, ok, win, ws);
    __syncthreads();
    if (c32 + 1 < nc32)
      ok = conv_stage_load<WH, WW, CT, UPIN>(a, img_base, row0, col0, co0, c32 + 1,
                                       wv, wt);
    conv_taps<MT, NT>(acc, &win[0][0][0], &ws[0][0][0], wave, g, i);
    __syncthreads();
  }

  conv_epilogue<MT, NT>(a, acc, reinterpret_cast<float*>(smem), wave, g, i,
                        img, row0, col0, co0, img_base);
}

// First layer (8 input channels, fp32 input): a block computes a (4 MT) x
// 16-pixel tile for 16 NT output channels.  The fp32 input window (tile
// plus a one-pixel halo, zero outside the image) is rounded to bf16 and
// staged in LDS once, one 16-byte word of 8 channels a pixel; the tile's
// own pixels also go to in_copy (the bf16 layer input, which the weight
// gradient reads).  K = 9 taps x 8 channels in the weights' order
// (k = 8 tap + e), padded to three MFMA steps: step s, lane group g takes
// tap 4 s + g, taps 9 .. 11 being zero.
template <int MT, int NT>
__global__ void __launch_bounds__(kBlock) conv3x3_in8_kernel(ConvArgs a) {
  constexpr int TH = 4 * MT, TW = 16, WH = TH + 2, WW = TW + 2;
  constexpr int CT = 16 * NT;
  constexpr int kWinJobs = WH * WW, kWinIt = (kWinJobs + kBlock - 1) / kBlock;
  constexpr int TS = CT + 4;
  constexpr int kOutBytes = TH * TW * TS * 4, kWinBytes = kWinJobs * 16;
  __shared__ __attribute__((aligned(16)))
  unsigned char smem[kWinBytes > kOutBytes ? kWinBytes : kOutBytes];
  u32x4* win = reinterpret_cast<u32x4*>(smem);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int g = lane >> 4, i = lane & 15;
  const int tiles_x = (a.w + TW - 1) / TW, tiles_y = (a.h + TH - 1) / TH;
  int tile, cob;
  xcd_group(blockIdx.x, a.cout / CT, tile, cob);
  if (tile >= a.n * tiles_x * tiles_y) return;  // whole block, before any barrier
  const int img = tile / (tiles_x * tiles_y);
  const int trem = tile % (tiles_x * tiles_y);
  const int row0 = (trem / tiles_x) * TH, col0 = (trem % tiles_x) * TW;
  const int co0 = cob * CT;
  const long long img_base = (long long)img * a.h * a.w;

  // window loads first (clamped addresses, all in flight together), then
  // the weights of the three steps straight into registers
  f32x4 lo[kWinIt], hi[kWinIt];
  uint32_t ok = 0;
#pragma unroll
  for (int k = 0; k < kWinIt; ++k) {
    const int jj = min((int)threadIdx.x + k * kBlock, kWinJobs - 1);
    const int y = row0 - 1 + jj / WW, x = col0 - 1 + jj % WW;
    if (y >= 0 && y < a.h && x >= 0 && x < a.w) ok |= 1u << k;
    const int yc = min(max(y, 0), a.h - 1), xc = min(max(x, 0), a.w - 1);
    const f32x4* src = reinterpret_cast<const f32x4*>(
        a.in_f32 + (img_base + (long long)yc * a.w + xc) * 8);
    lo[k] = src[0];
    hi[k] = src[1];
  }
  bf16x8 B[3][NT];
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int tap = min(4 * s + g, 8);
      B[s][nt] = 4 * s + g < 9
                     ? load8(a.wk + ((long long)tap * a.cout + co0 + nt * 16 +
                                     i) * 8)
                     : zero8();
    }
#pragma unroll
  for (int k = 0; k < kWinIt; ++k) {
    const int job = threadIdx.x + k * kBlock;
    if (job < kWinJobs) {
      u32x4 v = u32x4{0u, 0u, 0u, 0u};
      if ((ok >> k) & 1u) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          v[j] = (uint32_t)f2bf(lo[k][2 * j]) |
                 ((uint32_t)f2bf(lo[k][2 * j + 1]) << 16);
          v[j + 2] = (uint32_t)f2bf(hi[k][2 * j]) |
                     ((uint32_t)f2bf(hi[k][2 * j + 1]) << 16);
        }
        const int wy = job / WW, wx = job % WW;
        if (a.in_copy && wy >= 1 && wy <= TH && wx >= 1 && wx <= TW)
          *reinterpret_cast<u32x4*>(
              a.in_copy + (img_base + (long long)(row0 - 1 + wy) * a.w +
                           col0 - 1 + wx) * 8) = v;
      }
      win[job] = v;
    }
  }
  __syncthreads();

  f32x4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int tap = min(4 * s + g, 8);
    const int dy = tap / 3, dx = tap % 3;
    bf16x8 A[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
      A[mt] = 4 * s + g < 9
                  ? __builtin_bit_cast(
                        bf16x8, win[(wave * MT + mt + dy) * WW + i + dx])
                  : zero8();
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            A[mt], B[s][nt], acc[mt][nt], 0, 0, 0);
  }
  __syncthreads();  // the output tile reuses the window's LDS

  conv_epilogue<MT, NT>(a, acc, reinterpret_cast<float*>(smem), wave, g, i,
                        img, row0, col0, co0, img_base);
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

// Weight gradient over spatial tiles (input channels a multiple of 8, output
// channels of 32): a block owns 32 input x CO (32 or 64) output channels
// (input channels past cin read as zero) for all nine taps and sums over
// kWtH x 16-pixel tiles of the images.  Per tile it stages the input window
// (tile plus one-pixel halo) and the output gradient in LDS in their
// natural [pixel][channel] layout, straight from 16-byte global loads, and
// reads the MFMA operands (K = pixels) with gfx950's transposing LDS read:
// ds_read_b64_tr_b16 gives each lane of a 16-lane group one channel of four
// pixels, lanes 4j .. 4j + 3 naming pixel j's 16 channels, so a tap's
// shifted window is just other per-lane addresses.  One MFMA step covers
// two rows of the tile (lane group g: row g / 2, columns 8 (g % 2) .. +7,
// two transposing reads of four pixels each).  Wave (ta, tb) owns input
// channels 16 ta .. +15 and the CO / 32 output blocks of 16 starting at
// (CO / 32) tb.  The staging loads are unconditional (addresses clamped
// into the image, values outside selected to zero) and the next tile's
// loads are in flight while the MFMAs of the current one run.
constexpr int kWtW = 16;  // tile width; the height is a template argument

typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

// Two transposing reads (pixels p and p + 4 of a lane group's eight) as one
// 8-element MFMA operand.
__device__ __forceinline__ bf16x8 tr_pair(const bf16_t* lo, const bf16_t* hi) {
  const i16x4 a = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
      (__attribute__((address_space(3))) i16x4*)(lo));
  const i16x4 b = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
      (__attribute__((address_space(3))) i16x4*)(hi));
  const i16x8 c = __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, c);
}

template <int CO, int kWtH, bool UPIN>
__global__ void __launch_bounds__(kBlock)
    wgrad_lds_kernel(const bf16_t* __restrict__ x, int x_stride, int x_off,
                     int cin, int n, int h, int w,
                     const bf16_t* __restrict__ dy, int cout,
                     float* __restrict__ dw, float* __restrict__ db,
                     int tiles_per_block, const bf16_t* __restrict__ xu,
                     int xu_stride, int xu_off, int xu_ch) {
  constexpr int kWtWH = kWtH + 2, kWtWW = kWtW + 2;  // input window
  constexpr int NBW = CO / 32;                      // output blocks per wave
  constexpr int DG = CO / 8;                        // gradient channel groups
  constexpr int XJ = kWtWH * kWtWW * 4;             // window loads (16 B)
  constexpr int DJ = kWtH * kWtW * DG;              // gradient loads (16 B)
  constexpr int XIT = (XJ + kBlock - 1) / kBlock;
  static_assert(DJ % kBlock == 0, "whole rounds of gradient loads");
  constexpr int DIT = DJ / kBlock;
  __shared__ __attribute__((aligned(16))) bf16_t xs[kWtWH][kWtWW][32];
  __shared__ __attribute__((aligned(16))) bf16_t ds[kWtH][kWtW][CO];
  __shared__ float bsum[CO];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int g = lane >> 4, i = lane & 15;
  const int pj = i >> 2, pq = i & 3;  // transposing read: pixel, 4-channel chunk
  const int ta = wave >> 1, tb = wave & 1;
  const int nci = (cin + 31) / 32;
  int grp, mem;
  xcd_group(blockIdx.x, nci * (cout / CO), grp, mem);
  const int ci0 = (mem % nci) * 32, co0 = (mem / nci) * CO;
  // the block's input channels come from the half-resolution source
  // (pixel (y, x) <- (y / 2, x / 2)) when below xu_ch
  const bool half = UPIN && ci0 < xu_ch;
  const int tiles_x = (w + kWtW - 1) / kWtW, tiles_y = (h + kWtH - 1) / kWtH;
  const long long ntiles = (long long)n * tiles_x * tiles_y;
  const long long t_begin = (long long)grp * tiles_per_block;
  if (t_begin >= ntiles) return;  // whole block, before any barrier
  const bool bias = db && ci0 == 0;
  const int tid = threadIdx.x;
  const int dcg = tid % DG;  // the gradient channel group of all the
                             // thread's gradient loads (kBlock % DG == 0)
  if (tid < CO) bsum[tid] = 0.0f;
  float bpart[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) bpart[k] = 0.0f;

  u32x4 xv[XIT], dv[DIT];
  uint32_t okx = 0, okd = 0;
  auto load = [&](long long t) {
    const int img = (int)(t / (tiles_x * tiles_y));
    const int trem = (int)(t % (tiles_x * tiles_y));
    const int row0 = (trem / tiles_x) * kWtH, col0 = (trem % tiles_x) * kWtW;
    const long long ib = (long long)img * h;
    okx = okd = 0;
    // the image's base in the (wave-uniform) source pointer, 32-bit pixel
    // offsets per load
    const int sh = half ? 1 : 0, wr = w >> sh;
    const int xs_stride = half ? xu_stride : x_stride;
    const bf16_t* xsrc =
        half ? xu + xu_off + (long long)img * (h >> 1) * wr * xu_stride
             : x + x_off + ib * w * x_stride;
#pragma unroll
    for (int k = 0; k < XIT; ++k) {
      const int job = min(tid + k * kBlock, XJ - 1);
      const int pix = job >> 2, cg = job & 3;
      const int y = row0 - 1 + pix / kWtWW, xx = col0 - 1 + pix % kWtWW;
      const int ch = ci0 + cg * 8;
      if (tid + k * kBlock < XJ && ch < cin && y >= 0 && y < h && xx >= 0 &&
          xx < w)
        okx |= 1u << k;
      const int yc = min(max(y, 0), h - 1), xc = min(max(xx, 0), w - 1);
      const int q = (yc >> sh) * wr + (xc >> sh);
      xv[k] = *reinterpret_cast<const u32x4*>(
          xsrc + (long long)q * xs_stride + min(ch, cin - 8));
    }
#pragma unroll
    for (int k = 0; k < DIT; ++k) {
      const int job = tid + k * kBlock;
      const int pix = job / DG;
      const int y = row0 + pix / kWtW, xx = col0 + pix % kWtW;
      if (y < h && xx < w) okd |= 1u << k;
      const int yc = min(y, h - 1), xc = min(xx, w - 1);
      dv[k] = *reinterpret_cast<const u32x4*>(
          dy + ((ib + yc) * w + xc) * cout + co0 + dcg * 8);
    }
  };
  auto store = [&]() {
    const u32x4 zero = u32x4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < XIT; ++k) {
      const int job = tid + k * kBlock;
      if (job < XJ) {
        const int pix = job >> 2, cg = job & 3;
        *reinterpret_cast<u32x4*>(&xs[pix / kWtWW][pix % kWtWW][cg * 8]) =
            ((okx >> k) & 1u) ? xv[k] : zero;
      }
    }
#pragma unroll
    for (int k = 0; k < DIT; ++k) {
      const int pix = (tid + k * kBlock) / DG;
      const u32x4 v = ((okd >> k) & 1u) ? dv[k] : zero;
      if (bias) {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          bpart[e] += __uint_as_float((e & 1) ? (v[e >> 1] & 0xffff0000u)
                                              : (v[e >> 1] << 16));
      }
      *reinterpret_cast<u32x4*>(&ds[pix / kWtW][pix % kWtW][dcg * 8]) = v;
    }
  };

  f32x4 acc[NBW][9];
#pragma unroll
  for (int q = 0; q < NBW; ++q)
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[q][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  const long long t_end = min(ntiles, t_begin + tiles_per_block);
  load(t_begin);
  for (long long t = t_begin; t < t_end; ++t) {
    store();
    __syncthreads();
    if (t + 1 < t_end) load(t + 1);
#pragma unroll
    for (int s = 0; s < kWtH / 2; ++s) {
      const int row = 2 * s + (g >> 1), c0 = 8 * (g & 1);
      bf16x8 B[NBW];
#pragma unroll
      for (int q = 0; q < NBW; ++q) {
        const int cb = (tb * NBW + q) * 16 + 4 * pq;
        B[q] = tr_pair(&ds[row][c0 + pj][cb], &ds[row][c0 + 4 + pj][cb]);
      }
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int wy = row + 1 + tap / 3 - 1, wx = c0 + 1 + tap % 3 - 1;
        const int cb = ta * 16 + 4 * pq;
        const bf16x8 A =
            tr_pair(&xs[wy][wx + pj][cb], &xs[wy][wx + 4 + pj][cb]);
#pragma unroll
        for (int q = 0; q < NBW; ++q)
          acc[q][tap] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              A, B[q], acc[q][tap], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // bias: per-channel partial sums of each thread's gradient loads
  if (bias) {
#pragma unroll
    for (int e = 0; e < 8; ++e) atomicAdd(&bsum[dcg * 8 + e], bpart[e]);
  }
  __syncthreads();
  if (bias && tid < CO) atomicAdd(&db[co0 + tid], bsum[tid]);
  // D[ci][co]: lane holds col co = co0 + 16 (NBW tb + q) + i, rows
  // ci = ci0 + 16 ta + 4 g + r
#pragma unroll
  for (int q = 0; q < NBW; ++q) {
    const int co = co0 + (tb * NBW + q) * 16 + i;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ci = ci0 + ta * 16 + 4 * g + r;
        if (ci < cin)
          atomicAdd(&dw[((long long)tap * cin + ci) * cout + co],
                    acc[q][tap][r]);
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
                    bf16_t* __restrict__ dx) {
  // four threads per pixel, eight channels each: one 16-byte load of the
  // input, one 16-byte store of its gradient masked by input > 0 (the
  // backward of the ReLU that made the input), in bf16
  __shared__ float ws[32 * 3];
  __shared__ float wsum[kBlock / kWave][99];
  for (int k = threadIdx.x; k < 96; k += kBlock) ws[k] = bf2f(f2bf(w[k]));
  __syncthreads();
  const int c8 = threadIdx.x & 3;
  float gw[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) gw[k] = 0.0f;
  float gb[3] = {0.f, 0.f, 0.f};
  for (long long id = blockIdx.x * (long long)kBlock + threadIdx.x; id < 4 * p;
       id += (long long)gridDim.x * kBlock) {
    const long long q = id >> 2;
    const float d0 = dout[q * 3 + 0], d1 = dout[q * 3 + 1],
                d2 = dout[q * 3 + 2];
    if (c8 == 0) {
      gb[0] += d0;
      gb[1] += d1;
      gb[2] += d2;
    }
    const u32x4 xv =
        *reinterpret_cast<const u32x4*>(x + q * x_stride + x_off + c8 * 8);
    f32x4 o[2];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int c = c8 * 8 + k;
      const float v = __uint_as_float((k & 1) ? (xv[k >> 1] & 0xffff0000u)
                                              : (xv[k >> 1] << 16));
      gw[k * 3 + 0] += v * d0;
      gw[k * 3 + 1] += v * d1;
      gw[k * 3 + 2] += v * d2;
      o[k >> 2][k & 3] =
          d0 * ws[c * 3 + 0] + d1 * ws[c * 3 + 1] + d2 * ws[c * 3 + 2];
    }
    *reinterpret_cast<u32x4*>(dx + q * 32 + c8 * 8) =
        relu_mask_bf16(xv, o[0], o[1]);
  }
  // block sums of the 99 parameter gradients: butterfly sums over the lanes
  // of a wave that hold the same channels (fixed indices, so the
  // accumulators stay in registers), the four wave sums through LDS, one
  // atomic per parameter and block
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    float v = k < 24 ? gw[k] : gb[k - 24];
#pragma unroll
    for (int o = kWave / 2; o >= 4; o >>= 1) v += __shfl_xor(v, o);
    if (lane < 4) {
      if (k < 24)
        wsum[wave][(lane * 8 + k / 3) * 3 + k % 3] = v;
      else if (lane == 0)
        wsum[wave][96 + k - 24] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < 99) {
    float v = 0.0f;
#pragma unroll
    for (int q = 0; q < kBlock / kWave; ++q) v += wsum[q][threadIdx.x];
    atomicAdd(threadIdx.x < 96 ? &dw[threadIdx.x] : &db[threadIdx.x - 96], v);
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

namespace {
std::atomic<int> g_tile16_min_blocks{512};
}  // namespace

void set_tile16_min_blocks(int min_blocks) {
  g_tile16_min_blocks.store(min_blocks, std::memory_order_relaxed);
}

hipError_t conv3x3(const Slice& in, int n, int h, int w, const bf16_t* wk,
                   int cout, const float* bias, bool relu, const Slice* out_bf,
                   float* out_f32, int out_f32_stride, bool accumulate,
                   hipStream_t st, const ConvFolds& folds) {
  const Slice* mask = folds.mask;
  const UpSum* up = folds.up;
  const Slice* pool_out = folds.pool;
  const PoolBwd* pb = folds.pool_bwd;
  const Slice* inu = folds.in_up;
  if (in.ch % 8 || cout % 16 || in.stride % 8 || in.off % 8)
    return hipErrorInvalidValue;
  if (inu && (!inu->base || inu->ch <= 0 || inu->ch % 32 || inu->ch > in.ch ||
              inu->stride % 8 || inu->off % 8 || h % 2 || w % 2 ||
              in.ch % 32 || cout % 32))
    return hipErrorInvalidValue;
  if (in.off < 0 && (!inu || in.off + inu->ch < 0)) return hipErrorInvalidValue;
  if (pb && (!pb->out || !pb->gs || pb->y.ch != cout || pb->y.stride % 8 ||
             pb->y.off % 8 || pb->gs_stride % 8 || pb->gs_off % 8 ||
             pb->gs_off + cout > pb->gs_stride || out_bf || out_f32 || up ||
             in.ch % 32 || cout % 32))
    return hipErrorInvalidValue;
  if (mask && (!out_bf || mask->ch != cout || mask->stride % 8 ||
               mask->off % 8 || in.ch % 32 || cout % 32))
    return hipErrorInvalidValue;
  if (up && (!up->out || up->ch <= 0 || up->ch % 64 || up->ch > cout ||
             up->mask.ch != up->ch || up->mask.stride % 8 ||
             up->mask.off % 8 || h % 2 || w % 2 || in.ch % 32 ||
             cout % 32 || (up->ch < cout && !out_f32 == !out_bf) || mask ||
             accumulate))
    return hipErrorInvalidValue;
  if (pool_out && (!out_bf || pool_out->ch != cout || pool_out->stride % 8 ||
                   pool_out->off % 8 || in.ch % 32 || cout % 32 || h % 2 ||
                   w % 2))
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
  a.mask = mask ? static_cast<const bf16_t*>(mask->base) : nullptr;
  a.mask_stride = mask ? mask->stride : 0;
  a.mask_off = mask ? mask->off : 0;
  a.up_out = up ? up->out : nullptr;
  a.up_ch = up ? up->ch : 0;
  a.up_mask = up ? static_cast<const bf16_t*>(up->mask.base) : nullptr;
  a.up_mask_stride = up ? up->mask.stride : 0;
  a.up_mask_off = up ? up->mask.off : 0;
  a.pool_out = pool_out ? static_cast<bf16_t*>(pool_out->base) : nullptr;
  a.pool_stride = pool_out ? pool_out->stride : 0;
  a.pool_off = pool_out ? pool_out->off : 0;
  a.pb_out = pb ? pb->out : nullptr;
  a.pb_gs = pb ? pb->gs : nullptr;
  a.pb_gs_stride = pb ? pb->gs_stride : 0;
  a.pb_gs_off = pb ? pb->gs_off : 0;
  a.pb_y = pb ? static_cast<const bf16_t*>(pb->y.base) : nullptr;
  a.pb_y_stride = pb ? pb->y.stride : 0;
  a.pb_y_off = pb ? pb->y.off : 0;
  a.inu = inu ? static_cast<const bf16_t*>(inu->base) : nullptr;
  a.inu_stride = inu ? inu->stride : 0;
  a.inu_off = inu ? inu->off : 0;
  a.inu_ch = inu ? inu->ch : 0;
  const long long P = (long long)n * h * w;
  const bool vec_out = (!out_bf || (out_bf->stride % 8 == 0 &&
                                    out_bf->off % 8 == 0)) &&
                       (!out_f32 || out_f32_stride % 4 == 0);
  if (in.ch % 32 == 0 && cout % 32 == 0 && vec_out) {
    // LDS-tiled: 8 x 16-pixel tiles, 64 (or 32) output channels per block
    constexpr int LMT = 2;
    const unsigned tiles = (unsigned)n * ((h + 4 * LMT - 1) / (4 * LMT)) *
                           ((w + 15) / 16);
    const unsigned tiles4 = (unsigned)n * ((h + 15) / 16) * ((w + 15) / 16);
    const int min16 = g_tile16_min_blocks.load(std::memory_order_relaxed);
    if (cout % 64 == 0 && in.ch >= 128 && min16 >= 0 &&
        (long long)tiles4 * (cout / 64) >= min16) {
      // deep inputs: 16 x 16-pixel tiles, twice the MFMAs per staged weight
      // (when they still give two blocks a CU: small inputs, such as one
      // 256 x 256 image, keep the 8 x 16 tiles, whose extra blocks hide
      // more latency)
      if (inu)
        hipLaunchKernelGGL((conv3x3_lds_kernel<4, 4, true>),
                           dim3(xcd_grid(tiles4, cout / 64)), dim3(kBlock), 0,
                           st, a);
      else
        hipLaunchKernelGGL((conv3x3_lds_kernel<4, 4, false>),
                           dim3(xcd_grid(tiles4, cout / 64)), dim3(kBlock), 0,
                           st, a);
    } else if (cout % 64 == 0) {
      if (inu)
        hipLaunchKernelGGL((conv3x3_lds_kernel<LMT, 4, true>),
                           dim3(xcd_grid(tiles, cout / 64)), dim3(kBlock), 0,
                           st, a);
      else
        hipLaunchKernelGGL((conv3x3_lds_kernel<LMT, 4, false>),
                           dim3(xcd_grid(tiles, cout / 64)), dim3(kBlock), 0,
                           st, a);
    } else if (inu) {
      hipLaunchKernelGGL((conv3x3_lds_kernel<LMT, 2, true>),
                         dim3(xcd_grid(tiles, cout / 32)), dim3(kBlock), 0, st,
                         a);
    } else {
      hipLaunchKernelGGL((conv3x3_lds_kernel<LMT, 2, false>),
                         dim3(xcd_grid(tiles, cout / 32)), dim3(kBlock), 0, st,
                         a);
    }
    return hipGetLastError();
  }
  if (mask || up || pool_out || pb || inu)
    return hipErrorInvalidValue;  // unaligned output slice
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
                         hipStream_t st, const Slice* xu) {
  if (x.ch % 8 || cout % 8 || x.stride % 8 || x.off % 8)
    return hipErrorInvalidValue;
  if (xu && (xu->ch <= 0 || xu->ch % 32 || xu->ch > x.ch || xu->stride % 8 ||
             xu->off % 8 || h % 2 || w % 2 || cout % 32))
    return hipErrorInvalidValue;
  if (x.off < 0 && (!xu || x.off + xu->ch < 0)) return hipErrorInvalidValue;
  const bf16_t* xu_p = xu ? static_cast<const bf16_t*>(xu->base) : nullptr;
  const int xu_stride = xu ? xu->stride : 0, xu_off = xu ? xu->off : 0;
  const int xu_ch = xu ? xu->ch : 0;
  if (cout % 32 == 0) {
    constexpr int kWtH = 16;  // 16 x 16 tiles: 1.27x window halo, not 1.41x
    const long long ntiles = (long long)n * ((h + kWtH - 1) / kWtH) *
                             ((w + kWtW - 1) / kWtW);
    const int co = cout % 64 == 0 ? 64 : 32;
    const int tiles = ((x.ch + 31) / 32) * (cout / co);
    // aim at about 1024 blocks in all
    long long per = (ntiles * tiles + 1023) / 1024;
    if (per < 1) per = 1;
    const int xb = (int)((ntiles + per - 1) / per);
    const dim3 grid(xcd_grid(xb, tiles));
    auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, st,
                         static_cast<const bf16_t*>(x.base), x.stride, x.off,
                         x.ch, n, h, w, dy, cout, dw, db, (int)per, xu_p,
                         xu_stride, xu_off, xu_ch);
    };
    if (co == 64)
      launch(xu ? wgrad_lds_kernel<64, kWtH, true>
                : wgrad_lds_kernel<64, kWtH, false>);
    else
      launch(xu ? wgrad_lds_kernel<32, kWtH, true>
                : wgrad_lds_kernel<32, kWtH, false>);
    return hipGetLastError();
  }
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

hipError_t conv3x3_in8(const float* x, bf16_t* x_bf, int n, int h, int w,
                       const bf16_t* wk, int cout, const float* bias,
                       bool relu, const Slice& out, hipStream_t st) {
  if (!x || !wk || cout % 32 || out.stride % 8 || out.off % 8 ||
      reinterpret_cast<uintptr_t>(x) % 16 ||
      reinterpret_cast<uintptr_t>(x_bf) % 16)
    return hipErrorInvalidValue;
  ConvArgs a{};
  a.in_f32 = x;
  a.in_copy = x_bf;
  a.cin = 8;
  a.n = n;
  a.h = h;
  a.w = w;
  a.wk = wk;
  a.cout = cout;
  a.bias = bias;
  a.relu = relu;
  a.out_bf = static_cast<bf16_t*>(out.base);
  a.out_stride = out.stride;
  a.out_off = out.off;
  // 16 x 16-pixel tiles, 32 output channels a block
  constexpr int MT = 4, NT = 2;
  const long long tiles = (long long)n * ((h + 4 * MT - 1) / (4 * MT)) *
                          ((w + 15) / 16);
  hipLaunchKernelGGL((conv3x3_in8_kernel<MT, NT>),
                     dim3(xcd_grid(tiles, cout / (16 * NT))), dim3(kBlock), 0,
                     st, a);
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
                    float* dw, float* db, bf16_t* dx, hipStream_t st) {
  if (x.ch != 32 || x.stride % 8 || x.off % 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(head_bwd_kernel, dim3(blocks_for(4LL * p, 1024)),
                     dim3(kBlock),
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
