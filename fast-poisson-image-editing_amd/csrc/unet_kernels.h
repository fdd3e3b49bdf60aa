// Kernels of the warm-start U-Net (bf16 activations and weights, fp32
// accumulation on MFMA).  Activation tensors are NHWC bf16 with an explicit
// channel stride and offset, so a convolution can read from or write into a
// slice of a concatenation buffer.  Gradient tensors are NHWC fp32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pie {
namespace unet {

typedef uint16_t bf16_t;  // raw bfloat16 bits

// A channel slice of an NHWC tensor: element (pixel p, channel c) at
// base[p * stride + off + c].
struct Slice {
  void* base;
  int stride, off, ch;
};

// 3x3 convolution with zero padding as an implicit GEMM on
// v_mfma_f32_16x16x32_bf16: out[p][co] = sum_k A[p][k] B[k][co] with
// k = 8 * kb + e, kb = tap * (Cin / 8) + cb, A = in[p + offset(tap)][8cb+e].
// wk holds B as [kb][Cout][8] bf16.  Output: bf16 (+bias, optional ReLU)
// into out_bf, and/or fp32 into out_f32 (optionally accumulated).
//
// Epilogue folds (ConvFolds; the LDS-tiled path only: Cin and Cout
// multiples of 32, otherwise hipErrorInvalidValue; each rounds the fp32
// value once, as the separate pass it replaces did):
// - mask (Cout bf16 channels; needs out_bf): the bf16 output element (p, co)
//   is written as mask[p][co] > 0 ? v : 0, the backward of a ReLU whose
//   output is the mask;
// - up (data gradient of a convolution whose input is a nearest 2x upsample
//   in its first up->ch channels, concatenated with a skip connection): the
//   first up->ch output channels are not written to out_f32 or out_bf (one
//   of the two takes the skip channels, out_bf's slice offset then counting
//   from channel 0, so an offset of -up->ch starts them at 0) but summed over
//   each 2 x 2 pixel block (fine pixels (0,0), (0,1), (1,0), (1,1) in that
//   order), masked by up->mask > 0 of the half-resolution level and written
//   as bf16 [n][h/2][w/2][up->ch] to up->out (the upsampling backward with
//   the backward of the ReLU that made its input); up->ch a multiple of 64,
//   h and w even, no mask;
// - pool (Cout channels; needs out_bf): the bf16 output's 2 x 2 average
//   pool (the four bf16 values summed in the order (0,0), (0,1), (1,0),
//   (1,1), times 1/4, rounded; h and w even) is written there as well;
// - pool_bwd (data gradient of a convolution whose input is the 2 x 2
//   average pool of a skip connection; no out_bf, no out_f32): for each of
//   the four pixels q of the double-resolution level under output pixel p,
//   out[q][co] = bf16(y[q][co] > 0 ? gs[q][co] + v[p][co] / 4 : 0) (gs: the
//   skip channels' bf16 gradient from the decoder, y: the skip output),
//   out [n][2h][2w][Cout].
struct UpSum {
  bf16_t* out;
  int ch;
  Slice mask;
};
struct PoolBwd {
  bf16_t* out;
  const bf16_t* gs;
  int gs_stride, gs_off;
  Slice y;
};
struct ConvFolds {
  // not a fold: the input's first in_up->ch channels (a multiple of 32) are
  // the nearest 2x upsample of in_up, read at its half resolution (pixel
  // (y, x) of image n <- in_up's (n, y / 2, x / 2)); in's slice offset may
  // then be negative, as long as the channels it addresses are >= in_up->ch
  const Slice* in_up = nullptr;
  const Slice* mask = nullptr;
  const UpSum* up = nullptr;
  const Slice* pool = nullptr;
  const PoolBwd* pool_bwd = nullptr;
};
// Layers with 128+ input channels (and Cout a multiple of 64) use 16 x 16
// tiles when that gives at least this many blocks (default 512; 0 =
// always, -1 = never); the result does not depend on it.
void set_tile16_min_blocks(int min_blocks);
hipError_t conv3x3(const Slice& in, int n, int h, int w, const bf16_t* wk,
                   int cout, const float* bias, bool relu, const Slice* out_bf,
                   float* out_f32, int out_f32_stride, bool accumulate,
                   hipStream_t st, const ConvFolds& folds = ConvFolds{});

// dW[tap][ci][co] += sum_p x[p + offset(tap)][ci] * dy[p][co] (fp32
// atomics) and db[co] += sum_p dy[p][co]; x is a bf16 slice, dy bf16 with
// Cout contiguous channels; with xu, x's first xu->ch channels are read at
// half resolution from xu, as ConvFolds::in_up (Cout a multiple of 32).
hipError_t conv3x3_wgrad(const Slice& x, int n, int h, int w,
                         const bf16_t* dy, int cout, float* dw, float* db,
                         hipStream_t st, const Slice* xu = nullptr);

// Pack fp32 master weights w[tap][ci][co] into the forward operand
// wk[kb][co][8] (kb = tap*(ci/8)+cb) and the data-gradient operand
// wkt[kb'][ci][8] (kb' = tap'*(co/8)+ob, value w[8-tap'][ci][8ob+e]).
hipError_t pack_weights(const float* w, int cin, int cout, bf16_t* wk,
                        bf16_t* wkt, hipStream_t st);

// The first convolution (8 input channels) straight from the fp32 input
// x [P][8] (16-byte aligned): each value rounded to bf16 (nearest even;
// written to x_bf [P][8] unless x_bf is null), then as conv3x3 with a bf16
// output slice and no folds; Cout a multiple of 32.
hipError_t conv3x3_in8(const float* x, bf16_t* x_bf, int n, int h, int w,
                       const bf16_t* wk, int cout, const float* bias,
                       bool relu, const Slice& out, hipStream_t st);
// head: out[p][0..3) = sum_c x[p][c] * bf16(w[c][o]) + b[o]  (x bf16 32 ch)
hipError_t head_fwd(const Slice& x, int p, const float* w, const float* b,
                    float* out, hipStream_t st);
// head backward: dw, db (atomics), dx[p][c] = bf16(x[p][c] > 0 ?
// sum_o dout[p][o] bf16(w[c][o]) : 0) (with the backward of the ReLU that
// made x), dx [P][32]
hipError_t head_bwd(const Slice& x, int p, const float* w, const float* dout,
                    float* dw, float* db, bf16_t* dx, hipStream_t st);
// masked mean squared error over solved pixels: loss_sum[0] += sum d^2,
// count[0] += solved * 3, dout = 2 d (to be scaled by 1/count)
hipError_t mse(const float* out, const float* target, const uint8_t* mask,
               int p, float* dout, double* acc, hipStream_t st);
hipError_t scale(float* x, long long n, const double* acc, hipStream_t st);
// Adam on fp32 master weights
hipError_t adam(float* p, const float* g, float* m, float* v, long long n,
                float lr, float b1, float b2, float eps, int t,
                hipStream_t st);

}  // namespace unet
}  // namespace pie
