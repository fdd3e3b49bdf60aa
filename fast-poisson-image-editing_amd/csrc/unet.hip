// Warm-start U-Net runtime (C ABI pie_unet_*): owns the parameters, Adam
// state, bf16 operand copies of the weights and every activation and
// gradient buffer of one batch shape, and runs forward, backward and the
// optimiser step with the kernels of unet_kernels.hip.
//
// Network (no counterpart in the reference): 4 resolution levels, base 32
// channels, 3x3 convolutions with bias and ReLU, 2x2 average pooling,
// nearest 2x upsampling, skip concatenation (upsampled channels first), a
// 1x1 head to 3 channels.  Input: 8 channels (the warm start feeds the
// rms-normalised coarse residual of the correction problem in channels
// 0-2, zeros in 3-5, the mask in 6 and a zero channel); output: 3 channels
// (the shape of the coarse correction), trained with a masked mean squared
// error.
//
// Parameters are one fp32 array, layer by layer: conv w[tap][cin][cout]
// then b[cout]; the head w[32][3] then b[3].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_util.h"
#include "unet_kernels.h"

using pie::host::DevBuf;
using pie::host::fail;
using pie::unet::bf16_t;
using pie::unet::Slice;

namespace {

struct ConvSpec {
  const char* name;
  int cin, cout, level;
};

// name, cin, cout, resolution level (0 = full)
const ConvSpec kConvs[] = {
    {"e0a", 8, 32, 0},    {"e0b", 32, 32, 0},   {"e1a", 32, 64, 1},
    {"e1b", 64, 64, 1},   {"e2a", 64, 128, 2},  {"e2b", 128, 128, 2},
    {"bna", 128, 256, 3}, {"bnb", 256, 256, 3}, {"d2a", 384, 128, 2},
    {"d2b", 128, 128, 2}, {"d1a", 192, 64, 1},  {"d1b", 64, 64, 1},
    {"d0a", 96, 32, 0},   {"d0b", 32, 32, 0},
};
constexpr int kNumConv = sizeof(kConvs) / sizeof(kConvs[0]);
enum { E0A, E0B, E1A, E1B, E2A, E2B, BNA, BNB, D2A, D2B, D1A, D1B, D0A, D0B };

}  // namespace

struct pie_unet {
  int B = 0, H = 0, W = 0;
  long long P[4] = {0, 0, 0, 0};  // pixels per level
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // ordering against a caller's stream (pie::unet_forward_on_stream)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  long long nparam = 0;
  long long woff[kNumConv + 1] = {}, boff[kNumConv + 1] = {};
  DevBuf<float> param, grad, adam_m, adam_v;
  DevBuf<bf16_t> wk[kNumConv], wkt[kNumConv];
  int step = 0;
  // activations (bf16)
  DevBuf<bf16_t> in, e0a, c0, p0, e1a, c1, p1, e2a, c2, p2, bna, bnb, d2a,
      d2b, d1a, d1b, d0a, d0b;
  // fp32 host-facing buffers and gradients
  DevBuf<float> x32, out, dout, target;
  DevBuf<uint8_t> mask;
  DevBuf<bf16_t> gC0, gC1, gC2;
  DevBuf<bf16_t> dyb, dyb2;
  DevBuf<double> acc;
  double ms_total = 0.0, flops_total = 0.0;
  // pipelined training (pie_unet_staging / pie_unet_train_async /
  // pie_unet_finish): two slots, each a pinned host copy of a batch, its
  // device copy, its loss sums and the events that order the upload (on
  // the copy stream) against the steps that read it
  struct Slot {
    float* hx = nullptr;
    float* ht = nullptr;
    uint8_t* hm = nullptr;
    DevBuf<float> x, t;
    DevBuf<uint8_t> m;
    DevBuf<double> acc;
    hipEvent_t uploaded = nullptr, consumed = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;  // device time of the step
    bool used = false;     // an upload was enqueued
    bool pending = false;  // a step was enqueued and not yet finished
  } slot[2];
  hipStream_t copy = nullptr;
};

namespace {

Slice sl(DevBuf<bf16_t>& b, int stride, int off, int ch) {
  return Slice{b.p, stride, off, ch};
}

int alloc_all(pie_unet* u) {
  const long long P0 = u->P[0], P1 = u->P[1], P2 = u->P[2], P3 = u->P[3];
  PIE_HIP(u->in.alloc(P0 * 8));
  PIE_HIP(u->e0a.alloc(P0 * 32));
  PIE_HIP(u->c0.alloc(P0 * 32));
  PIE_HIP(u->p0.alloc(P1 * 32));
  PIE_HIP(u->e1a.alloc(P1 * 64));
  PIE_HIP(u->c1.alloc(P1 * 64));
  PIE_HIP(u->p1.alloc(P2 * 64));
  PIE_HIP(u->e2a.alloc(P2 * 128));
  PIE_HIP(u->c2.alloc(P2 * 128));
  PIE_HIP(u->p2.alloc(P3 * 128));
  PIE_HIP(u->bna.alloc(P3 * 256));
  PIE_HIP(u->bnb.alloc(P3 * 256));
  PIE_HIP(u->d2a.alloc(P2 * 128));
  PIE_HIP(u->d2b.alloc(P2 * 128));
  PIE_HIP(u->d1a.alloc(P1 * 64));
  PIE_HIP(u->d1b.alloc(P1 * 64));
  PIE_HIP(u->d0a.alloc(P0 * 32));
  PIE_HIP(u->d0b.alloc(P0 * 32));
  PIE_HIP(u->x32.alloc(P0 * 8));
  PIE_HIP(u->out.alloc(P0 * 3));
  PIE_HIP(u->dout.alloc(P0 * 3));
  PIE_HIP(u->target.alloc(P0 * 3));
  PIE_HIP(u->mask.alloc(P0));
  // gradient scratch: the largest per-level working set
  long long ga = std::max({P0 * 32, P1 * 64, P2 * 128, P3 * 256});
  PIE_HIP(u->gC0.alloc(P0 * 32));
  PIE_HIP(u->gC1.alloc(P1 * 64));
  PIE_HIP(u->gC2.alloc(P2 * 128));
  PIE_HIP(u->dyb.alloc(ga));
  PIE_HIP(u->dyb2.alloc(ga));
  PIE_HIP(u->acc.alloc(2));
  return PIE_OK;
}

int pack_all(pie_unet* u) {
  for (int l = 0; l < kNumConv; ++l)
    PIE_HIP(pie::unet::pack_weights(u->param.p + u->woff[l], kConvs[l].cin,
                                    kConvs[l].cout, u->wk[l].p, u->wkt[l].p,
                                    u->stream));
  return PIE_OK;
}

double conv_flops(const pie_unet* u, bool backward) {
  double f = 0.0;
  for (int l = 0; l < kNumConv; ++l) {
    const double one = 2.0 * u->P[kConvs[l].level] * 9.0 * kConvs[l].cin *
                       kConvs[l].cout;
    f += one;
    if (backward) f += one * (l == E0A ? 1.0 : 2.0);  // wgrad (+ dgrad)
  }
  return f;
}

// Forward of conv l (bias, ReLU) into the bf16 slice y; with pool, also
// y's 2 x 2 average pool; with xu, x's first xu->ch channels are the 2x
// nearest upsample of xu (read at its resolution).
int conv_fwd(pie_unet* u, int l, const Slice& x, const Slice& y,
             const Slice* pool = nullptr, const Slice* xu = nullptr) {
  const ConvSpec& c = kConvs[l];
  const int s = 1 << c.level;
  pie::unet::ConvFolds f;
  f.pool = pool;
  f.in_up = xu;
  PIE_HIP(pie::unet::conv3x3(x, u->B, u->H / s, u->W / s, u->wk[l].p, c.cout,
                             u->param.p + u->boff[l], true, &y, nullptr, 0,
                             false, u->stream, f));
  return PIE_OK;
}

int forward(pie_unet* u, const float* x) {
  hipStream_t st = u->stream;
  // the first layer reads the fp32 input and leaves its bf16 copy in u->in
  // for the weight gradient
  PIE_HIP(pie::unet::conv3x3_in8(x, u->in.p, u->B, u->H, u->W, u->wk[E0A].p,
                                 kConvs[E0A].cout, u->param.p + u->boff[E0A],
                                 true, sl(u->e0a, 32, 0, 32), st));
  // encoder outputs go to the skip buffers c0, c1, c2, with their average
  // pool for the next level.  A decoder level's first convolution reads
  // the concatenation (upsample of the coarser output) || skip: channels
  // below the upsample's width from the coarser output at its own
  // resolution, the rest from the skip buffer (a slice whose offset starts
  // the skip channels at its channel 0)
  const Slice p0 = sl(u->p0, 32, 0, 32), p1 = sl(u->p1, 64, 0, 64),
              p2 = sl(u->p2, 128, 0, 128);
  PIE_CHECK(conv_fwd(u, E0B, sl(u->e0a, 32, 0, 32), sl(u->c0, 32, 0, 32),
                     &p0));
  PIE_CHECK(conv_fwd(u, E1A, p0, sl(u->e1a, 64, 0, 64)));
  PIE_CHECK(conv_fwd(u, E1B, sl(u->e1a, 64, 0, 64), sl(u->c1, 64, 0, 64),
                     &p1));
  PIE_CHECK(conv_fwd(u, E2A, p1, sl(u->e2a, 128, 0, 128)));
  PIE_CHECK(conv_fwd(u, E2B, sl(u->e2a, 128, 0, 128), sl(u->c2, 128, 0, 128),
                     &p2));
  PIE_CHECK(conv_fwd(u, BNA, p2, sl(u->bna, 256, 0, 256)));
  const Slice bnb = sl(u->bnb, 256, 0, 256), d2b = sl(u->d2b, 128, 0, 128),
              d1b = sl(u->d1b, 64, 0, 64);
  PIE_CHECK(conv_fwd(u, BNB, sl(u->bna, 256, 0, 256), bnb));
  PIE_CHECK(conv_fwd(u, D2A, sl(u->c2, 128, -256, 384),
                     sl(u->d2a, 128, 0, 128), nullptr, &bnb));
  PIE_CHECK(conv_fwd(u, D2B, sl(u->d2a, 128, 0, 128), d2b));
  PIE_CHECK(conv_fwd(u, D1A, sl(u->c1, 64, -128, 192), sl(u->d1a, 64, 0, 64),
                     nullptr, &d2b));
  PIE_CHECK(conv_fwd(u, D1B, sl(u->d1a, 64, 0, 64), d1b));
  PIE_CHECK(conv_fwd(u, D0A, sl(u->c0, 32, -64, 96), sl(u->d0a, 32, 0, 32),
                     nullptr, &d1b));
  PIE_CHECK(conv_fwd(u, D0B, sl(u->d0a, 32, 0, 32), sl(u->d0b, 32, 0, 32)));
  const long long hw = u->woff[kNumConv];  // head weights
  PIE_HIP(pie::unet::head_fwd(sl(u->d0b, 32, 0, 32), (int)u->P[0],
                              u->param.p + hw, u->param.p + hw + 96,
                              u->out.p, st));
  return PIE_OK;
}

// Backward of a pair of convolutions la -> lb, lb's output gradient
// already in dyb (bf16, masked by lb's ReLU): lb's input is la's ReLU output
// xb, and nothing else reads it, so lb's data gradient is written by the
// convolution's epilogue as la's output gradient, masked by xb > 0 and
// rounded to bf16 once from the fp32 sum.  With up, la's data gradient's
// first up->ch channels (an upsampled input) go through the upsampling
// backward into up->out and the others (the skip channels) to dx as bf16
// [P][cin - up->ch]; with pb (an average-pooled input), all of it goes
// through the pooling backward into pb->out; with neither, it is not
// computed (la reads the network input).
int conv_pair_bwd(pie_unet* u, int la, int lb, const Slice& xb,
                  const Slice& xa, bf16_t* dx,
                  const pie::unet::UpSum* up = nullptr,
                  const pie::unet::PoolBwd* pb = nullptr,
                  const Slice* xa_up = nullptr) {
  const ConvSpec& ca = kConvs[la];
  const ConvSpec& cb = kConvs[lb];
  const int s = 1 << cb.level;
  const int h = u->H / s, w = u->W / s;
  hipStream_t st = u->stream;
  PIE_HIP(pie::unet::conv3x3_wgrad(xb, u->B, h, w, u->dyb.p, cb.cout,
                                   u->grad.p + u->woff[lb],
                                   u->grad.p + u->boff[lb], st));
  Slice dys{u->dyb.p, cb.cout, 0, cb.cout};
  Slice dya{u->dyb2.p, cb.cin, 0, cb.cin};
  pie::unet::ConvFolds fb;
  fb.mask = &xb;
  PIE_HIP(pie::unet::conv3x3(dys, u->B, h, w, u->wkt[lb].p, cb.cin, nullptr,
                             false, &dya, nullptr, 0, false, st, fb));
  PIE_HIP(pie::unet::conv3x3_wgrad(xa, u->B, h, w, u->dyb2.p, ca.cout,
                                   u->grad.p + u->woff[la],
                                   u->grad.p + u->boff[la], st, xa_up));
  if (up || pb) {
    pie::unet::ConvFolds fa;
    fa.up = up;
    fa.pool_bwd = pb;
    const int skip = up ? ca.cin - up->ch : 0;
    const Slice dxs{dx, skip, -(up ? up->ch : 0), skip};
    PIE_HIP(pie::unet::conv3x3(dya, u->B, h, w, u->wkt[la].p, ca.cin, nullptr,
                               false, up ? &dxs : nullptr, nullptr, 0, false,
                               st, fa));
  }
  return PIE_OK;
}

int backward(pie_unet* u) {
  hipStream_t st = u->stream;
  const long long hw = u->woff[kNumConv];
  bf16_t* dy = u->dyb.p;
  PIE_HIP(hipMemsetAsync(u->grad.p, 0, u->nparam * sizeof(float), st));
  // decoder: each level's output gradient comes from the head or from the
  // upsampling backward, with that level's ReLU folded in
  PIE_HIP(pie::unet::head_bwd(sl(u->d0b, 32, 0, 32), (int)u->P[0],
                              u->param.p + hw, u->dout.p, u->grad.p + hw,
                              u->grad.p + hw + 96, dy, st));
  // the decoder's first convolutions take an upsampled input: their data
  // gradient's epilogue makes the coarser level's output gradient
  const Slice bnb = sl(u->bnb, 256, 0, 256), d2b = sl(u->d2b, 128, 0, 128),
              d1b = sl(u->d1b, 64, 0, 64);
  const pie::unet::UpSum up0{dy, 64, d1b};
  PIE_CHECK(conv_pair_bwd(u, D0A, D0B, sl(u->d0a, 32, 0, 32),
                          sl(u->c0, 32, -64, 96), u->gC0.p, &up0, nullptr,
                          &d1b));
  const pie::unet::UpSum up1{dy, 128, d2b};
  PIE_CHECK(conv_pair_bwd(u, D1A, D1B, sl(u->d1a, 64, 0, 64),
                          sl(u->c1, 64, -128, 192), u->gC1.p, &up1, nullptr,
                          &d2b));
  const pie::unet::UpSum up2{dy, 256, bnb};
  PIE_CHECK(conv_pair_bwd(u, D2A, D2B, sl(u->d2a, 128, 0, 128),
                          sl(u->c2, 128, -256, 384), u->gC2.p, &up2, nullptr,
                          &bnb));
  // encoder: the first convolutions take an average-pooled input: their
  // data gradient's epilogue adds it (a quarter to each of the four pixels)
  // to the skip channels' gradient from the decoder and applies the ReLU of
  // the finer level's output
  const pie::unet::PoolBwd pb2{dy, u->gC2.p, 128, 0, sl(u->c2, 128, 0, 128)};
  PIE_CHECK(conv_pair_bwd(u, BNA, BNB, sl(u->bna, 256, 0, 256),
                          sl(u->p2, 128, 0, 128), nullptr, nullptr, &pb2));
  const pie::unet::PoolBwd pb1{dy, u->gC1.p, 64, 0, sl(u->c1, 64, 0, 64)};
  PIE_CHECK(conv_pair_bwd(u, E2A, E2B, sl(u->e2a, 128, 0, 128),
                          sl(u->p1, 64, 0, 64), nullptr, nullptr, &pb1));
  const pie::unet::PoolBwd pb0{dy, u->gC0.p, 32, 0, sl(u->c0, 32, 0, 32)};
  PIE_CHECK(conv_pair_bwd(u, E1A, E1B, sl(u->e1a, 64, 0, 64),
                          sl(u->p0, 32, 0, 32), nullptr, nullptr, &pb0));
  PIE_CHECK(conv_pair_bwd(u, E0A, E0B, sl(u->e0a, 32, 0, 32),
                          sl(u->in, 8, 0, 8), nullptr));
  return PIE_OK;
}

int load_batch(pie_unet* u, const float* x, const float* target,
               const uint8_t* mask) {
  const long long P0 = u->P[0];
  PIE_HIP(hipMemcpyAsync(u->x32.p, x, P0 * 8 * sizeof(float),
                         hipMemcpyHostToDevice, u->stream));
  if (target)
    PIE_HIP(hipMemcpyAsync(u->target.p, target, P0 * 3 * sizeof(float),
                           hipMemcpyHostToDevice, u->stream));
  if (mask)
    PIE_HIP(hipMemcpyAsync(u->mask.p, mask, P0, hipMemcpyHostToDevice,
                           u->stream));
  return PIE_OK;
}

// The masked loss's sums into acc (device) and its gradient back through
// the network; `consumed` (if given) is recorded once target and mask are
// read.
int loss_backward(pie_unet* u, const float* target, const uint8_t* mask,
                  double* acc, hipEvent_t consumed) {
  PIE_HIP(hipMemsetAsync(acc, 0, 2 * sizeof(double), u->stream));
  PIE_HIP(pie::unet::mse(u->out.p, target, mask, (int)u->P[0], u->dout.p,
                         acc, u->stream));
  if (consumed) PIE_HIP(hipEventRecord(consumed, u->stream));
  PIE_HIP(pie::unet::scale(u->dout.p, u->P[0] * 3, acc, u->stream));
  PIE_CHECK(backward(u));
  return PIE_OK;
}

int read_loss(pie_unet* u, const double* acc_dev, float* loss) {
  double acc[2];
  PIE_HIP(hipMemcpyAsync(acc, acc_dev, sizeof(acc), hipMemcpyDeviceToHost,
                         u->stream));
  PIE_HIP(hipStreamSynchronize(u->stream));
  if (loss) *loss = (float)(acc[0] / std::max(1.0, acc[1]));
  return PIE_OK;
}

int loss_and_grad(pie_unet* u, float* loss) {
  PIE_CHECK(loss_backward(u, u->target.p, u->mask.p, u->acc.p, nullptr));
  return read_loss(u, u->acc.p, loss);
}

// Adam on the gradient, then the bf16 operand copies of the new weights.
int optimiser_step(pie_unet* u, float lr) {
  ++u->step;
  PIE_HIP(pie::unet::adam(u->param.p, u->grad.p, u->adam_m.p, u->adam_v.p,
                          u->nparam, lr, 0.9f, 0.999f, 1e-8f, u->step,
                          u->stream));
  return pack_all(u);
}

int slot_alloc(pie_unet* u, int k) {
  pie_unet::Slot& sl = u->slot[k];
  if (sl.hx) return PIE_OK;
  const long long P0 = u->P[0];
  PIE_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.hx),
                        P0 * 8 * sizeof(float), hipHostMallocDefault));
  PIE_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.ht),
                        P0 * 3 * sizeof(float), hipHostMallocDefault));
  PIE_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.hm), P0,
                        hipHostMallocDefault));
  PIE_HIP(sl.x.alloc(P0 * 8));
  PIE_HIP(sl.t.alloc(P0 * 3));
  PIE_HIP(sl.m.alloc(P0));
  PIE_HIP(sl.acc.alloc(2));
  PIE_HIP(hipEventCreateWithFlags(&sl.uploaded, hipEventDisableTiming));
  PIE_HIP(hipEventCreateWithFlags(&sl.consumed, hipEventDisableTiming));
  PIE_HIP(hipEventCreate(&sl.t0));
  PIE_HIP(hipEventCreate(&sl.t1));
  if (!u->copy)
    PIE_HIP(hipStreamCreateWithFlags(&u->copy, hipStreamNonBlocking));
  return PIE_OK;
}

void slot_free(pie_unet* u) {
  for (auto& sl : u->slot) {
    if (sl.hx) (void)hipHostFree(sl.hx);
    if (sl.ht) (void)hipHostFree(sl.ht);
    if (sl.hm) (void)hipHostFree(sl.hm);
    for (hipEvent_t e : {sl.uploaded, sl.consumed, sl.t0, sl.t1})
      if (e) (void)hipEventDestroy(e);
  }
}

}  // namespace

namespace pie {

// The forward pass on device buffers, ordered after the work already
// queued on `st` and before the work queued on it afterwards (events
// between the caller's stream and the network's; no host wait).  st null:
// ordered on the network's own stream only.
int unet_forward_on_stream(pie_unet* u, int h, int w, const float* x,
                           float* out, hipStream_t st) {
  if (!u || !x || !out) return fail(PIE_EINVAL, "NULL argument");
  if (u->B != 1 || u->H != h || u->W != w)
    return fail(PIE_EINVAL, "network shape differs from (1, h, w)");
  const long long P0 = u->P[0];
  if (st) {
    PIE_HIP(hipEventRecord(u->ev_in, st));
    PIE_HIP(hipStreamWaitEvent(u->stream, u->ev_in, 0));
  }
  PIE_CHECK(forward(u, x));
  PIE_HIP(hipMemcpyAsync(out, u->out.p, P0 * 3 * sizeof(float),
                         hipMemcpyDeviceToDevice, u->stream));
  if (st) {
    PIE_HIP(hipEventRecord(u->ev_out, u->stream));
    PIE_HIP(hipStreamWaitEvent(st, u->ev_out, 0));
  }
  return PIE_OK;
}

}  // namespace pie

extern "C" {

int pie_unet_create(int batch, int h, int w, pie_unet** out) {
  if (!out) return fail(PIE_EINVAL, "out is NULL");
  *out = nullptr;
  if (batch < 1 || h < 8 || w < 8 || h % 8 || w % 8)
    return fail(PIE_EINVAL, "batch >= 1 and h, w positive multiples of 8");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(PIE_EHIP, "no HIP device");
  std::unique_ptr<pie_unet> u(new pie_unet());
  u->B = batch;
  u->H = h;
  u->W = w;
  for (int l = 0; l < 4; ++l)
    u->P[l] = (long long)batch * (h >> l) * (w >> l);
  long long off = 0;
  for (int l = 0; l < kNumConv; ++l) {
    u->woff[l] = off;
    off += 9LL * kConvs[l].cin * kConvs[l].cout;
    u->boff[l] = off;
    off += kConvs[l].cout;
  }
  u->woff[kNumConv] = off;  // head: w[32][3], b[3]
  u->boff[kNumConv] = off + 96;
  off += 99;
  u->nparam = off;
  PIE_HIP(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
  PIE_HIP(hipEventCreate(&u->ev0));
  PIE_HIP(hipEventCreate(&u->ev1));
  PIE_HIP(hipEventCreateWithFlags(&u->ev_in, hipEventDisableTiming));
  PIE_HIP(hipEventCreateWithFlags(&u->ev_out, hipEventDisableTiming));
  PIE_HIP(u->param.alloc(off));
  PIE_HIP(u->grad.alloc(off));
  PIE_HIP(u->adam_m.alloc(off));
  PIE_HIP(u->adam_v.alloc(off));
  PIE_HIP(hipMemsetAsync(u->param.p, 0, off * sizeof(float), u->stream));
  PIE_HIP(hipMemsetAsync(u->adam_m.p, 0, off * sizeof(float), u->stream));
  PIE_HIP(hipMemsetAsync(u->adam_v.p, 0, off * sizeof(float), u->stream));
  for (int l = 0; l < kNumConv; ++l) {
    PIE_HIP(u->wk[l].alloc(9LL * kConvs[l].cin * kConvs[l].cout));
    PIE_HIP(u->wkt[l].alloc(9LL * kConvs[l].cin * kConvs[l].cout));
  }
  PIE_CHECK(alloc_all(u.get()));
  PIE_HIP(hipStreamSynchronize(u->stream));
  *out = u.release();
  return PIE_OK;
}

int pie_unet_destroy(pie_unet* u) {
  if (!u) return PIE_OK;
  if (u->stream) (void)hipStreamSynchronize(u->stream);
  if (u->copy) (void)hipStreamSynchronize(u->copy);
  slot_free(u);
  if (u->copy) (void)hipStreamDestroy(u->copy);
  if (u->ev0) (void)hipEventDestroy(u->ev0);
  if (u->ev1) (void)hipEventDestroy(u->ev1);
  if (u->ev_in) (void)hipEventDestroy(u->ev_in);
  if (u->ev_out) (void)hipEventDestroy(u->ev_out);
  hipStream_t st = u->stream;
  delete u;
  if (st) (void)hipStreamDestroy(st);
  return PIE_OK;
}

int pie_unet_param_count(pie_unet* u, long long* n) {
  if (!u || !n) return fail(PIE_EINVAL, "NULL argument");
  *n = u->nparam;
  return PIE_OK;
}

int pie_unet_set_params(pie_unet* u, const float* p) {
  if (!u || !p) return fail(PIE_EINVAL, "NULL argument");
  PIE_HIP(hipMemcpyAsync(u->param.p, p, u->nparam * sizeof(float),
                         hipMemcpyHostToDevice, u->stream));
  PIE_HIP(hipMemsetAsync(u->adam_m.p, 0, u->nparam * sizeof(float),
                         u->stream));
  PIE_HIP(hipMemsetAsync(u->adam_v.p, 0, u->nparam * sizeof(float),
                         u->stream));
  u->step = 0;
  PIE_CHECK(pack_all(u));
  PIE_HIP(hipStreamSynchronize(u->stream));
  return PIE_OK;
}

int pie_unet_get_params(pie_unet* u, float* p) {
  if (!u || !p) return fail(PIE_EINVAL, "NULL argument");
  PIE_HIP(hipMemcpyAsync(p, u->param.p, u->nparam * sizeof(float),
                         hipMemcpyDeviceToHost, u->stream));
  PIE_HIP(hipStreamSynchronize(u->stream));
  return PIE_OK;
}

int pie_unet_forward(pie_unet* u, const float* x, float* out) {
  if (!u || !x || !out) return fail(PIE_EINVAL, "NULL argument");
  PIE_CHECK(load_batch(u, x, nullptr, nullptr));
  PIE_CHECK(forward(u, u->x32.p));
  PIE_HIP(hipMemcpyAsync(out, u->out.p, u->P[0] * 3 * sizeof(float),
                         hipMemcpyDeviceToHost, u->stream));
  PIE_HIP(hipStreamSynchronize(u->stream));
  return PIE_OK;
}

int pie_unet_forward_device(pie_unet* u, int h, int w, const float* x,
                            float* out) {
  PIE_CHECK(pie::unet_forward_on_stream(u, h, w, x, out, nullptr));
  PIE_HIP(hipStreamSynchronize(u->stream));
  return PIE_OK;
}

int pie_unet_loss_grad(pie_unet* u, const float* x, const float* target,
                       const uint8_t* mask, float* loss, float* grads) {
  if (!u || !x || !target || !mask) return fail(PIE_EINVAL, "NULL argument");
  PIE_CHECK(load_batch(u, x, target, mask));
  PIE_CHECK(forward(u, u->x32.p));
  PIE_CHECK(loss_and_grad(u, loss));
  if (grads) {
    PIE_HIP(hipMemcpyAsync(grads, u->grad.p, u->nparam * sizeof(float),
                           hipMemcpyDeviceToHost, u->stream));
    PIE_HIP(hipStreamSynchronize(u->stream));
  }
  return PIE_OK;
}

int pie_unet_train_step(pie_unet* u, const float* x, const float* target,
                        const uint8_t* mask, float lr, float* loss) {
  if (!u || !x || !target || !mask) return fail(PIE_EINVAL, "NULL argument");
  PIE_CHECK(load_batch(u, x, target, mask));
  PIE_HIP(hipEventRecord(u->ev0, u->stream));
  PIE_CHECK(forward(u, u->x32.p));
  PIE_CHECK(loss_and_grad(u, loss));  // synchronises
  PIE_HIP(hipEventRecord(u->ev1, u->stream));
  PIE_CHECK(optimiser_step(u, lr));
  PIE_HIP(hipStreamSynchronize(u->stream));
  float ms = 0.0f;
  PIE_HIP(hipEventElapsedTime(&ms, u->ev0, u->ev1));
  u->ms_total += ms;
  u->flops_total += conv_flops(u, true);
  return PIE_OK;
}

int pie_unet_staging(pie_unet* u, int k, float** x, float** target,
                     uint8_t** mask) {
  if (!u || !x || !target || !mask) return fail(PIE_EINVAL, "NULL argument");
  if (k < 0 || k > 1) return fail(PIE_EINVAL, "slot must be 0 or 1");
  PIE_CHECK(slot_alloc(u, k));
  pie_unet::Slot& sl = u->slot[k];
  // the host copy is free once its last upload has been read
  if (sl.used) PIE_HIP(hipEventSynchronize(sl.uploaded));
  *x = sl.hx;
  *target = sl.ht;
  *mask = sl.hm;
  return PIE_OK;
}

int pie_unet_train_async(pie_unet* u, int k, float lr) {
  if (!u) return fail(PIE_EINVAL, "NULL argument");
  if (k < 0 || k > 1 || !u->slot[k].hx)
    return fail(PIE_EINVAL, "slot not set up by pie_unet_staging");
  pie_unet::Slot& sl = u->slot[k];
  if (sl.pending)
    return fail(PIE_ESTATE, "the slot's last step is not finished");
  const long long P0 = u->P[0];
  // upload on the copy stream, once the step that last read the slot's
  // device copy has read it
  if (sl.used) PIE_HIP(hipStreamWaitEvent(u->copy, sl.consumed, 0));
  PIE_HIP(hipMemcpyAsync(sl.x.p, sl.hx, P0 * 8 * sizeof(float),
                         hipMemcpyHostToDevice, u->copy));
  PIE_HIP(hipMemcpyAsync(sl.t.p, sl.ht, P0 * 3 * sizeof(float),
                         hipMemcpyHostToDevice, u->copy));
  PIE_HIP(hipMemcpyAsync(sl.m.p, sl.hm, P0, hipMemcpyHostToDevice, u->copy));
  PIE_HIP(hipEventRecord(sl.uploaded, u->copy));
  sl.used = true;
  PIE_HIP(hipStreamWaitEvent(u->stream, sl.uploaded, 0));
  PIE_HIP(hipEventRecord(sl.t0, u->stream));
  PIE_CHECK(forward(u, sl.x.p));
  PIE_CHECK(loss_backward(u, sl.t.p, sl.m.p, sl.acc.p, sl.consumed));
  PIE_HIP(hipEventRecord(sl.t1, u->stream));
  PIE_CHECK(optimiser_step(u, lr));
  sl.pending = true;
  return PIE_OK;
}

int pie_unet_finish(pie_unet* u, int k, float* loss) {
  if (!u) return fail(PIE_EINVAL, "NULL argument");
  if (k < 0 || k > 1) return fail(PIE_EINVAL, "slot must be 0 or 1");
  pie_unet::Slot& sl = u->slot[k];
  if (!sl.pending) return fail(PIE_ESTATE, "no step in flight on the slot");
  PIE_HIP(hipEventSynchronize(sl.t1));
  float ms = 0.0f;
  PIE_HIP(hipEventElapsedTime(&ms, sl.t0, sl.t1));
  u->ms_total += ms;
  u->flops_total += conv_flops(u, true);
  double acc[2];
  PIE_HIP(hipMemcpy(acc, sl.acc.p, sizeof(acc), hipMemcpyDeviceToHost));
  if (loss) *loss = (float)(acc[0] / std::max(1.0, acc[1]));
  sl.pending = false;
  return PIE_OK;
}

int pie_unet_set_conv_tiles(int min_blocks16) {
  if (min_blocks16 < -1) return fail(PIE_EINVAL, "min_blocks16 below -1");
  pie::unet::set_tile16_min_blocks(min_blocks16);
  return PIE_OK;
}

int pie_unet_stats(pie_unet* u, double* ms, double* conv_flops_done) {
  if (!u) return fail(PIE_EINVAL, "NULL argument");
  if (ms) *ms = u->ms_total;
  if (conv_flops_done) *conv_flops_done = u->flops_total;
  return PIE_OK;
}

}  // extern "C"
