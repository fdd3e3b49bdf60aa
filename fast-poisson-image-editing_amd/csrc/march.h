// Temporally blocked sweep kernel (march_kernel) and its device helpers.
//
// Every 64-lane wavefront owns a column stripe of 64*C columns (C adjacent
// columns per lane) and walks down a range of rows.  At each row step it
// loads one new row of the state and pushes it through S levels held in
// registers: level k computes row (rho - k) of step t0 + k - 1 from three
// rows of level k - 1.  Up/down neighbours are registers of the same lane,
// left/right neighbours come from the adjacent lane through DPP wave shifts,
// so a sweep needs no LDS traffic and one launch advances the whole grid by
// S steps while reading and writing the state once.  The S outermost
// columns on each side of a stripe and the S rows above and below a row
// range are recomputed by the neighbouring tasks (the temporal-blocking
// halo).
//
// Arithmetic follows the reference update
//   x' = (b + n1 + n2 + n3 + n4) / 4.0            (fpie/core/gcc/equ.cc:48-53)
// left to right in float, with the neighbour order as a template argument,
// and the division by 4.0 as a multiplication by 0.25f, which rounds the
// same (a power-of-two scaling).  Built with -ffp-contract=off so no add is
// fused with the multiplication.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "kernels.h"

#pragma clang fp contract(off)

namespace pie {
namespace detail {

constexpr int kWave = 64;
constexpr int kBlockWaves = 4;
constexpr int kBlock = kWave * kBlockWaves;

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// DPP wave_shr:1 - lane l receives lane l-1 (lane 0 receives 0).
__device__ __forceinline__ float from_left(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138,
                                         0xF, 0xF, false));
}
// DPP wave_shl:1 - lane l receives lane l+1 (lane 63 receives 0).
__device__ __forceinline__ float from_right(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130,
                                         0xF, 0xF, false));
}

// Maximum of a 32-bit value over the wavefront.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1)
    v = max(v, (unsigned)__shfl_xor((int)v, off, kWave));
  return v;
}

// Workgroups are dealt to the 8 XCDs round-robin; renumber them so that each
// XCD gets one contiguous range of tasks, which keeps the halo columns and
// rows shared by neighbouring tasks inside one L2.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, i = blockIdx.x;
  const int x = i & 7, k = i >> 3, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + k;
}

// Rows are padded (kPadLeft zero columns before column 0, at least
// kPadRight after column W-1, see kernels.h), so a stripe's vector load at
// col0 in [-S, W + 64C) always stays inside the row: loads need no per-lane
// bounds test, and a prefetched value is not forced through a branch.
// Padding and clamped rows only ever feed halo lanes or unsolved pixels.
template <int C, int NCH>
__device__ __forceinline__ void load_row(const float* __restrict__ base,
                                         long long plane, int P, int row,
                                         int col0, float (&v)[C][NCH]) {
  const float* p = base + (long long)row * P + col0;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    if constexpr (C == 2) {
      const float2 q = *reinterpret_cast<const float2*>(p + ch * plane);
      v[0][ch] = q.x;
      v[1][ch] = q.y;
    } else if constexpr (C == 4) {
      const float4 q = *reinterpret_cast<const float4*>(p + ch * plane);
      v[0][ch] = q.x;
      v[1][ch] = q.y;
      v[2][ch] = q.z;
      v[3][ch] = q.w;
    } else {
      v[0][ch] = p[ch * plane];
    }
  }
}

// Mask bits of C adjacent columns (mask bytes are 0 or 1).
template <int C>
__device__ __forceinline__ uint32_t load_mask(const uint8_t* __restrict__ m,
                                              int P, int row, int col0) {
  const uint8_t* p = m + (long long)row * P + col0;
  if constexpr (C == 4) {
    const uint32_t v = *reinterpret_cast<const uint32_t*>(p);
    return (v & 1u) | ((v >> 7) & 2u) | ((v >> 14) & 4u) | ((v >> 21) & 8u);
  } else if constexpr (C == 2) {
    const uint32_t v = *reinterpret_cast<const uint16_t*>(p);
    return (v & 1u) | ((v >> 7) & 2u);
  } else {
    return *p;
  }
}

// Stores may land in the right padding (col >= W): those values are finite
// and only ever read by halo lanes.
template <int C, int NCH>
__device__ __forceinline__ void store_row(float* __restrict__ base,
                                          long long plane, int P, int row,
                                          int col0, const float (&v)[C][NCH]) {
  float* p = base + (long long)row * P + col0;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    if constexpr (C == 2) {
      *reinterpret_cast<float2*>(p + ch * plane) =
          make_float2(v[0][ch], v[1][ch]);
    } else if constexpr (C == 4) {
      *reinterpret_cast<float4*>(p + ch * plane) =
          make_float4(v[0][ch], v[1][ch], v[2][ch], v[3][ch]);
    } else {
      p[ch * plane] = v[0][ch];
    }
  }
}

// Row steps of look-ahead of the sweep kernel's loads (divides S, so the
// ring slot of every unrolled step is a compile-time constant).
#ifndef PIE_PREFETCH
template <int S>
constexpr int kPrefetch = S % 4 == 0 ? 4 : 2;
#else
template <int S>
constexpr int kPrefetch = PIE_PREFETCH;
#endif

// Guidance ring of the last S rows: registers (0) or LDS (1, each wave a
// private region, so no barrier is needed and fewer VGPRs are live).
#ifndef PIE_GQ_LDS
#define PIE_GQ_LDS 0
#endif

template <bool B, class T>
__device__ __forceinline__ T* pick(T* x, T* y) {
  if constexpr (B)
    return y;
  else
    return x;
}

// One wave task: column stripe `stripe`, row range `seg`, S steps.  When
// dmax is not null the largest |change| of the last step over the pixels
// the task writes is merged into *dmax (float bits, non-negative).
template <int S, int C, int NCH, int SCHED, int ORDER>
__device__ __forceinline__ void march_task(const MarchArgs& a, int stripe,
                                           int seg, int group,
                                           unsigned* dmax) {
  static_assert(S % 2 == 0 && S % C == 0, "S even and a multiple of C");
  static_assert(S * C <= 32, "mask queue is one 32-bit word");
  static_assert(NCH == 1 || NCH == 3, "one or three channels a task");
  // channels of this task: all three, or channel `group` only
  const long long choff = (long long)group * NCH * a.plane;
  const float* __restrict__ xin = a.xin + choff;
  float* __restrict__ xout = a.xout + choff;
  const float* __restrict__ gin = a.g + choff;
  constexpr int U = kWave * C - 2 * S;  // written columns per stripe
  const int lane = threadIdx.x & (kWave - 1);
  const int col0 = stripe * U - S + lane * C;
  const int r0 = seg * a.seg_rows;
  const int r1 = min(r0 + a.seg_rows, a.H);
  const bool lane_out = lane * C >= S && lane * C < kWave * C - S;

  if constexpr (SCHED == SCHED_GS_LEX) {
    // A task none of whose written pixels meets the wavefront band during
    // this launch only carries its rows over to the output buffer.
    const int t_last = min(a.t0 + S, a.t_end) - 1;
    const int c_lo = stripe * U, c_hi = min(c_lo + U, a.W) - 1;
    if (r0 + c_lo > t_last || (r1 - 1) + c_hi <= a.t0 - a.band) {
      for (int r = r0; r < r1; ++r) {
        float v[C][NCH];
        load_row<C, NCH>(xin, a.plane, a.P, r, col0, v);
        if (lane_out) store_row<C, NCH>(xout, a.plane, a.P, r, col0, v);
      }
      return;
    }
  }

  float dm = 0.0f;
  const int rb = max(0, r0 - S);
  const int nsteps = ((r1 - 1 + S) - rb + 1 + S - 1) / S * S;

  // Level j (0..S-1) keeps two rows: wa/wb alternate as the older and the
  // newer one, so a step renames instead of moving.  gq/mq hold the
  // guidance and mask of the last S rows (slot = row step mod S).
  float wa[S][C][NCH], wb[S][C][NCH];
#if PIE_GQ_LDS
  __shared__ float gl_all[kBlockWaves][S][NCH][kWave * C];
  float(&gl)[S][NCH][kWave * C] =
      gl_all[(threadIdx.x / kWave) % kBlockWaves];
#define PIE_GQ(slot, cc, ch) gl[slot][ch][lane * C + (cc)]
#else
  float gq[S][C][NCH];
#define PIE_GQ(slot, cc, ch) gq[slot][cc][ch]
#endif
  uint32_t mq = 0;
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int cc = 0; cc < C; ++cc)
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        wa[j][cc][ch] = wb[j][cc][ch] = 0.0f;
        PIE_GQ(j, cc, ch) = 0.0f;
      }

  // Rows are loaded PF row steps before they are used (a register ring of
  // PF slots), so each wave keeps PF rows of loads in flight instead of
  // waiting for every row it consumes.
  constexpr int PF = kPrefetch<S>;
  float px[PF][C][NCH], pg[PF][C][NCH];
  uint32_t pm[PF];
  static_for<0, PF>([&](auto pc) {
    constexpr int q = decltype(pc)::value;
    const int row = min(rb + q, a.H - 1);
    load_row<C, NCH>(xin, a.plane, a.P, row, col0, px[q]);
    load_row<C, NCH>(gin, a.plane, a.P, row, col0, pg[q]);
    pm[q] = rb + q < a.H ? load_mask<C>(a.m, a.P, row, col0) : 0u;
  });

  for (int base = 0; base < nsteps; base += S) {
    static_for<0, S>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      constexpr int q = u % PF;
      const int rho = rb + base + u;
      float fr[C][NCH], gn[C][NCH];
#pragma unroll
      for (int cc = 0; cc < C; ++cc)
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
          fr[cc][ch] = px[q][cc][ch];
          gn[cc][ch] = pg[q][cc][ch];
        }
      const uint32_t mn = pm[q];
      {
        const int row = min(rho + PF, a.H - 1);  // wave-uniform
        load_row<C, NCH>(xin, a.plane, a.P, row, col0, px[q]);
        load_row<C, NCH>(gin, a.plane, a.P, row, col0, pg[q]);
        pm[q] = rho + PF < a.H ? load_mask<C>(a.m, a.P, row, col0) : 0u;
      }

      static_for<1, S + 1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int j = k - 1;
        constexpr int slot = ((u - k) % S + S) % S;
        float(&older)[C][NCH] = *pick<(u & 1) != 0>(&wa[j], &wb[j]);
        float(&newer)[C][NCH] = *pick<(u & 1) != 0>(&wb[j], &wa[j]);
        const int r = rho - k;
        const int t = a.t0 + j;
        // Steps past t_end keep their value; folding that (wave-uniform)
        // test into the predicate keeps the level free of branches, so
        // the window registers are renamed rather than copied.
        const bool live = t < a.t_end;
        float out[C][NCH];
        const uint32_t mrow = mq >> (slot * C);
        bool upd[C];
#pragma unroll
        for (int cc = 0; cc < C; ++cc) {
          const int col = col0 + cc;
          upd[cc] = live && ((mrow >> cc) & 1u);
          if constexpr (SCHED == SCHED_REDBLACK) {
            upd[cc] = upd[cc] && ((r + col + t) & 1);
          } else if constexpr (SCHED == SCHED_GS_LEX) {
            const int d = t - r - col;
            upd[cc] = upd[cc] && d >= 0 && d < a.band && !(d & 1);
          }
        }
        // The C x NCH sums are independent chains; adding them in lockstep
        // (one term of every chain, then the next term) keeps other work
        // between an accumulator's write and the lane-shifted (DPP) add
        // that reads it, which otherwise stalls on the DPP read hazard.
        // Each chain still adds its own terms in the reference's order.
        float acc[C][NCH];
        auto left = [&](int cc, int ch) {
          return cc > 0 ? newer[cc > 0 ? cc - 1 : 0][ch]
                        : from_left(newer[C - 1][ch]);
        };
        auto right = [&](int cc, int ch) {
          return cc < C - 1 ? newer[cc < C - 1 ? cc + 1 : 0][ch]
                            : from_right(newer[0][ch]);
        };
#pragma unroll
        for (int cc = 0; cc < C; ++cc)
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch)
            acc[cc][ch] = PIE_GQ(slot, cc, ch) + older[cc][ch];
        if constexpr (ORDER == ORDER_UDLR) {
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              acc[cc][ch] = acc[cc][ch] + fr[cc][ch];
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              acc[cc][ch] = acc[cc][ch] + left(cc, ch);
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              acc[cc][ch] = acc[cc][ch] + right(cc, ch);
        } else {
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              acc[cc][ch] = acc[cc][ch] + left(cc, ch);
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              acc[cc][ch] = acc[cc][ch] + right(cc, ch);
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              acc[cc][ch] = acc[cc][ch] + fr[cc][ch];
        }
#pragma unroll
        for (int cc = 0; cc < C; ++cc)
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch)
            out[cc][ch] = upd[cc] ? acc[cc][ch] * 0.25f : newer[cc][ch];
#pragma unroll
        for (int cc = 0; cc < C; ++cc)
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch) {
            older[cc][ch] = fr[cc][ch];
            fr[cc][ch] = out[cc][ch];
          }
      });

      const int rs = rho - S;
      if (lane_out && rs >= r0 && rs < r1) {
        store_row<C, NCH>(xout, a.plane, a.P, rs, col0, fr);
        if (dmax) {
          // level S - 1 of row rs is still the newer row of that level
          const float(&prev)[C][NCH] =
              *pick<(u & 1) != 0>(&wb[S - 1], &wa[S - 1]);
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              dm = fmaxf(dm, fabsf(fr[cc][ch] - prev[cc][ch]));
        }
      }

      // the guidance and mask of row rho enter slot u only now: level S of
      // this step still needed the row that slot held (rho - S)
#pragma unroll
      for (int cc = 0; cc < C; ++cc)
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) PIE_GQ(u, cc, ch) = gn[cc][ch];
      mq = (mq & ~(((1u << C) - 1u) << (u * C))) | (mn << (u * C));
    });
  }
#undef PIE_GQ
  if (dmax) {
    // differences are non-negative, so float order is unsigned-int order
    const unsigned wave_max = wave_max_u32(__float_as_uint(dm));
    if (lane == 0 && wave_max) atomicMax(dmax, wave_max);
  }
}

#ifndef PIE_WAVES_PER_EU
#define PIE_WAVES_PER_EU 1
#endif
template <int S, int C, int NCH, int SCHED, int ORDER>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(PIE_WAVES_PER_EU)))
    march_kernel(MarchArgs a) {
  const int task = xcd_block() * kBlockWaves + (threadIdx.x / kWave);
  if (task >= a.ntask) return;  // whole wave leaves together
  // task = (channel group, segment, stripe), stripe fastest
  const int per_group = a.stripes * a.segs;
  const int group = task / per_group, rem = task % per_group;
  march_task<S, C, NCH, SCHED, ORDER>(a, rem % a.stripes, rem / a.stripes,
                                      group, a.dmax);
}

// Persistent sweep kernel of a batch: every wave takes tasks (tiles of any
// item) from one atomic work counter until the counter passes the total,
// so uneven items share the machine.  Items already stopped are skipped.
// Each item's launch count selects its input buffer and first step.
template <int S, int C, int NCH, int SCHED, int ORDER>
__global__ void __launch_bounds__(kBlock) march_batch_kernel(BatchArgs b) {
  const int lane = threadIdx.x & (kWave - 1);
  for (;;) {
    int t = 0;
    if (lane == 0) t = (int)atomicAdd(b.queue, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, kWave));
    if (t >= b.ntask) break;
    const int id = b.task_item[t];
    const BatchItem& it = b.items[id];
    if (it.done) continue;
    MarchArgs a{};
    const bool odd = it.launches & 1;
    a.xin = odd ? it.xb : it.xa;
    a.xout = odd ? it.xa : it.xb;
    a.g = it.g;
    a.m = it.m;
    a.plane = it.plane;
    a.H = it.H;
    a.W = it.W;
    a.P = it.P;
    a.t0 = it.launches * S;
    a.t_end = it.t_end;
    a.band = 0;
    a.stripes = it.stripes;
    a.segs = it.segs;
    a.seg_rows = b.seg_rows;
    const int local = t - it.task0;
    const int per_group = it.stripes * it.segs;
    const int group = local / per_group, rem = local % per_group;
    march_task<S, C, NCH, SCHED, ORDER>(a, rem % it.stripes, rem / it.stripes,
                                        group, &b.items[id].dmax_bits);
  }
}

}  // namespace detail
}  // namespace pie
