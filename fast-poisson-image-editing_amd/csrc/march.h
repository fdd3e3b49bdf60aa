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

// Workgroups are dealt to the 8 XCDs round-robin; renumber them so that each
// XCD gets one contiguous range of tasks, which keeps the halo columns and
// rows shared by neighbouring tasks inside one L2.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, i = blockIdx.x;
  const int x = i & 7, k = i >> 3, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + k;
}

template <int C>
__device__ __forceinline__ void load_row(const float* __restrict__ base,
                                         long long plane, int P, int H, int W,
                                         int row, int col0, float (&v)[C][3]) {
  if (row < H) {
    const float* p = base + (long long)row * P + col0;
    if (col0 >= 0 && col0 + C <= W) {
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
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
#pragma unroll
          for (int cc = 0; cc < C; ++cc) v[cc][ch] = p[ch * plane + cc];
        }
      }
      return;
    }
#pragma unroll
    for (int cc = 0; cc < C; ++cc) {
      const int col = col0 + cc;
      const bool ok = col >= 0 && col < W;
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) v[cc][ch] = ok ? p[ch * plane + cc] : 0.0f;
    }
    return;
  }
#pragma unroll
  for (int cc = 0; cc < C; ++cc)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) v[cc][ch] = 0.0f;
}

template <int C>
__device__ __forceinline__ uint32_t load_mask(const uint8_t* __restrict__ m,
                                              int P, int H, int W, int row,
                                              int col0) {
  uint32_t bits = 0;
  if (row < H) {
    const uint8_t* p = m + (long long)row * P + col0;
#pragma unroll
    for (int cc = 0; cc < C; ++cc) {
      const int col = col0 + cc;
      if (col >= 0 && col < W && p[cc]) bits |= 1u << cc;
    }
  }
  return bits;
}

template <int C>
__device__ __forceinline__ void store_row(float* __restrict__ base,
                                          long long plane, int P, int W,
                                          int row, int col0,
                                          const float (&v)[C][3]) {
  float* p = base + (long long)row * P + col0;
  if (col0 + C <= W) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      if constexpr (C == 2) {
        *reinterpret_cast<float2*>(p + ch * plane) =
            make_float2(v[0][ch], v[1][ch]);
      } else if constexpr (C == 4) {
        *reinterpret_cast<float4*>(p + ch * plane) =
            make_float4(v[0][ch], v[1][ch], v[2][ch], v[3][ch]);
      } else {
#pragma unroll
        for (int cc = 0; cc < C; ++cc) p[ch * plane + cc] = v[cc][ch];
      }
    }
    return;
  }
#pragma unroll
  for (int cc = 0; cc < C; ++cc)
    if (col0 + cc < W)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) p[ch * plane + cc] = v[cc][ch];
}

template <bool B, class T>
__device__ __forceinline__ T* pick(T* x, T* y) {
  if constexpr (B)
    return y;
  else
    return x;
}

template <int S, int C, int SCHED, int ORDER>
__global__ void __launch_bounds__(kBlock) march_kernel(MarchArgs a) {
  static_assert(S % 2 == 0 && S % C == 0, "S even and a multiple of C");
  static_assert(S * C <= 32, "mask queue is one 32-bit word");
  constexpr int U = kWave * C - 2 * S;  // written columns per stripe
  const int lane = threadIdx.x & (kWave - 1);
  const int task = xcd_block() * kBlockWaves + (threadIdx.x / kWave);
  if (task >= a.ntask) return;  // whole wave leaves together
  const int stripe = task % a.stripes;
  const int seg = task / a.stripes;
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
        float v[C][3];
        load_row<C>(a.xin, a.plane, a.P, a.H, a.W, r, col0, v);
        if (lane_out) store_row<C>(a.xout, a.plane, a.P, a.W, r, col0, v);
      }
      return;
    }
  }

  const int rb = max(0, r0 - S);
  const int nsteps = ((r1 - 1 + S) - rb + 1 + S - 1) / S * S;

  // Level j (0..S-1) keeps two rows: wa/wb alternate as the older and the
  // newer one, so a step renames instead of moving.  gq/mq hold the
  // guidance and mask of the last S rows (slot = row step mod S).
  float wa[S][C][3], wb[S][C][3], gq[S][C][3];
  uint32_t mq = 0;
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int cc = 0; cc < C; ++cc)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch)
        wa[j][cc][ch] = wb[j][cc][ch] = gq[j][cc][ch] = 0.0f;

  for (int base = 0; base < nsteps; base += S) {
    static_for<0, S>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int rho = rb + base + u;
      float fr[C][3];
      load_row<C>(a.xin, a.plane, a.P, a.H, a.W, rho, col0, fr);
      float gn[C][3];
      load_row<C>(a.g, a.plane, a.P, a.H, a.W, rho, col0, gn);
      const uint32_t mn = load_mask<C>(a.m, a.P, a.H, a.W, rho, col0);

      static_for<1, S + 1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int j = k - 1;
        constexpr int slot = ((u - k) % S + S) % S;
        float(&older)[C][3] = *pick<(u & 1) != 0>(&wa[j], &wb[j]);
        float(&newer)[C][3] = *pick<(u & 1) != 0>(&wb[j], &wa[j]);
        const int r = rho - k;
        const int t = a.t0 + j;
        float out[C][3];
        if (t < a.t_end) {
          const uint32_t mrow = mq >> (slot * C);
#pragma unroll
          for (int cc = 0; cc < C; ++cc) {
            const int col = col0 + cc;
            bool upd = (mrow >> cc) & 1u;
            if constexpr (SCHED == SCHED_REDBLACK) {
              upd = upd && ((r + col + t) & 1);
            } else if constexpr (SCHED == SCHED_GS_LEX) {
              const int d = t - r - col;
              upd = upd && d >= 0 && d < a.band && !(d & 1);
            }
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
              const float ctr = newer[cc][ch];
              const float up = older[cc][ch];
              const float dn = fr[cc][ch];
              const float lf =
                  cc > 0 ? newer[cc > 0 ? cc - 1 : 0][ch] : from_left(newer[C - 1][ch]);
              const float rt = cc < C - 1 ? newer[cc < C - 1 ? cc + 1 : 0][ch]
                                          : from_right(newer[0][ch]);
              float s = gq[slot][cc][ch];
              if constexpr (ORDER == ORDER_UDLR) {
                s = s + up;
                s = s + dn;
                s = s + lf;
                s = s + rt;
              } else {
                s = s + up;
                s = s + lf;
                s = s + rt;
                s = s + dn;
              }
              out[cc][ch] = upd ? s * 0.25f : ctr;
            }
          }
        } else {
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) out[cc][ch] = newer[cc][ch];
        }
#pragma unroll
        for (int cc = 0; cc < C; ++cc)
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            older[cc][ch] = fr[cc][ch];
            fr[cc][ch] = out[cc][ch];
          }
      });

      const int rs = rho - S;
      if (lane_out && rs >= r0 && rs < r1)
        store_row<C>(a.xout, a.plane, a.P, a.W, rs, col0, fr);

      // the guidance and mask of row rho enter slot u only now: level S of
      // this step still needed the row that slot held (rho - S)
#pragma unroll
      for (int cc = 0; cc < C; ++cc)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) gq[u][cc][ch] = gn[cc][ch];
      mq = (mq & ~(((1u << C) - 1u) << (u * C))) | (mn << (u * C));
    });
  }
}

}  // namespace detail
}  // namespace pie
