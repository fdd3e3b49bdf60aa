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
constexpr int kBlockWaves = kSweepBlockWaves;
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
// columns [cu, cu + 64C) with cu >= -S always stays inside the row: loads
// need no per-lane bounds test, and a prefetched value is not forced
// through a branch.  Padding and clamped rows only ever feed halo lanes or
// unsolved pixels.  `rowp` is the wave-uniform address of the stripe's
// first column in the row (scalar registers), `lo` the lane's offset.
template <int C, int NCH>
__device__ __forceinline__ void load_row(const float* __restrict__ rowp,
                                         long long plane, int lo,
                                         float (&v)[C][NCH]) {
  const float* p = rowp + lo;
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

// Stores may land in the right padding (col >= W): those values are finite
// and only ever read by halo lanes.
template <int C, int NCH>
__device__ __forceinline__ void store_row(float* __restrict__ rowp,
                                          long long plane, int lo,
                                          const float (&v)[C][NCH]) {
  float* p = rowp + lo;
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

// Prefetch ring of the sweep kernel in LDS.
//
// Each wave streams its stripe's rows (state and guidance, every channel of
// the task) into a private LDS ring of PF slots with gfx950's LDS-DMA loads
// (global_load_lds_dword: each lane's dword lands at slot base + 4 * lane),
// PF row steps ahead of their use.  Register prefetch rings did not work
// here: the register allocator copied some just-issued loads across the
// loop back-edge, and each copy waited for every load in flight.  LDS-DMA
// loads have no register destination, so nothing in flight is ever copied.
//
// A slot is read back by one asm block that first waits until the slot's
// loads have landed (vmcnt: all vector-memory operations issued after them
// may still be in flight, and at least (PF - 1) * D loads were issued after
// them, D = loads per row; stores only make the wait stricter), then reads
// the slot with ds_read and waits for the reads (lgkmcnt(0)) before the
// block ends, so its outputs are complete and the slot may be reloaded.
// The compiler's own wait insertion would wait for all LDS-DMA loads in
// flight before any LDS read it can see.
// Slots of the ring: a divisor of S (the slot of every unrolled step is a
// constant); three-channel and four-column tasks keep 4 rows in flight,
// the others S rows; at most pfmax (work-groups of many waves cap it to fit
// their rings into the LDS).
constexpr int ring_slots(int S, int C, int NCH, int pfmax) {
  int want = (NCH == 3 || C == 4) ? (S % 4 == 0 ? 4 : (S % 3 == 0 ? 3 : 2))
                                  : S;
  if (want > pfmax)
    for (int d = pfmax; d >= 1; --d)
      if (S % d == 0) return d;
  return want;
}

template <int S, int C, int NCH, int PFMAX = 64>
struct SweepRing {
  static constexpr int PF = ring_slots(S, C, NCH, PFMAX);
  static constexpr int kParts = C;                    // dword loads per row
  static constexpr int D = 2 * NCH * kParts;          // loads per row step
  static constexpr int kRowFloats = kWave * C;        // one channel row
  static constexpr int kSlotFloats = 2 * NCH * kRowFloats;
  static constexpr int kWaveFloats = PF * kSlotFloats;
  static constexpr int kWait = (PF - 1) * D;
  static_assert(S % PF == 0, "ring slots divide S");
  static_assert(kWait <= 63, "vmcnt field is 6 bits");
  static_assert(C == 1 || C == 2 || C == 4, "1, 2 or 4 columns per lane");
};

// LDS byte address of a __shared__ location: the low 32 bits of its flat
// address (the shared aperture maps LDS offset o to {aperture, o}).
__device__ __forceinline__ uint32_t lds_offset(const void* p) {
  return (uint32_t)(uintptr_t)p;
}

// Issue the LDS-DMA loads of one row into ring slot `slot` (wave-uniform
// LDS pointer).  Array 0 is the state, 1 the guidance; channel rows are
// kRowFloats apart.  With C columns per lane, load p of lane l fetches
// column 64p + l of the stripe, so the stripe's columns lie in LDS in order
// and lane l reads its columns Cl .. Cl + C - 1 with one ds_read.
template <int S, int C, int NCH>
__device__ __forceinline__ void ring_fill(float* slot, const float* xrow,
                                         const float* grow, long long plane,
                                         int lane) {
  using R = SweepRing<S, C, NCH>;  // (row layout only: any slot count)
#pragma unroll
  for (int arr = 0; arr < 2; ++arr)
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
      for (int part = 0; part < R::kParts; ++part) {
        const float* src = (arr ? grow : xrow) + ch * plane + part * kWave;
        __builtin_amdgcn_global_load_lds(
            src + lane, slot + (arr * NCH + ch) * R::kRowFloats + part * kWave,
            4, 0, 0);
      }
}

// Read one ring slot (byte offset `off` from the lane's address `addr` in
// the wave's ring) into x[C][NCH] and g[C][NCH]; WAIT = vmcnt bound.
template <int C, int NCH, int WAIT, int OFF>
__device__ __forceinline__ void ring_read(uint32_t addr, float (&x)[C][NCH],
                                          float (&g)[C][NCH]) {
  constexpr int RB = kWave * C * 4;  // bytes of one channel row
  if constexpr (C == 1 && NCH == 3) {
    asm volatile(
        "s_waitcnt vmcnt(%6)\n\t"
        "ds_read_b32 %0, %7 offset:%8\n\t"
        "ds_read_b32 %1, %7 offset:%9\n\t"
        "ds_read_b32 %2, %7 offset:%10\n\t"
        "ds_read_b32 %3, %7 offset:%11\n\t"
        "ds_read_b32 %4, %7 offset:%12\n\t"
        "ds_read_b32 %5, %7 offset:%13\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(x[0][0]), "=&v"(x[0][1]), "=&v"(x[0][2]), "=&v"(g[0][0]),
          "=&v"(g[0][1]), "=&v"(g[0][2])
        : "i"(WAIT), "v"(addr), "i"(OFF), "i"(OFF + RB), "i"(OFF + 2 * RB),
          "i"(OFF + 3 * RB), "i"(OFF + 4 * RB), "i"(OFF + 5 * RB)
        : "memory");
  } else if constexpr (C == 1 && NCH == 1) {
    asm volatile(
        "s_waitcnt vmcnt(%2)\n\t"
        "ds_read_b32 %0, %3 offset:%4\n\t"
        "ds_read_b32 %1, %3 offset:%5\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(x[0][0]), "=&v"(g[0][0])
        : "i"(WAIT), "v"(addr), "i"(OFF), "i"(OFF + RB)
        : "memory");
  } else if constexpr (C == 2 && NCH == 1) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    f2 xv, gv;
    asm volatile(
        "s_waitcnt vmcnt(%2)\n\t"
        "ds_read_b64 %0, %3 offset:%4\n\t"
        "ds_read_b64 %1, %3 offset:%5\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(xv), "=&v"(gv)
        : "i"(WAIT), "v"(addr), "i"(OFF), "i"(OFF + RB)
        : "memory");
    x[0][0] = xv.x;
    x[1][0] = xv.y;
    g[0][0] = gv.x;
    g[1][0] = gv.y;
  } else if constexpr (C == 2 && NCH == 3) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    f2 x0, x1, x2, g0, g1, g2;
    asm volatile(
        "s_waitcnt vmcnt(%6)\n\t"
        "ds_read_b64 %0, %7 offset:%8\n\t"
        "ds_read_b64 %1, %7 offset:%9\n\t"
        "ds_read_b64 %2, %7 offset:%10\n\t"
        "ds_read_b64 %3, %7 offset:%11\n\t"
        "ds_read_b64 %4, %7 offset:%12\n\t"
        "ds_read_b64 %5, %7 offset:%13\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(g0), "=&v"(g1), "=&v"(g2)
        : "i"(WAIT), "v"(addr), "i"(OFF), "i"(OFF + RB), "i"(OFF + 2 * RB),
          "i"(OFF + 3 * RB), "i"(OFF + 4 * RB), "i"(OFF + 5 * RB)
        : "memory");
    x[0][0] = x0.x;
    x[1][0] = x0.y;
    x[0][1] = x1.x;
    x[1][1] = x1.y;
    x[0][2] = x2.x;
    x[1][2] = x2.y;
    g[0][0] = g0.x;
    g[1][0] = g0.y;
    g[0][1] = g1.x;
    g[1][1] = g1.y;
    g[0][2] = g2.x;
    g[1][2] = g2.y;
  } else if constexpr (C == 4 && NCH == 1) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 xv, gv;
    asm volatile(
        "s_waitcnt vmcnt(%2)\n\t"
        "ds_read_b128 %0, %3 offset:%4\n\t"
        "ds_read_b128 %1, %3 offset:%5\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(xv), "=&v"(gv)
        : "i"(WAIT), "v"(addr), "i"(OFF), "i"(OFF + RB)
        : "memory");
    x[0][0] = xv.x;
    x[1][0] = xv.y;
    x[2][0] = xv.z;
    x[3][0] = xv.w;
    g[0][0] = gv.x;
    g[1][0] = gv.y;
    g[2][0] = gv.z;
    g[3][0] = gv.w;
  } else {
    static_assert(C == 0, "no ring read for this tile");
  }
}

template <bool B, class T>
__device__ __forceinline__ T* pick(T* x, T* y) {
  if constexpr (B)
    return y;
  else
    return x;
}

// Geometry of one wave task (all wave-uniform except lane, lo, col0 and
// lane_out).
struct TaskGeom {
  const float* xin;  // channel planes of the task
  float* xout;
  const float* gin;
  long long P;       // row pitch
  int lane, lo;      // lane, its column offset inside the stripe
  int cu, col0;      // first column of the stripe, of the lane
  int r0, r1;        // rows written
  bool lane_out;     // the lane writes its columns
};

// The rows of one wave task (see march_task).  COLOUR variants (two
// columns per lane, gs_lex or red-black) know at compile time which of the
// lane's two columns may change at each level: both schedules update one
// colour of the checkerboard per step, and with an even first column the
// colour is the same for every lane of a row.  PI is the task's parity
// (t0 - rb - 1) & 1; with BAND (1 or 2) the gs_lex band test
// 0 <= t - i - j < band is made per pixel, without it the whole task lies
// inside the band; BAND 2: every step of the launch lies inside the call
// (no per-level liveness masks: with the band test they no longer fit the
// scalar registers).
template <int S, int C, int NCH, int SCHED, int ORDER, bool DM, int PI,
          int BAND, int PFMAX>
__device__ __forceinline__ void march_rows(const MarchArgs& a,
                                           const TaskGeom& tg, float* ring,
                                           unsigned* dmax) {
  constexpr bool COLOUR = SCHED != SCHED_JACOBI && C == 2;
  const int lane = tg.lane, lo = tg.lo, cu = tg.cu, col0 = tg.col0;
  const int r0 = tg.r0, r1 = tg.r1;
  const bool lane_out = tg.lane_out;
  const long long P = tg.P;
  float* __restrict__ xout = tg.xout;
  auto xrow = [&](int row) { return tg.xin + row * P + cu; };
  auto grow = [&](int row) { return tg.gin + row * P + cu; };

  float dm = 0.0f;
  const int rb = max(0, r0 - S);
  const int nsteps = ((r1 - 1 + S) - rb + 1 + S - 1) / S * S;

  // Level j (0..S-1) keeps two rows: wa/wb alternate as the older and the
  // newer one, so a step renames instead of moving.  gq holds the guidance
  // of the last S rows and mb their solved flags (slot = row step mod S);
  // the flags are lane masks, held in scalar registers.  A pixel is solved
  // where its guidance is a number: the host stores NaN as the guidance of
  // every unsolved pixel and of the padding.
  float wa[S][C][NCH], wb[S][C][NCH];
  float gq[S][C][NCH];
  bool mb[S][C];
#pragma unroll
  for (int j = 0; j < S; ++j)
#pragma unroll
    for (int cc = 0; cc < C; ++cc) {
      mb[j][cc] = false;
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        wa[j][cc][ch] = wb[j][cc][ch] = 0.0f;
        gq[j][cc][ch] = 0.0f;
      }
    }

  // This wave's LDS ring; rows are loaded PF row steps before their use.
  // Rows past the grid are clamped to row H - 1, which is never solved, so
  // every load is unconditional.
  using R = SweepRing<S, C, NCH, PFMAX>;
  constexpr int PF = R::PF;
  const uint32_t ring_addr = lds_offset(ring + lo);
  static_for<0, PF>([&](auto pc) {
    constexpr int q = decltype(pc)::value;
    const int row = min(rb + q, a.H - 1);
    ring_fill<S, C, NCH>(ring + q * R::kSlotFloats, xrow(row), grow(row),
                         a.plane, lane);
  });

  // gs_lex band test (BAND): the lane's t - r - col0 at level k of row
  // step u of this iteration is dl + 2k - u; dlb = dl - band
  int dl = a.t0 - 1 - rb - col0, dlb = dl - a.band;
  for (int base = 0; base < nsteps; base += S, dl -= S, dlb -= S) {
    static_for<0, S>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      constexpr int q = u % PF;
      const int rho = rb + base + u;
      float fr[C][NCH], gn[C][NCH];
      ring_read<C, NCH, R::kWait, q * R::kSlotFloats * 4>(ring_addr, fr, gn);
      {
        const int row = min(rho + PF, a.H - 1);  // wave-uniform
        ring_fill<S, C, NCH>(ring + q * R::kSlotFloats, xrow(row), grow(row),
                             a.plane, lane);
      }

      static_for<1, S + 1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int j = k - 1;
        constexpr int slot = ((u - k) % S + S) % S;
        float(&older)[C][NCH] = *pick<(u & 1) != 0>(&wa[j], &wb[j]);
        float(&newer)[C][NCH] = *pick<(u & 1) != 0>(&wb[j], &wa[j]);
        const int r = rho - k;
        const int t = a.t0 + j;
        // Steps outside [t_lo, t_end) keep their value; folding that
        // (wave-uniform) test into the predicate keeps the level free of
        // branches, so the window registers are renamed rather than copied.
        const bool live = BAND == 2 || (t >= a.t_lo && t < a.t_end);
        float out[C][NCH];
        if constexpr (COLOUR) {
          // the column of the lane that may change at this level: t - r
          // has the parity PI ^ (u & 1) (base and 2k are even)
          constexpr int act = SCHED == SCHED_GS_LEX ? (PI ^ (u & 1))
                                                    : (1 ^ PI ^ (u & 1));
          constexpr int oth = 1 - act;
          bool upd = live && mb[slot][act];
          if constexpr (BAND) {
            // d = t - r - (col0 + act) = dl + 2k - u - act: compared with
            // a constant, so no per-level bound is held in a register
            constexpr int c = u + act - 2 * k;
            upd = upd && dl >= c && dlb < c;
          }
          float acc[NCH];
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch) {
            const float l = act == 0 ? from_left(newer[oth][ch])
                                     : newer[oth][ch];
            const float rr = act == 0 ? newer[oth][ch]
                                      : from_right(newer[oth][ch]);
            acc[ch] = gq[slot][act][ch] + older[act][ch];
            if constexpr (ORDER == ORDER_UDLR) {
              acc[ch] = acc[ch] + fr[act][ch];
              acc[ch] = acc[ch] + l;
              acc[ch] = acc[ch] + rr;
            } else {
              acc[ch] = acc[ch] + l;
              acc[ch] = acc[ch] + rr;
              acc[ch] = acc[ch] + fr[act][ch];
            }
            out[act][ch] = upd ? acc[ch] * 0.25f : newer[act][ch];
            out[oth][ch] = newer[oth][ch];
          }
        } else {
          bool upd[C];
#pragma unroll
          for (int cc = 0; cc < C; ++cc) {
            const int col = col0 + cc;
            upd[cc] = live && mb[slot][cc];
            if constexpr (SCHED == SCHED_REDBLACK) {
              upd[cc] = upd[cc] && ((r + col + t) & 1);
            } else if constexpr (SCHED == SCHED_GS_LEX) {
              const int d = t - r - col;
              upd[cc] = upd[cc] && d >= 0 && d < a.band && !(d & 1);
            }
          }
          // The C x NCH sums are independent chains; adding them in
          // lockstep (one term of every chain, then the next term) keeps
          // other work between an accumulator's write and the lane-shifted
          // (DPP) add that reads it, which otherwise stalls on the DPP read
          // hazard.  Each chain still adds its own terms in the reference's
          // order.
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
              acc[cc][ch] = gq[slot][cc][ch] + older[cc][ch];
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
        }
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
        store_row<C, NCH>(xout + rs * P + cu, a.plane, lo, fr);
        if constexpr (DM) {
          // the change made by the launch's last step: level S's output
          // against its input, which is still the newer row of level S - 1
          // (the last step is live: a launch that ends an item's budget is
          // aligned to end there)
          const float(&prev)[C][NCH] =
              *pick<(u & 1) != 0>(&wb[S - 1], &wa[S - 1]);
#pragma unroll
          for (int cc = 0; cc < C; ++cc)
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
              dm = fmaxf(dm, fabsf(fr[cc][ch] - prev[cc][ch]));
        }
      }

      // the guidance and flags of row rho enter slot u only now: level S of
      // this step still needed the row that slot held (rho - S)
#pragma unroll
      for (int cc = 0; cc < C; ++cc) {
        mb[u][cc] = gn[cc][0] == gn[cc][0];  // false for NaN
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) gq[u][cc][ch] = gn[cc][ch];
      }
    });
  }
  // no LDS-DMA load may still be in flight when the wave leaves the task
  // (the batch kernel reuses the ring for its next task)
  __builtin_amdgcn_s_waitcnt(0);
  if constexpr (DM) {
    // differences are non-negative, so float order is unsigned-int order
    const unsigned wave_max = wave_max_u32(__float_as_uint(dm));
    if (lane == 0 && wave_max) atomicMax(dmax, wave_max);
  }
}

// One wave task: column stripe `stripe`, row range `seg`, S steps.  With DM
// the largest |change| made by the launch's last step (step t0 + S - 1,
// live for every launch DM is used with) over the pixels the task writes is
// merged into *dmax (float bits, non-negative).
template <int S, int C, int NCH, int SCHED, int ORDER, bool DM,
          int WAVES = kBlockWaves, int PFMAX = 64>
__device__ __forceinline__ void march_task(const MarchArgs& a, int stripe,
                                           int seg, int group,
                                           unsigned* dmax) {
  static_assert(S % 2 == 0 && S % C == 0, "S even and a multiple of C");
  static_assert(NCH == 1 || NCH == 3, "one or three channels a task");
  // channels of this task: all three, or channel `group` only
  const long long choff = (long long)group * NCH * a.plane;
  constexpr int U = kWave * C - 2 * S;  // written columns per stripe
  TaskGeom tg;
  tg.xin = a.xin + choff;
  tg.xout = a.xout + choff;
  tg.gin = a.g + choff;
  tg.P = a.P;
  tg.lane = threadIdx.x & (kWave - 1);
  tg.lo = tg.lane * C;
  tg.cu = stripe * U - S;  // even: U and S are even
  tg.col0 = tg.cu + tg.lo;
  tg.r0 = seg * a.seg_rows;
  tg.r1 = min(tg.r0 + a.seg_rows, a.H);
  tg.lane_out = tg.lo >= S && tg.lo < kWave * C - S;
  const int c_lo = stripe * U, c_hi = min(c_lo + U, a.W) - 1;

  if constexpr (SCHED == SCHED_GS_LEX) {
    // A task none of whose written pixels meets the wavefront band during
    // this launch changes nothing.  Its rows only need to reach the output
    // buffer where the two buffers may differ: a region the band has not
    // reached is copied by the call's first launch (after which both
    // buffers hold it), a region the band has left is copied by the first
    // launch after its last update (pixel (i, j) is last updated at step
    // i + j + band - 2).  Every later launch leaves such a task out.
    const int t_last = min(a.t0 + S, a.t_end) - 1;
    const bool before = tg.r0 + c_lo > t_last;
    const bool after = (tg.r1 - 1) + c_hi <= a.t0 - a.band;
    if (before || after) {
      const bool agree = before ? a.t0 >= S
                                : (tg.r1 - 1) + c_hi <= a.t0 - S - a.band;
      if (agree) return;
      for (int r = tg.r0; r < tg.r1; ++r) {
        float v[C][NCH];
        load_row<C, NCH>(tg.xin + r * tg.P + tg.cu, a.plane, tg.lo, v);
        if (tg.lane_out)
          store_row<C, NCH>(tg.xout + r * tg.P + tg.cu, a.plane, tg.lo, v);
      }
      return;
    }
    // The same holds row by row: rows below `front` hold no computed pixel
    // the band reaches during the launch (and once t0 >= S both buffers
    // hold their first values), rows up to `back` hold only written pixels
    // last updated before the previous launch (both buffers hold their
    // final values).  Only the rows between are marched (with S halo rows
    // beyond each end, as for any task), which shortens the tasks at the
    // band's edges.  The range is not empty: the task is neither before
    // nor after the band.
    const int front = a.t0 >= S ? t_last - max(tg.cu, 0) : tg.r1 - 1;
    const int back = a.t0 - S - a.band - c_hi;
    const int r0 = max(tg.r0, back + 1), r1 = min(tg.r1, front + 1);
    if (r0 < r1) {
      tg.r0 = r0;
      tg.r1 = r1;
    }
  }

  // one ring for whichever variant of the rows runs (WAVES waves a group)
  using R = SweepRing<S, C, NCH, PFMAX>;
  __shared__ __attribute__((aligned(16)))
      float ring_all[WAVES * R::kWaveFloats];
  float* ring = ring_all + (int)__builtin_amdgcn_readfirstlane(
                               (int)(threadIdx.x / kWave)) * R::kWaveFloats;

  if constexpr (SCHED != SCHED_JACOBI && C == 2) {
    const int rb = max(0, tg.r0 - S);
    const int pi = (a.t0 - rb - 1) & 1;
    bool band = false;
    if constexpr (SCHED == SCHED_GS_LEX) {
      // every pixel the task computes (rows rb .. r1 + S - 1, the stripe's
      // 64C columns) inside the band at every step of the launch?
      const int d_min = a.t0 - (tg.r1 + S - 1) - (tg.cu + kWave * C - 1);
      const int d_max = a.t0 + S - 1 - rb - tg.cu;
      band = !(d_min >= 0 && d_max < a.band);
    }
    const bool all_live = a.t_lo <= a.t0 && a.t0 + S <= a.t_end;
    auto rows = [&](auto pc, auto bc) {
      march_rows<S, C, NCH, SCHED, ORDER, DM, decltype(pc)::value,
                 decltype(bc)::value, PFMAX>(a, tg, ring, dmax);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    if (pi) {
      if (!band)
        rows(I1{}, I0{});
      else if (all_live)
        rows(I1{}, I2{});
      else
        rows(I1{}, I1{});
    } else {
      if (!band)
        rows(I0{}, I0{});
      else if (all_live)
        rows(I0{}, I2{});
      else
        rows(I0{}, I1{});
    }
  } else {
    march_rows<S, C, NCH, SCHED, ORDER, DM, 0,
               SCHED == SCHED_GS_LEX ? 1 : 0, PFMAX>(a, tg, ring, dmax);
  }
}

#ifndef PIE_WAVES_PER_EU
#define PIE_WAVES_PER_EU 1
#endif
// Most LDS ring slots a wave of the sweep kernel uses.  Four rows of
// prefetch: 16 KiB of LDS per 4-wave group instead of 48 KiB with one slot
// per level at S = 12, so occupancy is set by the registers (4 waves a SIMD
// at (12, 2, 1) instead of 3), and fewer LDS addresses are kept in scalar
// registers; measured 3 % faster on an 8192^2 grid, the same on 4096^2.
#ifndef PIE_RING_MAX
#define PIE_RING_MAX 4
#endif
template <int S, int C, int NCH, int SCHED, int ORDER>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(PIE_WAVES_PER_EU)))
    march_kernel(MarchArgs a) {
  // wave-uniform task index (scalar registers for everything derived)
  const int idx = __builtin_amdgcn_readfirstlane(
      xcd_block() * kBlockWaves + (int)(threadIdx.x / kWave));
  if (idx >= a.ntask) return;  // whole wave leaves together
  const int task = a.tasks ? a.tasks[idx] : idx;
  // task = (channel group, segment, stripe), stripe fastest
  const int per_group = a.stripes * a.segs;
  const int group = task / per_group, rem = task % per_group;
  march_task<S, C, NCH, SCHED, ORDER, false, kBlockWaves, PIE_RING_MAX>(
      a, rem % a.stripes,
                                             rem / a.stripes, group, nullptr);
}

// Work-groups of the persistent batch kernel: 16 waves, with four ring
// slots a wave (as the sweep kernel) so that 16 rings fit the LDS.
constexpr int kBatchWaves = kBatchGroupWaves;
constexpr int kBatchBlock = kWave * kBatchWaves;
constexpr int kBatchPF = 4;

// Sweep-kernel arguments of one launch of a batch item.
__device__ __forceinline__ MarchArgs batch_launch(const BatchItem& it, int S,
                                                  int seg_rows) {
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
  a.t_lo = a.t0;
  a.t_end = it.t_end;
  if (a.t0 + S > a.t_end) {
    // the last launch of the budget: its live steps are its last ones, so
    // the change of the item's last sweep is measured at the store
    a.t0 = a.t_end - S;
  }
  a.band = 0;
  a.stripes = it.stripes;
  a.segs = it.segs;
  a.seg_rows = seg_rows;
  return a;
}

// After a launch of a live item: count it, apply the stop rules (the
// launch's last sweep changed no value by tol or more, or the sweep budget
// is used up) and clear the change maximum for the next launch.
__device__ __forceinline__ void batch_advance(BatchItem& it, int S,
                                              unsigned long long now) {
  it.launches += 1;
  it.sweeps = min(it.sweeps + S, it.t_end);
  const float d = __uint_as_float(atomicExch(&it.dmax_bits, 0u));
  it.last_dmax = d;
  if (d < it.tol || it.sweeps >= it.max_sweeps) {
    it.done = 1;
    it.t_done = now;
  }
}

// Persistent sweep kernel of a batch (one round).  Items come in two kinds:
//  * solo items (few tasks) are each taken by one work-group, whose waves
//    split the item's tasks and which runs launch after launch, with a
//    work-group barrier between launches, for as long as tasks of the
//    shared items remain unclaimed (at least one launch a claim), or to
//    the item's end once no shared item is live.  A small item thus
//    advances many launches a round and finishes early;
//  * shared items: every wave takes tasks (tiles of any live shared item)
//    from one atomic work counter until it passes the total, so uneven
//    items share the machine; the epilogue kernel advances them by one
//    launch a round.
// Items already stopped are skipped; an item's launch count selects its
// input buffer and first step.
template <int S, int C, int NCH, int SCHED, int ORDER>
__global__ void __launch_bounds__(kBatchBlock) march_batch_kernel(BatchArgs b) {
  constexpr int kW = kBatchWaves;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  __shared__ int s_item, s_go;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(b.solo_queue, 1u);
    __syncthreads();
    const int si = s_item;
    __syncthreads();  // s_item is rewritten by the next claim
    if (si >= b.nsolo) break;
    BatchItem& it = b.items[b.solo_items[si]];
    // it.* written by thread 0 is read by every thread after a barrier
    while (!it.done) {
      const MarchArgs a = batch_launch(it, S, b.seg_rows);
      const int per_group = it.stripes * it.segs;
      for (int t = wave; t < it.ntask; t += kW) {
        const int id = it.tasks[t];
        const int group = id / per_group, rem = id % per_group;
        march_task<S, C, NCH, SCHED, ORDER, true, kW, kBatchPF>(
            a, rem % it.stripes, rem / it.stripes, group, &it.dmax_bits);
      }
      __syncthreads();  // every task of the launch has stored its rows
      if (threadIdx.x == 0) {
        batch_advance(it, S, wall_clock64());
        // go on while shared tasks wait to be claimed, or to the item's
        // end once no shared item is live (nothing else to wait for)
        s_go = atomicAdd(b.queue, 0u) < (unsigned)b.ntask ||
               atomicAdd(const_cast<unsigned*>(b.shared_live), 0u) == 0u;
      }
      __syncthreads();
      if (!s_go) break;
    }
  }
  for (;;) {
    int t = 0;
    if (lane == 0) t = (int)atomicAdd(b.queue, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, kWave));
    if (t >= b.ntask) break;
    const int item = b.task_item[t];
    const BatchItem& it = b.items[item];
    if (it.done) continue;
    const MarchArgs a = batch_launch(it, S, b.seg_rows);
    const int id = it.tasks[t - it.task0];
    const int per_group = it.stripes * it.segs;
    const int group = id / per_group, rem = id % per_group;
    march_task<S, C, NCH, SCHED, ORDER, true, kW, kBatchPF>(
        a, rem % it.stripes, rem / it.stripes, group,
        &b.items[item].dmax_bits);
  }
}

}  // namespace detail
}  // namespace pie
