// Device kernels of the Poisson solver and their host-side launchers.
//
// Layout in HBM: a solve works on a pitched grid of H rows x W columns.
// Every channel row is a segment of kPadLeft padding elements, the W values
// and at least kPadRight more padding; the base pointers the kernels get
// point at column 0 of channel 0 of row 0, so element (i, j) of channel c is
// base[c * plane + i * P + j] for j in [-kPadLeft, segment - kPadLeft).
// The host picks the layout (host_util.h: grid_layout): row-interleaved
// (plane = one segment, P = three segments) or planar (plane = H
// segments, P = one segment).  The float32 RGB state, the guidance (grad
// for the GridSolver, b for the EquSolver) and the uint8 mask (row pitch P,
// first segment) all use it.  Channel rows make every row load of a
// wavefront one contiguous 256-byte (or 512-byte) transaction, and the
// padding lets the sweep kernel load whole stripes without bounds tests.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pie {

constexpr int kPadLeft = 64;
constexpr int kPadRight = 256;  // >= 64 * (largest columns per lane)
constexpr int kSweepBlockWaves = 4;  // waves per work-group of the sweep kernel
constexpr int kBatchGroupWaves = 16;  // ... and of the persistent batch kernel

enum Sched { SCHED_JACOBI = 0, SCHED_GS_LEX = 1, SCHED_REDBLACK = 2 };
enum Order { ORDER_UDLR = 0, ORDER_ULRD = 1 };

// One launch of the temporally blocked sweep kernel: levels 1..S of the
// launch are global steps t0 .. t0+S-1; steps before t_lo or from t_end on
// leave the state as it is.  For SCHED_GS_LEX a pixel (i, j) is updated at
// step t when d = t - i - j is even and 0 <= d < band (band = 2 * sweeps of
// the call).
struct MarchArgs {
  const float* xin;
  float* xout;
  const float* g;
  const uint8_t* m;
  long long plane;  // elements per plane (H * P)
  int H, W, P;
  int t0, t_end, band;
  int t_lo;         // first live step (t0, or later in an end-aligned launch)
  int stripes, segs, seg_rows, ntask;
  // optional list of the tasks to run (ntask entries, each a task number
  // group * stripes * segs + seg * stripes + stripe); null = every task
  const int* tasks;
};

// One composite of a batch.  The host fills the geometry; the device keeps
// the solve state (launches, sweeps, dmax_bits, done, t_done) up to date.
struct BatchItem {
  float* xa;  // column-0 base pointers of the two state buffers
  float* xb;
  const float* g;
  const uint8_t* m;
  long long plane;
  int H, W, P;
  int stripes, segs, task0;  // tile geometry, first global (shared) task
  const int* tasks;          // the item's wave tasks with work: ids group *
  int ntask;                 // stripes * segs + seg * stripes + stripe
  int t_end;                 // sweeps never run past this step count
  int max_sweeps;            // stop rule: sweep budget
  float tol;                 // stop rule: max |update| of the last sweep
  int launches;              // sweep launches done (parity of the state)
  int sweeps;                // sweeps done
  int done;                  // 1 once a stop rule fired
  unsigned dmax_bits;        // max |update| of the launch in flight (bits)
  float last_dmax;           // max |update| of the last finished launch
  unsigned long long t_done; // device wall clock when the item stopped
  int solo;                  // 1: advanced by one work-group inside the
                             // sweep kernel, not by the epilogue
};

// columns of a wave (lanes) in the batch's task geometry
constexpr int kBatchLaneColumns = 64;

struct BatchArgs {
  BatchItem* items;
  const int* task_item;   // item of every shared task
  unsigned* queue;        // shared-task counter, reset by the epilogue
  unsigned* solo_queue;   // solo-item counter, reset by the epilogue
  const int* solo_items;  // the solo items, smallest first
  const unsigned* shared_live;  // live shared items, kept by the epilogue
  int ntask, seg_rows, nsolo;
};

// Temporal-blocking tile: S sweeps per launch, C columns per lane
// (a wavefront covers 64*C columns, 64*C - 2S of them written), seg_rows
// rows written per wave task.
struct TileCfg {
  int S, C, seg_rows;
  int N = 0;  // channels per wave task: 3, 1, or 0 for the tile's default
};

// Compiled (sweeps, columns per lane, channels per task) variants of the
// sweep kernel; each is built from march_inst.hip with -DPIE_S=s -DPIE_C=c
// -DPIE_N=n.  A wave task carries all three colour channels (n = 3) or one
// (n = 1: three task groups per tile, the channels are independent
// problems).  Tiles of two or four columns per lane keep that many values
// per level in registers and take one channel; one-column tiles come in both
// forms (one channel a task: a third of the registers and three times the
// tasks, which fills the machine on small grids).
#define PIE_TILES(X)                                                       \
  X(4, 1, 3) X(6, 1, 3) X(8, 1, 3) X(8, 1, 1) X(4, 2, 1) X(6, 2, 1)        \
  X(8, 2, 1) X(12, 2, 1) X(4, 4, 1) X(8, 4, 1)
// the channels per task of cfg (its N, or the default of its column count)
constexpr int tile_nch(const TileCfg& c) {
  return c.N ? c.N : (c.C >= 2 ? 1 : 3);
}
constexpr int tile_groups(const TileCfg& c) { return 3 / tile_nch(c); }
#define PIE_DECL(s, c, n)                                                  \
  hipError_t launch_march_##s##_##c##_##n(int sched, int order,           \
                                          const MarchArgs& a,             \
                                          hipStream_t st);                \
  hipError_t launch_batch_##s##_##c##_##n(int sched, int order,           \
                                          const BatchArgs& b, int blocks, \
                                          hipStream_t st);                \
  int batch_blocks_per_cu_##s##_##c##_##n(int sched, int order);         \
  int march_blocks_per_cu_##s##_##c##_##n(int sched, int order);
PIE_TILES(PIE_DECL)
#undef PIE_DECL

bool tile_supported(const TileCfg& cfg);
// The compiled (S, C, N) triples, for the autotuner.
int tile_variants(int (*out)[3], int max);
// Wave tasks (stripes x segments) of a tile on an H x W grid.
void tile_tasks(const TileCfg& cfg, int H, int W, int* stripes, int* segs);

hipError_t launch_march(const TileCfg& cfg, int sched, int order,
                        MarchArgs a, hipStream_t st);
// occ[stripe * H + r] = 1 if row r holds a solved pixel in the columns of
// stripe `stripe` ([stripe * U, min(stripe * U + U, W)), U the written
// columns of a stripe), else 0 (mask m: column-0 base, row pitch P).
hipError_t launch_stripe_rows(const uint8_t* m, int H, int W, int P, int U,
                              int stripes, uint8_t* occ, hipStream_t st);
// Resident work-groups (kSweepBlockWaves waves each) of the sweep kernel
// per CU.
int march_blocks_per_cu(const TileCfg& cfg, int sched, int order);

// Persistent batched sweep kernel (a round: S Jacobi sweeps of every live
// shared item, several launches of solo items) and its epilogue (advance
// the live shared items, apply the stop rules, reset the work counters,
// count live items into *live).
hipError_t launch_batch(const TileCfg& cfg, int sched, int order,
                        const BatchArgs& b, int blocks, hipStream_t st);
int batch_blocks_per_cu(const TileCfg& cfg, int sched, int order);
hipError_t launch_batch_epilogue(BatchItem* items, int n, int S,
                                 unsigned* queue, unsigned* solo_queue,
                                 unsigned* live, unsigned* shared_live,
                                 hipStream_t st);
// *out = device wall clock (for per-item latencies)
hipError_t launch_clock(unsigned long long* out, hipStream_t st);

// float[H][W][3] (interleaved) -> three pitched planes
hipError_t launch_deinterleave(const float* src, float* dst, long long plane,
                               int H, int W, int P, hipStream_t st);
// three pitched planes -> float[H][W][3]
hipError_t launch_interleave(const float* src, float* dst, long long plane,
                             int H, int W, int P, hipStream_t st);
// int32 mask[H][W] -> uint8 plane (0/1); counts[0] += solved pixels,
// counts[1] += solved pixels on the outermost rows/columns.
hipError_t launch_mask(const int32_t* src, uint8_t* dst, int H, int W, int P,
                       unsigned long long* counts, hipStream_t st);
// img[H][W][3] = tgt[H][W][3] at the unsolved pixels of the mask plane.
hipError_t launch_keep_unsolved(uint8_t* img, const uint8_t* tgt,
                                const uint8_t* m, int H, int W, int P,
                                hipStream_t st);
// GridSolver inputs from uint8 crops src/tgt [H][W][3] and the mask plane:
// x = tgt, g = sum over up, down, left, right of mixgrad(src_c - src_n,
// tgt_c - tgt_n) at solved pixels (mode 0 src, 1 avg, 2 max), 0 elsewhere.
// equ != 0 gives the EquSolver system on the grid instead: x = tgt at solved
// pixels and 0 (X[0]) elsewhere, g (= B) adds tgt of each unsolved neighbour.
hipError_t launch_divergence(const uint8_t* src, const uint8_t* tgt,
                             const uint8_t* m, float* x, float* g,
                             long long plane, int H, int W, int P, int mode,
                             int equ, hipStream_t st);
// The processors' mask preparation: out[i][j] = 1 where the channel mean
// of mask uint8[H][W][channels] is >= 128, except on the outermost rows and
// columns; rows[i] / cols[j] (zeroed by the caller) become 1 where a row /
// column holds a solved pixel.
hipError_t launch_prep_mask(const uint8_t* mask, int H, int W, int channels,
                            uint8_t* out, unsigned* rows, unsigned* cols,
                            hipStream_t st);
// crop int32[h][w] = m[x0 + r][y0 + c] (m uint8 with row length W)
hipError_t launch_crop_mask(const uint8_t* m, int W, int x0, int y0, int h,
                            int w, int32_t* crop, hipStream_t st);
// Guidance for the sweep kernel: NaN at every unsolved pixel and in the
// padding of the buffer g (allocation start; mask m likewise, both with row
// pitch P, channel stride plane, segment length seg); solved pixels whose
// guidance is NaN are counted into *nan_solved.
hipError_t launch_box_guidance(float* g, const uint8_t* m, long long plane,
                               int H, int P, int seg,
                               unsigned long long* nan_solved,
                               hipStream_t st);
// fill the n elements of a three-channel buffer (row pitch P, channel
// stride plane) with a per-channel constant
hipError_t launch_fill3(float* dst, long long n, int P, long long plane,
                        float v0, float v1, float v2, hipStream_t st);
// Equ system on the grid layout: X[id], B[id] (float[N][3]) to the pixel
// lin[id] = i * P + j of each id >= 1.
hipError_t launch_equ_scatter(const float* X, const float* B,
                              const int32_t* lin, int N, float* x, float* g,
                              long long plane, hipStream_t st);
// uint8[N][3] (out_u8) or float[N][3] (out_f32) in id order; row 0 = x0.
hipError_t launch_equ_gather(const float* x, const int32_t* lin, int N,
                             long long plane, float x00, float x01, float x02,
                             uint8_t* out_u8, float* out_f32,
                             hipStream_t st);
// Residual sum over the solved pixels into partial[nblocks][3] (double) and,
// if img != nullptr, the clamped/truncated image uint8[H][W][3].
int finalize_blocks(int H, int W);
hipError_t launch_finalize(int order, const float* x, const float* g,
                           const uint8_t* m, long long plane, int H, int W,
                           int P, uint8_t* img, double* partial, int nblocks,
                           hipStream_t st);
// Fixed-order sum of partial[n][3] -> err[3] (float).
hipError_t launch_reduce(const double* partial, int n, float* err,
                         hipStream_t st);

// ---- coarse-grid correction (warm start) -------------------------------
// out[hc][wc][3] = per f x f block, the sum over solved pixels of the
// signed residual (g + neighbours) - 4 x, in double.
hipError_t launch_restrict(const float* x, const float* g, const uint8_t* m,
                           long long plane, int P, int f, int hc, int wc,
                           float* out, hipStream_t st);
// x += cell-centred bilinear upsampling of corr[hc][wc][3] at solved pixels
hipError_t launch_prolong(float* x, const uint8_t* m, long long plane, int H,
                          int W, int P, int f, int hc, int wc,
                          const float* corr, hipStream_t st);
// Line-search sums of the prolonged correction d of corr: partial[b][0..2]
// = sum r . d (exact residual of x), partial[b][3..5] = sum d . A d, per
// channel, for blocks b < nblocks (prolong_dots_blocks()).
int prolong_dots_blocks(int H, int W);
hipError_t launch_prolong_dots(const float* x, const float* g,
                               const uint8_t* m, long long plane, int H,
                               int W, int P, int f, int hc, int wc,
                               const float* corr, double* partial,
                               int nblocks, hipStream_t st);
// one weighted Jacobi sweep xin -> xout (weight omega)
hipError_t launch_weighted(int order, const float* xin, float* xout,
                           const float* g, const uint8_t* m, long long plane,
                           int H, int W, int P, float omega, hipStream_t st);

// ---- EquSolver on an arbitrary A (no grid geometry) -------------------
hipError_t launch_equ_pack(const float* X, float4* out, int N,
                           hipStream_t st);
hipError_t launch_equ_jacobi(const int4* A, const float4* B,
                             const float4* xin, float4* xout, int N,
                             hipStream_t st);
int equ_generic_blocks(int N);
hipError_t launch_equ_finalize(const int4* A, const float4* B,
                               const float4* x, int N, uint8_t* img,
                               float* state, double* partial, int nblocks,
                               hipStream_t st);

// ---- coarse-grid algebra of the learned warm start (coarse.hip) ----
// Fields double [hc][wc][3]; mc uint8 [hc][wc] with an unsolved outer ring.
constexpr int kMaxDots = 4;
struct CoarseDots {  // per-channel sums of a[q] * b[q], q < count
  const double* a[kMaxDots];
  const double* b[kMaxDots];
  int count;
};
int coarse_dots_blocks(int n);
// Device scalars of the learned coarse correction (doubles, offsets):
// dot results (kMaxDots x 3), sig (3), any (1), the orthogonalisation
// coefficient (3), the step and its negative (3 + 3), the fine line-search
// alpha (3), <p_k, A p_k> per step (kWarmMaxSteps x 3).
constexpr int kWarmMaxSteps = 8;
constexpr int kWsDots = 0, kWsSig = kWsDots + kMaxDots * 3, kWsAny = kWsSig + 3,
              kWsCoef = kWsAny + 1, kWsStep = kWsCoef + 3,
              kWsNegStep = kWsStep + 3, kWsAlpha = kWsNegStep + 3,
              kWsPap = kWsAlpha + 3, kWsCount = kWsPap + 3 * kWarmMaxSteps;
constexpr int kWarmSig = 0, kWarmOrth = 1, kWarmStep = 2, kWarmAlpha = 3;
// out[k] = sum over b < nb of partial[b * width + k] (width <= 12; a fixed
// summation tree: deterministic)
hipError_t launch_sum_partials(const double* partial, int nb, int width,
                               double* out, hipStream_t st);
// one scalar step on sc (op kWarmSig: sig/any from the dots of (r, r) and
// `cells`; kWarmOrth: coefficient from the dot (z, A p_k) and pap_k;
// kWarmStep: step and pap_k from (z, Az), (z, z), (r, z); kWarmAlpha:
// alpha from the six line-search sums)
hipError_t launch_warm_scalar(double* sc, int op, int k, double cells,
                              hipStream_t st);
// r = rc (float) on solved cells, 0 elsewhere; c = 0
hipError_t launch_coarse_load(const float* rc, const uint8_t* mc, int n,
                              double* r, double* c, hipStream_t st);
// network input x [hp][wp][8]: r / sig in channels 0..2, the mask in 6
// (sig: three doubles in device memory)
hipError_t launch_net_input(const double* r, const uint8_t* mc, int hc,
                            int wc, int hp, int wp, const double* sig,
                            float* x, hipStream_t st);
// z = network output [hp][wp][3] on solved cells (0 if *any is 0)
hipError_t launch_net_output(const float* out, const uint8_t* mc, int hc,
                             int wc, int wp, const double* any, double* z,
                             hipStream_t st);
// one damped (weight 1/2) sweep of the homogeneous coarse problem
hipError_t launch_lowpass(const double* q, const uint8_t* mc, int hc, int wc,
                          double* out, hipStream_t st);
// out = A_c p = 4 p - sum of the four neighbours, on solved cells
hipError_t launch_coarse_apply(const double* p, const uint8_t* mc, int hc,
                               int wc, double* out, hipStream_t st);
// partial[block][q][3] for blocks < coarse_dots_blocks(n)
hipError_t launch_coarse_dots(const CoarseDots& d, int n, double* partial,
                              hipStream_t st);
// y += a_c x (a: three doubles in device memory)
hipError_t launch_coarse_axpy(double* y, const double* x, const double* a,
                              int n, hipStream_t st);
hipError_t launch_coarse_to_float(const double* c, int n3, float* out,
                                  hipStream_t st);
// v = float(double(v) * alpha_c); *maxbits = max(*maxbits, bits of
// max |v|) (alpha: three doubles in device memory)
hipError_t launch_coarse_scale(float* v, const double* alpha, int n,
                               unsigned* maxbits, hipStream_t st);

}  // namespace pie
