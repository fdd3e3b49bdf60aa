// C ABI of the gfx950 Poisson solver (declared in include/fpie_hip.h).
//
// A pie_solver owns its device buffers and one HIP stream.  reset() deep-
// copies the caller's arrays to HBM (as fpie/core/base_solver.h:27-59,
// 101-140 copies numpy arrays to new[] buffers) and lays them out as
// pitched float planes; step() runs the sweeps on the device and copies
// back only the uint8 image and the three residual sums.
//
// EquSolver systems whose A matches the geometry recorded by partition()
// (what fpie/process.py:248-266 always builds) are solved on the grid
// layout by the same kernels as the GridSolver: b takes the place of grad
// and every unsolved pixel holds X[0] (the value the reference reads through
// id 0).  Any other A is solved by a gather kernel (Jacobi only).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <initializer_list>
#include <map>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_util.h"

using pie::MarchArgs;
using pie::TileCfg;
using pie::host::DevBuf;
using pie::host::fail;
using pie::host::pitch_of;
using pie::host::round_up;
using pie::host::upload;

namespace {
thread_local std::string g_err;
}  // namespace

int pie_fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct pie_solver {
  int kind = PIE_GRID, sched = PIE_JACOBI, order = PIE_ORDER_ULRD;
  int param_a = 0, param_b = 0;
  int device = 0, cus = 1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  bool loaded = false;
  bool grid_layout = false;  // Grid, or Equ solved on the grid layout
  int H = 0, W = 0, P = 0;
  long long plane = 0, masked = 0;
  long long bufsz = 0, mbytes = 0;  // elements per state buffer, mask bytes
  int cur = 0;  // which of xa/xb holds the state
  DevBuf<float> xa, xb, g;
  DevBuf<uint8_t> m, img;
  // uint8 target crop [H][W][3] kept by pie_reset_equ_images for the crop
  // output of an EquSolver (unsolved pixels keep the target)
  DevBuf<uint8_t> tgt8;
  DevBuf<double> partial;
  DevBuf<float> err;
  DevBuf<float> coarse;  // restriction / prolongation staging
  // learned coarse correction (pie_warm_net_start): coarse mask, double
  // fields (residual, correction, smoothing scratch, directions p_k and
  // A_c p_k), network input / output, dot-product partials; warm_factor is
  // the factor of the correction left in `coarse` (0: none)
  DevBuf<uint8_t> wmc;
  DevBuf<double> wf, wpart, wsc;  // fields, dot partials, scalars
  DevBuf<float> wx, wout;
  DevBuf<unsigned> wmax;
  int warm_factor = 0;

  // EquSolver
  int N = 0;
  float x0[3] = {0, 0, 0};
  std::vector<int32_t> part_ids;
  int ph = 0, pw = 0, part_count = -1;
  DevBuf<int32_t> lin;
  DevBuf<int4> A4;
  DevBuf<float4> B4, X4a, X4b;

  TileCfg cfg{8, 1, 64};
  bool tile_fixed = false, tuned = false;
  // Per tile (S, C, channels, rows): the tasks whose region holds a solved
  // pixel (the others have nothing to do), and their count.
  struct TaskList {
    DevBuf<int> ids;
    int count = 0;
  };
  std::map<std::array<int, 4>, std::unique_ptr<TaskList>> task_lists;
  // gs_lex: the tasks with work ordered by their first anti-diagonal
  // r0 + c_lo (the launches whose band covers little of the grid run one
  // window of this order: the tasks that may have work there)
  struct DiagList {
    DevBuf<int> ids;
    std::vector<int> d0;  // r0 + c_lo of each entry, ascending
    int span = 0;         // largest (r1 - 1 + c_hi) - (r0 + c_lo)
  };
  std::map<std::array<int, 4>, std::unique_ptr<DiagList>> diag_lists;
  // per stripe width U: cum[stripe * (H + 1) + r] = rows < r of the stripe
  // that hold a solved pixel
  std::map<int, std::vector<int>> stripe_cum;
  // Launch sequences of step(n) calls, captured as hipGraphs and replayed
  // when a call repeats (same sweeps, tile and starting buffer): kernel
  // launches of small grids cost more to issue one by one than to run.
  // Key: iters, cur, S, C, channels, rows, listed tasks.
  struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::map<std::array<int, 7>, Graph> graphs;

  bool timing = false;
  double t_ms = 0.0, t_updates = 0.0;
  long long t_launches = 0;

  // column-0 base pointers of the padded planes
  float* xa0() { return xa.p + pie::kPadLeft; }
  float* xb0() { return xb.p + pie::kPadLeft; }
  float* g0() { return g.p + pie::kPadLeft; }
  uint8_t* m0() { return m.p + pie::kPadLeft; }
  float* state() { return cur ? xb0() : xa0(); }
  float* other() { return cur ? xa0() : xb0(); }
  int eff_order() const {
    return kind == PIE_EQU ? PIE_ORDER_UDLR : order;  // gcc/equ.cc:48-53
  }
};

namespace {

void drop_graph(pie_solver* s) {
  for (auto& kv : s->graphs) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
  }
  s->graphs.clear();
}

void free_problem(pie_solver* s) {
  s->loaded = false;
  s->grid_layout = false;
  s->xa.release();
  s->xb.release();
  s->g.release();
  s->m.release();
  s->img.release();
  s->tgt8.release();
  s->partial.release();
  s->err.release();
  s->coarse.release();
  s->wmc.release();
  s->wf.release();
  s->wpart.release();
  s->wsc.release();
  s->wx.release();
  s->wout.release();
  s->wmax.release();
  s->warm_factor = 0;
  s->lin.release();
  s->A4.release();
  s->B4.release();
  s->X4a.release();
  s->X4b.release();
  s->H = s->W = s->P = 0;
  s->plane = s->masked = 0;
  s->bufsz = s->mbytes = 0;
  s->N = 0;
  s->cur = 0;
  s->tuned = false;
  s->task_lists.clear();
  s->diag_lists.clear();
  s->stripe_cum.clear();
  drop_graph(s);
}

// Total global steps of one call of `iters` sweeps and the band width.
void schedule_steps(const pie_solver* s, int iters, int* steps, int* band) {
  *band = 2 * iters;
  if (s->sched == PIE_JACOBI)
    *steps = iters;
  else if (s->sched == PIE_REDBLACK)
    *steps = 2 * iters;
  else  // sweep k of pixel (i, j) runs at step 2k + i + j
    *steps = iters > 0 ? 2 * (iters - 1) + (s->H - 1) + (s->W - 1) + 1 : 0;
}

MarchArgs march_args(pie_solver* s, const float* xin, float* xout, int t0,
                     int t_end, int band) {
  MarchArgs a{};
  a.xin = xin;
  a.xout = xout;
  a.g = s->g0();
  a.m = s->m0();
  a.plane = s->plane;
  a.H = s->H;
  a.W = s->W;
  a.P = s->P;
  a.t0 = t0;
  a.t_lo = t0;
  a.t_end = t_end;
  a.band = band;
  return a;
}

int alloc_grid(pie_solver* s, int H, int W) {
  if (H < 1 || W < 1) return fail(PIE_EINVAL, "empty grid");
  s->H = H;
  s->W = W;
  const pie::host::GridLayout L = pie::host::grid_layout(H, W);
  if ((long long)H * L.P > (1LL << 31) - 1)
    return fail(PIE_EINVAL, "grid larger than 2^31 elements");
  s->P = L.P;
  s->plane = L.plane;
  s->bufsz = L.floats;
  s->mbytes = L.mask;
  PIE_HIP(s->xa.alloc(s->bufsz));
  PIE_HIP(s->xb.alloc(s->bufsz));
  PIE_HIP(s->g.alloc(s->bufsz));
  PIE_HIP(s->m.alloc(s->mbytes));
  // padding must hold finite values (zeros) and an unsolved mask
  PIE_HIP(hipMemsetAsync(s->xa.p, 0, s->bufsz * sizeof(float), s->stream));
  PIE_HIP(hipMemsetAsync(s->xb.p, 0, s->bufsz * sizeof(float), s->stream));
  PIE_HIP(hipMemsetAsync(s->g.p, 0, s->bufsz * sizeof(float), s->stream));
  PIE_HIP(hipMemsetAsync(s->m.p, 0, s->mbytes, s->stream));
  const int nb = pie::finalize_blocks(H, W);
  PIE_HIP(s->partial.alloc(3 * (size_t)nb));
  PIE_HIP(s->err.alloc(3));
  s->grid_layout = true;
  return PIE_OK;
}

std::array<int, 4> tile_key(const TileCfg& c) {
  return {c.S, c.C, pie::tile_nch(c), c.seg_rows};
}

// Per-stripe prefix counts of the rows holding a solved pixel, for the
// stripe width of tile c (one mask scan per width, cached).
int stripe_cum(pie_solver* s, const TileCfg& c, const std::vector<int>** out) {
  const int U = 64 * c.C - 2 * c.S;  // written columns of a stripe
  auto it = s->stripe_cum.find(U);
  if (it == s->stripe_cum.end()) {
    const int stripes = (s->W + U - 1) / U, H = s->H;
    DevBuf<uint8_t> occ;
    PIE_HIP(occ.alloc((size_t)stripes * H));
    PIE_HIP(pie::launch_stripe_rows(s->m0(), H, s->W, s->P, U, stripes, occ.p,
                                    s->stream));
    std::vector<uint8_t> h((size_t)stripes * H);
    PIE_HIP(hipMemcpyAsync(h.data(), occ.p, h.size(), hipMemcpyDeviceToHost,
                           s->stream));
    PIE_HIP(hipStreamSynchronize(s->stream));
    std::vector<int> cum((size_t)stripes * (H + 1), 0);
    for (int st = 0; st < stripes; ++st)
      for (int r = 0; r < H; ++r)
        cum[(size_t)st * (H + 1) + r + 1] =
            cum[(size_t)st * (H + 1) + r] + h[(size_t)st * H + r];
    it = s->stripe_cum.emplace(U, std::move(cum)).first;
  }
  *out = &it->second;
  return PIE_OK;
}

// Which task regions of tile c hold a solved pixel: flags[region] on the
// host (region = seg * stripes + stripe).
int region_flags(pie_solver* s, const TileCfg& c, std::vector<int>* flags) {
  const std::vector<int>* cum = nullptr;
  PIE_CHECK(stripe_cum(s, c, &cum));
  int stripes = 0, segs = 0;
  pie::tile_tasks(c, s->H, s->W, &stripes, &segs);
  flags->assign((size_t)stripes * segs, 0);
  const int H = s->H;
  for (int seg = 0; seg < segs; ++seg) {
    const int r0 = seg * c.seg_rows, r1 = std::min(r0 + c.seg_rows, H);
    for (int st = 0; st < stripes; ++st) {
      const int* row = cum->data() + (size_t)st * (H + 1);
      (*flags)[(size_t)seg * stripes + st] = row[r1] > row[r0];
    }
  }
  return PIE_OK;
}

// Number of tasks of tile c with work.
int task_count(pie_solver* s, const TileCfg& c, int* count) {
  std::vector<int> flags;
  PIE_CHECK(region_flags(s, c, &flags));
  int n = 0;
  for (int f : flags) n += f != 0;
  *count = n * pie::tile_groups(c);
  return PIE_OK;
}

// The list of tasks of tile c with work, on the device (cached per tile):
// the sweep kernel runs only these.  Tasks keep their natural order, so
// neighbouring tasks still land on one XCD.
int task_list(pie_solver* s, const TileCfg& c, const int** ids, int* count) {
  const auto key = tile_key(c);
  auto it = s->task_lists.find(key);
  if (it == s->task_lists.end()) {
    std::vector<int> flags;
    PIE_CHECK(region_flags(s, c, &flags));
    const int regions = (int)flags.size();
    std::vector<int> list;
    for (int g = 0; g < pie::tile_groups(c); ++g)
      for (int r = 0; r < regions; ++r)
        if (flags[r]) list.push_back(g * regions + r);
    std::unique_ptr<pie_solver::TaskList> tl(new pie_solver::TaskList());
    tl->count = (int)list.size();
    PIE_HIP(tl->ids.alloc(std::max<size_t>(list.size(), 1)));
    if (!list.empty()) {
      PIE_CHECK(upload(list.data(), list.size() * sizeof(int), tl->ids.p,
                       s->stream));
      PIE_HIP(hipStreamSynchronize(s->stream));  // list is host memory
    }
    it = s->task_lists.emplace(key, std::move(tl)).first;
  }
  *ids = it->second->ids.p;
  *count = it->second->count;
  return PIE_OK;
}

// Tile c's tasks with work in the order of their first anti-diagonal
// (cached per tile; see pie_solver::DiagList).
int diag_list(pie_solver* s, const TileCfg& c,
              const pie_solver::DiagList** out) {
  const auto key = tile_key(c);
  auto it = s->diag_lists.find(key);
  if (it == s->diag_lists.end()) {
    std::vector<int> flags;
    PIE_CHECK(region_flags(s, c, &flags));
    int stripes = 0, segs = 0;
    pie::tile_tasks(c, s->H, s->W, &stripes, &segs);
    const int U = 64 * c.C - 2 * c.S;  // written columns of a stripe
    const int regions = (int)flags.size();
    std::vector<std::pair<int, int>> order;  // (r0 + c_lo, task id)
    std::unique_ptr<pie_solver::DiagList> dl(new pie_solver::DiagList());
    for (int g = 0; g < pie::tile_groups(c); ++g)
      for (int r = 0; r < regions; ++r) {
        if (!flags[r]) continue;
        const int seg = r / stripes, st = r % stripes;
        const int r0 = seg * c.seg_rows, r1 = std::min(r0 + c.seg_rows, s->H);
        const int c_lo = st * U, c_hi = std::min(c_lo + U, s->W) - 1;
        order.push_back({r0 + c_lo, g * regions + r});
        dl->span = std::max(dl->span, (r1 - 1 + c_hi) - (r0 + c_lo));
      }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int, int>& x,
                        const std::pair<int, int>& y) {
                       return x.first < y.first;
                     });
    std::vector<int> ids(order.size());
    dl->d0.resize(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
      dl->d0[k] = order[k].first;
      ids[k] = order[k].second;
    }
    PIE_HIP(dl->ids.alloc(std::max<size_t>(ids.size(), 1)));
    if (!ids.empty()) {
      PIE_CHECK(upload(ids.data(), ids.size() * sizeof(int), dl->ids.p,
                       s->stream));
      PIE_HIP(hipStreamSynchronize(s->stream));  // ids is host memory
    }
    it = s->diag_lists.emplace(key, std::move(dl)).first;
  }
  *out = it->second.get();
  return PIE_OK;
}

// Rows per wave task that make one launch of tile cfg `rounds` rounds of
// resident waves: the smallest row count whose tasks with work fit into
// every resident wave slot of the chip (the kernel's occupancy x CUs) that
// many times, so no round is left partly empty and the rows per task
// (hence the share of recomputed halo rows) are as large as that allows.
int round_rows(pie_solver* s, TileCfg cfg, int rounds, int* rows) {
  const int slots = pie::march_blocks_per_cu(cfg, s->sched, s->eff_order()) *
                    s->cus * pie::kSweepBlockWaves * rounds;
  int lo = 2 * cfg.S, hi = std::max(2 * cfg.S, s->H);
  while (lo < hi) {  // task counts fall as the rows per task grow
    cfg.seg_rows = (lo + hi) / 2;
    int n = 0;
    PIE_CHECK(task_count(s, cfg, &n));
    if (n <= slots)
      hi = cfg.seg_rows;
    else
      lo = cfg.seg_rows + 1;
  }
  *rows = hi;
  return PIE_OK;
}

// NaN guidance at the unsolved pixels and in the padding (the sweep kernel
// tells solved pixels by it); a solved pixel whose guidance is NaN is an
// error.  Call once the guidance and the mask are in place.
int box_guidance(pie_solver* s) {
  DevBuf<unsigned long long> bad;
  PIE_HIP(bad.alloc(1));
  PIE_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), s->stream));
  PIE_HIP(pie::launch_box_guidance(s->g.p, s->m.p, s->plane, s->H, s->P,
                                   pitch_of(s->W), bad.p, s->stream));
  unsigned long long n = 0;
  PIE_HIP(hipMemcpyAsync(&n, bad.p, sizeof(n), hipMemcpyDeviceToHost,
                         s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  if (n) {
    free_problem(s);
    return fail(PIE_EINVAL, std::to_string(n) +
                                " guidance values of solved pixels are NaN");
  }
  return PIE_OK;
}

int default_tile(pie_solver* s) {
  if (s->tile_fixed) return PIE_OK;
  // Below 2^22 pixels (where pie_iterate does not autotune): the measured
  // small-grid tiles.  From 2^22 pixels on the autotuner picks among all
  // tiles; its starting point is (8, 1, 3) with one round of resident
  // waves.
  const long long px = (long long)s->H * s->W;
  if (px < (1LL << 22)) {
    s->cfg = pie::host::small_grid_tile(px);
  } else {
    TileCfg c{8, 1, 64, 3};
    PIE_CHECK(round_rows(s, c, 1, &c.seg_rows));
    s->cfg = c;
  }
  return PIE_OK;
}

int run_sweeps(pie_solver* s, int iters);

// step(n) calls of at least this many launches on grids below 2^22 pixels
// replay a captured graph of their launches.
constexpr long long kGraphMinLaunches = 16;
// graphs kept per solver (all are dropped when one more is needed)
constexpr size_t kMaxGraphs = 8;

// Time `launches` launches of tile c on the loaded problem (the state buffer
// is only read; the spare buffer takes the output): milliseconds per step.
// The launches start at steps spread evenly over a call of `steps` steps,
// so that a gs_lex call's fill and drain (tasks outside the wavefront band
// skip their rows) weigh as they do in the call.
int time_tile(pie_solver* s, const TileCfg& c, int launches, int steps,
              int band, float* per_sweep) {
  const int last = std::max(0, steps - c.S);
  auto start = [&](int l) {
    const long long t = (2LL * l + 1) * steps / (2LL * launches);
    return (int)std::min<long long>(t / c.S * c.S, last);
  };
  MarchArgs a = march_args(s, s->state(), s->other(), start(0),
                           std::max(steps, c.S), band);
  PIE_CHECK(task_list(s, c, &a.tasks, &a.ntask));
  PIE_HIP(pie::launch_march(c, s->sched, s->eff_order(), a, s->stream));
  PIE_HIP(hipEventRecord(s->ev0, s->stream));
  for (int l = 0; l < launches; ++l) {
    a.t0 = a.t_lo = start(l);
    PIE_HIP(pie::launch_march(c, s->sched, s->eff_order(), a, s->stream));
  }
  PIE_HIP(hipEventRecord(s->ev1, s->stream));
  PIE_HIP(hipEventSynchronize(s->ev1));
  float ms = 0.0f;
  PIE_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  *per_sweep = ms / (launches * c.S);
  return PIE_OK;
}

// Autotuner: every compiled tile with rows per task from a fixed list and
// the counts that fill one or two rounds of resident waves, timed once;
// then the fastest kBest of them timed kReps more times each (launch times
// vary by several per cent from run to run), keeping each one's best time.
// Timed as launches of a call of `iters` sweeps (the call that triggers
// the autotuner).
int autotune(pie_solver* s, int launches, int iters) {
  if (!s->grid_layout) return PIE_OK;
  constexpr int kBest = 6, kReps = 3;
  int var[32][3];
  const int nv = std::min(pie::tile_variants(var, 32), 32);
  int steps, band;
  schedule_steps(s, std::max(iters, 1), &steps, &band);
  std::vector<std::pair<float, TileCfg>> cand;
  for (int v = 0; v < nv; ++v) {
    TileCfg c0{var[v][0], var[v][1], 16, var[v][2]};
    int r1 = 0, r2 = 0;
    PIE_CHECK(round_rows(s, c0, 1, &r1));
    PIE_CHECK(round_rows(s, c0, 2, &r2));
    const int rows[] = {8, 16, 32, 64, 96, 128, 192, 256, 384, r1, r2};
    const int nr = (int)(sizeof(rows) / sizeof(rows[0]));
    for (int i = 0; i < nr; ++i) {
      const int rr = rows[i];
      if (rr > 2 * s->H && rr != rows[0]) continue;
      bool seen = false;  // a count listed twice is timed once
      for (int k = 0; k < i; ++k) seen = seen || rows[k] == rr;
      if (seen) continue;
      TileCfg c{var[v][0], var[v][1], rr, var[v][2]};
      float t = 0.0f;
      PIE_CHECK(time_tile(s, c, std::max(2, launches / 2), steps, band, &t));
      cand.push_back({t, c});
    }
  }
  std::sort(cand.begin(), cand.end(),
            [](const auto& x, const auto& y) { return x.first < y.first; });
  if ((int)cand.size() > kBest) cand.resize(kBest);
  for (auto& c : cand)
    for (int r = 0; r < kReps; ++r) {
      float t = 0.0f;
      PIE_CHECK(time_tile(s, c.second, launches, steps, band, &t));
      c.first = std::min(c.first, t);
    }
  std::sort(cand.begin(), cand.end(),
            [](const auto& x, const auto& y) { return x.first < y.first; });
  if (!cand.empty()) s->cfg = cand[0].second;
  s->tuned = true;
  return PIE_OK;
}

// Pixels (i, j) of an H x W grid with i + j <= k.
long long diag_count(long long k, long long H, long long W) {
  auto tri = [](long long n) { return n > 0 ? n * (n + 1) / 2 : 0LL; };
  return tri(k + 1) - tri(k - H + 1) - tri(k - W + 1) + tri(k - H - W + 1);
}

// gs_lex launches whose wavefront band covers less than this share of the
// grid (the fill and the drain of a call) take a tile with a quarter of the
// rows per task: such a launch lasts one task's march, not its work
// (measured at 4096^2: 0.4 slightly ahead of 0.25).
#ifndef PIE_FINE_SHARE
#define PIE_FINE_SHARE 0.4
#endif

int run_sweeps(pie_solver* s, int iters) {
  if (iters <= 0) return PIE_OK;
  if (!s->grid_layout) {  // EquSolver on an arbitrary A: one sweep a launch
    if (s->timing) PIE_HIP(hipEventRecord(s->ev0, s->stream));
    for (int k = 0; k < iters; ++k) {
      const float4* in = s->cur ? s->X4b.p : s->X4a.p;
      float4* out = s->cur ? s->X4a.p : s->X4b.p;
      PIE_HIP(pie::launch_equ_jacobi(s->A4.p, s->B4.p, in, out, s->N,
                                     s->stream));
      s->cur ^= 1;
    }
    if (s->timing) {
      PIE_HIP(hipEventRecord(s->ev1, s->stream));
      PIE_HIP(hipEventSynchronize(s->ev1));
      float ms = 0.0f;
      PIE_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
      s->t_ms += ms;
      s->t_launches += iters;
      s->t_updates += (double)iters * s->masked;
    }
    return PIE_OK;
  }
  int steps, band;
  schedule_steps(s, iters, &steps, &band);
  const TileCfg c = s->cfg;
  const int* tasks = nullptr;
  int ntask = 0;
  PIE_CHECK(task_list(s, c, &tasks, &ntask));
  const long long launches = (steps + c.S - 1) / c.S;
  // gs_lex: the same tile with a quarter of the rows per task, for the
  // launches whose band covers little of the grid (the decomposition does
  // not change the result: every launch advances every pixel S steps)
  TileCfg cf = c;
  const pie_solver::DiagList* dl = nullptr;
  const bool fine = s->sched == PIE_GS_LEX && c.seg_rows >= 32;
  if (fine) {
    cf.seg_rows = (c.seg_rows + 3) / 4;
    PIE_CHECK(diag_list(s, cf, &dl));
  }
  const double area = (double)s->H * s->W;
  auto sparse = [&](int t0) {
    // pixels the band covers at some step of the launch (i + j in
    // (t0 - band - S, t0 + S)), as a share of the grid
    const long long hi = diag_count(t0 + c.S - 1, s->H, s->W);
    const long long lo = diag_count((long long)t0 - band - c.S, s->H, s->W);
    return (double)(hi - lo) < PIE_FINE_SHARE * area;
  };
  // issue the launches, or capture them into a graph: the pointers, step
  // numbers and tiles of every launch follow from (iters, cur, tile)
  auto issue = [&](int cur0) -> int {
    int cur = cur0;
    for (int t0 = 0; t0 < steps; t0 += c.S) {
      MarchArgs a = march_args(s, cur ? s->xb0() : s->xa0(),
                               cur ? s->xa0() : s->xb0(), t0, steps, band);
      const bool f = fine && sparse(t0);
      a.tasks = tasks;
      a.ntask = ntask;
      if (f) {
        // the window of the diagonal order holding every task that may have
        // work: not before the band once t0 >= S (r0 + c_lo <= last step),
        // not after it with both buffers agreeing ((r1 - 1) + c_hi >
        // t0 - S - band, so r0 + c_lo > t0 - S - band - span)
        const std::vector<int>& d0 = dl->d0;
        const long long x = (long long)t0 - c.S - band - dl->span;
        const int t_last = std::min(t0 + c.S, steps) - 1;
        const size_t lo =
            std::upper_bound(d0.begin(), d0.end(), x) - d0.begin();
        const size_t hi =
            t0 < c.S ? d0.size()
                     : std::upper_bound(d0.begin(), d0.end(), t_last) -
                           d0.begin();
        a.tasks = dl->ids.p + lo;
        a.ntask = hi > lo ? (int)(hi - lo) : 0;
      }
      PIE_HIP(pie::launch_march(f ? cf : c, s->sched, s->eff_order(), a,
                                s->stream));
      cur ^= 1;
    }
    return PIE_OK;
  };
  const bool use_graph = launches >= kGraphMinLaunches &&
                         (long long)s->H * s->W < (1LL << 22);
  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    const std::array<int, 7> key{iters, s->cur, c.S, c.C, pie::tile_nch(c),
                                 c.seg_rows, ntask};
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
      if (s->graphs.size() >= kMaxGraphs) drop_graph(s);
      PIE_HIP(hipStreamBeginCapture(s->stream,
                                    hipStreamCaptureModeThreadLocal));
      const int rc = issue(s->cur);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s->stream, &g);
      if (rc != PIE_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc != PIE_OK ? rc
                            : fail(PIE_EHIP, std::string("sweep graph: ") +
                                                 hipGetErrorString(e));
      }
      pie_solver::Graph entry;
      entry.graph = g;
      const hipError_t ei =
          hipGraphInstantiate(&entry.exec, g, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(PIE_EHIP, std::string("sweep graph instantiate: ") +
                                  hipGetErrorString(ei));
      }
      it = s->graphs.emplace(key, entry).first;
    }
    exec = it->second.exec;
  }
  if (s->timing) PIE_HIP(hipEventRecord(s->ev0, s->stream));
  if (use_graph)
    PIE_HIP(hipGraphLaunch(exec, s->stream));
  else
    PIE_CHECK(issue(s->cur));
  s->cur ^= (int)(launches & 1);
  if (s->timing) {
    PIE_HIP(hipEventRecord(s->ev1, s->stream));
    PIE_HIP(hipEventSynchronize(s->ev1));
    float ms = 0.0f;
    PIE_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->t_ms += ms;
    s->t_launches += launches;
    s->t_updates += (double)iters * s->masked;
  }
  return PIE_OK;
}

int read_outputs(pie_solver* s, uint8_t* host_img, float* host_err,
                 float* host_state) {
  if (!s->loaded) return fail(PIE_ESTATE, "step before reset");
  if (s->grid_layout) {
    const int nb = pie::finalize_blocks(s->H, s->W);
    const bool grid_img = host_img && s->kind == PIE_GRID;
    if (grid_img && !s->img.p)
      PIE_HIP(s->img.alloc((size_t)s->H * s->W * 3));
    PIE_HIP(pie::launch_finalize(s->eff_order(), s->state(), s->g0(), s->m0(),
                                 s->plane, s->H, s->W, s->P,
                                 grid_img ? s->img.p : nullptr, s->partial.p,
                                 nb, s->stream));
    PIE_HIP(pie::launch_reduce(s->partial.p, nb, s->err.p, s->stream));
    if (s->kind == PIE_EQU && (host_img || host_state)) {
      if (host_img && !s->img.p) PIE_HIP(s->img.alloc((size_t)s->N * 3));
      DevBuf<float> st;
      if (host_state) PIE_HIP(st.alloc((size_t)s->N * 3));
      PIE_HIP(pie::launch_equ_gather(s->state(), s->lin.p, s->N, s->plane,
                                     s->x0[0], s->x0[1], s->x0[2],
                                     host_img ? s->img.p : nullptr,
                                     host_state ? st.p : nullptr, s->stream));
      if (host_img)
        PIE_HIP(hipMemcpyAsync(host_img, s->img.p, (size_t)s->N * 3,
                               hipMemcpyDeviceToHost, s->stream));
      if (host_state)
        PIE_HIP(hipMemcpyAsync(host_state, st.p, (size_t)s->N * 3 * 4,
                               hipMemcpyDeviceToHost, s->stream));
      PIE_HIP(hipStreamSynchronize(s->stream));
    } else if (s->kind == PIE_GRID) {
      if (grid_img)
        PIE_HIP(hipMemcpyAsync(host_img, s->img.p, (size_t)s->H * s->W * 3,
                               hipMemcpyDeviceToHost, s->stream));
      if (host_state) {
        DevBuf<float> st;
        PIE_HIP(st.alloc((size_t)s->H * s->W * 3));
        PIE_HIP(pie::launch_interleave(s->state(), st.p, s->plane, s->H, s->W,
                                       s->P, s->stream));
        PIE_HIP(hipMemcpyAsync(host_state, st.p, (size_t)s->H * s->W * 3 * 4,
                               hipMemcpyDeviceToHost, s->stream));
        PIE_HIP(hipStreamSynchronize(s->stream));
      }
    }
  } else {
    const int nb = pie::equ_generic_blocks(s->N);
    if (host_img && !s->img.p) PIE_HIP(s->img.alloc((size_t)s->N * 3));
    DevBuf<float> st;
    if (host_state) PIE_HIP(st.alloc((size_t)s->N * 3));
    const float4* x = s->cur ? s->X4b.p : s->X4a.p;
    PIE_HIP(pie::launch_equ_finalize(s->A4.p, s->B4.p, x, s->N,
                                     host_img ? s->img.p : nullptr,
                                     host_state ? st.p : nullptr, s->partial.p,
                                     nb, s->stream));
    PIE_HIP(pie::launch_reduce(s->partial.p, nb, s->err.p, s->stream));
    if (host_img)
      PIE_HIP(hipMemcpyAsync(host_img, s->img.p, (size_t)s->N * 3,
                             hipMemcpyDeviceToHost, s->stream));
    if (host_state)
      PIE_HIP(hipMemcpyAsync(host_state, st.p, (size_t)s->N * 3 * 4,
                             hipMemcpyDeviceToHost, s->stream));
    PIE_HIP(hipStreamSynchronize(s->stream));
  }
  float e[3];
  PIE_HIP(hipMemcpyAsync(e, s->err.p, sizeof(e), hipMemcpyDeviceToHost,
                         s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  if (host_err) std::memcpy(host_err, e, sizeof(e));
  return PIE_OK;
}

}  // namespace

extern "C" {

const char* pie_last_error(void) { return g_err.c_str(); }

int pie_create(int kind, int schedule, int order, int param_a, int param_b,
               pie_solver** out) {
  if (!out) return fail(PIE_EINVAL, "out is NULL");
  *out = nullptr;
  if (kind != PIE_EQU && kind != PIE_GRID)
    return fail(PIE_EINVAL, "kind must be PIE_EQU or PIE_GRID");
  if (schedule != PIE_JACOBI && schedule != PIE_GS_LEX &&
      schedule != PIE_REDBLACK)
    return fail(PIE_EINVAL, "unknown schedule");
  if (order != PIE_ORDER_UDLR && order != PIE_ORDER_ULRD)
    return fail(PIE_EINVAL, "unknown neighbour order");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(PIE_EHIP, "no HIP device");
  pie_solver* s = new pie_solver();
  s->kind = kind;
  s->sched = schedule;
  s->order = order;
  s->param_a = param_a;
  s->param_b = param_b;
  hipError_t e = hipGetDevice(&s->device);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount,
                              s->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&s->ev0);
  if (e == hipSuccess) e = hipEventCreate(&s->ev1);
  if (e != hipSuccess) {
    pie_destroy(s);
    return fail(PIE_EHIP, std::string("solver set-up: ") + hipGetErrorString(e));
  }
  *out = s;
  return PIE_OK;
}

int pie_destroy(pie_solver* s) {
  if (!s) return PIE_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_problem(s);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return PIE_OK;
}

int pie_partition(pie_solver* s, const int32_t* mask, int h, int w,
                  int32_t* ids) {
  if (!s || !mask || !ids) return fail(PIE_EINVAL, "NULL argument");
  if (s->kind != PIE_EQU) return fail(PIE_EINVAL, "partition is an EquSolver call");
  if (h < 1 || w < 1) return fail(PIE_EINVAL, "empty mask");
  // fpie/core/gcc/equ.cc:14-32: row-major ids from 1 over mask > 0
  s->part_ids.assign((size_t)h * w, 0);
  int cnt = 0;
  for (long long i = 0; i < (long long)h * w; ++i)
    s->part_ids[i] = mask[i] > 0 ? ++cnt : 0;
  s->ph = h;
  s->pw = w;
  s->part_count = cnt;
  std::memcpy(ids, s->part_ids.data(), sizeof(int32_t) * (size_t)h * w);
  return PIE_OK;
}

int pie_reset_equ(pie_solver* s, int n, const int32_t* A, const float* X,
                  const float* B) {
  if (!s || !A || !X || !B) return fail(PIE_EINVAL, "NULL argument");
  if (s->kind != PIE_EQU) return fail(PIE_EINVAL, "reset_equ on a GridSolver");
  if (n < 1) return fail(PIE_EINVAL, "N must be >= 1");
  for (long long k = 0; k < 4LL * n; ++k)
    if (A[k] < 0 || A[k] >= n)
      return fail(PIE_EINVAL, "A[" + std::to_string(k / 4) + "][" +
                                  std::to_string(k % 4) + "] = " +
                                  std::to_string(A[k]) + " outside [0, N)");
  PIE_HIP(hipStreamSynchronize(s->stream));
  free_problem(s);
  s->N = n;
  for (int c = 0; c < 3; ++c) s->x0[c] = X[c];

  // Does A describe the 4-neighbourhood of the partitioned mask?
  std::vector<int32_t> lin;
  bool structured = s->part_count == n - 1 && s->ph >= 3 && s->pw >= 3;
  if (structured) {
    const int h = s->ph, w = s->pw, P = pie::host::grid_layout(h, w).P;
    const int32_t* ids = s->part_ids.data();
    lin.assign(n, 0);
    for (int i = 0; i < h && structured; ++i)
      for (int j = 0; j < w; ++j) {
        const int id = ids[(long long)i * w + j];
        if (!id) continue;
        if (i == 0 || j == 0 || i == h - 1 || j == w - 1) {
          structured = false;
          break;
        }
        const int32_t* a = A + 4LL * id;
        if (a[0] != ids[(long long)(i - 1) * w + j] ||
            a[1] != ids[(long long)(i + 1) * w + j] ||
            a[2] != ids[(long long)i * w + j - 1] ||
            a[3] != ids[(long long)i * w + j + 1]) {
          structured = false;
          break;
        }
        lin[id] = i * P + j;
      }
  }

  if (structured) {
    PIE_CHECK(alloc_grid(s, s->ph, s->pw));
    std::vector<uint8_t> mh((size_t)s->mbytes, 0);
    for (int i = 0; i < s->H; ++i)
      for (int j = 0; j < s->W; ++j)
        mh[(size_t)i * s->P + pie::kPadLeft + j] =
            s->part_ids[(size_t)i * s->W + j] ? 1 : 0;
    DevBuf<float> Xd, Bd;
    PIE_HIP(Xd.alloc(3 * (size_t)n));
    PIE_HIP(Bd.alloc(3 * (size_t)n));
    PIE_HIP(s->lin.alloc(n));
    PIE_CHECK(upload(mh.data(), mh.size(), s->m.p, s->stream));
    PIE_CHECK(upload(X, 12 * (size_t)n, Xd.p, s->stream));
    PIE_CHECK(upload(B, 12 * (size_t)n, Bd.p, s->stream));
    PIE_CHECK(upload(lin.data(), 4 * (size_t)n, s->lin.p, s->stream));
    // unsolved pixels hold X[0]: the value the reference reads via id 0
    PIE_HIP(pie::launch_fill3(s->xa.p, s->bufsz, s->P, s->plane, X[0], X[1],
                              X[2], s->stream));
    PIE_HIP(pie::launch_equ_scatter(Xd.p, Bd.p, s->lin.p, n, s->xa0(),
                                    s->g0(), s->plane, s->stream));
    PIE_HIP(hipMemcpyAsync(s->xb.p, s->xa.p, s->bufsz * sizeof(float),
                           hipMemcpyDeviceToDevice, s->stream));
    PIE_CHECK(box_guidance(s));
    s->masked = n - 1;
    PIE_CHECK(default_tile(s));
  } else {
    if (s->sched != PIE_JACOBI)
      return fail(PIE_EUNSUPPORTED,
                  "gs_lex and redblack need the system built from partition()");
    PIE_HIP(s->A4.alloc(n));
    PIE_HIP(s->B4.alloc(n));
    PIE_HIP(s->X4a.alloc(n));
    PIE_HIP(s->X4b.alloc(n));
    DevBuf<float> Xd, Bd;
    PIE_HIP(Xd.alloc(3 * (size_t)n));
    PIE_HIP(Bd.alloc(3 * (size_t)n));
    PIE_CHECK(upload(A, 16 * (size_t)n, s->A4.p, s->stream));
    PIE_CHECK(upload(X, 12 * (size_t)n, Xd.p, s->stream));
    PIE_CHECK(upload(B, 12 * (size_t)n, Bd.p, s->stream));
    PIE_HIP(pie::launch_equ_pack(Xd.p, s->X4a.p, n, s->stream));
    PIE_HIP(pie::launch_equ_pack(Bd.p, s->B4.p, n, s->stream));
    PIE_HIP(hipMemcpyAsync(s->X4b.p, s->X4a.p, (size_t)n * sizeof(float4),
                           hipMemcpyDeviceToDevice, s->stream));
    const int nb = pie::equ_generic_blocks(n);
    PIE_HIP(s->partial.alloc(3 * (size_t)nb));
    PIE_HIP(s->err.alloc(3));
    PIE_HIP(hipStreamSynchronize(s->stream));
    s->masked = n - 1;
  }
  s->loaded = true;
  return PIE_OK;
}

namespace {

// Allocate the grid of a GridSolver problem and load its mask (staged
// through `stage`, at least h * w * 4 bytes); rejects solved pixels on the
// outermost rows/columns.  *masked gets the solved-pixel count.
int load_grid_mask(pie_solver* s, int h, int w, const int32_t* mask,
                   DevBuf<char>& stage, long long* masked) {
  PIE_CHECK(alloc_grid(s, h, w));
  const size_t px = (size_t)h * w;
  DevBuf<unsigned long long> counts;
  PIE_HIP(counts.alloc(2));
  PIE_HIP(hipMemsetAsync(counts.p, 0, 2 * sizeof(unsigned long long), s->stream));
  PIE_CHECK(upload(mask, px * 4, stage.p, s->stream));
  PIE_HIP(pie::launch_mask(reinterpret_cast<const int32_t*>(stage.p),
                           s->m0(), h, w, s->P, counts.p, s->stream));
  unsigned long long cnt[2] = {0, 0};
  PIE_HIP(hipMemcpyAsync(cnt, counts.p, sizeof(cnt), hipMemcpyDeviceToHost,
                         s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  if (cnt[1]) {
    free_problem(s);
    return fail(PIE_EINVAL,
                "mask has solved pixels on the outermost rows/columns");
  }
  *masked = (long long)cnt[0];
  return PIE_OK;
}

}  // namespace

int pie_reset_grid(pie_solver* s, int h, int w, const int32_t* mask,
                   const float* tgt, const float* grad) {
  if (!s || !mask || !tgt || !grad) return fail(PIE_EINVAL, "NULL argument");
  if (s->kind != PIE_GRID) return fail(PIE_EINVAL, "reset_grid on an EquSolver");
  if (h < 1 || w < 1) return fail(PIE_EINVAL, "empty grid");
  PIE_HIP(hipStreamSynchronize(s->stream));
  free_problem(s);
  const size_t px = (size_t)h * w;
  DevBuf<char> stage;
  PIE_HIP(stage.alloc(px * 12));
  long long masked = 0;
  PIE_CHECK(load_grid_mask(s, h, w, mask, stage, &masked));
  PIE_CHECK(upload(tgt, px * 12, stage.p, s->stream));
  PIE_HIP(pie::launch_deinterleave(reinterpret_cast<const float*>(stage.p),
                                   s->xa0(), s->plane, h, w, s->P, s->stream));
  PIE_HIP(hipMemcpyAsync(s->xb.p, s->xa.p, s->bufsz * sizeof(float),
                         hipMemcpyDeviceToDevice, s->stream));
  PIE_CHECK(upload(grad, px * 12, stage.p, s->stream));
  PIE_HIP(pie::launch_deinterleave(reinterpret_cast<const float*>(stage.p),
                                   s->g0(), s->plane, h, w, s->P, s->stream));
  PIE_CHECK(box_guidance(s));
  s->masked = masked;
  s->N = h * w;
  PIE_CHECK(default_tile(s));
  s->loaded = true;
  return PIE_OK;
}

int pie_reset_grid_images(pie_solver* s, int h, int w, const int32_t* mask,
                          const uint8_t* src, const uint8_t* tgt,
                          int gradient) {
  if (!s || !mask || !src || !tgt) return fail(PIE_EINVAL, "NULL argument");
  if (s->kind != PIE_GRID) return fail(PIE_EINVAL, "reset_grid on an EquSolver");
  if (h < 1 || w < 1) return fail(PIE_EINVAL, "empty grid");
  if (gradient < PIE_GRAD_SRC || gradient > PIE_GRAD_MAX)
    return fail(PIE_EINVAL, "gradient must be PIE_GRAD_SRC, _AVG or _MAX");
  PIE_HIP(hipStreamSynchronize(s->stream));
  free_problem(s);
  const size_t px = (size_t)h * w;
  DevBuf<char> stage;  // the int32 mask, then both uint8 images
  PIE_HIP(stage.alloc(px * 6 > px * 4 ? px * 6 : px * 4));
  long long masked = 0;
  PIE_CHECK(load_grid_mask(s, h, w, mask, stage, &masked));
  uint8_t* su = reinterpret_cast<uint8_t*>(stage.p);
  PIE_CHECK(upload(src, px * 3, su, s->stream));
  PIE_CHECK(upload(tgt, px * 3, su + px * 3, s->stream));
  PIE_HIP(pie::launch_divergence(su, su + px * 3, s->m0(), s->xa0(), s->g0(),
                                 s->plane, h, w, s->P, gradient, 0,
                                 s->stream));
  PIE_HIP(hipMemcpyAsync(s->xb.p, s->xa.p, s->bufsz * sizeof(float),
                         hipMemcpyDeviceToDevice, s->stream));
  PIE_CHECK(box_guidance(s));
  s->masked = masked;
  s->N = h * w;
  PIE_CHECK(default_tile(s));
  s->loaded = true;
  return PIE_OK;
}

int pie_reset_equ_images(pie_solver* s, int h, int w, const int32_t* mask,
                         const uint8_t* src, const uint8_t* tgt,
                         int gradient) {
  if (!s || !mask || !src || !tgt) return fail(PIE_EINVAL, "NULL argument");
  if (s->kind != PIE_EQU) return fail(PIE_EINVAL, "reset_equ on a GridSolver");
  if (h < 3 || w < 3) return fail(PIE_EINVAL, "mask crop smaller than 3 x 3");
  if (gradient < PIE_GRAD_SRC || gradient > PIE_GRAD_MAX)
    return fail(PIE_EINVAL, "gradient must be PIE_GRAD_SRC, _AVG or _MAX");
  PIE_HIP(hipStreamSynchronize(s->stream));
  free_problem(s);
  const size_t px = (size_t)h * w;
  // fpie/core/gcc/equ.cc:14-32: row-major ids from 1 over mask > 0, as
  // pie_partition; the system is then held on the grid (id -> i P + j)
  s->part_ids.assign(px, 0);
  int cnt = 0;
  for (size_t k = 0; k < px; ++k) s->part_ids[k] = mask[k] > 0 ? ++cnt : 0;
  s->ph = h;
  s->pw = w;
  s->part_count = cnt;
  DevBuf<char> stage;  // the int32 mask, then both uint8 images
  PIE_HIP(stage.alloc(px * 6 > px * 4 ? px * 6 : px * 4));
  long long masked = 0;
  PIE_CHECK(load_grid_mask(s, h, w, mask, stage, &masked));
  const int n = cnt + 1;
  std::vector<int32_t> lin(n, 0);
  for (int i = 0; i < h; ++i)
    for (int j = 0; j < w; ++j) {
      const int id = s->part_ids[(size_t)i * w + j];
      if (id) lin[id] = i * s->P + j;
    }
  PIE_HIP(s->lin.alloc(n));
  PIE_CHECK(upload(lin.data(), 4 * (size_t)n, s->lin.p, s->stream));
  uint8_t* su = reinterpret_cast<uint8_t*>(stage.p);
  PIE_CHECK(upload(src, px * 3, su, s->stream));
  PIE_CHECK(upload(tgt, px * 3, su + px * 3, s->stream));
  PIE_HIP(s->tgt8.alloc(px * 3));
  PIE_HIP(hipMemcpyAsync(s->tgt8.p, su + px * 3, px * 3,
                         hipMemcpyDeviceToDevice, s->stream));
  // fpie/process.py:227-270: X[0] = 0, X[id] = tgt, B = grad + boundary
  PIE_HIP(pie::launch_fill3(s->xa.p, s->bufsz, s->P, s->plane, 0.0f, 0.0f,
                            0.0f, s->stream));
  PIE_HIP(pie::launch_divergence(su, su + px * 3, s->m0(), s->xa0(), s->g0(),
                                 s->plane, h, w, s->P, gradient, 1,
                                 s->stream));
  PIE_HIP(hipMemcpyAsync(s->xb.p, s->xa.p, s->bufsz * sizeof(float),
                         hipMemcpyDeviceToDevice, s->stream));
  PIE_CHECK(box_guidance(s));
  for (int c = 0; c < 3; ++c) s->x0[c] = 0.0f;
  s->N = n;
  s->masked = masked;
  PIE_CHECK(default_tile(s));
  s->loaded = true;
  return PIE_OK;
}

int pie_sync(pie_solver* s) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  PIE_HIP(hipStreamSynchronize(s->stream));
  return PIE_OK;
}

int pie_iterate(pie_solver* s, int iterations) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (!s->loaded) return fail(PIE_ESTATE, "step before reset");
  if (iterations < 0) return fail(PIE_EINVAL, "negative iteration count");
  // autotune once per problem: at once from 2^22 pixels on, and on smaller
  // grids for the first call long enough to repay it (2^29 pixel sweeps)
  const long long px = (long long)s->H * s->W;
  if (s->grid_layout && !s->tuned && !s->tile_fixed &&
      (px >= (1LL << 22) || (long long)iterations * px >= (1LL << 29)))
    PIE_CHECK(autotune(s, 6, iterations));
  return run_sweeps(s, iterations);
}

int pie_read(pie_solver* s, uint8_t* img, float* err) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  return read_outputs(s, img, err, nullptr);
}

int pie_read_crop(pie_solver* s, uint8_t* crop, float* err) {
  if (!s || !crop) return fail(PIE_EINVAL, "NULL argument");
  if (!s->loaded) return fail(PIE_ESTATE, "step before reset");
  if (s->kind == PIE_GRID) return read_outputs(s, crop, err, nullptr);
  if (!s->tgt8.p)
    return fail(PIE_ESTATE, "the crop output needs pie_reset_equ_images");
  // the clamped state on the grid (fpie/core/gcc/equ.cc:82-85), the target
  // crop at unsolved pixels, and the residual as pie_read
  const int nb = pie::finalize_blocks(s->H, s->W);
  const size_t px = (size_t)s->H * s->W;
  if (!s->img.p || s->img.n < px * 3) {
    s->img.release();
    PIE_HIP(s->img.alloc(px * 3));
  }
  PIE_HIP(pie::launch_finalize(s->eff_order(), s->state(), s->g0(), s->m0(),
                               s->plane, s->H, s->W, s->P, s->img.p,
                               s->partial.p, nb, s->stream));
  PIE_HIP(pie::launch_reduce(s->partial.p, nb, s->err.p, s->stream));
  PIE_HIP(pie::launch_keep_unsolved(s->img.p, s->tgt8.p, s->m0(), s->H, s->W,
                                    s->P, s->stream));
  PIE_HIP(hipMemcpyAsync(crop, s->img.p, px * 3, hipMemcpyDeviceToHost,
                         s->stream));
  float e[3];
  PIE_HIP(hipMemcpyAsync(e, s->err.p, sizeof(e), hipMemcpyDeviceToHost,
                         s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  if (err) std::memcpy(err, e, sizeof(e));
  return PIE_OK;
}

int pie_step(pie_solver* s, int iterations, uint8_t* img, float* err) {
  PIE_CHECK(pie_iterate(s, iterations));
  return pie_read(s, img, err);
}

int pie_get_state(pie_solver* s, float* out) {
  if (!s || !out) return fail(PIE_EINVAL, "NULL argument");
  return read_outputs(s, nullptr, nullptr, out);
}

int pie_set_state(pie_solver* s, const float* x) {
  if (!s || !x) return fail(PIE_EINVAL, "NULL argument");
  if (!s->loaded) return fail(PIE_ESTATE, "set_state before reset");
  if (s->kind != PIE_GRID)
    return fail(PIE_EUNSUPPORTED, "set_state is a GridSolver call");
  const size_t px = (size_t)s->H * s->W;
  DevBuf<float> stage;
  PIE_HIP(stage.alloc(px * 3));
  PIE_CHECK(upload(x, px * 12, stage.p, s->stream));
  // both buffers: padding and unsolved pixels keep their values
  PIE_HIP(pie::launch_deinterleave(stage.p, s->xa0(), s->plane, s->H, s->W,
                                   s->P, s->stream));
  PIE_HIP(hipMemcpyAsync(s->xb.p, s->xa.p, s->bufsz * sizeof(float),
                         hipMemcpyDeviceToDevice, s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  s->cur = 0;
  return PIE_OK;
}

int pie_solve_to_residual(pie_solver* s, double tol_mean, int check_every,
                          long long max_sweeps, long long* sweeps,
                          double* mean_residual) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (!s->loaded) return fail(PIE_ESTATE, "solve before reset");
  if (check_every < 1 || max_sweeps < 0)
    return fail(PIE_EINVAL, "check_every >= 1 and max_sweeps >= 0");
  const double denom = 3.0 * (double)std::max(1LL, s->masked);
  long long done = 0;
  float e[3];
  PIE_CHECK(read_outputs(s, nullptr, e, nullptr));
  double mean = ((double)e[0] + e[1] + e[2]) / denom;
  while (mean > tol_mean && done < max_sweeps) {
    const int k = (int)std::min<long long>(check_every, max_sweeps - done);
    PIE_CHECK(pie_iterate(s, k));
    done += k;
    PIE_CHECK(read_outputs(s, nullptr, e, nullptr));
    mean = ((double)e[0] + e[1] + e[2]) / denom;
  }
  if (sweeps) *sweeps = done;
  if (mean_residual) *mean_residual = mean;
  return PIE_OK;
}

namespace {

// Checks shared by the coarse-grid correction calls; *hc, *wc = the coarse
// grid (crop size // factor).
int coarse_grid(pie_solver* s, int factor, int* hc, int* wc) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (!s->loaded) return fail(PIE_ESTATE, "coarse-grid call before reset");
  if (s->kind != PIE_GRID)
    return fail(PIE_EUNSUPPORTED, "coarse-grid calls are GridSolver calls");
  if (factor < 1) return fail(PIE_EINVAL, "factor must be >= 1");
  *hc = s->H / factor;
  *wc = s->W / factor;
  if (*hc < 1 || *wc < 1)
    return fail(PIE_EINVAL, "factor exceeds the crop size");
  const size_t n = (size_t)*hc * *wc * 3;
  if (s->coarse.n < n) PIE_HIP(s->coarse.alloc(n));
  return PIE_OK;
}

}  // namespace

int pie_restrict(pie_solver* s, int factor, float* out) {
  int hc = 0, wc = 0;
  if (!out) return fail(PIE_EINVAL, "NULL argument");
  PIE_CHECK(coarse_grid(s, factor, &hc, &wc));
  PIE_HIP(pie::launch_restrict(s->state(), s->g0(), s->m0(), s->plane, s->P,
                               factor, hc, wc, s->coarse.p, s->stream));
  PIE_HIP(hipMemcpyAsync(out, s->coarse.p, (size_t)hc * wc * 3 * sizeof(float),
                         hipMemcpyDeviceToHost, s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  return PIE_OK;
}

int pie_prolong(pie_solver* s, int factor, int hc, int wc, const float* corr) {
  int ec = 0, ew = 0;
  if (!corr) return fail(PIE_EINVAL, "NULL argument");
  PIE_CHECK(coarse_grid(s, factor, &ec, &ew));
  if (hc != ec || wc != ew)
    return fail(PIE_EINVAL, "coarse grid must be (h // factor, w // factor)");
  PIE_CHECK(upload(corr, (size_t)hc * wc * 3 * sizeof(float), s->coarse.p,
                   s->stream));
  PIE_HIP(pie::launch_prolong(s->state(), s->m0(), s->plane, s->H, s->W, s->P,
                              factor, hc, wc, s->coarse.p, s->stream));
  // the upload reads pageable host memory: finish before returning
  PIE_HIP(hipStreamSynchronize(s->stream));
  return PIE_OK;
}

namespace {

// pie_prolong_dots of the correction in s->coarse
int prolong_dots_device(pie_solver* s, int factor, int hc, int wc,
                        double* out) {
  const int nb = pie::prolong_dots_blocks(s->H, s->W);
  DevBuf<double> part;
  PIE_HIP(part.alloc((size_t)nb * 6));
  PIE_HIP(pie::launch_prolong_dots(s->state(), s->g0(), s->m0(), s->plane,
                                   s->H, s->W, s->P, factor, hc, wc,
                                   s->coarse.p, part.p, nb, s->stream));
  std::vector<double> h((size_t)nb * 6);
  PIE_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double),
                         hipMemcpyDeviceToHost, s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  for (int k = 0; k < 6; ++k) out[k] = 0.0;
  for (int b = 0; b < nb; ++b)  // fixed order: deterministic
    for (int k = 0; k < 6; ++k) out[k] += h[(size_t)b * 6 + k];
  return PIE_OK;
}

// Per-channel inner products of up to kMaxDots pairs of coarse fields into
// the device scalars sc[kWsDots..] (blocks added in order: deterministic).
int coarse_dots(pie_solver* s, int n, std::initializer_list<
                std::pair<const double*, const double*>> pairs) {
  pie::CoarseDots d{};
  d.count = 0;
  for (const auto& pr : pairs) {
    d.a[d.count] = pr.first;
    d.b[d.count] = pr.second;
    ++d.count;
  }
  const int nb = pie::coarse_dots_blocks(n);
  PIE_HIP(pie::launch_coarse_dots(d, n, s->wpart.p, s->stream));
  PIE_HIP(pie::launch_sum_partials(s->wpart.p, nb, pie::kMaxDots * 3,
                                   s->wsc.p + pie::kWsDots, s->stream));
  return PIE_OK;
}

}  // namespace

int pie_prolong_dots(pie_solver* s, int factor, int hc, int wc,
                     const float* corr, double* out) {
  int ec = 0, ew = 0;
  if (!corr || !out) return fail(PIE_EINVAL, "NULL argument");
  PIE_CHECK(coarse_grid(s, factor, &ec, &ew));
  if (hc != ec || wc != ew)
    return fail(PIE_EINVAL, "coarse grid must be (h // factor, w // factor)");
  s->warm_factor = 0;
  PIE_CHECK(upload(corr, (size_t)hc * wc * 3 * sizeof(float), s->coarse.p,
                   s->stream));
  return prolong_dots_device(s, factor, hc, wc, out);
}

int pie_warm_net_start(pie_solver* s, pie_unet* net, int factor,
                       const uint8_t* mc, int net_steps, int lowpass_sweeps,
                       float* corr_out) {
  constexpr int kMaxSteps = pie::kWarmMaxSteps;
  int hc = 0, wc = 0;
  if (!net || !mc) return fail(PIE_EINVAL, "NULL argument");
  if (net_steps < 0 || net_steps > kMaxSteps)
    return fail(PIE_EINVAL, "net_steps must be in [0, 8]");
  if (lowpass_sweeps < 0) return fail(PIE_EINVAL, "negative lowpass_sweeps");
  PIE_CHECK(coarse_grid(s, factor, &hc, &wc));
  const int n = hc * wc;
  const int hp = (hc + 7) / 8 * 8, wp = (wc + 7) / 8 * 8;
  const size_t fs = (size_t)n * 3;  // doubles per field
  // fields: r, c, t, then p_k and A_c p_k for each step
  const size_t nfields = 3 + 2 * (size_t)kMaxSteps;
  if (s->wf.n < nfields * fs) PIE_HIP(s->wf.alloc(nfields * fs));
  if (s->wmc.n < (size_t)n) PIE_HIP(s->wmc.alloc(n));
  // partials of the coarse dots here and of the line search in
  // pie_warm_apply (sized for both: no reallocation between the two)
  const size_t np = std::max(
      (size_t)pie::coarse_dots_blocks(n) * pie::kMaxDots * 3,
      (size_t)pie::prolong_dots_blocks(s->H, s->W) * 6);
  if (s->wpart.n < np) PIE_HIP(s->wpart.alloc(np));
  if (s->wsc.n < (size_t)pie::kWsCount) PIE_HIP(s->wsc.alloc(pie::kWsCount));
  if (s->wx.n < (size_t)hp * wp * 8) PIE_HIP(s->wx.alloc((size_t)hp * wp * 8));
  if (s->wout.n < (size_t)hp * wp * 3) PIE_HIP(s->wout.alloc((size_t)hp * wp * 3));
  double* r = s->wf.p;
  double* c = r + fs;
  double* t = c + fs;
  auto P = [&](int k) { return t + fs * (1 + 2 * (size_t)k); };
  auto AP = [&](int k) { return t + fs * (2 + 2 * (size_t)k); };
  double* sc = s->wsc.p;
  int cells = 0;
  for (int k = 0; k < n; ++k) cells += mc[k] != 0;

  s->warm_factor = 0;
  // mc is pageable host memory: copy it in stream order (after any work
  // still queued that reads the previous mask) and wait for the copy
  PIE_CHECK(upload(mc, (size_t)n, s->wmc.p, s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  // r_c: the restricted residual on the solved cells
  PIE_HIP(pie::launch_restrict(s->state(), s->g0(), s->m0(), s->plane, s->P,
                               factor, hc, wc, s->coarse.p, s->stream));
  PIE_HIP(pie::launch_coarse_load(s->coarse.p, s->wmc.p, n, r, c, s->stream));
  // every scalar of the steps stays on the device: no host round trip
  for (int k = 0; k < net_steps; ++k) {
    double* z = P(k);
    double* az = AP(k);
    // network direction for the current residual (zero if r is zero)
    PIE_CHECK(coarse_dots(s, n, {{r, r}}));
    PIE_HIP(pie::launch_warm_scalar(sc, pie::kWarmSig, k, (double)cells,
                                    s->stream));
    PIE_HIP(pie::launch_net_input(r, s->wmc.p, hc, wc, hp, wp,
                                  sc + pie::kWsSig, s->wx.p, s->stream));
    PIE_CHECK(pie::unet_forward_on_stream(net, hp, wp, s->wx.p, s->wout.p,
                                          s->stream));
    PIE_HIP(pie::launch_net_output(s->wout.p, s->wmc.p, hc, wc, wp,
                                   sc + pie::kWsAny, z, s->stream));
    // low-pass: damped homogeneous sweeps (z -> t -> z)
    for (int q = 0; q < lowpass_sweeps; ++q) {
      PIE_HIP(pie::launch_lowpass(z, s->wmc.p, hc, wc, t, s->stream));
      PIE_HIP(hipMemcpyAsync(z, t, fs * sizeof(double),
                             hipMemcpyDeviceToDevice, s->stream));
    }
    // A_c-orthogonal to the earlier directions, one after the other
    for (int j = 0; j < k; ++j) {
      PIE_CHECK(coarse_dots(s, n, {{z, AP(j)}}));
      PIE_HIP(pie::launch_warm_scalar(sc, pie::kWarmOrth, j, 0.0, s->stream));
      PIE_HIP(pie::launch_coarse_axpy(z, P(j), sc + pie::kWsCoef, n,
                                      s->stream));
    }
    PIE_HIP(pie::launch_coarse_apply(z, s->wmc.p, hc, wc, az, s->stream));
    // <z, A z>, <z, z>, <r, z>: the energy-optimal step along z
    PIE_CHECK(coarse_dots(s, n, {{z, az}, {z, z}, {r, z}}));
    PIE_HIP(pie::launch_warm_scalar(sc, pie::kWarmStep, k, 0.0, s->stream));
    PIE_HIP(pie::launch_coarse_axpy(c, z, sc + pie::kWsStep, n, s->stream));
    PIE_HIP(pie::launch_coarse_axpy(r, az, sc + pie::kWsNegStep, n,
                                    s->stream));
  }
  PIE_HIP(pie::launch_coarse_to_float(c, (int)fs, s->coarse.p, s->stream));
  if (corr_out) {
    PIE_HIP(hipMemcpyAsync(corr_out, s->coarse.p, fs * sizeof(float),
                           hipMemcpyDeviceToHost, s->stream));
    PIE_HIP(hipStreamSynchronize(s->stream));
  }
  s->warm_factor = factor;
  return PIE_OK;
}

int pie_warm_apply(pie_solver* s, int factor, const float* corr,
                   double* alpha_out, double* max_corr) {
  int hc = 0, wc = 0;
  PIE_CHECK(coarse_grid(s, factor, &hc, &wc));
  const size_t fs = (size_t)hc * wc * 3;
  if (corr) {
    PIE_CHECK(upload(corr, fs * sizeof(float), s->coarse.p, s->stream));
  } else if (s->warm_factor != factor) {
    return fail(PIE_ESTATE, "no device correction of this factor: call "
                            "pie_warm_net_start first or pass corr");
  }
  s->warm_factor = 0;
  // (the buffers are already large enough after pie_warm_net_start; with a
  // host correction they may be allocated here, nothing queued uses them)
  if (s->wsc.n < (size_t)pie::kWsCount) PIE_HIP(s->wsc.alloc(pie::kWsCount));
  if (!s->wmax.p) PIE_HIP(s->wmax.alloc(1));
  double* sc = s->wsc.p;
  // line search on the device: alpha_c = max(0, <r, d> / <d, A d>)
  const int nb = pie::prolong_dots_blocks(s->H, s->W);
  if (s->wpart.n < (size_t)nb * 6) PIE_HIP(s->wpart.alloc((size_t)nb * 6));
  PIE_HIP(pie::launch_prolong_dots(s->state(), s->g0(), s->m0(), s->plane,
                                   s->H, s->W, s->P, factor, hc, wc,
                                   s->coarse.p, s->wpart.p, nb, s->stream));
  PIE_HIP(pie::launch_sum_partials(s->wpart.p, nb, 6, sc + pie::kWsDots,
                                   s->stream));
  PIE_HIP(pie::launch_warm_scalar(sc, pie::kWarmAlpha, 0, 0.0, s->stream));
  PIE_HIP(hipMemsetAsync(s->wmax.p, 0, sizeof(unsigned), s->stream));
  PIE_HIP(pie::launch_coarse_scale(s->coarse.p, sc + pie::kWsAlpha, hc * wc,
                                   s->wmax.p, s->stream));
  PIE_HIP(pie::launch_prolong(s->state(), s->m0(), s->plane, s->H, s->W, s->P,
                              factor, hc, wc, s->coarse.p, s->stream));
  double alpha[3];
  unsigned mb = 0;
  PIE_HIP(hipMemcpyAsync(alpha, sc + pie::kWsAlpha, sizeof(alpha),
                         hipMemcpyDeviceToHost, s->stream));
  PIE_HIP(hipMemcpyAsync(&mb, s->wmax.p, sizeof(mb), hipMemcpyDeviceToHost,
                         s->stream));
  PIE_HIP(hipStreamSynchronize(s->stream));
  if (alpha_out)
    for (int ch = 0; ch < 3; ++ch) alpha_out[ch] = alpha[ch];
  if (max_corr) {
    float m;
    std::memcpy(&m, &mb, sizeof(m));
    *max_corr = m;
  }
  return PIE_OK;
}

int pie_relax(pie_solver* s, int sweeps, float omega) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (!s->loaded) return fail(PIE_ESTATE, "relax before reset");
  if (s->kind != PIE_GRID)
    return fail(PIE_EUNSUPPORTED, "relax is a GridSolver call");
  if (sweeps < 0) return fail(PIE_EINVAL, "negative sweep count");
  if (!(omega > 0.0f && omega <= 1.0f))
    return fail(PIE_EINVAL, "omega must be in (0, 1]");
  for (int k = 0; k < sweeps; ++k) {
    PIE_HIP(pie::launch_weighted(s->eff_order(), s->state(), s->other(),
                                 s->g0(), s->m0(), s->plane, s->H, s->W, s->P,
                                 omega, s->stream));
    s->cur ^= 1;
  }
  return PIE_OK;
}

int pie_set_tile(pie_solver* s, int sweeps, int cols_per_lane, int seg_rows) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  TileCfg c = s->cfg;
  if (sweeps > 0) c.S = sweeps;
  if (cols_per_lane > 0) c.C = cols_per_lane;
  if (seg_rows > 0) c.seg_rows = seg_rows;
  c.N = 0;  // the column count's default; pie_set_tile_channels overrides
  if (!pie::tile_supported(c))
    return fail(PIE_EUNSUPPORTED, "tile (sweeps " + std::to_string(c.S) +
                                      ", columns/lane " + std::to_string(c.C) +
                                      ") not compiled");
  s->cfg = c;
  s->tile_fixed = true;
  return PIE_OK;
}

int pie_set_tile_channels(pie_solver* s, int channels) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (channels != 0 && channels != 1 && channels != 3)
    return fail(PIE_EINVAL, "channels per task must be 0, 1 or 3");
  TileCfg c = s->cfg;
  c.N = channels;
  if (!pie::tile_supported(c))
    return fail(PIE_EUNSUPPORTED,
                "tile (sweeps " + std::to_string(c.S) + ", columns/lane " +
                    std::to_string(c.C) + ", channels " +
                    std::to_string(pie::tile_nch(c)) + ") not compiled");
  s->cfg = c;
  s->tile_fixed = true;
  return PIE_OK;
}

int pie_get_tile_channels(pie_solver* s, int* channels) {
  if (!s || !channels) return fail(PIE_EINVAL, "NULL argument");
  *channels = pie::tile_nch(s->cfg);
  return PIE_OK;
}

int pie_get_tile(pie_solver* s, int* sweeps, int* cols_per_lane,
                 int* seg_rows) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (sweeps) *sweeps = s->cfg.S;
  if (cols_per_lane) *cols_per_lane = s->cfg.C;
  if (seg_rows) *seg_rows = s->cfg.seg_rows;
  return PIE_OK;
}

int pie_autotune(pie_solver* s, int launches) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (!s->loaded) return fail(PIE_ESTATE, "autotune before reset");
  if (launches < 1) launches = 1;
  s->tile_fixed = false;
  PIE_CHECK(autotune(s, launches, 5000));
  return PIE_OK;
}

int pie_set_timing(pie_solver* s, int enable) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  s->timing = enable != 0;
  s->t_ms = 0.0;
  s->t_updates = 0.0;
  s->t_launches = 0;
  return PIE_OK;
}

int pie_get_timing(pie_solver* s, double* kernel_ms, long long* launches,
                   double* pixel_updates) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (kernel_ms) *kernel_ms = s->t_ms;
  if (launches) *launches = s->t_launches;
  if (pixel_updates) *pixel_updates = s->t_updates;
  return PIE_OK;
}

int pie_shape(pie_solver* s, int* h, int* w, int* n, long long* masked) {
  if (!s) return fail(PIE_EINVAL, "NULL solver");
  if (h) *h = s->H;
  if (w) *w = s->W;
  if (n) *n = s->N;
  if (masked) *masked = s->masked;
  return PIE_OK;
}

int pie_prepare_mask(const uint8_t* mask, int h, int w, int channels,
                     int* box, int32_t* crop, long long crop_capacity) {
  if (!mask || !box) return fail(PIE_EINVAL, "NULL argument");
  if (h < 1 || w < 1) return fail(PIE_EINVAL, "empty mask");
  if (channels < 1 || channels > 4)
    return fail(PIE_EINVAL, "channels must be 1 to 4");
  const size_t px = (size_t)h * w;
  hipStream_t st = nullptr;
  PIE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{st};
  DevBuf<uint8_t> dmask, dm;
  DevBuf<unsigned> rc;
  PIE_HIP(dmask.alloc(px * channels));
  PIE_HIP(dm.alloc(px));
  PIE_HIP(rc.alloc((size_t)h + w));
  PIE_CHECK(upload(mask, px * channels, dmask.p, st));
  PIE_HIP(hipMemsetAsync(rc.p, 0, sizeof(unsigned) * ((size_t)h + w), st));
  PIE_HIP(pie::launch_prep_mask(dmask.p, h, w, channels, dm.p, rc.p, rc.p + h,
                                st));
  std::vector<unsigned> any((size_t)h + w);
  PIE_HIP(hipMemcpyAsync(any.data(), rc.p, sizeof(unsigned) * any.size(),
                         hipMemcpyDeviceToHost, st));
  PIE_HIP(hipStreamSynchronize(st));
  int r_lo = -1, r_hi = -1, c_lo = -1, c_hi = -1;
  for (int i = 0; i < h; ++i)
    if (any[i]) {
      if (r_lo < 0) r_lo = i;
      r_hi = i;
    }
  for (int j = 0; j < w; ++j)
    if (any[(size_t)h + j]) {
      if (c_lo < 0) c_lo = j;
      c_hi = j;
    }
  if (r_lo < 0)
    return fail(PIE_EINVAL, "mask selects no pixel inside the image edge");
  // fpie/process.py:219-221: the bounding box plus a 1-pixel margin
  box[0] = r_lo - 1;
  box[1] = r_hi + 2;
  box[2] = c_lo - 1;
  box[3] = c_hi + 2;
  if (crop) {
    const int ch = box[1] - box[0], cw = box[3] - box[2];
    if ((long long)ch * cw > crop_capacity)
      return fail(PIE_EINVAL, "crop buffer smaller than the " +
                                  std::to_string(ch) + " x " +
                                  std::to_string(cw) + " crop");
    DevBuf<int32_t> dcrop;
    PIE_HIP(dcrop.alloc((size_t)ch * cw));
    PIE_HIP(pie::launch_crop_mask(dm.p, w, box[0], box[2], ch, cw, dcrop.p,
                                  st));
    PIE_HIP(hipMemcpyAsync(crop, dcrop.p, sizeof(int32_t) * (size_t)ch * cw,
                           hipMemcpyDeviceToHost, st));
    PIE_HIP(hipStreamSynchronize(st));
  }
  return PIE_OK;
}

int pie_device_info(char* buf, int len) {
  if (!buf || len < 1) return fail(PIE_EINVAL, "NULL buffer");
  int dev = 0;
  PIE_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  PIE_HIP(hipGetDeviceProperties(&p, dev));
  std::snprintf(buf, (size_t)len,
                "device %d: %s (%s), %d CUs, %.1f GiB HBM, %d MHz, "
                "L2 %d KiB, LDS/block %zu KiB",
                dev, p.name, p.gcnArchName, p.multiProcessorCount,
                p.totalGlobalMem / 1073741824.0, p.clockRate / 1000,
                p.l2CacheSize / 1024, p.sharedMemPerBlock / 1024);
  return PIE_OK;
}

}  // extern "C"
