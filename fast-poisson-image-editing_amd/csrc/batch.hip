// Batched GridSolver (C ABI pie_batch_*): many composites of uneven size
// and mask, each with its own stop rule, solved together.
//
// Every item keeps its own padded planes (same layout as pie_solver).  One
// round is two kernels: the persistent sweep kernel and an epilogue.  In
// the sweep kernel, each small ("solo") item -- few tasks, or a whole solve
// of at most half of one work-group's even share of the batch's work -- is
// taken by one work-group that runs launch after launch of it (S Jacobi
// sweeps each, the stop rules after each) while the shared items' tasks
// last, or to its end once no shared item is live; the waves then take tiles
// of any live shared item from one atomic work counter and advance them by
// S sweeps.  The epilogue applies the stop rules to the shared items (last
// sweep's max |update| below the item's tolerance, or its sweep budget
// used) and resets the counters.  Small items therefore progress several
// launches a round and finish early, while large ones share the machine.
// R rounds are captured once into a hipGraph and the graph is replayed
// until no item is live, so the host only looks at one counter every R
// rounds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_util.h"

using pie::BatchArgs;
using pie::BatchItem;
using pie::TileCfg;
using pie::host::DevBuf;
using pie::host::fail;
using pie::host::pitch_of;
using pie::host::upload;

namespace {

// Items of at most this many wave tasks are always solved by one
// work-group (its waves take them in four passes a launch).
constexpr int kSoloTasks = 4 * pie::kBatchGroupWaves;
// Larger items are too when their whole solve (tasks x launches of their
// sweep budget) is at most this fraction of one work-group's even share of
// the batch: they then finish well before the batch does, in order of
// size, while the large items still spread over the machine at the end.
constexpr double kSoloShare = 0.5;

struct ItemBuffers {
  DevBuf<float> xa, xb, g, x0;  // x0: initial state, for restarts
  DevBuf<uint8_t> m;
  // solved-pixel runs of each row, [begin, end) column pairs: row r's runs
  // are run[run0[r] .. run0[r + 1]) (for the task lists of any tile)
  std::vector<int> run0, run;
  int H = 0, W = 0, P = 0;
  long long plane = 0, masked = 0;
  long long bufsz = 0, mbytes = 0;
  int max_sweeps = 0;
  float tol = 0.0f;
};

}  // namespace

struct pie_batch {
  int order = PIE_ORDER_ULRD;
  TileCfg cfg{8, 1, 64};
  hipStream_t stream = nullptr;
  int device = 0;
  double wall_hz = 1e8;
  std::vector<std::unique_ptr<ItemBuffers>> items;
  std::vector<BatchItem> desc;  // descriptors as last read back
  DevBuf<BatchItem> d_items;
  DevBuf<int> d_task_item, d_solo_items;
  DevBuf<int> d_tlist;  // every item's task list
  int ntask_all = 0;    // wave tasks with work, all items
  // [0] shared-task queue, [1] live items, [2] solo-item queue,
  // [3] live shared items
  DevBuf<unsigned> d_counters;
  int nsolo = 0;
  unsigned nshared = 0;  // shared (not solo) items
  DevBuf<unsigned long long> d_clock;
  unsigned* h_live = nullptr;   // pinned
  int ntask = 0;
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_rounds = 0;
  int blocks = 0;
  unsigned long long t_start = 0;
  DevBuf<double> partial;
  DevBuf<float> err;
  DevBuf<uint8_t> img;
};

namespace {

void drop_graph(pie_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_rounds = 0;
}

// Descriptor table and task table for the current item list.
int prepare(pie_batch* b) {
  if (!b->dirty) return PIE_OK;
  const int n = (int)b->items.size();
  b->desc.assign(n, BatchItem{});
  std::vector<int> task_item;
  std::vector<std::pair<long long, int>> solo;  // (work, item)
  int per_cu = pie::batch_blocks_per_cu(b->cfg, PIE_JACOBI, b->order);
  hipDeviceProp_t prop;
  PIE_HIP(hipGetDeviceProperties(&prop, b->device));
  b->blocks = std::max(1, per_cu * prop.multiProcessorCount);
  // Task lists: the wave tasks whose written region (rows seg * R ..,
  // columns stripe * U ..) holds a solved pixel, as ids group * stripes *
  // segs + seg * stripes + stripe; the others have nothing to change
  // (both state buffers hold the initial state at unsolved pixels).  All
  // items' lists are concatenated, item i's at tlist0[i].
  const int U = pie::kBatchLaneColumns * b->cfg.C - 2 * b->cfg.S;
  const int R = b->cfg.seg_rows;
  const int groups = pie::tile_groups(b->cfg);
  std::vector<int> tlist, tlist0(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const ItemBuffers& it = *b->items[i];
    int stripes = 0, segs = 0;
    pie::tile_tasks(b->cfg, it.H, it.W, &stripes, &segs);
    std::vector<uint8_t> flag((size_t)stripes * segs, 0);
    for (int r = 0; r < it.H; ++r) {
      uint8_t* f = flag.data() + (size_t)(r / R) * stripes;
      for (int k = it.run0[r]; k < it.run0[r + 1]; k += 2)
        for (int st = it.run[k] / U; st <= (it.run[k + 1] - 1) / U; ++st)
          f[st] = 1;
    }
    const int per_group = stripes * segs;
    for (int g = 0; g < groups; ++g)
      for (int region = 0; region < per_group; ++region)
        if (flag[region]) tlist.push_back(g * per_group + region);
    tlist0[i + 1] = (int)tlist.size();
  }
  PIE_HIP(b->d_tlist.alloc(std::max<size_t>(tlist.size(), 1)));
  b->ntask_all = (int)tlist.size();
  // work of an item: wave tasks x launches of its sweep budget
  std::vector<long long> work(n);
  double total = 0.0;
  for (int i = 0; i < n; ++i) {
    const long long launches =
        (b->items[i]->max_sweeps + b->cfg.S - 1) / b->cfg.S;
    work[i] = (long long)(tlist0[i + 1] - tlist0[i]) * launches;
    total += (double)work[i];
  }
  const double solo_work = kSoloShare * total / b->blocks;
  for (int i = 0; i < n; ++i) {
    ItemBuffers& it = *b->items[i];
    BatchItem& d = b->desc[i];
    d.xa = it.xa.p + pie::kPadLeft;
    d.xb = it.xb.p + pie::kPadLeft;
    d.g = it.g.p + pie::kPadLeft;
    d.m = it.m.p + pie::kPadLeft;
    d.plane = it.plane;
    d.H = it.H;
    d.W = it.W;
    d.P = it.P;
    pie::tile_tasks(b->cfg, it.H, it.W, &d.stripes, &d.segs);
    d.task0 = (int)task_item.size();
    d.tasks = b->d_tlist.p + tlist0[i];
    d.ntask = tlist0[i + 1] - tlist0[i];
    d.t_end = it.max_sweeps;
    d.max_sweeps = it.max_sweeps;
    d.tol = it.tol;
    const int tasks = d.ntask;
    d.solo = tasks <= kSoloTasks || (double)work[i] <= solo_work;
    if (d.solo) {
      solo.push_back({work[i], i});
    } else {
      task_item.insert(task_item.end(), (size_t)tasks, i);
    }
  }
  // smallest first: the first claims of a round run longest
  std::sort(solo.begin(), solo.end());
  std::vector<int> solo_items;
  for (const auto& sp : solo) solo_items.push_back(sp.second);
  b->nsolo = (int)solo_items.size();
  b->nshared = (unsigned)(n - b->nsolo);
  b->ntask = (int)task_item.size();
  PIE_HIP(b->d_items.alloc(std::max(n, 1)));
  PIE_HIP(b->d_task_item.alloc(std::max(b->ntask, 1)));
  PIE_HIP(b->d_solo_items.alloc(std::max(b->nsolo, 1)));
  if (b->ntask)
    PIE_CHECK(upload(task_item.data(), sizeof(int) * task_item.size(),
                     b->d_task_item.p, b->stream));
  if (!tlist.empty())
    PIE_CHECK(upload(tlist.data(), sizeof(int) * tlist.size(), b->d_tlist.p,
                     b->stream));
  if (b->nsolo)
    PIE_CHECK(upload(solo_items.data(), sizeof(int) * solo_items.size(),
                     b->d_solo_items.p, b->stream));
  // the uploads read host vectors: finish them before they go away
  PIE_HIP(hipStreamSynchronize(b->stream));
  drop_graph(b);
  b->dirty = false;
  return PIE_OK;
}

// Every item back to its initial state and stop-rule bookkeeping.
int restart(pie_batch* b) {
  for (size_t i = 0; i < b->items.size(); ++i) {
    ItemBuffers& it = *b->items[i];
    PIE_HIP(hipMemcpyAsync(it.xa.p, it.x0.p, it.bufsz * sizeof(float),
                           hipMemcpyDeviceToDevice, b->stream));
    BatchItem& d = b->desc[i];
    d.launches = d.sweeps = d.done = 0;
    d.dmax_bits = 0u;
    d.last_dmax = 0.0f;
    d.t_done = 0ull;
  }
  if (!b->desc.empty())
    PIE_CHECK(upload(b->desc.data(), sizeof(BatchItem) * b->desc.size(),
                     b->d_items.p, b->stream));
  PIE_HIP(hipMemsetAsync(b->d_counters.p, 0, 3 * sizeof(unsigned), b->stream));
  PIE_CHECK(upload(&b->nshared, sizeof(unsigned), b->d_counters.p + 3,
                   b->stream));
  return PIE_OK;
}

int capture(pie_batch* b, int rounds) {
  if (b->exec && b->graph_rounds == rounds) return PIE_OK;
  drop_graph(b);
  BatchArgs a{};
  a.items = b->d_items.p;
  a.task_item = b->d_task_item.p;
  a.queue = b->d_counters.p;
  a.solo_queue = b->d_counters.p + 2;
  a.solo_items = b->d_solo_items.p;
  a.shared_live = b->d_counters.p + 3;
  a.nsolo = b->nsolo;
  a.ntask = b->ntask;
  a.seg_rows = b->cfg.seg_rows;
  PIE_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int r = 0; r < rounds && e == hipSuccess; ++r) {
    e = pie::launch_batch(b->cfg, PIE_JACOBI, b->order, a, b->blocks,
                          b->stream);
    if (e == hipSuccess)
      e = pie::launch_batch_epilogue(b->d_items.p, (int)b->items.size(),
                                     b->cfg.S, b->d_counters.p,
                                     b->d_counters.p + 2,
                                     b->d_counters.p + 1,
                                     b->d_counters.p + 3, b->stream);
  }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(PIE_EHIP, std::string("batch graph capture: ") +
                              hipGetErrorString(e != hipSuccess ? e : e2));
  }
  b->graph = g;
  PIE_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_rounds = rounds;
  return PIE_OK;
}

}  // namespace

extern "C" {

int pie_batch_create(int order, int sweeps_per_launch, int cols_per_lane,
                     int seg_rows, pie_batch** out) {
  if (!out) return fail(PIE_EINVAL, "out is NULL");
  *out = nullptr;
  if (order != PIE_ORDER_UDLR && order != PIE_ORDER_ULRD)
    return fail(PIE_EINVAL, "unknown neighbour order");
  TileCfg cfg{sweeps_per_launch > 0 ? sweeps_per_launch : 8,
              cols_per_lane > 0 ? cols_per_lane : 1,
              seg_rows > 0 ? seg_rows : 64};
  if (!pie::tile_supported(cfg))
    return fail(PIE_EUNSUPPORTED, "tile not compiled");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(PIE_EHIP, "no HIP device");
  std::unique_ptr<pie_batch> b(new pie_batch());
  b->order = order;
  b->cfg = cfg;
  PIE_HIP(hipGetDevice(&b->device));
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                            b->device) == hipSuccess && khz > 0)
    b->wall_hz = khz * 1e3;
  PIE_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  PIE_HIP(b->d_counters.alloc(4));
  PIE_HIP(b->d_clock.alloc(1));
  PIE_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_live), sizeof(unsigned),
                        hipHostMallocDefault));
  PIE_HIP(b->err.alloc(3));
  *out = b.release();
  return PIE_OK;
}

int pie_batch_set_tile_channels(pie_batch* b, int channels) {
  if (!b) return fail(PIE_EINVAL, "NULL batch");
  if (channels != 0 && channels != 1 && channels != 3)
    return fail(PIE_EINVAL, "channels per task must be 0, 1 or 3");
  TileCfg c = b->cfg;
  c.N = channels;
  if (!pie::tile_supported(c))
    return fail(PIE_EUNSUPPORTED,
                "tile (sweeps " + std::to_string(c.S) + ", columns/lane " +
                    std::to_string(c.C) + ", channels " +
                    std::to_string(pie::tile_nch(c)) + ") not compiled");
  b->cfg = c;
  b->dirty = true;  // the task table depends on the channel split
  drop_graph(b);
  return PIE_OK;
}

int pie_batch_destroy(pie_batch* b) {
  if (!b) return PIE_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->h_live) (void)hipHostFree(b->h_live);
  b->items.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PIE_OK;
}

int pie_batch_add(pie_batch* b, int h, int w, const int32_t* mask,
                  const float* tgt, const float* grad, int max_sweeps,
                  float tol, int* index) {
  if (!b || !mask || !tgt || !grad) return fail(PIE_EINVAL, "NULL argument");
  if (h < 3 || w < 3) return fail(PIE_EINVAL, "grid smaller than 3 x 3");
  if (max_sweeps < 0) return fail(PIE_EINVAL, "negative sweep budget");
  std::unique_ptr<ItemBuffers> it(new ItemBuffers());
  it->H = h;
  it->W = w;
  const pie::host::GridLayout L = pie::host::grid_layout(h, w);
  if ((long long)h * L.P > (1LL << 31) - 1)
    return fail(PIE_EINVAL, "grid larger than 2^31 elements");
  it->P = L.P;
  it->plane = L.plane;
  it->bufsz = L.floats;
  it->mbytes = L.mask;
  it->max_sweeps = max_sweeps;
  it->tol = tol;
  const size_t fl = (size_t)it->bufsz;
  PIE_HIP(it->xa.alloc(fl));
  PIE_HIP(it->xb.alloc(fl));
  PIE_HIP(it->g.alloc(fl));
  PIE_HIP(it->x0.alloc(fl));
  PIE_HIP(it->m.alloc(it->mbytes));
  PIE_HIP(hipMemsetAsync(it->x0.p, 0, fl * sizeof(float), b->stream));
  PIE_HIP(hipMemsetAsync(it->g.p, 0, fl * sizeof(float), b->stream));
  PIE_HIP(hipMemsetAsync(it->m.p, 0, it->mbytes, b->stream));
  const size_t px = (size_t)h * w;
  DevBuf<char> stage;
  PIE_HIP(stage.alloc(px * 12));
  DevBuf<unsigned long long> counts;
  PIE_HIP(counts.alloc(2));
  PIE_HIP(hipMemsetAsync(counts.p, 0, 2 * sizeof(unsigned long long),
                         b->stream));
  PIE_CHECK(upload(mask, px * 4, stage.p, b->stream));
  PIE_HIP(pie::launch_mask(reinterpret_cast<const int32_t*>(stage.p),
                           it->m.p + pie::kPadLeft, h, w, it->P, counts.p,
                           b->stream));
  unsigned long long cnt[2] = {0, 0};
  PIE_HIP(hipMemcpyAsync(cnt, counts.p, sizeof(cnt), hipMemcpyDeviceToHost,
                         b->stream));
  PIE_HIP(hipStreamSynchronize(b->stream));
  if (cnt[1])
    return fail(PIE_EINVAL,
                "mask has solved pixels on the outermost rows/columns");
  it->masked = (long long)cnt[0];
  it->run0.assign(h + 1, 0);
  for (int r = 0; r < h; ++r) {
    const int32_t* row = mask + (size_t)r * w;
    for (int c = 0; c < w;) {
      if (!row[c]) {
        ++c;
        continue;
      }
      const int c0 = c;
      while (c < w && row[c]) ++c;
      it->run.push_back(c0);
      it->run.push_back(c);
    }
    it->run0[r + 1] = (int)it->run.size();
  }
  PIE_CHECK(upload(tgt, px * 12, stage.p, b->stream));
  PIE_HIP(pie::launch_deinterleave(reinterpret_cast<const float*>(stage.p),
                                   it->x0.p + pie::kPadLeft, it->plane, h, w,
                                   it->P, b->stream));
  PIE_CHECK(upload(grad, px * 12, stage.p, b->stream));
  PIE_HIP(pie::launch_deinterleave(reinterpret_cast<const float*>(stage.p),
                                   it->g.p + pie::kPadLeft, it->plane, h, w,
                                   it->P, b->stream));
  // NaN guidance at unsolved pixels and in the padding (the sweep kernel
  // tells solved pixels by it)
  PIE_HIP(hipMemsetAsync(counts.p, 0, sizeof(unsigned long long), b->stream));
  PIE_HIP(pie::launch_box_guidance(it->g.p, it->m.p, it->plane, h, it->P,
                                   pitch_of(w), counts.p, b->stream));
  // both state buffers start as the initial state (padding included)
  PIE_HIP(hipMemcpyAsync(it->xa.p, it->x0.p, fl * sizeof(float),
                         hipMemcpyDeviceToDevice, b->stream));
  PIE_HIP(hipMemcpyAsync(it->xb.p, it->x0.p, fl * sizeof(float),
                         hipMemcpyDeviceToDevice, b->stream));
  unsigned long long nan_solved = 0;
  PIE_HIP(hipMemcpyAsync(&nan_solved, counts.p, sizeof(nan_solved),
                         hipMemcpyDeviceToHost, b->stream));
  PIE_HIP(hipStreamSynchronize(b->stream));
  if (nan_solved)
    return fail(PIE_EINVAL, std::to_string(nan_solved) +
                                " guidance values of solved pixels are NaN");
  if (index) *index = (int)b->items.size();
  b->items.push_back(std::move(it));
  b->dirty = true;
  return PIE_OK;
}

int pie_batch_autotune(pie_batch* b, int rounds) {
  if (!b) return fail(PIE_EINVAL, "NULL batch");
  if (b->items.empty()) return fail(PIE_ESTATE, "empty batch");
  if (rounds < 1) rounds = 1;
  int var[32][3];
  const int nv = std::min(pie::tile_variants(var, 32), 32);
  const int rows[] = {32, 64, 128, 256};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PIE_HIP(hipEventCreate(&e0));
  PIE_HIP(hipEventCreate(&e1));
  struct Events {
    hipEvent_t a, b;
    ~Events() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } guard{e0, e1};
  const TileCfg start = b->cfg;
  TileCfg best = start;
  float best_ms = 1e30f;
  for (int v = 0; v < nv; ++v) {
    // four columns a lane would spill in the 16-wave batch work-groups
    if (var[v][1] > 2) continue;
    for (int rr : rows) {
      b->cfg = TileCfg{var[v][0], var[v][1], rr, var[v][2]};
      b->dirty = true;
      PIE_CHECK(prepare(b));
      PIE_CHECK(restart(b));
      BatchArgs a{};
      a.items = b->d_items.p;
      a.task_item = b->d_task_item.p;
      a.queue = b->d_counters.p;
      a.solo_queue = b->d_counters.p + 2;
      a.solo_items = b->d_solo_items.p;
      a.shared_live = b->d_counters.p + 3;
      a.nsolo = b->nsolo;
      a.ntask = b->ntask;
      a.seg_rows = b->cfg.seg_rows;
      PIE_HIP(hipEventRecord(e0, b->stream));
      for (int r = 0; r < rounds; ++r) {
        PIE_HIP(pie::launch_batch(b->cfg, PIE_JACOBI, b->order, a, b->blocks,
                                  b->stream));
        PIE_HIP(pie::launch_batch_epilogue(
            b->d_items.p, (int)b->items.size(), b->cfg.S, b->d_counters.p,
            b->d_counters.p + 2, b->d_counters.p + 1, b->d_counters.p + 3,
            b->stream));
      }
      PIE_HIP(hipEventRecord(e1, b->stream));
      PIE_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      PIE_HIP(hipEventElapsedTime(&ms, e0, e1));
      // solved-pixel sweeps done (solo items run a tile-dependent number
      // of launches a round): time per pixel sweep
      PIE_HIP(hipMemcpyAsync(b->desc.data(), b->d_items.p,
                             sizeof(BatchItem) * b->desc.size(),
                             hipMemcpyDeviceToHost, b->stream));
      PIE_HIP(hipStreamSynchronize(b->stream));
      double done = 0.0;
      for (size_t i = 0; i < b->desc.size(); ++i)
        done += (double)b->desc[i].sweeps * (double)b->items[i]->masked;
      const float per_sweep = done > 0.0 ? (float)(ms / done) : 1e30f;
      if (per_sweep < best_ms) {
        best_ms = per_sweep;
        best = b->cfg;
      }
    }
  }
  b->cfg = best;
  b->dirty = true;
  drop_graph(b);
  return PIE_OK;
}

int pie_batch_tile(pie_batch* b, int* sweeps, int* cols_per_lane,
                   int* seg_rows, int* channels) {
  if (!b) return fail(PIE_EINVAL, "NULL batch");
  if (sweeps) *sweeps = b->cfg.S;
  if (cols_per_lane) *cols_per_lane = b->cfg.C;
  if (seg_rows) *seg_rows = b->cfg.seg_rows;
  if (channels) *channels = pie::tile_nch(b->cfg);
  return PIE_OK;
}

int pie_batch_solve(pie_batch* b, int rounds_per_graph, double* seconds,
                    int* rounds_run) {
  if (!b) return fail(PIE_EINVAL, "NULL batch");
  if (b->items.empty()) return fail(PIE_ESTATE, "empty batch");
  const int R = rounds_per_graph > 0 ? rounds_per_graph : 16;
  PIE_CHECK(prepare(b));
  PIE_CHECK(restart(b));
  PIE_CHECK(capture(b, R));
  int max_rounds = 1;
  for (auto& it : b->items)
    max_rounds = std::max(max_rounds,
                          (it->max_sweeps + b->cfg.S - 1) / b->cfg.S);
  PIE_HIP(pie::launch_clock(b->d_clock.p, b->stream));
  PIE_HIP(hipStreamSynchronize(b->stream));
  const auto t0 = std::chrono::steady_clock::now();
  int rounds = 0;
  // every item stops within max_rounds rounds (sweep budget), so the loop
  // ends even if no tolerance is ever met
  while (rounds < max_rounds) {
    PIE_HIP(hipGraphLaunch(b->exec, b->stream));
    rounds += R;
    PIE_HIP(hipMemcpyAsync(b->h_live, b->d_counters.p + 1, sizeof(unsigned),
                           hipMemcpyDeviceToHost, b->stream));
    PIE_HIP(hipStreamSynchronize(b->stream));
    if (*b->h_live == 0) break;
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  if (rounds_run) *rounds_run = rounds;
  PIE_HIP(hipMemcpyAsync(b->desc.data(), b->d_items.p,
                         sizeof(BatchItem) * b->desc.size(),
                         hipMemcpyDeviceToHost, b->stream));
  PIE_HIP(hipMemcpyAsync(&b->t_start, b->d_clock.p, sizeof(b->t_start),
                         hipMemcpyDeviceToHost, b->stream));
  PIE_HIP(hipStreamSynchronize(b->stream));
  return PIE_OK;
}

int pie_batch_item(pie_batch* b, int i, int* sweeps, float* last_dmax,
                   double* latency_s) {
  if (!b) return fail(PIE_EINVAL, "NULL batch");
  if (i < 0 || i >= (int)b->desc.size())
    return fail(PIE_EINVAL, "item index out of range (or not solved yet)");
  const BatchItem& d = b->desc[i];
  if (sweeps) *sweeps = d.sweeps;
  if (last_dmax) *last_dmax = d.last_dmax;
  if (latency_s)
    *latency_s = d.done ? (double)(d.t_done - b->t_start) / b->wall_hz : -1.0;
  return PIE_OK;
}

int pie_batch_read(pie_batch* b, int i, uint8_t* img, float* err,
                   float* state) {
  if (!b) return fail(PIE_EINVAL, "NULL batch");
  if (i < 0 || i >= (int)b->items.size())
    return fail(PIE_EINVAL, "item index out of range");
  ItemBuffers& it = *b->items[i];
  const bool odd = i < (int)b->desc.size() && (b->desc[i].launches & 1);
  const float* x = (odd ? it.xb.p : it.xa.p) + pie::kPadLeft;
  const int nb = pie::finalize_blocks(it.H, it.W);
  PIE_HIP(b->partial.alloc(3 * (size_t)nb));
  DevBuf<uint8_t> dimg;
  if (img) PIE_HIP(dimg.alloc((size_t)it.H * it.W * 3));
  PIE_HIP(pie::launch_finalize(b->order, x, it.g.p + pie::kPadLeft,
                               it.m.p + pie::kPadLeft, it.plane, it.H, it.W,
                               it.P, img ? dimg.p : nullptr, b->partial.p, nb,
                               b->stream));
  PIE_HIP(pie::launch_reduce(b->partial.p, nb, b->err.p, b->stream));
  if (img)
    PIE_HIP(hipMemcpyAsync(img, dimg.p, (size_t)it.H * it.W * 3,
                           hipMemcpyDeviceToHost, b->stream));
  DevBuf<float> st;
  if (state) {
    PIE_HIP(st.alloc((size_t)it.H * it.W * 3));
    PIE_HIP(pie::launch_interleave(x, st.p, it.plane, it.H, it.W, it.P,
                                   b->stream));
    PIE_HIP(hipMemcpyAsync(state, st.p, (size_t)it.H * it.W * 3 * 4,
                           hipMemcpyDeviceToHost, b->stream));
  }
  float e[3];
  PIE_HIP(hipMemcpyAsync(e, b->err.p, sizeof(e), hipMemcpyDeviceToHost,
                         b->stream));
  PIE_HIP(hipStreamSynchronize(b->stream));
  if (err) std::memcpy(err, e, sizeof(e));
  return PIE_OK;
}

int pie_batch_size(pie_batch* b, int* items, int* tasks, long long* masked) {
  if (!b) return fail(PIE_EINVAL, "NULL batch");
  PIE_CHECK(prepare(b));
  long long total = 0;
  for (auto& it : b->items) total += it->masked;
  if (items) *items = (int)b->items.size();
  if (tasks) *tasks = b->ntask_all;
  if (masked) *masked = total;
  return PIE_OK;
}

}  // extern "C"
