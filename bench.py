"""Benchmark of the hip Poisson solver on one MI355X.

Default workload (``grid8192``, the roofline run): GridSolver on an 8192 x
8192 x 3 uint8 blob composite whose mask is the union of 6 seeded ellipses
(60 % coverage; a 7183 x 7546 crop), gradient ``max``, float32 state, 10000
Jacobi sweeps per step (the g+U+L+R+D order of the openmp GridSolver).  A
step is one ``GridSolver.step(10000)``: the sweeps, then the uint8 crop and
the residual sums copied back.  Inputs are resident in HBM (``reset`` is
outside the timed region).

Printed: one JSON line with the metric (solved-pixel updates per second),
the roofline of the sweep kernel (algorithmic bytes per launch over its
event-timed average launch duration, against the 8 TB/s HBM peak, and the
rocprofv3-measured HBM bytes of the same kernel and tile when recorded in
profiles/traffic.json), the 4096 x 4096 workload under Jacobi and under
the gcc backend's Gauss-Seidel order (``extra``), the time to a mean
residual of 1e-3 on an unseen 1024 x 1024 composite with and without the
learned warm start (``time_to_residual``), the MFMA utilisation of the warm
start's U-Net in a training step (``unet_mfma``) and the reference's CPU
path timed on this host (its openmp GridSolver compiled from the reference
sources, or the oracle's restatement when that is absent) on a bounded
sample of the same workload, on all host threads and on 8.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fast-poisson-image-editing_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBPS = 8000.0
BYTES_PER_UPDATE = 36  # read x (12 B) + read grad (12 B) + write x (12 B)


def build_inputs(name: str):
  from fpie_amd import process, synth

  class Grab:
    def reset(self, *a):
      self.a = a

    def partition(self, mask):
      flat = mask > 0
      return (np.cumsum(flat.reshape(-1)).reshape(mask.shape) *
              flat).astype(np.int32)

  w = synth.WORKLOADS[name]()
  process.register_backend("grab", lambda **kw: Grab(), lambda **kw: Grab())
  cls = process.EquProcessor if w.method == "equ" else process.GridProcessor
  p = cls(w.gradient, "grab")
  p.reset(w.src, w.mask, w.tgt, w.mask_on_src, w.mask_on_tgt)
  return w, p.core.a


def bench_batch(a):
  """Batched workload: 256 composites, each stopped after 5000 sweeps or
  once a sweep changes no value by 1e-3; one step = one batch solve."""
  from fpie_amd import core_hip, native, process, synth

  class Grab:
    def reset(self, *args):
      self.a = args

  print(f"# {native.device_info()}", file=sys.stderr, flush=True)
  t_in = time.perf_counter()
  items = synth.batch(256, seed=0)
  process.register_backend("grab", lambda **kw: Grab(), lambda **kw: Grab())
  tile = ([int(v) for v in a.tile.split(",")] if a.tile else [8, 1, 64])
  b = core_hip.BatchSolver(sweeps_per_launch=tile[0], cols_per_lane=tile[1],
                           rows_per_task=tile[2],
                           channels=tile[3] if len(tile) > 3 else 0)
  masked = []
  for w in items:
    p = process.GridProcessor(w.gradient, "grab")
    p.reset(w.src, w.mask, w.tgt, w.mask_on_src, w.mask_on_tgt)
    _, m, t, g = p.core.a
    b.add(m, t, g, max_sweeps=w.iters, tol=w.tol)
    masked.append(int(np.count_nonzero(m)))
  if not a.tile:
    b.autotune(6)
  tile = list(b.tile())
  size = b.size()
  print(f"# batch of {size['items']} built in "
        f"{time.perf_counter() - t_in:.1f} s: {size['tasks']} tiles, "
        f"{size['masked']} solved pixels", file=sys.stderr, flush=True)
  for _ in range(a.warmup):
    b.solve(16)
  secs, lats, upd = [], [], 0.0
  for _ in range(a.steps):
    sec, _ = b.solve(16)
    secs.append(sec)
    info = [b.item(i) for i in range(len(items))]
    lats.extend(x["latency_s"] for x in info)
    upd += sum(x["sweeps"] * mk for x, mk in zip(info, masked))
  sweeps = [x["sweeps"] for x in info]
  total = sum(secs)
  line = {
      "metric": "composites/s (batched, per-item stop rule)",
      "value": len(items) * a.steps / total,
      "unit": "composites/s",
      "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
      "ms_per_step": total * 1e3 / a.steps,
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
      "dtype": "f32",
      "data": "synthetic (seeded blob composites, star-polygon masks)",
      "config": {"workload": "batch256", "items": len(items),
                 "solved_pixels": size["masked"], "tiles": size["tasks"],
                 "max_sweeps": 5000, "tol_max_abs_update": 1e-3,
                 "tile": {"sweeps_per_launch": tile[0],
                          "cols_per_lane": tile[1], "rows_per_task": tile[2],
                          "channels_per_task": tile[3]},
                 "rounds_per_graph": 16},
      "latency_p50_ms": float(np.percentile(lats, 50) * 1e3),
      "latency_p99_ms": float(np.percentile(lats, 99) * 1e3),
      "pixel_updates_per_s": upd / total,
      "sweeps_min_median_max": [int(np.min(sweeps)), int(np.median(sweeps)),
                                int(np.max(sweeps))],
  }
  print(json.dumps(line), flush=True)


def kernel_name(tile, schedule: str, order: str = "ulrd") -> str:
  """Name of the sweep-kernel instantiation a tile runs (as rocprofv3
  prints it)."""
  S, C, rows, nch = tile
  sched = {"jacobi": 0, "gs_lex": 1, "redblack": 2}[schedule]
  return (f"void pie::detail::march_kernel<{S}, {C}, {nch}, {sched}, "
          f"{1 if order == 'ulrd' else 0}>(pie::MarchArgs)")


def traffic_from_profiles(workload: str, schedule: str, tile):
  """HBM bytes per launch of the sweep kernel measured with rocprofv3 PMC
  counters (profiles/traffic.json) for exactly this workload, schedule,
  tile (sweeps, columns per lane, rows per task, channels per task) and
  kernel; None when no such record exists."""
  path = os.path.join(ROOT, "profiles", "traffic.json")
  if not os.path.exists(path):
    return None
  with open(path) as f:
    records = json.load(f)["records"]
  for r in records:
    if (r["workload"] == workload and r["schedule"] == schedule and
        list(r["tile"]) == list(tile) and
        r["kernel"] == kernel_name(tile, schedule)):
      return r
  return None


def cpu_model() -> str:
  try:
    with open("/proc/cpuinfo") as f:
      for line in f:
        if line.startswith("model name"):
          return line.split(":", 1)[1].strip()
  except OSError:
    pass
  return "unknown"


def cpu_baseline(w, budget_s: float):
  """The reference's openmp solver core (compiled from the reference sources
  into oracle/_ref) driven by the processor on the same composite, timed on
  a number of sweeps sized to about budget_s / 2 seconds, once on all host
  threads (OMP_NUM_THREADS, else the CPUs this process may run on) and once
  on 8 (the reference's default cap, fpie/process.py:11,26); the oracle's
  port of the same loops, one thread, when oracle/_ref is absent."""
  sys.path.insert(0, os.path.join(ROOT, "oracle"))
  import oracle as O
  from fpie_amd import process

  avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") \
      else (os.cpu_count() or 1)
  threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
  ref = O.ref_core("core_openmp")
  masked = int(process.prepare_mask(w.mask)[0].sum())
  cls = process.EquProcessor if w.method == "equ" else process.GridProcessor

  def timed(make_equ, make_grid, seconds):
    process.register_backend("cpu-baseline", make_equ, make_grid)
    p = cls(w.gradient, "cpu-baseline")
    p.reset(w.src, w.mask, w.tgt, w.mask_on_src, w.mask_on_tgt)
    p.core.step(1)  # warm-up
    t0 = time.perf_counter()
    p.core.step(2)
    dt = (time.perf_counter() - t0) / 2
    k = max(2, int(seconds / max(dt, 1e-6)))
    t0 = time.perf_counter()
    p.core.step(k)
    dt = time.perf_counter() - t0
    return masked * k / dt, k, dt

  if ref is not None:
    # openmp flags of docs/benchmark.md (GridSolver --grid-x 2 --grid-y 16)
    runs = {}
    for n in sorted({threads, 8}, reverse=True):
      runs[n] = timed(lambda **kw: ref.EquSolver(n),
                      lambda **kw: ref.GridSolver(2, 16, n), budget_s / 2)
    v, k, dt = runs[threads]
    out = {"value": v, "unit": "pixel-updates/s", "cores": threads,
           "kind": "reference",
           "sample": f"fpie/core/openmp compiled from the reference sources: "
                     f"{k} sweeps of the {w.name} problem ({masked} solved "
                     f"pixels), {dt:.2f} s, {threads} threads",
           "cpu_model": cpu_model(), "host_cpus": avail}
    if 8 in runs and threads != 8:
      v8, k8, dt8 = runs[8]
      out["threads_8"] = {"value": v8, "sweeps": k8, "seconds": dt8}
    return out
  v, k, dt = timed(lambda **kw: O.EquCore("openmp"),
                   lambda **kw: O.GridCore(2, 16, "openmp"), budget_s)
  return {"value": v, "unit": "pixel-updates/s", "cores": 1, "kind": "port",
          "sample": f"oracle restatement of fpie/core/openmp (1 thread): {k} "
                    f"sweeps of the {w.name} problem ({masked} solved "
                    f"pixels), {dt:.2f} s",
          "cpu_model": cpu_model(), "host_cpus": avail}


def bench_warm_start(size: int = 1024, seed: int = 900_001,
                     tol: float = 1e-3) -> dict:
  """Time to a mean absolute residual of `tol` on an unseen size x size
  composite: plain Jacobi from the target (cold) vs the learned warm start
  (fpie_amd.warmstart with the packaged network) followed by plain Jacobi.
  Accuracy against the exact solution is measured by
  `python -m fpie_amd.warmstart` (it needs a float64 reference solve)."""
  from fpie_amd import synth, warmstart
  from fpie_amd.core_hip import GridSolver
  if not os.path.exists(warmstart.WEIGHTS):
    return {"skipped": "no packaged network weights"}
  mask, tgt, grad, mode = synth.composite((size, size), seed)

  def solver():
    s = GridSolver(schedule="jacobi")
    s.reset(0, mask, tgt, grad)
    s.sync()
    return s

  ws = warmstart.WarmStart(warmstart.load_params())
  ws.run(solver(), mask, tgt)  # untimed: kernel loading, network set-up
  s = solver()
  t = time.perf_counter()
  n_cold, r_cold = s.solve_to_residual(tol, 512, 10**8)
  t_cold = time.perf_counter() - t
  del s
  s = solver()
  t = time.perf_counter()
  info = ws.run(s, mask, tgt)
  n_warm, r_warm = s.solve_to_residual(tol, 512, 10**8)
  t_warm = time.perf_counter() - t
  return {
      "workload": f"composite {size}x{size} seed {seed} (blob source and "
                  f"target, star mask, gradient {mode})",
      "solved_pixels": int(np.count_nonzero(mask)),
      "tol_mean_residual": tol,
      "cold": {"seconds": t_cold, "sweeps": n_cold, "residual": r_cold},
      "warm": {"seconds": t_warm, "cycles": info["cycles"],
               "smoothing_sweeps": info["sweeps"], "jacobi_sweeps": n_warm,
               "residual": r_warm, "prediction_used": info["used"]},
      "speedup": t_cold / max(t_warm, 1e-9),
  }


def run_grid(name: str, schedule: str, steps: int, warmup: int, tile: str,
             iters: int = 0) -> dict:
  """K timed steps of one workload on the GPU (after W untimed ones), with
  the sweep kernel's roofline; the solver is freed before returning."""
  from fpie_amd import core_hip

  t_in = time.perf_counter()
  w, args = build_inputs(name)
  iters = iters or w.iters
  print(f"# inputs {w.name} built in {time.perf_counter() - t_in:.1f} s",
        file=sys.stderr, flush=True)
  if w.method == "grid":
    _, m, t, g = args
    core = core_hip.GridSolver(schedule=schedule, order="ulrd")
    core.reset(0, m, t, g)
  else:
    N, A, X, B = args
    core = core_hip.EquSolver(schedule=schedule)
    from fpie_amd.process import prepare_mask
    pm, *_ = prepare_mask(w.mask)
    core.partition(pm)
    core.reset(N, A, X, B)
  masked = core.shape()["masked"]
  if tile:
    core.set_tile(*[int(x) for x in tile.split(",")])
  else:
    core.autotune(8)
  t4 = list(core.tile()) + [core.tile_channels()]
  print(f"# {name} {schedule}: tile S={t4[0]} C={t4[1]} rows={t4[2]} "
        f"channels/task {t4[3]}", file=sys.stderr, flush=True)
  for _ in range(warmup):
    core.step(iters)
  core.sync()
  core.set_timing(True)
  t0 = time.perf_counter()
  for _ in range(steps):
    core.step(iters)
  core.sync()
  dt = time.perf_counter() - t0
  tm = core.timing()
  del core
  updates = masked * iters * steps
  launch_ms = tm["kernel_ms"] / max(tm["launches"], 1)
  launches_per_step = tm["launches"] / max(steps, 1)
  sweeps_per_launch = iters / max(launches_per_step, 1)
  alg_bytes = BYTES_PER_UPDATE * masked * sweeps_per_launch
  achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
  traffic = traffic_from_profiles(w.name, schedule, t4)
  hbm = (traffic["bytes_per_launch"] / (launch_ms * 1e-3) / 1e9
         if traffic else None)
  return {
      "workload": w,
      "value": updates / dt,
      "ms_per_step": dt * 1e3 / steps,
      "iters": iters,
      "masked": masked,
      "tile": t4,
      "roofline": {
          "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
          "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
          "traffic": traffic["bytes_per_launch"] if traffic else None,
          # measured HBM bytes per launch over the launch time
          "hbm_achieved": hbm,
          "hbm_frac": hbm / HBM_PEAK_GBPS if hbm else None,
          "kernel": kernel_name(t4, schedule),
          "bytes_per_update": BYTES_PER_UPDATE,
          "sweeps_per_launch": sweeps_per_launch,
          "avg_launch_ms": launch_ms,
          "launches_timed": tm["launches"],
          "traffic_profile": traffic["summary"] if traffic else None},
  }


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=3)
  ap.add_argument("--warmup", type=int, default=1)
  ap.add_argument("--workload", default="grid8192",
                  choices=["grid8192", "grid4096", "equ512", "batch256"])
  ap.add_argument("--schedule", default="jacobi",
                  choices=["jacobi", "gs_lex", "redblack"])
  ap.add_argument("--iters", type=int, default=0,
                  help="sweeps per step (default: the workload's)")
  ap.add_argument("--tile", type=str, default="",
                  help="S,C,rows[,channels] to skip the autotuner")
  ap.add_argument("--cpu-budget", type=float, default=16.0)
  ap.add_argument("--no-cpu-baseline", action="store_true")
  ap.add_argument("--no-extra", action="store_true",
                  help="skip the 4096 x 4096 jacobi / gs_lex lines")
  ap.add_argument("--no-warm-start", action="store_true",
                  help="skip the time-to-residual comparison")
  ap.add_argument("--no-unet", action="store_true",
                  help="skip the U-Net training-step MFMA measurement")
  a = ap.parse_args()
  if a.gpus != 1:
    print(f"bench.py measures one GPU: this build has no multi-GPU path "
          f"(--gpus {a.gpus} requested)", file=sys.stderr)
    sys.exit(2)

  if a.workload == "batch256":
    return bench_batch(a)

  from fpie_amd import native

  print(f"# {native.device_info()}", file=sys.stderr, flush=True)
  r = run_grid(a.workload, a.schedule, a.steps, a.warmup, a.tile, a.iters)
  w = r["workload"]
  line = {
      "metric": "Jacobi pixel-updates/s + % HBM peak; time-to-residual "
                "warm-start vs plain",
      "value": r["value"],
      "unit": "pixel-updates/s",
      "n_gpus": 1,
      "steps": a.steps,
      "warmup": a.warmup,
      "ms_per_step": r["ms_per_step"],
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "f32",
      "data": "synthetic (seeded blob composites, fpie_amd.synth)",
      "config": {"workload": w.name, "method": w.method,
                 "image": list(w.src.shape), "solved_pixels": r["masked"],
                 "gradient": w.gradient, "sweeps_per_step": r["iters"],
                 "schedule": a.schedule,
                 "order": "ulrd" if w.method == "grid" else "udlr",
                 "tile": {"sweeps_per_launch": r["tile"][0],
                          "cols_per_lane": r["tile"][1],
                          "rows_per_task": r["tile"][2],
                          "channels_per_task": r["tile"][3]},
                 "parallelism": "single GPU"},
      "roofline": r["roofline"],
  }
  if not a.no_extra and a.workload == "grid8192":
    # the 4096 x 4096 square workload: Jacobi, and the gcc backend's
    # Gauss-Seidel order (bit-exact to fpie/core/gcc/grid.cc)
    extra = {}
    for sched, steps in (("jacobi", 3), ("gs_lex", 2)):
      e = run_grid("grid4096", sched, steps, 1, "")
      extra[f"grid4096_{sched}"] = {
          "value": e["value"], "unit": "pixel-updates/s",
          "ms_per_step": e["ms_per_step"], "sweeps_per_step": e["iters"],
          "steps": steps, "tile": e["tile"],
          "avg_launch_ms": e["roofline"]["avg_launch_ms"],
          "frac": e["roofline"]["frac"],
          "hbm_frac": e["roofline"]["hbm_frac"]}
    line["extra"] = extra
  if not a.no_warm_start and w.method == "grid":
    line["time_to_residual"] = bench_warm_start()
  if not a.no_unet:
    # the warm start's network: one Adam step of UNet(64, 256, 256) in bf16
    # on the MFMA units (dense bf16 peak 2.5 PFLOP/s)
    from fpie_amd import unet
    line["unet_mfma"] = unet.measure(64, 256, 20)
  if not a.no_cpu_baseline:
    line["cpu_baseline"] = cpu_baseline(w, a.cpu_budget)
  print(json.dumps(line), flush=True)


if __name__ == "__main__":
  main()
