"""Train the warm-start U-Net and measure what it gains (GPU).

    python -m fpie_amd.warmstart [--train N] [--epochs E] [--sizes 1024,2048]

1. N seeded training composites (blob images, star-polygon masks) drawn at
   1024 x 1024 and coarsened to 256 x 256 (the form in which large problems
   are shown to the network), labelled by the batched GPU Jacobi solver
   (20000 sweeps from the target).
2. Adam training (lr 1e-3, batch 64, masked mean squared error) with the
   hand-written forward/backward kernels.
3. On composites with unseen seeds at each evaluation size: plain Jacobi from
   the target state until the mean absolute residual is <= 1e-3 (cold), and
   the same from the network's prediction (warm; network inference, the
   residual comparison that decides the fallback and the state upload are
   inside the timed region).  Problems larger than the training size are
   predicted on a coarsened grid (fpie_amd.unet.warm_start).  Both final
   states are compared with a reference: the cold run continued for as many
   sweeps again.

Prints one JSON line per evaluation and a summary line.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import unet
from .core_hip import GridSolver


def evaluate(params: np.ndarray, size: int, seed: int, tol: float = 1e-3,
             check_every: int = 512, max_sweeps: int = 4_000_000,
             log=print) -> dict:
  mask, tgt, grad, mode = unet.composite((size, size), seed)
  solved = int(np.count_nonzero(mask))
  unet.warm_start(params, mask, tgt, grad)  # kernel loading, not timed

  def solver():
    s = GridSolver(schedule="jacobi")
    s.reset(0, mask, tgt, grad)
    s.set_tile(8, 1, 128 if size >= 2048 else 64)
    s.iterate(0)
    s.sync()
    return s

  def solve(s, tol_):
    # in chunks, with a progress line per chunk
    n, r, chunk = 0, float("inf"), 200_000
    while n < max_sweeps:
      k, r = s.solve_to_residual(tol_, check_every, min(chunk, max_sweeps - n))
      n += k
      log(f"#   {size}^2 seed {seed}: {n} sweeps, mean residual {r:.3g}")
      if r <= tol_ or k == 0:
        break
    return n, r

  # cold start
  s = solver()
  _, e0 = s.step(0)
  t = time.perf_counter()
  n_cold, r_cold = solve(s, tol)
  t_cold = time.perf_counter() - t
  x_cold = s.state()
  # reference: the same run continued for as many sweeps again
  s.iterate(max(n_cold, 1))
  _, e_ref = s.step(0)
  x_ref = s.state()
  r_ref = float(e_ref.sum()) / (3 * solved)
  del s
  # warm start: prediction, fallback test and upload inside the timing
  s = solver()
  t = time.perf_counter()
  x_pred = unet.warm_start(params, mask, tgt, grad)
  s.set_state(x_pred)
  _, e_pred = s.step(0)
  used = bool(e_pred.sum() < e0.sum())
  if not used:  # fallback: the prediction is worse than the target state
    s.set_state(tgt)
  t_pred = time.perf_counter() - t
  n_warm, r_warm = solve(s, tol)
  t_warm = time.perf_counter() - t
  x_warm = s.state()
  m = mask != 0
  rec = {
      "size": size, "seed": seed, "gradient": mode, "solved": solved,
      "tol_mean_residual": tol,
      "cold": {"seconds": t_cold, "sweeps": n_cold, "residual": r_cold,
               "start_residual": float(e0.sum()) / (3 * solved),
               "start_max_abs_diff_vs_reference": float(np.abs(
                   tgt[m] - x_ref[m]).max()),
               "max_abs_diff_vs_reference": float(np.abs(
                   x_cold[m] - x_ref[m]).max())},
      "warm": {"seconds": t_warm, "prediction_seconds": t_pred,
               "sweeps": n_warm, "residual": r_warm,
               "start_residual": float(e_pred.sum()) / (3 * solved),
               "start_max_abs_diff_vs_reference": float(np.abs(
                   x_pred[m] - x_ref[m]).max()),
               "start_mean_abs_diff_vs_reference": float(np.abs(
                   x_pred[m] - x_ref[m]).mean()),
               "prediction_used": used,
               "max_abs_diff_vs_reference": float(np.abs(
                   x_warm[m] - x_ref[m]).max())},
      "reference": {"sweeps": 2 * n_cold, "residual": r_ref},
      "speedup": t_cold / max(t_warm, 1e-9),
  }
  log(json.dumps(rec))
  return rec


def main(argv=None) -> None:
  ap = argparse.ArgumentParser()
  ap.add_argument("--train", type=int, default=4096)
  ap.add_argument("--size", type=int, default=256)
  ap.add_argument("--render", type=int, default=1024,
                  help="draw training composites at this size, then coarsen")
  ap.add_argument("--label-sweeps", type=int, default=20000)
  ap.add_argument("--epochs", type=int, default=8)
  ap.add_argument("--batch", type=int, default=64)
  ap.add_argument("--lr", type=float, default=1e-3)
  ap.add_argument("--sizes", default="1024,2048")
  ap.add_argument("--seeds", default="900001,900002")
  ap.add_argument("--tol", type=float, default=1e-3)
  ap.add_argument("--save", default="")
  ap.add_argument("--load", default="")
  a = ap.parse_args(argv)

  def log(msg):
    print(msg, flush=True)

  if a.load:
    with np.load(a.load) as f:
      params = f["params"].astype(np.float32)
    tr = {}
  else:
    M, T, G2, Y = unet.training_set(a.train, a.size, sweeps=a.label_sweeps,
                                    render=a.render, log=log)
    net, tr = unet.train(M, T, G2, Y, epochs=a.epochs, batch=a.batch,
                         lr=a.lr, log=log)
    params = net.params()
    del net, M, T, G2, Y
    if a.save:
      np.savez_compressed(a.save, params=params.astype(np.float16))
  evals = [evaluate(params, int(sz), int(sd), a.tol, log=log)
           for sz in a.sizes.split(",") for sd in a.seeds.split(",")]
  summary = {
      "train_samples": a.train, "train_size": a.size,
      "label_sweeps": a.label_sweeps, "epochs": a.epochs,
      "train_seconds": tr.get("seconds"), "train_steps": tr.get("steps"),
      "final_loss": float(np.mean(tr["losses"][-20:])) if tr else None,
      "conv_tflops": tr.get("tflops"),
      "mfma_utilisation_vs_2500_tflops": (tr.get("tflops") or 0) / 2500.0,
      "speedup_by_size": {
          sz: float(np.mean([e["speedup"] for e in evals
                             if e["size"] == int(sz)]))
          for sz in a.sizes.split(",")},
  }
  log(json.dumps({"summary": summary}))


if __name__ == "__main__":
  main(sys.argv[1:])
