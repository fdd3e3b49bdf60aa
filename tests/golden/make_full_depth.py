"""Write the full-depth fixtures of the benchmark workloads.

The 4096 x 4096 and 8192 x 8192 workloads are too large to run through the
oracle inside a GPU test at full depth (minutes of CPU per case), so this
script runs them once, in a container that holds the reference checkout,
and records SHA-256 digests of the results; the GPU tests recompute the
same inputs from ``fpie_amd.synth`` and compare digests:

    python tests/golden/make_full_depth.py grid8192_jacobi
    python tests/golden/make_full_depth.py grid4096_jacobi
    python tests/golden/make_full_depth.py grid4096_gs

Each case writes ``tests/golden/full_depth_<case>.json`` holding

* digests of the solver inputs (cropped mask, target crop and guidance from
  the GridProcessor), so a test can tell a changed generator from a changed
  solver;
* per checkpoint (sweep count): digests of the oracle's float32 state and
  of its clamped uint8 crop, its residual sums (``err``, float32, in the
  reference core's own accumulation order) and the exactly rounded sums of
  the same terms (``err_exact``, float64);
* the reference core compiled from the reference sources
  (``oracle/build_ref.sh``): its uint8 crop and err after the same sweep
  count.  The script fails unless they equal the oracle's (uint8 bit for
  bit, err to rtol 1e-5, the tolerance of fpie's tests/test_smoke.py:66).

Schedules: ``jacobi`` is fpie/core/openmp/grid.cc:75-103 (g+U+L+R+D), checked
against core_openmp; ``gs`` is fpie/core/gcc/grid.cc:57-75 (lexicographic
Gauss-Seidel in place), checked against core_gcc.  Only data is written.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "fast-poisson-image-editing_amd"))

import oracle as O  # noqa: E402

CASES = {
    # name: (workload, schedule, checkpoints)
    "grid8192_jacobi": ("grid8192", "jacobi", (24, 10000)),
    "grid4096_jacobi": ("grid4096", "jacobi", (24, 5000)),
    "grid4096_gs": ("grid4096", "gs", (3, 5000)),
}


def digest(a: np.ndarray) -> str:
  return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


class _Grab:
  def reset(self, *a):
    self.a = a


def grid_inputs(workload: str):
  """(mask int32, tgt float32, grad float32) of the GridProcessor."""
  from fpie_amd import process, synth
  w = synth.WORKLOADS[workload]()
  process.register_backend("grab", lambda **kw: _Grab(), lambda **kw: _Grab())
  p = process.GridProcessor(w.gradient, "grab")
  p.reset(w.src, w.mask, w.tgt, w.mask_on_src, w.mask_on_tgt)
  _, mask, tgt, grad = p.core.a
  return (np.ascontiguousarray(mask, np.int32),
          np.ascontiguousarray(tgt, np.float32),
          np.ascontiguousarray(grad, np.float32))


def run(case: str, threads: int) -> dict:
  workload, sched, points = CASES[case]
  t0 = time.time()
  mask, tgt, grad = grid_inputs(workload)
  out = {"case": case, "workload": workload, "schedule": sched,
         "shape": list(mask.shape), "solved": int(np.count_nonzero(mask)),
         "inputs": {"mask": digest(mask), "tgt": digest(tgt),
                    "grad": digest(grad)},
         "checkpoints": {}}
  print(f"{case}: inputs {mask.shape} in {time.time() - t0:.1f} s",
        flush=True)
  x, done = tgt, 0
  kind = "openmp" if sched == "jacobi" else "gcc"
  for n in points:
    t = time.time()
    if sched == "jacobi":
      x = O.grid_jacobi(mask, x, grad, n - done, "ulrd", threads=threads)
    else:
      x = O.grid_gs(mask, x, grad, n - done)
    done = n
    u8 = O.clamp_u8(x)
    err = O.grid_err(mask, x, grad, kind)
    exact = O.grid_terms(mask, x, grad, "ulrd").reshape(-1, 3).sum(0)
    out["checkpoints"][str(n)] = {
        "state": digest(x), "u8": digest(u8),
        "err": [float(v) for v in err],
        "err_exact": [float(v) for v in exact]}
    print(f"{case}: oracle {n} sweeps in {time.time() - t:.1f} s, err "
          f"{err}", flush=True)

  # the reference core, compiled from the reference sources, on the same
  # problem and the last sweep count
  core = O.ref_core("core_openmp" if sched == "jacobi" else "core_gcc")
  if core is None:
    raise SystemExit("oracle/_ref is not built: run oracle/build_ref.sh")
  t = time.time()
  if sched == "jacobi":
    solver = core.GridSolver(2, 16, threads)  # docs/benchmark.md flags
  else:
    solver = core.GridSolver(8, 8)
  solver.reset(0, mask, tgt, grad)
  u8_ref, err_ref = solver.step(done)
  if digest(np.ascontiguousarray(u8_ref, np.uint8)) != digest(u8):
    raise SystemExit(f"{case}: the reference core's uint8 crop differs "
                     "from the oracle's")
  np.testing.assert_allclose(np.asarray(err_ref, np.float32), err, rtol=1e-5)
  out["reference_core"] = {
      "sweeps": done, "u8_equal": True,
      "err": [float(v) for v in np.asarray(err_ref, np.float32)]}
  print(f"{case}: reference core {done} sweeps in {time.time() - t:.1f} s, "
        f"uint8 identical, err {err_ref}", flush=True)
  return out


def main():
  threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (
      os.cpu_count() or 1)
  for case in sys.argv[1:] or list(CASES):
    res = run(case, threads)
    path = os.path.join(HERE, f"full_depth_{case}.json")
    with open(path, "w") as f:
      json.dump(res, f, indent=1)
    print(f"wrote {path}", flush=True)


if __name__ == "__main__":
  main()
