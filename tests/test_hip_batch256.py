"""The batched workload at full size (GPU): 256 composites, sides
log-uniform in [128, 2048], star-polygon masks, each stopped after 5000
sweeps or once a sweep changes no value by 1e-3 (fpie_amd.synth.batch).

The batch is solved at its autotuned tile; the smallest, median and
largest items (by solved pixels) must equal the oracle's Jacobi loop
(oracle/pie_oracle.c: po_grid_jacobi_ulrd, fpie/core/openmp/grid.cc:75-103)
run for the sweeps each reports, bit for bit, and the items must not all
finish together: the median item's latency is below the largest one's.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Grab:
  def reset(self, *a):
    self.a = a


def test_batch256_against_oracle(oracle):
  from fpie_amd import core_hip, process, synth
  process.register_backend("grab", lambda **kw: _Grab(), lambda **kw: _Grab())
  print("\nbatch256: building 256 composites", flush=True)
  items = []
  b = core_hip.BatchSolver()
  for w in synth.batch(256, seed=0):
    p = process.GridProcessor(w.gradient, "grab")
    p.reset(w.src, w.mask, w.tgt, w.mask_on_src, w.mask_on_tgt)
    _, m, t, g = p.core.a
    b.add(m, t, g, max_sweeps=w.iters, tol=w.tol)
    items.append((m, t, g))
  print(f"batch256: autotuned tile {b.autotune(3)}", flush=True)
  sec, rounds = b.solve(16)
  print(f"batch256: solved in {sec:.2f} s, {rounds} rounds", flush=True)
  sizes = np.array([int(np.count_nonzero(m)) for m, _, _ in items])
  order = np.argsort(sizes, kind="stable")
  pick = {"smallest": int(order[0]), "median": int(order[len(order) // 2]),
          "largest": int(order[-1])}
  info = {k: b.item(i) for k, i in pick.items()}
  for name, i in pick.items():
    m, t, g = items[i]
    k = info[name]["sweeps"]
    print(f"batch256: {name} item {i}: {sizes[i]} px, {k} sweeps, "
          f"latency {info[name]['latency_s']:.3f} s", flush=True)
    want = oracle.grid_jacobi(m, t, g, k, "ulrd", threads=16)
    got = b.read(i, state=True)[2]
    assert np.array_equal(np.ascontiguousarray(got, np.float32).view(np.uint32),
                          want.view(np.uint32)), name
  assert info["median"]["latency_s"] < info["largest"]["latency_s"]
