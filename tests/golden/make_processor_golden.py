"""Write tests/golden/processor_cases.npz from the reference processor itself.

The processors' preprocessing (mask preparation, crop offsets, the three
gradient modes, the EquSolver system A/X/B with its boundary terms, the
GridSolver crops and guidance: fpie/process.py:192-271, 321-386) is the
input every solver core receives.  The reference's fpie.process imports with
numpy alone, so this script, run once in a container that holds the
reference checkout (default /root/reference, override with FPIE_REFERENCE),
drives the reference EquProcessor / GridProcessor with a recording core and
stores exactly what they pass to the core:

    python tests/golden/make_processor_golden.py

Per case: the processor inputs (src, mask, tgt, offsets, gradient mode) and
the recorded core calls: Equ ``partition`` input mask and ``reset(N, A, X,
B)``; Grid ``reset(N, mask, tgt, grad)``.  Only data is written.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("FPIE_REFERENCE", "/root/reference")


class RecordingEqu:
  """Stands in for a solver core; row-major ids as every reference core
  but openmp numbers them (fpie/core/gcc/equ.cc:14-32)."""

  def partition(self, mask):
    self.part_mask = np.array(mask, np.int32)
    flat = (mask > 0).reshape(-1)
    return np.cumsum(flat).reshape(mask.shape)

  def reset(self, N, A, X, B):
    self.N, self.A, self.X, self.B = int(N), np.array(A), np.array(X), \
        np.array(B)

  def sync(self):
    pass


class RecordingGrid:
  def reset(self, N, mask, tgt, grad):
    self.N, self.mask, self.tgt, self.grad = int(N), np.array(mask), \
        np.array(tgt), np.array(grad)

  def sync(self):
    pass


def cases():
  rng = np.random.default_rng(2024)
  out = []
  # (name, mask channels, src/tgt extra size, offsets, mode, what)
  spec = [("src3", 3, (6, 4), (3, 2), (5, 4), "src", "both"),
          ("avg2d", 0, (9, 8), (0, 7), (9, 1), "avg", "both"),
          ("max3", 3, (7, 9), (4, 4), (2, 6), "max", "both"),
          ("max1", 1, (3, 3), (1, 0), (0, 2), "max", "both"),
          ("srcwide", 3, (2, 30), (1, 11), (0, 25), "src", "both"),
          # a source offset that makes numpy indices negative: the
          # reference reads wrapped rows (no bounds checks, process.py:
          # 202-207); the Equ system pins that behaviour
          ("wrap", 3, (5, 5), (-1, 0), (2, 1), "max", "equ")]
  for name, ch, extra, mos, mot, mode, what in spec:
    h, w = int(rng.integers(18, 40)), int(rng.integers(18, 40))
    shape = (h, w) if ch == 0 else (h, w, ch)
    vals = rng.integers(110, 256, shape)
    keep = rng.random((h, w)) < 0.6
    mask = np.where(keep if ch == 0 else keep[..., None], vals, 0).astype(
        np.uint8)
    if name == "wrap":
      mask[:] = 0
      mask[1:4, 3:12] = 255  # solved pixels on row 1: x0 = 0
    src = rng.integers(0, 256, (h + extra[0], w + extra[1], 3), np.uint8)
    tgt = rng.integers(0, 256, (h + extra[0] + 6, w + extra[1] + 6, 3),
                       np.uint8)
    out.append((name, src, mask, tgt, mos, mot, mode, what))
  return out


def main():
  sys.path.insert(0, REF)
  from fpie import process  # the reference processor (read-only tree)

  rec = {}
  names = []
  for name, src, mask, tgt, mos, mot, mode, what in cases():
    names.append(name)
    rec[f"{name}_src"] = src
    rec[f"{name}_mask"] = mask
    rec[f"{name}_tgt"] = tgt
    rec[f"{name}_mask_on_src"] = np.array(mos)
    rec[f"{name}_mask_on_tgt"] = np.array(mot)
    rec[f"{name}_mode"] = np.array(mode)
    rec[f"{name}_what"] = np.array(what)
    p = process.EquProcessor(mode, "numpy")
    p.core = RecordingEqu()
    n = p.reset(src, mask, tgt, mos, mot)
    rec[f"{name}_equ_n"] = np.array(n)
    rec[f"{name}_equ_mask"] = p.core.part_mask
    rec[f"{name}_equ_A"] = p.core.A.astype(np.int32)
    rec[f"{name}_equ_X"] = p.core.X.astype(np.float32)
    rec[f"{name}_equ_B"] = p.core.B.astype(np.float32)
    assert p.core.A.dtype == np.int32 and p.core.X.dtype == np.float32
    if what == "both":
      g = process.GridProcessor(mode, "numpy")
      g.core = RecordingGrid()
      n = g.reset(src, mask, tgt, mos, mot)
      rec[f"{name}_grid_n"] = np.array(n)
      rec[f"{name}_grid_mask"] = g.core.mask.astype(np.int32)
      rec[f"{name}_grid_tgt"] = g.core.tgt.astype(np.float32)
      rec[f"{name}_grid_grad"] = g.core.grad.astype(np.float32)
      assert g.core.tgt.dtype == np.float32 and g.core.grad.dtype == np.float32
    print(f"{name}: equ N={rec[f'{name}_equ_n']}", flush=True)
  rec["names"] = np.array(names)
  path = os.path.join(HERE, "processor_cases.npz")
  np.savez_compressed(path, **rec)
  print(f"wrote {path} ({os.path.getsize(path)} B)")


if __name__ == "__main__":
  main()
