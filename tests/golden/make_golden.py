"""Write the golden fixtures under tests/golden/ from the reference itself.

Run once in a container that holds the reference checkout (default
/root/reference, override with FPIE_REFERENCE) after ``oracle/build_ref.sh``:

    python tests/golden/make_golden.py

Sources of the expected outputs:

* ``core_gcc`` / ``core_openmp``: the reference's C++ solver cores compiled
  from their own sources by oracle/build_ref.sh (uint8 output + err of
  ``step``; the pybind11 interface exposes nothing else);
* ``fpie.np_solver``: the reference's numpy Jacobi solver, imported from the
  reference checkout; its raw float32 state (``.X`` / ``.tgt``) after
  ``step`` is recorded, because no other check pins the bits of numpy's
  g+U+D+L+R order.

The solver inputs (A, X, B / mask, tgt, grad) are assembled by the oracle's
pixel-loop restatement of fpie/process.py and stored beside the processor
inputs (src, mask, tgt, offsets, gradient mode), so tests can check both the
product's processor and its solver against them.  Only data is written:
every fixture is an .npz of arrays.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

REF = os.environ.get("FPIE_REFERENCE", "/root/reference")


def blobs(h, w, seed):
  rng = np.random.default_rng(seed)
  yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
  img = np.full((h, w, 3), 128.0)
  for _ in range(4):
    cy, cx = rng.uniform(0, h), rng.uniform(0, w)
    s = rng.uniform(h / 8, h / 3)
    amp = rng.uniform(-96, 96, 3)
    img += np.exp(-((yy - cy)**2 + (xx - cx)**2) / (2 * s * s))[..., None] * amp
  img += rng.integers(-8, 9, (h, w, 3))
  return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def cases():
  out = []
  # tests/test_smoke.py:19-24 (6x6, 2x2 mask, target 10), step(2)
  src = np.zeros((6, 6, 3), np.uint8)
  mask = np.zeros((6, 6), np.uint8)
  mask[2:4, 2:4] = 255
  tgt = np.ones((6, 6, 3), np.uint8) * 10
  out.append(("smoke6", src, mask, tgt, (0, 0), (0, 0), "max", 2))
  # tests/test_smoke.py:51-55 (seeded 24x24, 65 % mask), step(5)
  rng = np.random.default_rng(0)
  src = rng.integers(0, 256, size=(24, 24, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, size=(24, 24, 3), dtype=np.uint8)
  mask = np.zeros((24, 24), dtype=np.uint8)
  mask[2:-2, 2:-2] = (rng.random((20, 20)) > 0.35).astype(np.uint8) * 255
  out.append(("rng24", src, mask, tgt, (0, 0), (0, 0), "max", 5))
  # random sizes and offsets, every gradient mode, 3-channel masks
  rng = np.random.default_rng(7)
  for k, (h, w, mode, iters) in enumerate([(17, 23, "src", 7),
                                          (32, 29, "avg", 40),
                                          (40, 40, "max", 100)]):
    src = rng.integers(0, 256, (h + 6, w + 4, 3), dtype=np.uint8)
    tgt = rng.integers(0, 256, (h + 9, w + 7, 3), dtype=np.uint8)
    mask = np.zeros((h, w, 3), np.uint8)
    inner = rng.random((h - 2, w - 2)) > 0.4
    mask[1:-1, 1:-1] = (inner[..., None] * rng.integers(150, 256, 3)).astype(
        np.uint8)
    out.append((f"rand{k}", src, mask, tgt, (3, 2), (5, 4), mode, iters))
  # smooth blob images with a disc mask, 300 iterations
  h = w = 48
  src, tgt = blobs(h, w, 1001), blobs(h, w, 2001)
  yy, xx = np.mgrid[0:h, 0:w]
  mask = (((yy - 24)**2 + (xx - 24)**2) <= 18**2).astype(np.uint8) * 255
  out.append(("disc48", src, mask, tgt, (0, 0), (0, 0), "max", 300))
  return out


# "# of vars" published for the shipped benchmark images (EquSolver counts
# N = masked + 1, GridSolver counts the cropped bbox):
# docs/benchmark.md:70,84 (EquSolver) and :131,145 (GridSolver).
PUBLISHED = {
    "square": {"equ": [4097, 16385, 65537, 262145, 1048577],
               "grid": [4356, 16900, 66564, 264196, 1052676]},
    "circle": {"equ": [4256, 16676, 65972, 262338, 1049486],
               "grid": [5476, 21316, 84100, 335241, 1338649]},
}


def benchmark_masks():
  """Pack docs/_static/images/{square,circle}{6..10}.png (pure 0/255, equal
  channels) as bits, with the published variable counts."""
  from PIL import Image
  rec = {}
  for shape in ("square", "circle"):
    for k, x in enumerate(range(6, 11)):
      im = np.asarray(Image.open(
          os.path.join(REF, "docs", "_static", "images", f"{shape}{x}.png")))
      assert im.ndim == 3 and set(np.unique(im)) <= {0, 255}
      assert (im == im[..., :1]).all()
      rec[f"{shape}{x}_bits"] = np.packbits(im[..., 0] == 255)
      rec[f"{shape}{x}_shape"] = np.array(im.shape)
      rec[f"{shape}{x}_equ_vars"] = np.array(PUBLISHED[shape]["equ"][k])
      rec[f"{shape}{x}_grid_vars"] = np.array(PUBLISHED[shape]["grid"][k])
  path = os.path.join(HERE, "benchmark_masks.npz")
  np.savez_compressed(path, **rec)
  print(f"wrote {path} ({os.path.getsize(path)} B)")


def main():
  benchmark_masks()
  gcc, omp = O.ref_core("core_gcc"), O.ref_core("core_openmp")
  if gcc is None or omp is None:
    raise SystemExit("run oracle/build_ref.sh first")
  sys.path.insert(0, REF)
  from fpie import np_solver  # the reference's numpy solver (read-only tree)

  for name, src, mask, tgt, mos, mot, mode, iters in cases():
    rec = dict(src=src, mask_in=mask, tgt_in=tgt, mask_on_src=np.array(mos),
               mask_on_tgt=np.array(mot), mode=np.array(mode),
               iters=np.array(iters))
    # ---- Equ: gcc (row-major ids) and numpy share the system
    e = O.build_equ(src, mask, tgt, mos, mot, mode)
    rec.update(equ_mask=e["mask"], equ_ids=e["ids"], equ_A=e["A"],
               equ_X=e["X"], equ_B=e["B"])
    c = gcc.EquSolver()
    ids = np.asarray(c.partition(e["mask"]))
    assert (np.where(e["mask"] == 0, 0, ids) == e["ids"]).all()
    c.reset(e["max_id"], e["A"], e["X"], e["B"])
    x, err = c.step(iters)
    rec.update(equ_gcc_u8=np.array(x), equ_gcc_err=np.array(err))
    n = np_solver.EquSolver()
    n.reset(e["max_id"], e["A"].copy(), e["X"].copy(), e["B"].copy())
    xn, errn = n.step(iters)
    rec.update(equ_numpy_X=np.asarray(n.X, np.float32),
               equ_numpy_err=np.asarray(errn, np.float32))
    # ---- Equ: openmp (red-black ids)
    e2 = O.build_equ(src, mask, tgt, mos, mot, mode,
                     partition=O.partition_redblack)
    c = omp.EquSolver(4)
    c.partition(e2["mask"])
    c.reset(e2["max_id"], e2["A"], e2["X"], e2["B"])
    x, err = c.step(iters)
    rec.update(equ_rb_ids=e2["ids"], equ_rb_A=e2["A"], equ_rb_X=e2["X"],
               equ_rb_B=e2["B"], equ_openmp_u8=np.array(x),
               equ_openmp_err=np.array(err))
    # ---- Grid
    g = O.build_grid(src, mask, tgt, mos, mot, mode)
    rec.update(grid_mask=g["mask"], grid_tgt=g["tgt"], grid_grad=g["grad"],
               grid_window=np.array(g["window"]))
    c = gcc.GridSolver(8, 8)
    c.reset(0, g["mask"], g["tgt"], g["grad"])
    x, err = c.step(iters)
    rec.update(grid_gcc_u8=np.array(x), grid_gcc_err=np.array(err))
    c = omp.GridSolver(2, 3, 4)
    c.reset(0, g["mask"], g["tgt"], g["grad"])
    x, err = c.step(iters)
    rec.update(grid_openmp_u8=np.array(x), grid_openmp_err=np.array(err))
    n = np_solver.GridSolver()
    n.reset(0, g["mask"].copy(), g["tgt"].copy(), g["grad"].copy())
    xn, errn = n.step(iters)
    rec.update(grid_numpy_tgt=np.asarray(n.tgt, np.float32),
               grid_numpy_err=np.asarray(errn, np.float32))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **rec)
    print(f"wrote {path} ({os.path.getsize(path)} B)")


if __name__ == "__main__":
  main()
