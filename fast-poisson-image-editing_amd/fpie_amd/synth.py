"""Seeded synthetic composites for tests and benchmarks.

``blobs``: 128 plus 16 Gaussian blobs (centre uniform in the image, sigma
uniform in [20, 80] px, per-channel amplitude uniform in [-96, 96], each cut
off 4 sigma from its centre) plus integer noise uniform in [-8, 8], rounded
and clipped to uint8.  The blobs are separable outer products, so an image
costs O(H * W) work.

Workloads (seeds: source 1000 + i, target 2000 + i):

* ``equ512``   512 x 512, filled disc of radius 200 at the centre (125 629
  solved pixels), gradient ``max``, EquSolver, 5000 iterations;
* ``grid4096`` 4096 x 4096, centred square of side 3800 (14 440 000 solved
  pixels in a 3802 x 3802 crop), gradient ``src``, GridSolver, 5000
  iterations;
* ``grid8192`` 8192 x 8192, union of 6 random ellipses covering 60 +- 2 %
  of the image, gradient ``max``, GridSolver, 10000 iterations.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def blobs(h: int, w: int, seed: int, n_blobs: int = 16) -> np.ndarray:
  rng = np.random.default_rng(seed)
  planes = np.full((3, h, w), 128.0, np.float32)
  for _ in range(n_blobs):
    cy, cx = rng.uniform(0, h), rng.uniform(0, w)
    sigma = rng.uniform(20, 80)
    amp = rng.uniform(-96, 96, 3)
    # each blob is cut off 4 sigma from its centre on both axes
    r0, r1 = max(0, int(cy - 4 * sigma)), min(h, int(cy + 4 * sigma) + 1)
    c0, c1 = max(0, int(cx - 4 * sigma)), min(w, int(cx + 4 * sigma) + 1)
    if r0 >= r1 or c0 >= c1:
      continue
    gy = np.exp(-(np.arange(r0, r1) - cy)**2 / (2 * sigma * sigma))
    gx = np.exp(-(np.arange(c0, c1) - cx)**2 / (2 * sigma * sigma))
    g = np.outer(gy.astype(np.float32), gx.astype(np.float32))
    for c in range(3):
      planes[c, r0:r1, c0:c1] += np.float32(amp[c]) * g
  planes += rng.integers(-8, 9, (3, h, w), dtype=np.int8)
  np.rint(planes, out=planes)
  np.clip(planes, 0, 255, out=planes)
  return np.ascontiguousarray(planes.astype(np.uint8).transpose(1, 2, 0))


def disc_mask(h: int, w: int, r: float, cy: float | None = None,
              cx: float | None = None) -> np.ndarray:
  cy = h / 2 if cy is None else cy
  cx = w / 2 if cx is None else cx
  yy, xx = np.ogrid[0:h, 0:w]
  return (((yy - cy)**2 + (xx - cx)**2) <= r * r).astype(np.uint8) * 255


def square_mask(h: int, w: int, side: int) -> np.ndarray:
  m = np.zeros((h, w), np.uint8)
  r0, c0 = (h - side) // 2, (w - side) // 2
  m[r0:r0 + side, c0:c0 + side] = 255
  return m


def ellipse_mask(h: int, w: int, seed: int, n: int = 6,
                 coverage: float = 0.60, tol: float = 0.02) -> np.ndarray:
  """Union of n random ellipses, redrawn until the covered fraction is
  within coverage +- tol."""
  rng = np.random.default_rng(seed)
  yy, xx = np.ogrid[0:h, 0:w]
  for _ in range(1000):
    m = np.zeros((h, w), bool)
    for _ in range(n):
      cy, cx = rng.uniform(0.15, 0.85) * h, rng.uniform(0.15, 0.85) * w
      ay, ax = rng.uniform(0.12, 0.45) * h, rng.uniform(0.12, 0.45) * w
      th = rng.uniform(0, np.pi)
      c, s = np.cos(th), np.sin(th)
      u = (yy - cy) * c + (xx - cx) * s
      v = -(yy - cy) * s + (xx - cx) * c
      m |= (u / ay)**2 + (v / ax)**2 <= 1.0
    if abs(m.mean() - coverage) <= tol:
      return m.astype(np.uint8) * 255
  raise RuntimeError("no ellipse union within the coverage band")


@dataclass
class Workload:
  name: str
  method: str  # "equ" or "grid"
  src: np.ndarray
  mask: np.ndarray
  tgt: np.ndarray
  gradient: str
  iters: int
  mask_on_src: tuple[int, int] = (0, 0)
  mask_on_tgt: tuple[int, int] = (0, 0)


def equ512(i: int = 0) -> Workload:
  return Workload("equ512", "equ", blobs(512, 512, 1000 + i),
                  disc_mask(512, 512, 200, 256, 256),
                  blobs(512, 512, 2000 + i), "max", 5000)


def grid4096(i: int = 0) -> Workload:
  return Workload("grid4096", "grid", blobs(4096, 4096, 1000 + i),
                  square_mask(4096, 4096, 3800), blobs(4096, 4096, 2000 + i),
                  "src", 5000)


def grid8192(i: int = 0) -> Workload:
  return Workload("grid8192", "grid", blobs(8192, 8192, 1000 + i),
                  ellipse_mask(8192, 8192, 3000 + i),
                  blobs(8192, 8192, 2000 + i), "max", 10000)


WORKLOADS = {"equ512": equ512, "grid4096": grid4096, "grid8192": grid8192}
