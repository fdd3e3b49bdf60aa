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
  of the image, gradient ``max``, GridSolver, 10000 iterations;
* ``batch256`` 256 composites, side log-uniform in [128, 2048], aspect
  ratio log-uniform in [0.5, 2], star polygon masks of 8-24 vertices
  covering 20-80 % of the image, gradient ``src`` or ``max`` per item, each
  stopped after 5000 sweeps or once a sweep changes no value by 1e-3.
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


def _ellipses(h, w, params):
  yy = np.arange(h, dtype=np.float32)[:, None]
  xx = np.arange(w, dtype=np.float32)[None, :]
  m = np.zeros((h, w), bool)
  for cy, cx, ay, ax, th in params:
    c, s = np.float32(np.cos(th)), np.float32(np.sin(th))
    dy, dx = yy - np.float32(cy * h), xx - np.float32(cx * w)
    u = (dy * c + dx * s) / np.float32(ay * h)
    v = (dx * c - dy * s) / np.float32(ax * w)
    m |= u * u + v * v <= 1.0
  return m


def ellipse_mask(h: int, w: int, seed: int, n: int = 6,
                 coverage: float = 0.60, tol: float = 0.02) -> np.ndarray:
  """Union of n random ellipses (centres in the middle 70 %, semi-axes 12-45
  % of the side, any orientation), parameters redrawn until the covered
  fraction, measured on a 256-px copy and then on the full grid, is within
  coverage +- tol."""
  rng = np.random.default_rng(seed)
  f = max(1, max(h, w) // 256)
  for _ in range(1000):
    params = [(rng.uniform(0.15, 0.85), rng.uniform(0.15, 0.85),
               rng.uniform(0.12, 0.45), rng.uniform(0.12, 0.45),
               rng.uniform(0, np.pi)) for _ in range(n)]
    if abs(_ellipses(max(8, h // f), max(8, w // f), params).mean() -
           coverage) > tol:
      continue
    m = _ellipses(h, w, params)
    if abs(m.mean() - coverage) <= tol:
      return m.astype(np.uint8) * 255
  raise RuntimeError("no ellipse union within the coverage band")


def _polygon_raster(h: int, w: int, py: np.ndarray,
                    px: np.ndarray) -> np.ndarray:
  """Even-odd fill of a closed polygon: pixel (y, x) is inside when an odd
  number of edges cross row y to the right of x.  Each edge crossing of a
  row becomes a count at column ceil(x_cross); a reverse cumulative sum
  along the row gives, per pixel, the crossings to its right."""
  y = np.arange(h, dtype=np.float64)[:, None]
  y0, x0 = py[None, :], px[None, :]
  y1, x1 = np.roll(py, -1)[None, :], np.roll(px, -1)[None, :]
  dy = np.where(y1 == y0, 1.0, y1 - y0)
  cross = (y >= np.minimum(y0, y1)) & (y < np.maximum(y0, y1)) & (y1 != y0)
  xs = x0 + (y - y0) * (x1 - x0) / dy
  col = np.clip(np.ceil(xs), 0, w).astype(np.int64)
  cnt = np.zeros((h, w + 1), np.int32)
  rows = np.broadcast_to(np.arange(h)[:, None], col.shape)
  np.add.at(cnt, (rows[cross], col[cross]), 1)
  right = np.cumsum(cnt[:, ::-1], axis=1)[:, ::-1]  # crossings at >= col
  return (right[:, 1:] & 1).astype(bool)  # pixel x: crossings at > x


def star_mask(h: int, w: int, rng: np.random.Generator,
              coverage: tuple[float, float] = (0.2, 0.8)) -> np.ndarray:
  """Filled star-shaped polygon (8-24 vertices at sorted random angles
  around the centre, radii uniform in [0.35, 1] of the largest radius),
  scaled so that it covers a fraction drawn uniformly from `coverage`."""
  n = int(rng.integers(8, 25))
  ang = np.sort(rng.uniform(0, 2 * np.pi, n))
  rad = rng.uniform(0.35, 1.0, n)
  want = rng.uniform(*coverage)

  def raster(hh, ww, scale):
    py = hh / 2 + scale * (hh / 2) * rad * np.sin(ang)
    px = ww / 2 + scale * (ww / 2) * rad * np.cos(ang)
    return _polygon_raster(hh, ww, py, px)

  # bisection on the star's size over a coarse copy of the image
  f = max(1, max(h, w) // 256)
  hc, wc = max(8, h // f), max(8, w // f)
  lo, hi = 0.05, 3.0
  for _ in range(24):
    mid = (lo + hi) / 2
    if raster(hc, wc, mid).mean() < want:
      lo = mid
    else:
      hi = mid
  return raster(h, w, (lo + hi) / 2).astype(np.uint8) * 255


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


def batch(n: int = 256, seed: int = 0, side: tuple[int, int] = (128, 2048),
          max_sweeps: int = 5000, tol: float = 1e-3) -> list[Workload]:
  """The batched workload: n independent GridSolver composites."""
  rng = np.random.default_rng(seed)
  out = []
  for i in range(n):
    s = float(np.exp(rng.uniform(np.log(side[0]), np.log(side[1]))))
    aspect = float(np.exp(rng.uniform(np.log(0.5), np.log(2.0))))
    h = max(16, int(round(s * np.sqrt(aspect))))
    w = max(16, int(round(s / np.sqrt(aspect))))
    mode = "src" if rng.random() < 0.5 else "max"
    item_rng = np.random.default_rng([seed, i])
    wl = Workload(f"batch{n}[{i}]", "grid", blobs(h, w, 1000 + i),
                  star_mask(h, w, item_rng), blobs(h, w, 2000 + i), mode,
                  max_sweeps)
    wl.tol = tol
    out.append(wl)
  return out


def composite(size: tuple[int, int], seed: int, gradient: str | None = None):
  """One GridSolver problem on a whole image (no crop): blob source and
  target, a star-polygon mask inside a 1-px border, gradient mode drawn per
  seed when not given.  Returns (mask int32 [h, w], target float32
  [h, w, 3], divergence float32 [h, w, 3], gradient mode); the divergence is
  the GridProcessor guidance (fpie/process.py:362-378)."""
  from . import process
  h, w = size
  rng = np.random.default_rng([7, seed])
  if gradient is None:
    gradient = "src" if rng.random() < 0.5 else "max"
  src = blobs(h, w, 10_000 + seed).astype(np.float32)
  tgt = blobs(h, w, 20_000 + seed).astype(np.float32)
  mask = (star_mask(h, w, rng) > 0).astype(np.int32)
  mask[[0, -1], :] = 0
  mask[:, [0, -1]] = 0
  grad = process.divergence(src, tgt, mask,
                            lambda a, b: process.mixgrad(a, b, gradient))
  return mask, tgt, grad, gradient


WORKLOADS = {"equ512": equ512, "grid4096": grid4096, "grid8192": grid8192}
