"""Processor layer: the caller of the solver cores.

Mirrors the interface of fpie/process.py (``EquProcessor``, ``GridProcessor``,
``BaseProcessor.mixgrad/reset/sync/step``, ``ALL_BACKEND``,
``DEFAULT_BACKEND``, ``CPU_COUNT``) with a backend registry in place of the
reference's chain of try-imports (process.py:34-83): the ``hip`` backend
registers itself when ``libfpie_hip.so`` loads, and other solver cores with
the same duck-typed interface can be added with ``register_backend``.

The preprocessing is vectorised numpy and produces exactly the reference's
solver inputs:

* mask: channel mean >= 128, 1-px image edge cleared, bbox crop with a
  1-px margin (process.py:209-224, :338-351);
* gradient mixing ``src`` / ``avg`` / ``max`` (|src| < |tgt| strictly picks
  the target gradient; process.py:113-122);
* EquSolver system: A = ids of the up/down/left/right neighbours, X = target
  pixels, B = mixed divergence plus the unsolved neighbours' target values
  added in the order up, left, right, down (process.py:227-270);
* GridSolver inputs: cropped mask, float32 target crop and divergence with
  zeros off the mask (process.py:352-386).

All values are integers or halves below 2^11, so every float32 sum here is
exact; only the solver iterations round.
"""

from __future__ import annotations

import os
from abc import ABC, abstractmethod
from typing import Any, Callable

import numpy as np

from . import native

DEFAULT_CPU_CAP = 8


def _default_cpu_count() -> int:
  """min(cpu_count, affinity, 8), as fpie/process.py:14-26."""
  counts = [c for c in (os.cpu_count(),) if c]
  try:
    counts.append(len(os.sched_getaffinity(0)))
  except (AttributeError, OSError):
    pass
  return max(1, min(min(counts), DEFAULT_CPU_CAP)) if counts else 1


CPU_COUNT = _default_cpu_count()

# name -> (EquSolver factory, GridSolver factory, install hint)
_BACKENDS: dict[str, tuple[Callable[..., Any], Callable[..., Any], str]] = {}
ALL_BACKEND: list[str] = []
DEFAULT_BACKEND = "hip"


def register_backend(name: str, equ_factory: Callable[..., Any],
                     grid_factory: Callable[..., Any],
                     hint: str = "") -> None:
  """Add a solver core.  Factories receive the processor's keyword
  arguments (n_cpu, min_interval, block_size, grid_x, grid_y, schedule,
  order) and return an object with the reference core interface."""
  _BACKENDS[name] = (equ_factory, grid_factory, hint)
  if name not in ALL_BACKEND:
    ALL_BACKEND.append(name)


def _register_hip() -> None:
  if not native.available():
    return
  from . import core_hip

  register_backend(
      "hip",
      lambda **kw: core_hip.EquSolver(kw.get("block_size", 1024),
                                      kw.get("schedule", "jacobi")),
      lambda **kw: core_hip.GridSolver(kw.get("grid_x", 8),
                                       kw.get("grid_y", 8),
                                       kw.get("schedule", "jacobi"),
                                       kw.get("order", "ulrd")),
      "Please build libfpie_hip.so "
      "(python fast-poisson-image-editing_amd/build.py) on a ROCm host.")


_register_hip()


def prepare_mask(mask: np.ndarray):
  """Binarise, clear the image edge and crop to the bbox plus a 1-px margin.

  Returns (cropped int32 mask, x0, x1, y0, y1) with the crop window
  [x0, x1) x [y0, y1) in mask coordinates.
  """
  m = mask.mean(-1) if mask.ndim == 3 else mask
  m = (m >= 128).astype(np.int32)
  m[[0, -1], :] = 0
  m[:, [0, -1]] = 0
  rows = np.flatnonzero(m.any(axis=1))
  cols = np.flatnonzero(m.any(axis=0))
  if rows.size == 0:
    raise ValueError("mask selects no pixel inside the image edge")
  x0, x1 = int(rows[0]) - 1, int(rows[-1]) + 2
  y0, y1 = int(cols[0]) - 1, int(cols[-1]) + 2
  return m[x0:x1, y0:y1], x0, x1, y0, y1


def mixgrad(a: np.ndarray, b: np.ndarray, gradient: str) -> np.ndarray:
  """Source gradient ('src'), mean of both ('avg'), or the larger in
  magnitude ('max', ties keep the source), elementwise in float32
  (fpie/process.py:113-122)."""
  if gradient == "src":
    return a
  if gradient == "avg":
    return (a + b) / 2
  return np.where(np.abs(a) < np.abs(b), b, a)


# neighbour offsets (drow, dcol)
_UP, _DOWN, _LEFT, _RIGHT = (-1, 0), (1, 0), (0, -1), (0, 1)


def divergence(src: np.ndarray, tgt: np.ndarray, mask: np.ndarray,
               mixgrad) -> np.ndarray:
  """GridSolver guidance (fpie/process.py:362-378): at every solved pixel
  the sum over its four neighbours of mixgrad(src_c - src_n, tgt_c - tgt_n),
  accumulated up, down, left, right; zero off the mask.  src and tgt are
  float32 [h, w, 3]; pixels without a neighbour in a direction get no term
  for it."""
  h, w = mask.shape
  grad = np.zeros((h, w, 3), np.float32)
  for d in (_UP, _DOWN, _LEFT, _RIGHT):
    rs = slice(max(0, -d[0]), h - max(0, d[0]))
    cs = slice(max(0, -d[1]), w - max(0, d[1]))
    rn = slice(rs.start + d[0], rs.stop + d[0])
    cn = slice(cs.start + d[1], cs.stop + d[1])
    grad[rs, cs] += mixgrad(src[rs, cs] - src[rn, cn],
                            tgt[rs, cs] - tgt[rn, cn])
  grad[mask == 0] = 0
  return grad


class BaseProcessor(ABC):
  """Common processor interface (fpie/process.py:86-143)."""

  def __init__(self, gradient: str, rank: int, backend: str,
               core: Any | None):
    if core is None:
      hint = _BACKENDS.get(backend, (None, None, ""))[2]
      print(hint or f"Backend {backend!r} is not available; "
            f"available: {ALL_BACKEND}")
      raise AssertionError(f"Invalid backend {backend}.")
    if gradient not in ("max", "src", "avg"):
      raise ValueError(f"gradient must be max, src or avg, not {gradient!r}")
    self.gradient = gradient
    self.rank = rank
    self.backend = backend
    self.core = core
    self.root = rank == 0

  @staticmethod
  def _make_core(backend: str, which: int, **kw) -> Any | None:
    entry = _BACKENDS.get(backend)
    return entry[which](**kw) if entry is not None else None

  def mixgrad(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return mixgrad(a, b, self.gradient)

  def prepare_mask(self, mask: np.ndarray):
    """prepare_mask on the core's device when it offers it (uint8 masks,
    the images the command line reads), else on the host; same result."""
    if (hasattr(self.core, "prepare_mask") and
        np.asarray(mask).dtype == np.uint8 and np.ndim(mask) in (2, 3)):
      return self.core.prepare_mask(mask)
    return prepare_mask(mask)

  @abstractmethod
  def reset(self, src: np.ndarray, mask: np.ndarray, tgt: np.ndarray,
            mask_on_src: tuple[int, int],
            mask_on_tgt: tuple[int, int]) -> int:
    """Load a new blend; returns the number of variables."""

  def sync(self) -> None:
    self.core.sync()

  @abstractmethod
  def step(self, iteration: int) -> tuple[np.ndarray, np.ndarray] | None:
    """Run iterations; returns (blended uint8 image, float32[3] err)."""


class EquProcessor(BaseProcessor):
  """Equation-form processor (fpie/process.py:146-280)."""

  def __init__(self, gradient: str = "max", backend: str = DEFAULT_BACKEND,
               n_cpu: int = CPU_COUNT, min_interval: int = 100,
               block_size: int = 1024, schedule: str = "jacobi"):
    core = self._make_core(backend, 0, n_cpu=n_cpu,
                           min_interval=min_interval, block_size=block_size,
                           schedule=schedule)
    super().__init__(gradient, 0, backend, core)

  def mask2index(self, mask: np.ndarray):
    """ids from the core's partition (0 off the mask) and the pixel
    coordinates sorted by id (fpie/process.py:180-190)."""
    xs, ys = np.nonzero(mask)
    ids = np.asarray(self.core.partition(mask)).copy()
    ids[mask == 0] = 0
    by_id = np.argsort(ids[xs, ys], kind="stable")
    return ids, xs.size + 1, xs[by_id], ys[by_id]

  def reset(self, src, mask, tgt, mask_on_src, mask_on_tgt) -> int:
    m, x0, x1, y0, y1 = self.prepare_mask(mask)
    so = (x0 + mask_on_src[0], y0 + mask_on_src[1])
    to = (x0 + mask_on_tgt[0], y0 + mask_on_tgt[1])
    if (hasattr(self.core, "reset_images") and src.dtype == np.uint8
        and tgt.dtype == np.uint8 and src.ndim == 3 and tgt.ndim == 3
        and min(so + to) >= 0):
      # the core numbers the pixels and builds X and B on the device from
      # the uint8 crops: the same system as below (ids row-major, so the
      # coordinates in id order are np.nonzero's)
      h, w = m.shape
      s_crop = src[so[0]:so[0] + h, so[1]:so[1] + w]
      t_crop = tgt[to[0]:to[0] + h, to[1]:to[1] + w]
      if s_crop.shape[:2] == (h, w) and t_crop.shape[:2] == (h, w):
        n = self.core.reset_images(m, s_crop, t_crop, self.gradient)
        ix, iy = np.nonzero(m)
        self.tgt = tgt.copy()
        self.tgt_index = (ix + to[0], iy + to[1])
        # step() writes the crop the device assembles (a slice copy, not a
        # scatter through the index arrays)
        self.crop_window = (to[0], to[0] + h, to[1], to[1] + w)
        return int(n)
    self.crop_window = None
    ids, n, ix, iy = self.mask2index(m)
    sx, sy = ix + so[0], iy + so[1]
    tx, ty = ix + to[0], iy + to[1]

    def at(img, px, py, d):
      return img[px + d[0], py + d[1]].astype(np.float32)

    s_c, t_c = at(src, sx, sy, (0, 0)), at(tgt, tx, ty, (0, 0))
    grad = np.zeros((n - 1, 3), np.float32)
    for d in (_LEFT, _RIGHT, _UP, _DOWN):  # process.py:241-246
      grad = grad + self.mixgrad(s_c - at(src, sx, sy, d),
                                 t_c - at(tgt, tx, ty, d))

    A = np.zeros((n, 4), np.int32)
    X = np.zeros((n, 3), np.float32)
    B = np.zeros((n, 3), np.float32)
    X[1:] = tgt[tx, ty]
    for k, d in enumerate((_UP, _DOWN, _LEFT, _RIGHT)):
      A[1:, k] = ids[ix + d[0], iy + d[1]]
    B[1:] = grad
    for d in (_UP, _LEFT, _RIGHT, _DOWN):  # process.py:259-266
      outside = (m[ix + d[0], iy + d[1]] == 0).astype(np.float64)[:, None]
      B[1:] = (B[1:].astype(np.float64) +
               outside * tgt[tx + d[0], ty + d[1]]).astype(np.float32)

    self.tgt = tgt.copy()
    self.tgt_index = (tx, ty)
    self.core.reset(n, A, X, B)
    return n

  def step(self, iteration: int):
    if getattr(self, "crop_window", None) is not None:
      crop, err = self.core.step_crop(iteration)
      x0, x1, y0, y1 = self.crop_window
      self.tgt[x0:x1, y0:y1] = crop
      return self.tgt, err
    x, err = self.core.step(iteration)
    self.tgt[self.tgt_index] = x[1:]
    return self.tgt, err


class GridProcessor(BaseProcessor):
  """Grid-form processor (fpie/process.py:283-395)."""

  def __init__(self, gradient: str = "max", backend: str = DEFAULT_BACKEND,
               n_cpu: int = CPU_COUNT, min_interval: int = 100,
               block_size: int = 1024, grid_x: int = 8, grid_y: int = 8,
               schedule: str = "jacobi", order: str = "ulrd"):
    core = self._make_core(backend, 1, n_cpu=n_cpu,
                           min_interval=min_interval, block_size=block_size,
                           grid_x=grid_x, grid_y=grid_y, schedule=schedule,
                           order=order)
    super().__init__(gradient, 0, backend, core)

  def reset(self, src, mask, tgt, mask_on_src, mask_on_tgt) -> int:
    m, x0, x1, y0, y1 = self.prepare_mask(mask)
    h, w = m.shape
    s_crop = src[mask_on_src[0] + x0:mask_on_src[0] + x1,
                 mask_on_src[1] + y0:mask_on_src[1] + y1]
    t_crop = tgt[mask_on_tgt[0] + x0:mask_on_tgt[0] + x1,
                 mask_on_tgt[1] + y0:mask_on_tgt[1] + y1]
    self.x0, self.x1 = mask_on_tgt[0] + x0, mask_on_tgt[0] + x1
    self.y0, self.y1 = mask_on_tgt[1] + y0, mask_on_tgt[1] + y1
    self.tgt = tgt.copy()
    n = h * w
    if (hasattr(self.core, "reset_images") and s_crop.dtype == np.uint8
        and t_crop.dtype == np.uint8 and s_crop.ndim == 3):
      # the core builds target and divergence on the device from the uint8
      # crops: the same values as below, a quarter of the bytes to copy
      self.core.reset_images(m, s_crop, t_crop, self.gradient)
      return int(n)
    sc = s_crop.astype(np.float32)
    tc = t_crop.astype(np.float32)
    grad = divergence(sc, tc, m, self.mixgrad)
    self.core.reset(n, m, tc, grad)
    return int(n)

  def step(self, iteration: int):
    crop, err = self.core.step(iteration)
    self.tgt[self.x0:self.x1, self.y0:self.y1] = crop
    return self.tgt, err
