"""``hip`` solver cores with the reference's solver-core interface.

``EquSolver`` and ``GridSolver`` expose the methods that
fpie/process.py calls on ``core_gcc`` / ``core_cuda`` objects
(fpie/core/gcc/solver.cc:3-15, fpie/core/base_solver.h):

* ``EquSolver.partition(mask) -> int32[h, w]`` (gcc/equ.cc:14-32)
* ``EquSolver.reset(N, A, X, B)`` (base_solver.h:27-59)
* ``GridSolver.reset(N, mask, tgt, grad)`` (base_solver.h:101-140)
* ``sync()`` and ``step(iteration) -> (uint8 image, float32[3] err)``
  (gcc/equ.cc:74-86, gcc/grid.cc:57-81)

Arguments are converted to C-contiguous arrays of the reference's dtypes (the
pybind11 ``array_t`` force-cast) and handed to the C ABI; every array
returned is a new numpy array, as the pybind11 cores return copies.

Beyond the reference, the constructors take ``schedule`` ("jacobi",
"gs_lex" = the gcc backend's lexicographic Gauss-Seidel, "redblack" = the
openmp EquSolver's order) and, for the GridSolver, ``order`` ("ulrd" =
grad+up+left+right+down as gcc/openmp, "udlr" = grad+up+down+left+right as
numpy).  The EquSolver always adds b+up+down+left+right (gcc/equ.cc:48-53).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import native


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
  return ctypes.c_void_p(a.ctypes.data)


class _Core:
  _kind = native.PIE_GRID

  def __init__(self, schedule: str, order: str, a: int, b: int):
    if schedule not in native.SCHEDULES:
      raise ValueError(f"schedule must be one of {sorted(native.SCHEDULES)}")
    if order not in native.ORDERS:
      raise ValueError(f"order must be one of {sorted(native.ORDERS)}")
    lib = native.load()
    h = ctypes.c_void_p()
    native.check(lib.pie_create(self._kind, native.SCHEDULES[schedule],
                                native.ORDERS[order], int(a), int(b),
                                ctypes.byref(h)))
    self._h = h
    self._lib = lib
    self.schedule = schedule
    self.order = order

  def __del__(self):
    h = getattr(self, "_h", None)
    if h is not None and h.value:
      self._lib.pie_destroy(h)
      self._h = None

  def sync(self) -> None:
    native.check(self._lib.pie_sync(self._h))

  # ---- extensions (no counterpart in the reference) --------------------
  def iterate(self, iteration: int) -> None:
    """Queue sweeps on the device without reading anything back."""
    native.check(self._lib.pie_iterate(self._h, int(iteration)))

  def set_tile(self, sweeps: int = 0, cols_per_lane: int = 0,
               seg_rows: int = 0, channels: int = 0) -> None:
    """Fix the sweep tile (0 keeps a value); channels per wave task 3 or
    1 (0: the default of the column count)."""
    native.check(self._lib.pie_set_tile(self._h, sweeps, cols_per_lane,
                                        seg_rows))
    if channels:
      native.check(self._lib.pie_set_tile_channels(self._h, int(channels)))

  def tile_channels(self) -> int:
    """Colour channels per wave task of the current tile (3 or 1)."""
    v = ctypes.c_int()
    native.check(self._lib.pie_get_tile_channels(self._h, ctypes.byref(v)))
    return v.value

  def tile(self) -> tuple[int, int, int]:
    v = [ctypes.c_int() for _ in range(3)]
    native.check(self._lib.pie_get_tile(self._h, *(ctypes.byref(x)
                                                   for x in v)))
    return tuple(x.value for x in v)

  def autotune(self, launches: int = 3) -> tuple[int, int, int]:
    native.check(self._lib.pie_autotune(self._h, int(launches)))
    return self.tile()

  def set_timing(self, enable: bool = True) -> None:
    native.check(self._lib.pie_set_timing(self._h, int(bool(enable))))

  def timing(self) -> dict:
    ms = ctypes.c_double()
    n = ctypes.c_longlong()
    upd = ctypes.c_double()
    native.check(self._lib.pie_get_timing(self._h, ctypes.byref(ms),
                                          ctypes.byref(n), ctypes.byref(upd)))
    return {"kernel_ms": ms.value, "launches": n.value,
            "pixel_updates": upd.value}

  def shape(self) -> dict:
    h, w, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    masked = ctypes.c_longlong()
    native.check(self._lib.pie_shape(self._h, ctypes.byref(h),
                                     ctypes.byref(w), ctypes.byref(n),
                                     ctypes.byref(masked)))
    return {"h": h.value, "w": w.value, "n": n.value, "masked": masked.value}


# mixed-gradient modes of the device resets (fpie/process.py:113-122)
GRADIENT_MODES = {"src": 0, "avg": 1, "max": 2}


def prepare_mask(mask: np.ndarray):
  """The processors' mask preparation on the device (pie_prepare_mask):
  channel mean >= 128, outermost rows/columns cleared, crop to the bounding
  box plus a 1-px margin (fpie/process.py:209-221, 338-349).  mask: uint8
  [h, w] or [h, w, channels].  Returns (cropped int32 mask, x0, x1, y0, y1)
  like fpie_amd.process.prepare_mask; ValueError if nothing is solved."""
  m = np.ascontiguousarray(mask, dtype=np.uint8)
  if m.ndim not in (2, 3):
    raise ValueError("mask must be 2-D or 3-D")
  h, w = m.shape[:2]
  channels = m.shape[2] if m.ndim == 3 else 1
  lib = native.load()
  box = (ctypes.c_int * 4)()
  crop = np.empty(h * w, np.int32)
  rc = lib.pie_prepare_mask(_ptr(m), h, w, channels, box, _ptr(crop),
                            crop.size)
  if rc == native.PIE_EINVAL and b"no pixel" in lib.pie_last_error():
    raise ValueError("mask selects no pixel inside the image edge")
  native.check(rc)
  x0, x1, y0, y1 = (int(v) for v in box)
  return (crop[:(x1 - x0) * (y1 - y0)].reshape(x1 - x0, y1 - y0), x0, x1, y0,
          y1)


class EquSolver(_Core):
  """Equation-form solver (fpie/core/base_solver.h:12-75)."""

  _kind = native.PIE_EQU
  prepare_mask = staticmethod(prepare_mask)

  def __init__(self, block_size: int = 1024, schedule: str = "jacobi"):
    super().__init__(schedule, "udlr", block_size, 0)
    self.N = 0
    self.hw = None  # crop shape after reset_images

  def partition(self, mask: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(mask, dtype=np.int32)
    if m.ndim != 2:
      raise ValueError("mask must be 2-D")
    ids = np.empty_like(m)
    native.check(self._lib.pie_partition(self._h, _ptr(m), m.shape[0],
                                         m.shape[1], _ptr(ids)))
    return ids

  def reset(self, N: int, A: np.ndarray, X: np.ndarray,
            B: np.ndarray) -> None:
    N = int(N)
    A = np.ascontiguousarray(A, dtype=np.int32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    if A.shape != (N, 4) or X.shape != (N, 3) or B.shape != (N, 3):
      raise ValueError(f"expected A[{N},4], X[{N},3], B[{N},3]; got "
                       f"{A.shape}, {X.shape}, {B.shape}")
    native.check(self._lib.pie_reset_equ(self._h, N, _ptr(A), _ptr(X),
                                         _ptr(B)))
    self.N = N
    self.hw = None

  def reset_images(self, mask: np.ndarray, src: np.ndarray, tgt: np.ndarray,
                   gradient: str) -> int:
    """partition() + reset() built on the device from the uint8 source and
    target crops [h, w, 3] and the prepared crop mask [h, w] (extension):
    ids row-major from 1, X = the target, B = the mixed-gradient divergence
    plus the target of unsolved neighbours, exactly the system that
    EquProcessor.reset builds on the host (fpie/process.py:180-190,
    227-270).  Returns N (solved pixels + 1); step() then returns the N x 3
    values in id order."""
    m = np.ascontiguousarray(mask, dtype=np.int32)
    s8 = np.ascontiguousarray(src, dtype=np.uint8)
    t8 = np.ascontiguousarray(tgt, dtype=np.uint8)
    if m.ndim != 2 or s8.shape != m.shape + (3,) or t8.shape != m.shape + (3,):
      raise ValueError("src and tgt must be uint8 [h, w, 3] like the mask")
    if gradient not in GRADIENT_MODES:
      raise ValueError(f"gradient must be one of {list(GRADIENT_MODES)}")
    native.check(self._lib.pie_reset_equ_images(
        self._h, m.shape[0], m.shape[1], _ptr(m), _ptr(s8), _ptr(t8),
        GRADIENT_MODES[gradient]))
    self.N = int(np.count_nonzero(m > 0)) + 1
    self.hw = m.shape
    return self.N

  def step_crop(self, iteration: int) -> tuple[np.ndarray, np.ndarray]:
    """step() after reset_images, with the output as the uint8 crop
    [h, w, 3] assembled on the device (extension): the clamped values at
    solved pixels, the target crop elsewhere.  Returns (crop, err[3])."""
    if self.hw is None:
      raise RuntimeError("the crop output needs reset_images "
                         "(pie_reset_equ_images)")
    native.check(self._lib.pie_iterate(self._h, int(iteration)))
    crop = np.empty(self.hw + (3,), np.uint8)
    err = np.empty(3, np.float32)
    native.check(self._lib.pie_read_crop(self._h, _ptr(crop), _ptr(err)))
    return crop, err

  def step(self, iteration: int) -> tuple[np.ndarray, np.ndarray]:
    img = np.empty((self.N, 3), np.uint8)
    err = np.empty(3, np.float32)
    native.check(self._lib.pie_step(self._h, int(iteration), _ptr(img),
                                    _ptr(err)))
    return img, err

  def state(self) -> np.ndarray:
    """Float32 X[N][3] (extension)."""
    out = np.empty((self.N, 3), np.float32)
    native.check(self._lib.pie_get_state(self._h, _ptr(out)))
    return out


class GridSolver(_Core):
  """Grid-form solver (fpie/core/base_solver.h:77-152)."""

  _kind = native.PIE_GRID
  prepare_mask = staticmethod(prepare_mask)

  def __init__(self, grid_x: int = 8, grid_y: int = 8,
               schedule: str = "jacobi", order: str = "ulrd"):
    super().__init__(schedule, order, grid_x, grid_y)
    self.grid_x, self.grid_y = grid_x, grid_y
    self.hw = (0, 0)

  def reset(self, N: int, mask: np.ndarray, tgt: np.ndarray,
            grad: np.ndarray) -> None:
    # N is ignored; the shape comes from mask (base_solver.h:112-113)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    t = np.ascontiguousarray(tgt, dtype=np.float32)
    g = np.ascontiguousarray(grad, dtype=np.float32)
    if m.ndim != 2 or t.shape != m.shape + (3,) or g.shape != t.shape:
      raise ValueError(f"expected mask[h,w], tgt[h,w,3], grad[h,w,3]; got "
                       f"{m.shape}, {t.shape}, {g.shape}")
    native.check(self._lib.pie_reset_grid(self._h, m.shape[0], m.shape[1],
                                          _ptr(m), _ptr(t), _ptr(g)))
    self.hw = m.shape

  def step(self, iteration: int) -> tuple[np.ndarray, np.ndarray]:
    img = np.empty(self.hw + (3,), np.uint8)
    err = np.empty(3, np.float32)
    native.check(self._lib.pie_step(self._h, int(iteration), _ptr(img),
                                    _ptr(err)))
    return img, err

  def state(self) -> np.ndarray:
    """Float32 tgt[h][w][3] (extension)."""
    out = np.empty(self.hw + (3,), np.float32)
    native.check(self._lib.pie_get_state(self._h, _ptr(out)))
    return out

  GRADIENT_MODES = GRADIENT_MODES

  def reset_images(self, mask: np.ndarray, src: np.ndarray, tgt: np.ndarray,
                   gradient: str) -> None:
    """reset() built on the device from the uint8 source and target crops
    [h, w, 3] and the prepared crop mask [h, w] (extension): the target as
    the state, the mixed-gradient divergence as the guidance, exactly what
    GridProcessor.reset computes on the host (fpie/process.py:352-386)."""
    m = np.ascontiguousarray(mask, dtype=np.int32)
    s8 = np.ascontiguousarray(src, dtype=np.uint8)
    t8 = np.ascontiguousarray(tgt, dtype=np.uint8)
    if s8.shape != m.shape + (3,) or t8.shape != m.shape + (3,):
      raise ValueError("src and tgt must be uint8 [h, w, 3] like the mask")
    if gradient not in self.GRADIENT_MODES:
      raise ValueError(f"gradient must be one of {list(self.GRADIENT_MODES)}")
    native.check(self._lib.pie_reset_grid_images(
        self._h, m.shape[0], m.shape[1], _ptr(m), _ptr(s8), _ptr(t8),
        self.GRADIENT_MODES[gradient]))
    self.hw = m.shape

  def set_state(self, x: np.ndarray) -> None:
    """Replace the state (a warm start) by float32 x[h][w][3] (extension)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.shape != self.hw + (3,):
      raise ValueError(f"expected state of shape {self.hw + (3,)}")
    native.check(self._lib.pie_set_state(self._h, _ptr(x)))

  def solve_to_residual(self, tol_mean: float, check_every: int = 256,
                        max_sweeps: int = 10**7) -> tuple[int, float]:
    """Sweep until the mean absolute residual is <= tol_mean (extension).
    Returns (sweeps run, final mean residual)."""
    n = ctypes.c_longlong()
    r = ctypes.c_double()
    native.check(self._lib.pie_solve_to_residual(
        self._h, float(tol_mean), int(check_every), int(max_sweeps),
        ctypes.byref(n), ctypes.byref(r)))
    return n.value, r.value

  # coarse-grid correction (extensions used by the learned warm start)

  def coarse_shape(self, factor: int) -> tuple[int, int]:
    if int(factor) < 1:
      raise ValueError("factor must be >= 1")
    h, w = self.hw
    return h // factor, w // factor

  def restrict(self, factor: int) -> np.ndarray:
    """float32 [h // factor, w // factor, 3]: per factor x factor block, the
    sum of the signed residual (g + neighbours - 4 x) of the current state
    over the block's solved pixels."""
    out = np.empty(self.coarse_shape(factor) + (3,), np.float32)
    native.check(self._lib.pie_restrict(self._h, int(factor), _ptr(out)))
    return out

  def prolong(self, factor: int, corr: np.ndarray) -> None:
    """state += bilinear upsampling of corr [h // factor, w // factor, 3] at
    the solved pixels."""
    corr = np.ascontiguousarray(corr, dtype=np.float32)
    hc, wc = self.coarse_shape(factor)
    if corr.shape != (hc, wc, 3):
      raise ValueError(f"correction must be {(hc, wc, 3)}, not {corr.shape}")
    native.check(self._lib.pie_prolong(self._h, int(factor), hc, wc,
                                       _ptr(corr)))

  def prolong_dots(self, factor: int, corr: np.ndarray) -> np.ndarray:
    """float64 [2, 3]: per channel, sum r . d and sum d . A d for the
    correction d that prolong(factor, corr) would add (r the exact residual
    of the state); the energy-optimal scale of corr is row 0 / row 1."""
    corr = np.ascontiguousarray(corr, dtype=np.float32)
    hc, wc = self.coarse_shape(factor)
    if corr.shape != (hc, wc, 3):
      raise ValueError(f"correction must be {(hc, wc, 3)}, not {corr.shape}")
    out = np.zeros(6, np.float64)
    native.check(self._lib.pie_prolong_dots(self._h, int(factor), hc, wc,
                                            _ptr(corr), _ptr(out)))
    return out.reshape(2, 3)

  def warm_net_start(self, factor: int, mc: np.ndarray, net,
                     net_steps: int = 2, lowpass_sweeps: int = 4,
                     want: bool = False):
    """The learned coarse correction on the device (extension,
    pie_warm_net_start): restrict the residual, then `net_steps` flexible
    conjugate-gradient steps over the network's low-passed directions.
    mc: bool [hc, wc] coarse mask; net: a callable returning the batch-1
    fpie_amd.unet.UNet of the padded coarse shape.  The correction stays on
    the device for warm_apply(); with want=True a float32 copy
    [hc, wc, 3] is returned too."""
    hc, wc = self.coarse_shape(factor)
    m = np.ascontiguousarray(mc, dtype=np.uint8)
    if m.shape != (hc, wc):
      raise ValueError(f"coarse mask must be {(hc, wc)}, not {m.shape}")
    u = net()
    corr = np.empty((hc, wc, 3), np.float32) if want else None
    native.check(self._lib.pie_warm_net_start(
        self._h, u._h, int(factor), _ptr(m), int(net_steps),
        int(lowpass_sweeps), _ptr(corr) if want else None))
    return corr

  def warm_apply(self, factor: int, corr: np.ndarray | None = None):
    """state += alpha * prolong(corr) with the per-channel energy-optimal
    alpha >= 0 (extension, pie_warm_apply); corr None uses the correction
    warm_net_start left on the device.  Returns (alpha [3], max |alpha
    corr|)."""
    c = None
    if corr is not None:
      c = np.ascontiguousarray(corr, dtype=np.float32)
      if c.shape != self.coarse_shape(factor) + (3,):
        raise ValueError("correction must be [hc, wc, 3] of the factor")
    alpha = np.zeros(3, np.float64)
    size = ctypes.c_double()
    native.check(self._lib.pie_warm_apply(self._h, int(factor),
                                          _ptr(c) if c is not None else None,
                                          _ptr(alpha), ctypes.byref(size)))
    return alpha, size.value

  def relax(self, sweeps: int = 1, omega: float = 0.5) -> None:
    """`sweeps` weighted Jacobi sweeps x <- x + omega ((g + sum) / 4 - x)."""
    native.check(self._lib.pie_relax(self._h, int(sweeps), float(omega)))


class BatchSolver:
  """Batched GridSolver (extension): composites of uneven size and mask,
  each with its own stop rule, solved by one persistent Jacobi kernel (small
  items each by one work-group, several launches a round; the others over a
  shared tile queue) with the round loop in a hipGraph."""

  def __init__(self, order: str = "ulrd", sweeps_per_launch: int = 8,
               cols_per_lane: int = 1, rows_per_task: int = 64,
               channels: int = 0):
    lib = native.load()
    h = ctypes.c_void_p()
    native.check(lib.pie_batch_create(native.ORDERS[order], sweeps_per_launch,
                                      cols_per_lane, rows_per_task,
                                      ctypes.byref(h)))
    self._h, self._lib = h, lib
    self.shapes: list[tuple[int, int]] = []
    if channels:
      self.set_tile_channels(channels)

  def set_tile_channels(self, channels: int) -> None:
    """Channels per wave task (1 or 3; 0 = the tile's default); the results
    do not depend on it."""
    native.check(self._lib.pie_batch_set_tile_channels(self._h, channels))

  def autotune(self, rounds: int = 6) -> tuple[int, int, int, int]:
    """Time every tile for a few rounds from the initial state and keep the
    fastest; returns it (see tile())."""
    native.check(self._lib.pie_batch_autotune(self._h, int(rounds)))
    return self.tile()

  def tile(self) -> tuple[int, int, int, int]:
    """(sweeps per launch, columns per lane, rows per task, channels)."""
    v = [ctypes.c_int() for _ in range(4)]
    native.check(self._lib.pie_batch_tile(self._h, *(ctypes.byref(x)
                                                     for x in v)))
    return tuple(x.value for x in v)

  def __del__(self):
    h = getattr(self, "_h", None)
    if h is not None and h.value:
      self._lib.pie_batch_destroy(h)
      self._h = None

  def add(self, mask: np.ndarray, tgt: np.ndarray, grad: np.ndarray,
          max_sweeps: int = 5000, tol: float = 1e-3) -> int:
    m = np.ascontiguousarray(mask, dtype=np.int32)
    t = np.ascontiguousarray(tgt, dtype=np.float32)
    g = np.ascontiguousarray(grad, dtype=np.float32)
    if m.ndim != 2 or t.shape != m.shape + (3,) or g.shape != t.shape:
      raise ValueError("expected mask[h,w], tgt[h,w,3], grad[h,w,3]")
    idx = ctypes.c_int()
    native.check(self._lib.pie_batch_add(self._h, m.shape[0], m.shape[1],
                                         _ptr(m), _ptr(t), _ptr(g),
                                         int(max_sweeps), float(tol),
                                         ctypes.byref(idx)))
    self.shapes.append(m.shape)
    return idx.value

  def solve(self, rounds_per_graph: int = 16) -> tuple[float, int]:
    sec = ctypes.c_double()
    rounds = ctypes.c_int()
    native.check(self._lib.pie_batch_solve(self._h, int(rounds_per_graph),
                                           ctypes.byref(sec),
                                           ctypes.byref(rounds)))
    return sec.value, rounds.value

  def item(self, i: int) -> dict:
    sw, dm, lat = ctypes.c_int(), ctypes.c_float(), ctypes.c_double()
    native.check(self._lib.pie_batch_item(self._h, i, ctypes.byref(sw),
                                          ctypes.byref(dm),
                                          ctypes.byref(lat)))
    return {"sweeps": sw.value, "last_dmax": dm.value,
            "latency_s": lat.value}

  def read(self, i: int, state: bool = False):
    h, w = self.shapes[i]
    img = np.empty((h, w, 3), np.uint8)
    err = np.empty(3, np.float32)
    st = np.empty((h, w, 3), np.float32) if state else None
    native.check(self._lib.pie_batch_read(
        self._h, i, _ptr(img), _ptr(err),
        _ptr(st) if state else ctypes.c_void_p()))
    return (img, err, st) if state else (img, err)

  def size(self) -> dict:
    n, t, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    native.check(self._lib.pie_batch_size(self._h, ctypes.byref(n),
                                          ctypes.byref(t), ctypes.byref(m)))
    return {"items": n.value, "tasks": t.value, "masked": m.value}
