"""CPU oracle for the Poisson image-editing solver path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product package never imports it.

Two parts:

* ctypes bindings of ``pie_oracle.c`` (the C restatement of the reference
  solver cores: gcc lexicographic Gauss-Seidel, openmp red-black / Jacobi,
  numpy Jacobi), and
* a loop-level restatement of the reference processor's preprocessing
  (mask preparation, gradient mixing, linear-system assembly), written pixel
  by pixel so that it shares no code with the product's vectorised version.

Reference citations are ``file:line`` in Trinkle23897/Fast-Poisson-Image-Editing
(fpie 0.3.2).  ``*Core`` classes expose the pybind11 solver-core interface of
``fpie/core/gcc/solver.cc:3-15`` so they plug into the product's processor
registry next to the ``hip`` backend.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpie_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build(force: bool = False) -> str:
  """Compile ``pie_oracle.c`` with gcc (oracle/Makefile)."""
  if force or not os.path.exists(LIB_PATH):
    subprocess.run(["make", "-s", "-B" if force else "all"], cwd=HERE,
                   check=True)
  return LIB_PATH


def lib() -> ctypes.CDLL:
  global _lib
  if _lib is None:
    if not os.path.exists(LIB_PATH):
      build()
    _lib = ctypes.CDLL(LIB_PATH)
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    ci = ctypes.c_int
    sig = {
        "po_partition_rowmajor": ([i32p, ci, ci, i32p], ci),
        "po_partition_redblack": ([i32p, ci, ci, i32p], ci),
        "po_equ_gs": ([ci, i32p, f32p, f32p, ci], None),
        "po_equ_redblack": ([ci, ci, i32p, f32p, f32p, ci, ci], None),
        "po_equ_jacobi": ([ci, i32p, f32p, f32p, ci, f32p], None),
        "po_equ_err_gcc": ([ci, i32p, f32p, f32p, f32p], None),
        "po_equ_err_openmp": ([ci, i32p, f32p, f32p, f32p], None),
        "po_equ_terms": ([ci, i32p, f32p, f32p, f64p], None),
        "po_grid_gs": ([ci, ci, i32p, f32p, f32p, ci, ci, ci], None),
        "po_grid_jacobi_ulrd": ([ci, ci, i32p, f32p, f32p, ci, f32p, ci],
                                None),
        "po_grid_jacobi_udlr": ([ci, ci, i32p, f32p, f32p, ci, f32p], None),
        "po_grid_err_gcc": ([ci, ci, i32p, f32p, f32p, f32p], None),
        "po_grid_err_openmp": ([ci, ci, i32p, f32p, f32p, f32p], None),
        "po_grid_terms": ([ci, ci, i32p, f32p, f32p, ci, f64p], None),
        "po_clamp_u8": ([f32p, ctypes.c_long, u8p], None),
        "po_grid_wavefront": ([ci, ci, i32p, f32p, f32p, ci, ci], None),
        "po_grid_colour": ([ci, ci, i32p, f32p, f32p, ci, ci], None),
    }
    for name, (args, res) in sig.items():
      fn = getattr(_lib, name)
      fn.argtypes = args
      fn.restype = res
  return _lib


def _c(a, dtype):
  return np.ascontiguousarray(a, dtype=dtype)


# --------------------------------------------------------------- solvers ---


def partition_rowmajor(mask: np.ndarray) -> np.ndarray:
  """fpie/core/gcc/equ.cc:14-32."""
  m = _c(mask, np.int32)
  ids = np.zeros_like(m)
  lib().po_partition_rowmajor(m, m.shape[0], m.shape[1], ids)
  return ids


def partition_redblack(mask: np.ndarray) -> tuple[np.ndarray, int]:
  """fpie/core/openmp/equ.cc:22-56; returns (ids, n_mid)."""
  m = _c(mask, np.int32)
  ids = np.zeros_like(m)
  n_mid = lib().po_partition_redblack(m, m.shape[0], m.shape[1], ids)
  return ids, int(n_mid)


def equ_gs(A, X, B, iters):
  """fpie/core/gcc/equ.cc:74-80 (returns the updated copy of X)."""
  X = _c(X, np.float32).copy()
  lib().po_equ_gs(X.shape[0], _c(A, np.int32), X, _c(B, np.float32), iters)
  return X


def equ_redblack(A, X, B, n_mid, iters, threads=1):
  """fpie/core/openmp/equ.cc:107-118."""
  X = _c(X, np.float32).copy()
  lib().po_equ_redblack(X.shape[0], n_mid, _c(A, np.int32), X,
                        _c(B, np.float32), iters, threads)
  return X


def equ_jacobi(A, X, B, iters):
  """fpie/np_solver.py:31-41."""
  X = _c(X, np.float32).copy()
  tmp = np.empty_like(X)
  lib().po_equ_jacobi(X.shape[0], _c(A, np.int32), X, _c(B, np.float32),
                      iters, tmp)
  return X


def equ_err(A, X, B, kind="gcc"):
  """fpie/core/gcc/equ.cc:56-72 (kind 'gcc') or openmp/equ.cc:82-105."""
  err = np.zeros(3, np.float32)
  fn = lib().po_equ_err_gcc if kind == "gcc" else lib().po_equ_err_openmp
  fn(len(X), _c(A, np.int32), _c(X, np.float32), _c(B, np.float32), err)
  return err


def equ_terms(A, X, B):
  out = np.zeros((len(X), 3), np.float64)
  lib().po_equ_terms(len(X), _c(A, np.int32), _c(X, np.float32),
                     _c(B, np.float32), out)
  return out


def grid_gs(mask, tgt, grad, iters, grid_x=8, grid_y=8):
  """fpie/core/gcc/grid.cc:57-75 (returns the updated copy of tgt)."""
  t = _c(tgt, np.float32).copy()
  m = _c(mask, np.int32)
  lib().po_grid_gs(m.shape[0], m.shape[1], m, t, _c(grad, np.float32), iters,
                   grid_x, grid_y)
  return t


def grid_jacobi(mask, tgt, grad, iters, order="ulrd", threads=1):
  """openmp/grid.cc:75-103 (order 'ulrd') or np_solver.py:79-88 ('udlr')."""
  t = _c(tgt, np.float32).copy()
  m = _c(mask, np.int32)
  tmp = np.zeros_like(t)
  if order == "ulrd":
    lib().po_grid_jacobi_ulrd(m.shape[0], m.shape[1], m, t,
                              _c(grad, np.float32), iters, tmp, threads)
  else:
    lib().po_grid_jacobi_udlr(m.shape[0], m.shape[1], m, t,
                              _c(grad, np.float32), iters, tmp)
  return t


def grid_err(mask, tgt, grad, kind="gcc"):
  """fpie/core/gcc/grid.cc:38-55 (kind 'gcc') or openmp/grid.cc:46-73."""
  err = np.zeros(3, np.float32)
  m = _c(mask, np.int32)
  fn = lib().po_grid_err_gcc if kind == "gcc" else lib().po_grid_err_openmp
  fn(m.shape[0], m.shape[1], m, _c(tgt, np.float32), _c(grad, np.float32),
     err)
  return err


def grid_terms(mask, tgt, grad, order="ulrd"):
  m = _c(mask, np.int32)
  out = np.zeros(m.shape + (3,), np.float64)
  lib().po_grid_terms(m.shape[0], m.shape[1], m, _c(tgt, np.float32),
                      _c(grad, np.float32), 1 if order == "ulrd" else 0, out)
  return out


def grid_wavefront(mask, tgt, grad, iters, order="ulrd"):
  t = _c(tgt, np.float32).copy()
  m = _c(mask, np.int32)
  lib().po_grid_wavefront(m.shape[0], m.shape[1], m, t, _c(grad, np.float32),
                          iters, 1 if order == "ulrd" else 0)
  return t


def grid_colour(mask, tgt, grad, iters, order="ulrd"):
  t = _c(tgt, np.float32).copy()
  m = _c(mask, np.int32)
  lib().po_grid_colour(m.shape[0], m.shape[1], m, t, _c(grad, np.float32),
                       iters, 1 if order == "ulrd" else 0)
  return t


def clamp_u8(x):
  """fpie/core/gcc/equ.cc:82-84: clamp to [0, 255], truncate."""
  x = _c(x, np.float32)
  out = np.empty(x.shape, np.uint8)
  lib().po_clamp_u8(x.reshape(-1), x.size, out.reshape(-1))
  return out


# ---------------------------------------- processor (loop restatement) ---


def prep_mask(mask: np.ndarray):
  """fpie/process.py:209-224 (Equ) and :338-351 (Grid).

  Channel mean >= 128, 1-px image edge cleared, bbox crop with a 1-px
  margin.  Returns (cropped int32 mask, x0, x1, y0, y1).
  """
  h, w = mask.shape[:2]
  m = np.zeros((h, w), np.int32)
  for i in range(h):
    for j in range(w):
      if mask.ndim == 3:
        v = sum(float(x) for x in mask[i, j]) / mask.shape[2]
      else:
        v = float(mask[i, j])
      m[i, j] = 1 if v >= 128 else 0
  for i in range(h):
    m[i, 0] = m[i, w - 1] = 0
  for j in range(w):
    m[0, j] = m[h - 1, j] = 0
  rows = [i for i in range(h) for j in range(w) if m[i, j]]
  cols = [j for i in range(h) for j in range(w) if m[i, j]]
  if not rows:
    raise ValueError("empty mask")
  x0, x1 = min(rows) - 1, max(rows) + 2
  y0, y1 = min(cols) - 1, max(cols) + 2
  return m[x0:x1, y0:y1].copy(), x0, x1, y0, y1


def mixgrad(a: float, b: float, mode: str) -> float:
  """fpie/process.py:113-122 for one element (float32 arithmetic)."""
  a, b = np.float32(a), np.float32(b)
  if mode == "src":
    return a
  if mode == "avg":
    return np.float32((a + b) / np.float32(2))
  return b if abs(a) < abs(b) else a


def build_equ(src, mask, tgt, mask_on_src, mask_on_tgt, mode,
              partition=partition_rowmajor):
  """fpie/process.py:192-271, one pixel at a time.

  Returns dict with ids, A, X, B, max_id and the target coordinates of each
  id (row-major id order of ``partition``).
  """
  m, x0, x1, y0, y1 = prep_mask(mask)
  so = (x0 + mask_on_src[0], y0 + mask_on_src[1])
  to = (x0 + mask_on_tgt[0], y0 + mask_on_tgt[1])
  ids = partition(m)
  if isinstance(ids, tuple):
    ids = ids[0]
  ids = np.where(m == 0, 0, ids).astype(np.int32)
  n = int(m.sum()) + 1
  coords = [None] * n
  for i in range(m.shape[0]):
    for j in range(m.shape[1]):
      if m[i, j]:
        coords[ids[i, j]] = (i, j)
  A = np.zeros((n, 4), np.int32)
  X = np.zeros((n, 3), np.float32)
  B = np.zeros((n, 3), np.float32)
  h_s, w_s = src.shape[:2]
  h_t, w_t = tgt.shape[:2]

  def px(img, i, j, hh, ww, c):
    # numpy fancy indexing wraps negative indices (fpie/process.py:230-239)
    return np.float32(img[i % hh if i < 0 else i, j % ww if j < 0 else j, c])

  for k in range(1, n):
    i, j = coords[k]
    si, sj = i + so[0], j + so[1]
    ti, tj = i + to[0], j + to[1]
    for c in range(3):
      g = np.float32(0)
      for di, dj in ((0, -1), (0, 1), (-1, 0), (1, 0)):  # L, R, U, D
        a = px(src, si, sj, h_s, w_s, c) - px(src, si + di, sj + dj, h_s, w_s,
                                              c)
        b = px(tgt, ti, tj, h_t, w_t, c) - px(tgt, ti + di, tj + dj, h_t, w_t,
                                              c)
        g = np.float32(g + mixgrad(a, b, mode))
      bnd = np.float64(g)
      for di, dj in ((-1, 0), (0, -1), (0, 1), (1, 0)):  # U, L, R, D
        if m[i + di, j + dj] == 0:
          bnd = np.float32(bnd + np.float64(px(tgt, ti + di, tj + dj, h_t,
                                               w_t, c)))
      B[k, c] = bnd
      X[k, c] = px(tgt, ti, tj, h_t, w_t, c)
    A[k] = [ids[i - 1, j], ids[i + 1, j], ids[i, j - 1], ids[i, j + 1]]
  tgt_index = (np.array([coords[k][0] + to[0] for k in range(1, n)]),
               np.array([coords[k][1] + to[1] for k in range(1, n)]))
  return dict(ids=ids, A=A, X=X, B=B, max_id=n, mask=m,
              tgt_index=tgt_index)


def build_grid(src, mask, tgt, mask_on_src, mask_on_tgt, mode):
  """fpie/process.py:321-386, one pixel at a time.

  Returns dict with mask (cropped int32), tgt (float32 crop), grad and the
  target window (x0, x1, y0, y1).
  """
  m, x0, x1, y0, y1 = prep_mask(mask)
  h, w = m.shape
  sc = src[mask_on_src[0] + x0:mask_on_src[0] + x1,
           mask_on_src[1] + y0:mask_on_src[1] + y1].astype(np.float32)
  tc = tgt[mask_on_tgt[0] + x0:mask_on_tgt[0] + x1,
           mask_on_tgt[1] + y0:mask_on_tgt[1] + y1].astype(np.float32)
  grad = np.zeros((h, w, 3), np.float32)
  for i in range(h):
    for j in range(w):
      for c in range(3):
        g = np.float32(0)
        # the four directional terms in the order of process.py:363-376
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
          ni, nj = i + di, j + dj
          if 0 <= ni < h and 0 <= nj < w:
            g = np.float32(g + mixgrad(sc[i, j, c] - sc[ni, nj, c],
                                       tc[i, j, c] - tc[ni, nj, c], mode))
        grad[i, j, c] = g if m[i, j] else np.float32(0)
  return dict(mask=m, tgt=tc, grad=grad,
              window=(mask_on_tgt[0] + x0, mask_on_tgt[0] + x1,
                      mask_on_tgt[1] + y0, mask_on_tgt[1] + y1))


# ------------------------------------------------ solver-core adapters ---


class EquCore:
  """Oracle EquSolver with the pybind11 interface of fpie/core/gcc/solver.cc.

  kind: 'gcc' (lexicographic GS, gcc/equ.cc), 'openmp' (red-black,
  openmp/equ.cc) or 'numpy' (Jacobi, np_solver.py).
  """

  def __init__(self, kind: str = "gcc"):
    self.kind = kind
    self.n_mid = 0

  def partition(self, mask):
    if self.kind == "openmp":
      ids, self.n_mid = partition_redblack(mask)
      return ids
    return partition_rowmajor(mask)

  def reset(self, N, A, X, B):
    self.N = int(N)
    self.A = _c(A, np.int32).copy()
    self.X = _c(X, np.float32).copy()
    self.B = _c(B, np.float32).copy()

  def sync(self):
    pass

  def step(self, iteration):
    if self.kind == "gcc":
      self.X = equ_gs(self.A, self.X, self.B, iteration)
    elif self.kind == "openmp":
      self.X = equ_redblack(self.A, self.X, self.B, self.n_mid, iteration)
    else:
      self.X = equ_jacobi(self.A, self.X, self.B, iteration)
    kind = "openmp" if self.kind == "openmp" else "gcc"
    return clamp_u8(self.X), equ_err(self.A, self.X, self.B, kind)


class GridCore:
  """Oracle GridSolver (pybind11 interface of fpie/core/gcc/solver.cc).

  kind: 'gcc' (lexicographic GS, gcc/grid.cc), 'openmp' (Jacobi, g+U+L+R+D,
  openmp/grid.cc) or 'numpy' (Jacobi, g+U+D+L+R, np_solver.py).
  """

  def __init__(self, grid_x: int = 8, grid_y: int = 8, kind: str = "gcc"):
    self.kind = kind
    self.grid_x, self.grid_y = grid_x, grid_y

  def reset(self, N, mask, tgt, grad):
    self.mask = _c(mask, np.int32).copy()
    self.tgt = _c(tgt, np.float32).copy()
    self.grad = _c(grad, np.float32).copy()

  def sync(self):
    pass

  def step(self, iteration):
    if self.kind == "gcc":
      self.tgt = grid_gs(self.mask, self.tgt, self.grad, iteration,
                         self.grid_x, self.grid_y)
    elif self.kind == "openmp":
      self.tgt = grid_jacobi(self.mask, self.tgt, self.grad, iteration,
                             "ulrd")
    else:
      self.tgt = grid_jacobi(self.mask, self.tgt, self.grad, iteration,
                             "udlr")
    kind = "openmp" if self.kind == "openmp" else "gcc"
    return clamp_u8(self.tgt), grid_err(self.mask, self.tgt, self.grad, kind)


def ref_core(name: str):
  """Load a reference core compiled by oracle/build_ref.sh, or None."""
  import importlib.util
  import glob
  hits = glob.glob(os.path.join(REF_DIR, f"{name}.*.so"))
  if not hits:
    return None
  spec = importlib.util.spec_from_file_location(name, hits[0])
  mod = importlib.util.module_from_spec(spec)
  spec.loader.exec_module(mod)
  return mod
