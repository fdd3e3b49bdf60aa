"""Learned warm start for the GridSolver (no counterpart in the reference).

A 4-level U-Net (base 32 channels, 3x3 convolutions, bf16 weights and
activations, fp32 accumulation on MFMA; kernels in csrc/unet_kernels.hip,
runtime in csrc/unet.hip) predicts the solution of a Poisson blend from

  channel 0-2  divergence (GridSolver guidance) / 64
  channel 3-5  (target - 128) / 64   (the boundary-extended target, which is
               also the cold-start state)
  channel 6    mask (1 = solved)
  channel 7    0

as a correction: x0 = target + 64 * net(input) on the solved pixels, target
elsewhere.  It is trained (Adam, masked mean squared error) on composites
whose labels are the Jacobi solver's own long runs, and used as the start
state of plain Jacobi, which then finishes to the requested residual.
Fallback: if the predicted start has a larger residual than the cold start,
the cold start is used.

Layers (name, in, out channels), resolution level in parentheses:
  e0a 8-32, e0b 32-32 (0) | pool | e1a 32-64, e1b 64-64 (1) | pool |
  e2a 64-128, e2b 128-128 (2) | pool | bna 128-256, bnb 256-256 (3)
  up(bnb) ++ e2b: d2a 384-128, d2b 128-128 (2)
  up(d2b) ++ e1b: d1a 192-64,  d1b 64-64   (1)
  up(d1b) ++ e0b: d0a 96-32,   d0b 32-32   (0)   | head 1x1 32-3
"""

from __future__ import annotations

import ctypes
import os
import time

import numpy as np

from . import native

LAYERS = [("e0a", 8, 32), ("e0b", 32, 32), ("e1a", 32, 64), ("e1b", 64, 64),
          ("e2a", 64, 128), ("e2b", 128, 128), ("bna", 128, 256),
          ("bnb", 256, 256), ("d2a", 384, 128), ("d2b", 128, 128),
          ("d1a", 192, 64), ("d1b", 64, 64), ("d0a", 96, 32),
          ("d0b", 32, 32)]
SCALE = 64.0


def param_shapes() -> list[tuple[str, tuple[int, ...]]]:
  out = []
  for name, ci, co in LAYERS:
    out += [(name + ".w", (9, ci, co)), (name + ".b", (co,))]
  out += [("out.w", (32, 3)), ("out.b", (3,))]
  return out


def param_count() -> int:
  return int(sum(np.prod(s) for _, s in param_shapes()))


def unflatten(flat: np.ndarray) -> dict:
  d, off = {}, 0
  for name, shape in param_shapes():
    n = int(np.prod(shape))
    d[name] = flat[off:off + n].reshape(shape)
    off += n
  return d


def flatten(d: dict) -> np.ndarray:
  return np.concatenate([np.asarray(d[name], np.float32).reshape(-1)
                         for name, _ in param_shapes()]).astype(np.float32)


def init_params(seed: int = 0) -> np.ndarray:
  """He-normal 3x3 weights (std sqrt(2 / (9 cin))), head std
  sqrt(1 / 32) scaled by 0.1 so the first prediction is close to the cold
  start, zero biases."""
  rng = np.random.default_rng(seed)
  d = {}
  for name, ci, co in LAYERS:
    d[name + ".w"] = rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), (9, ci, co))
    d[name + ".b"] = np.zeros(co)
  d["out.w"] = rng.normal(0.0, 0.1 * np.sqrt(1.0 / 32), (32, 3))
  d["out.b"] = np.zeros(3)
  return flatten(d)


def make_input(mask: np.ndarray, tgt: np.ndarray,
               grad: np.ndarray) -> np.ndarray:
  """[h, w, 8] float32 network input of one GridSolver problem."""
  h, w = mask.shape
  x = np.zeros((h, w, 8), np.float32)
  x[..., 0:3] = grad / SCALE
  x[..., 3:6] = (tgt - 128.0) / SCALE
  x[..., 6] = mask != 0
  return x


def pad8(a: np.ndarray, mode: str = "edge") -> np.ndarray:
  """Pad the first two axes at the bottom/right to multiples of 8."""
  h, w = a.shape[:2]
  ph, pw = (-h) % 8, (-w) % 8
  if not ph and not pw:
    return a
  width = [(0, ph), (0, pw)] + [(0, 0)] * (a.ndim - 2)
  return np.pad(a, width, mode=mode)


class UNet:
  """The network on the GPU for one batch shape (pie_unet_* C ABI)."""

  def __init__(self, batch: int, h: int, w: int, params=None, seed: int = 0):
    self._lib = native.load()
    hnd = ctypes.c_void_p()
    native.check(self._lib.pie_unet_create(batch, h, w, ctypes.byref(hnd)))
    self._h = hnd
    self.shape = (batch, h, w)
    n = ctypes.c_longlong()
    native.check(self._lib.pie_unet_param_count(self._h, ctypes.byref(n)))
    if n.value != param_count():
      raise RuntimeError("parameter layout of the library and module differ")
    self.set_params(init_params(seed) if params is None else params)

  def __del__(self):
    h = getattr(self, "_h", None)
    if h is not None and h.value:
      self._lib.pie_unet_destroy(h)
      self._h = None

  @staticmethod
  def _p(a):
    return ctypes.c_void_p(a.ctypes.data)

  def set_params(self, p: np.ndarray) -> None:
    p = np.ascontiguousarray(p, np.float32)
    if p.size != param_count():
      raise ValueError("wrong number of parameters")
    native.check(self._lib.pie_unet_set_params(self._h, self._p(p)))

  def params(self) -> np.ndarray:
    p = np.empty(param_count(), np.float32)
    native.check(self._lib.pie_unet_get_params(self._h, self._p(p)))
    return p

  def _x(self, x):
    x = np.ascontiguousarray(x, np.float32)
    if x.shape != self.shape + (8,):
      raise ValueError(f"input must be {self.shape + (8,)}")
    return x

  def forward(self, x: np.ndarray) -> np.ndarray:
    x = self._x(x)
    out = np.empty(self.shape + (3,), np.float32)
    native.check(self._lib.pie_unet_forward(self._h, self._p(x),
                                            self._p(out)))
    return out

  def loss_grad(self, x, target, mask):
    x = self._x(x)
    t = np.ascontiguousarray(target, np.float32)
    m = np.ascontiguousarray(mask, np.uint8)
    loss = ctypes.c_float()
    g = np.empty(param_count(), np.float32)
    native.check(self._lib.pie_unet_loss_grad(self._h, self._p(x), self._p(t),
                                              self._p(m), ctypes.byref(loss),
                                              self._p(g)))
    return loss.value, g

  def train_step(self, x, target, mask, lr: float) -> float:
    x = self._x(x)
    t = np.ascontiguousarray(target, np.float32)
    m = np.ascontiguousarray(mask, np.uint8)
    loss = ctypes.c_float()
    native.check(self._lib.pie_unet_train_step(self._h, self._p(x),
                                               self._p(t), self._p(m),
                                               float(lr), ctypes.byref(loss)))
    return loss.value

  def stats(self) -> dict:
    ms, fl = ctypes.c_double(), ctypes.c_double()
    native.check(self._lib.pie_unet_stats(self._h, ctypes.byref(ms),
                                          ctypes.byref(fl)))
    return {"ms": ms.value, "conv_flops": fl.value,
            "tflops": fl.value / max(ms.value, 1e-9) / 1e9}


# ------------------------------------------------------------------ data ---


def composite(size: tuple[int, int], seed: int, gradient: str | None = None):
  """One synthetic GridSolver problem on the whole image (no crop): blob
  source and target, star-polygon mask inside a 1-px border.  Returns
  (mask int32, tgt float32 [h, w, 3], grad float32 [h, w, 3], gradient)."""
  from . import process, synth
  h, w = size
  rng = np.random.default_rng([7, seed])
  if gradient is None:
    gradient = "src" if rng.random() < 0.5 else "max"
  src = synth.blobs(h, w, 10_000 + seed).astype(np.float32)
  tgt = synth.blobs(h, w, 20_000 + seed).astype(np.float32)
  mask = (synth.star_mask(h, w, rng) > 0).astype(np.int32)
  mask[[0, -1], :] = 0
  mask[:, [0, -1]] = 0
  grad = process.divergence(src, tgt, mask,
                            lambda a, b: process.mixgrad(a, b, gradient))
  return mask, tgt, grad, gradient


def _composite_job(args):
  size, render, seed = args
  m, t, g, _ = composite((render, render), seed)
  if render != size:
    m, t, g = coarsen(m, t, g, render // size)
  return m.astype(np.uint8), t.astype(np.float16), g.astype(np.float32)


def training_set(n: int, size: int = 256, seed0: int = 0,
                 sweeps: int = 20000, workers: int = 16, chunk: int = 512,
                 render: int | None = None, log=print):
  """n composites of size x size and their Jacobi labels (sweeps sweeps from
  the target state, solved by the batched GPU solver).  With render > size
  each composite is drawn at render x render and coarsened (coarsen()), the
  form in which warm_start() shows large problems to the network.  Stored
  compactly: mask uint8, target float16, divergence float32, label
  correction float16 ((x - target) / 64)."""
  render = render or size
  from concurrent.futures import ProcessPoolExecutor

  from .core_hip import BatchSolver

  t0 = time.perf_counter()
  jobs = [(size, render, seed0 + i) for i in range(n)]
  with ProcessPoolExecutor(max_workers=workers) as ex:
    parts = list(ex.map(_composite_job, jobs, chunksize=16))
  M = np.stack([p[0] for p in parts])
  T = np.stack([p[1] for p in parts])
  G2 = np.stack([p[2] for p in parts])
  del parts
  log(f"composites: {n} x {size}^2 (drawn at {render}^2) in "
      f"{time.perf_counter() - t0:.1f} s")
  Y = np.zeros((n, size, size, 3), np.float16)
  t1 = time.perf_counter()
  for c0 in range(0, n, chunk):
    b = BatchSolver(sweeps_per_launch=8, rows_per_task=64)
    idx = range(c0, min(n, c0 + chunk))
    for i in idx:
      b.add(M[i], T[i].astype(np.float32), G2[i], max_sweeps=sweeps,
            tol=0.0)
    b.solve(32)
    for k, i in enumerate(idx):
      _, _, st = b.read(k, state=True)
      Y[i] = ((st - T[i].astype(np.float32)) / SCALE *
              (M[i][..., None] > 0)).astype(np.float16)
    del b
  log(f"labels: {sweeps} Jacobi sweeps each in {time.perf_counter() - t1:.1f} s")
  return M, T, G2, Y


def batch_arrays(M, T, G2, Y, idx):
  x = np.zeros((len(idx), M.shape[1], M.shape[2], 8), np.float32)
  x[..., 0:3] = G2[idx] / SCALE
  x[..., 3:6] = (T[idx].astype(np.float32) - 128.0) / SCALE
  x[..., 6] = M[idx]
  return x, Y[idx].astype(np.float32), M[idx]


def train(M, T, G2, Y, epochs: int = 4, batch: int = 64, lr: float = 1e-3,
          seed: int = 0, log=print) -> tuple[UNet, dict]:
  n, h, w = M.shape
  net = UNet(batch, h, w, seed=seed)
  rng = np.random.default_rng(seed)
  losses = []
  t0 = time.perf_counter()
  for ep in range(epochs):
    order = rng.permutation(n)
    for s in range(0, n - batch + 1, batch):
      x, y, m = batch_arrays(M, T, G2, Y, order[s:s + batch])
      losses.append(net.train_step(x, y, m, lr))
    log(f"epoch {ep + 1}/{epochs}: loss {np.mean(losses[-(n // batch):]):.5f} "
        f"({time.perf_counter() - t0:.1f} s)")
  st = net.stats()
  return net, {"losses": losses, "seconds": time.perf_counter() - t0,
               "steps": len(losses), **st}


# ----------------------------------------------------------- warm start ---


def coarsen(mask, tgt, grad, f: int):
  """The problem on a grid f times coarser: the mask where most of a block
  is solved, the target averaged, the divergence summed over the block (the
  discrete Laplacian of a field sampled every f pixels is f^2 times larger,
  which the block sum matches).  Sizes are cut to multiples of f."""
  h, w = (mask.shape[0] // f) * f, (mask.shape[1] // f) * f
  def blocks(a):
    return a[:h, :w].reshape(h // f, f, w // f, f, *a.shape[2:])
  m = (blocks(mask != 0).mean((1, 3)) > 0.5).astype(np.int32)
  m[[0, -1], :] = 0
  m[:, [0, -1]] = 0
  t = blocks(tgt).mean((1, 3)).astype(np.float32)
  g = (blocks(grad).sum((1, 3)) * (m[..., None] != 0)).astype(np.float32)
  return m, t, g


def _upsample(a: np.ndarray, f: int, h: int, w: int) -> np.ndarray:
  """Bilinear upsampling of [hc, wc, c] by f (cell-centred), cut to h x w."""
  hc, wc = a.shape[:2]
  ys = (np.arange(h) + 0.5) / f - 0.5
  xs = (np.arange(w) + 0.5) / f - 0.5
  y0 = np.clip(np.floor(ys).astype(int), 0, hc - 1)
  x0 = np.clip(np.floor(xs).astype(int), 0, wc - 1)
  y1, x1 = np.minimum(y0 + 1, hc - 1), np.minimum(x0 + 1, wc - 1)
  fy = np.clip(ys - y0, 0, 1)[:, None, None]
  fx = np.clip(xs - x0, 0, 1)[None, :, None]
  top = a[y0][:, x0] * (1 - fx) + a[y0][:, x1] * fx
  bot = a[y1][:, x0] * (1 - fx) + a[y1][:, x1] * fx
  return (top * (1 - fy) + bot * fy).astype(np.float32)


def predict(params: np.ndarray, mask, tgt, grad, net: UNet | None = None):
  """Network correction (x - target) in pixel units, [h, w, 3]."""
  h, w = mask.shape
  x = pad8(make_input(mask, tgt, grad), "edge")
  x[h:, :, 6] = 0
  x[:, w:, 6] = 0
  if net is None or net.shape != (1,) + x.shape[:2]:
    net = UNet(1, x.shape[0], x.shape[1], params=params)
  return SCALE * net.forward(x[None])[0, :h, :w]


def warm_start(params: np.ndarray, mask, tgt, grad, coarse: int = 256,
               net: UNet | None = None) -> np.ndarray:
  """Predicted start state x0 [h, w, 3] for one problem.

  Problems larger than `coarse` pixels on a side are predicted on a grid
  coarsened by the power of two that brings them to about that size (the
  scale the network is trained at), and the predicted correction is
  upsampled bilinearly: the network supplies the smooth, image-wide part of
  the solution, which is what Jacobi removes slowest, and Jacobi removes the
  local remainder quickly."""
  h, w = mask.shape
  f = 1
  while max(h, w) // (2 * f) >= coarse:
    f *= 2
  if f == 1:
    corr = predict(params, mask, tgt, grad, net)
  else:
    mc, tc, gc = coarsen(mask, tgt, grad, f)
    corr = _upsample(predict(params, mc, tc, gc, net), f, h, w)
  return np.where(mask[..., None] != 0, tgt + corr, tgt).astype(np.float32)


def time_to_residual(mask, tgt, grad, x0=None, tol: float = 1e-3,
                     check_every: int = 256, max_sweeps: int = 10**7):
  """Plain Jacobi (g+U+L+R+D) from x0 (default: the target) until the mean
  absolute residual is <= tol.  Returns (seconds, sweeps, residual, solver)."""
  from .core_hip import GridSolver
  s = GridSolver(schedule="jacobi")
  s.reset(0, mask, tgt, grad)
  s.set_tile(8, 1, 64 if mask.shape[0] < 2048 else 128)
  if x0 is not None:
    s.set_state(x0)
  s.sync()
  t = time.perf_counter()
  n, r = s.solve_to_residual(tol, check_every, max_sweeps)
  return time.perf_counter() - t, n, r, s


def mean_residual(mask, tgt, grad, x) -> float:
  from .core_hip import GridSolver
  s = GridSolver(schedule="jacobi")
  s.reset(0, mask, tgt, grad)
  s.set_state(x)
  _, err = s.step(0)
  return float(err.sum()) / (3 * max(1, int(np.count_nonzero(mask))))


def choose_start(mask, tgt, grad, x_pred) -> np.ndarray:
  """The predicted start if its mean residual is below the cold start's,
  otherwise the cold start (the target)."""
  if mean_residual(mask, tgt, grad, x_pred) < mean_residual(mask, tgt, grad,
                                                            tgt):
    return np.asarray(x_pred, np.float32)
  return np.asarray(tgt, np.float32).copy()
