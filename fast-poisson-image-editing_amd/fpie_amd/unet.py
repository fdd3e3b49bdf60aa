"""The warm-start network (no counterpart in the reference).

A 4-level U-Net (base 32 channels, 3x3 convolutions, bf16 weights and
activations, fp32 accumulation on MFMA; kernels in csrc/unet_kernels.hip,
runtime in csrc/unet.hip, forward, backward and Adam all on the GPU).

It approximates the solution operator of the GridSolver system on a coarse
grid, in the correction (defect) form used by fpie_amd.warmstart: for the
system A x = b of a composite (A: 4 on the diagonal, -1 per solved
neighbour) and a state x, the correction c = x* - x solves A c = r with
r = b - A x and c = 0 on the unsolved pixels.  Input channels:

  channel 0-2  r / sigma, sigma = rms of r over the solved pixels (the
               guidance divergence of the correction problem; for the
               target state r = divergence + discrete Laplacian of the
               target, so the target's own guidance enters here)
  channel 3-5  boundary values of the correction: 0
  channel 6    mask (1 = solved)
  channel 7    0

Output: the correction's shape, c / rms(c) on the solved pixels (its
amplitude is fitted by fpie_amd.warmstart with an energy-optimal step).

Layers (name, in, out channels), resolution level in parentheses:
  e0a 8-32, e0b 32-32 (0) | pool | e1a 32-64, e1b 64-64 (1) | pool |
  e2a 64-128, e2b 128-128 (2) | pool | bna 128-256, bnb 256-256 (3)
  up(bnb) ++ e2b: d2a 384-128, d2b 128-128 (2)
  up(d2b) ++ e1b: d1a 192-64,  d1b 64-64   (1)
  up(d1b) ++ e0b: d0a 96-32,   d0b 32-32   (0)   | head 1x1 32-3
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from . import native

LAYERS = [("e0a", 8, 32), ("e0b", 32, 32), ("e1a", 32, 64), ("e1b", 64, 64),
          ("e2a", 64, 128), ("e2b", 128, 128), ("bna", 128, 256),
          ("bnb", 256, 256), ("d2a", 384, 128), ("d2b", 128, 128),
          ("d1a", 192, 64), ("d1b", 64, 64), ("d0a", 96, 32),
          ("d0b", 32, 32)]


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
  sqrt(1 / 32) scaled by 0.1 (small first outputs), zero biases."""
  rng = np.random.default_rng(seed)
  d = {}
  for name, ci, co in LAYERS:
    d[name + ".w"] = rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), (9, ci, co))
    d[name + ".b"] = np.zeros(co)
  d["out.w"] = rng.normal(0.0, 0.1 * np.sqrt(1.0 / 32), (32, 3))
  d["out.b"] = np.zeros(3)
  return flatten(d)


def pad8(a: np.ndarray, mode: str = "edge") -> np.ndarray:
  """Pad the first two axes at the bottom/right to multiples of 8."""
  h, w = a.shape[:2]
  ph, pw = (-h) % 8, (-w) % 8
  if not ph and not pw:
    return a
  width = [(0, ph), (0, pw)] + [(0, 0)] * (a.ndim - 2)
  return np.pad(a, width, mode=mode)


def set_conv_tiles(min_blocks16: int = 512) -> None:
  """Convolution tile policy of this process: layers with 128+ input
  channels use 16 x 16-pixel tiles when that gives at least min_blocks16
  blocks (0 = always, -1 = never).  Results do not depend on it."""
  native.check(native.load().pie_unet_set_conv_tiles(int(min_blocks16)))


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

  def staging(self, k: int):
    """Slot k's pinned host batch as numpy views (x [B, H, W, 8], target
    [B, H, W, 3] float32, mask [B, H, W] uint8), once its last upload has
    finished.  Fill them, then call train_async(k)."""
    px, pt, pm = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    native.check(self._lib.pie_unet_staging(self._h, int(k), ctypes.byref(px),
                                            ctypes.byref(pt),
                                            ctypes.byref(pm)))
    b, h, w = self.shape

    def view(ptr, ctype, shape):
      n = int(np.prod(shape))
      buf = (ctype * n).from_address(ptr.value)
      return np.frombuffer(buf, dtype=np.dtype(ctype)).reshape(shape)

    return (view(px, ctypes.c_float, (b, h, w, 8)),
            view(pt, ctypes.c_float, (b, h, w, 3)),
            view(pm, ctypes.c_uint8, (b, h, w)))

  def train_async(self, k: int, lr: float) -> None:
    """Enqueue the upload of slot k and an Adam step on it; returns at
    once (the loss comes from finish(k))."""
    native.check(self._lib.pie_unet_train_async(self._h, int(k), float(lr)))

  def finish(self, k: int) -> float:
    """Wait for slot k's step; its loss."""
    loss = ctypes.c_float()
    native.check(self._lib.pie_unet_finish(self._h, int(k),
                                           ctypes.byref(loss)))
    return loss.value

  def stats(self) -> dict:
    ms, fl = ctypes.c_double(), ctypes.c_double()
    native.check(self._lib.pie_unet_stats(self._h, ctypes.byref(ms),
                                          ctypes.byref(fl)))
    return {"ms": ms.value, "conv_flops": fl.value,
            "tflops": fl.value / max(ms.value, 1e-9) / 1e9}


# -------------------------------------------------------------- training ---


def batch_arrays(M, X, Y, idx, out=None):
  """Network input, label and mask of the samples idx: M uint8 [n, h, w]
  masks, X [n, h, w, 3] normalised residuals, Y [n, h, w, 3] normalised
  corrections.  out: (x, y, m) arrays to fill instead of new ones."""
  if out is None:
    out = (np.empty((len(idx), M.shape[1], M.shape[2], 8), np.float32),
           np.empty((len(idx),) + Y.shape[1:], np.float32),
           np.empty((len(idx),) + M.shape[1:], np.uint8))
  x, y, m = out
  x.fill(0.0)
  x[..., 0:3] = X[idx]
  x[..., 6] = M[idx]
  y[...] = Y[idx]
  m[...] = M[idx]
  return x, y, m


def train(M, X, Y, epochs: int = 4, batch: int = 64, lr: float = 1e-3,
          seed: int = 0, log=print) -> tuple[UNet, dict]:
  """Adam on the masked mean squared error over shuffled minibatches (the
  last partial batch of an epoch is dropped).  Pipelined: the next batch is
  gathered into one pinned staging slot and uploaded while the GPU runs
  the step on the other."""
  n, h, w = M.shape
  net = UNet(batch, h, w, seed=seed)
  rng = np.random.default_rng(seed)
  losses = []
  t0 = time.perf_counter()
  inflight = None  # slot of the step on the GPU
  for ep in range(epochs):
    order = rng.permutation(n)
    for s in range(0, n - batch + 1, batch):
      slot = 0 if inflight is None else 1 - inflight
      batch_arrays(M, X, Y, order[s:s + batch], out=net.staging(slot))
      net.train_async(slot, lr)
      if inflight is not None:
        losses.append(net.finish(inflight))
      inflight = slot
    if inflight is not None:
      losses.append(net.finish(inflight))
      inflight = None
    log(f"epoch {ep + 1}/{epochs}: loss {np.mean(losses[-(n // batch):]):.5f} "
        f"({time.perf_counter() - t0:.1f} s)")
  st = net.stats()
  return net, {"losses": losses, "seconds": time.perf_counter() - t0,
               "steps": len(losses), **st}


def measure(batch: int = 64, size: int = 256, steps: int = 5) -> dict:
  """Time Adam steps of UNet(batch, size, size) on random inputs (one
  untimed step first), pipelined as in train(): each step's batch is
  uploaded from a pinned staging slot on a copy stream while the other
  slot's step runs.  device_ms_per_step: forward + backward between HIP
  events on the network's stream; ms_per_step: wall time per step of the
  loop in steady state, from the moment the first step's batch is on the
  device (uploads of the later steps, Adam and the weight packing
  included; the batches are written into the two slots once, before the
  loop).  ms_per_step_sync:
  the same step through the unpipelined train_step (upload from pageable
  memory on the compute stream, then the step).  conv_tflops counts
  2 H W 9 Cin Cout per convolution and pass."""
  rng = np.random.default_rng(0)
  b, s = batch, size
  net = UNet(b, s, s, seed=0)
  for k in range(2):
    x, y, m = net.staging(k)
    x.fill(0.0)
    x[..., :3] = rng.normal(0, 1, (b, s, s, 3))
    x[..., 6] = 1.0
    y[...] = rng.normal(0, 1, (b, s, s, 3))
    m.fill(1)
  net.train_async(0, 1e-3)  # untimed: kernel loading
  net.finish(0)
  before = net.stats()
  # steady state: the clock starts once the first timed step's upload has
  # landed (staging waits for it), so the one-off pipeline fill is not
  # counted; the loop then keeps one step in flight while it enqueues the
  # next
  net.train_async(0, 1e-3)
  net.staging(0)
  t0 = time.perf_counter()
  for k in range(1, steps):
    net.staging(k % 2)  # the slot's host copy is free again
    net.train_async(k % 2, 1e-3)
    net.finish((k - 1) % 2)
  net.finish((steps - 1) % 2)
  dt = time.perf_counter() - t0
  after = net.stats()
  flops = after["conv_flops"] - before["conv_flops"]
  dev = (after["ms"] - before["ms"]) / 1e3
  x, y, m = (np.array(a) for a in net.staging(0))
  t1 = time.perf_counter()
  net.train_step(x, y, m, 1e-3)
  sync_ms = (time.perf_counter() - t1) * 1e3
  return {"batch": b, "size": s, "steps": steps,
          "ms_per_step": dt * 1e3 / steps,
          "device_ms_per_step": dev * 1e3 / steps,
          "ms_per_step_sync": sync_ms,
          "conv_tflops": flops / dev / 1e12,
          "conv_tflops_wall": flops / dt / 1e12,
          "mfma_utilisation_vs_2500_tflops": flops / dev / 2.5e15}


def main(argv=None) -> None:
  """Time Adam steps of the network on random inputs:
  python -m fpie_amd.unet [--batch 64] [--size 256] [--steps 5]"""
  import argparse
  import json
  ap = argparse.ArgumentParser()
  ap.add_argument("--batch", type=int, default=64)
  ap.add_argument("--size", type=int, default=256)
  ap.add_argument("--steps", type=int, default=5)
  a = ap.parse_args(argv)
  print(json.dumps(measure(a.batch, a.size, a.steps)), flush=True)


if __name__ == "__main__":
  import sys
  main(sys.argv[1:])
