"""CPU oracle of the warm-start U-Net (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the HIP forward and
backward kernels; the product never imports it.  The network has no
counterpart in the reference; its definition is the one in
fast-poisson-image-editing_amd/fpie_amd/unet.py (restated here, not
imported):

  input 8 channels (the warm start's features: normalised coarse
  residual, zeros, mask, zero);
  encoder  e0a 8->32, e0b 32->32 | pool | e1a 32->64, e1b 64->64 | pool |
           e2a 64->128, e2b 128->128 | pool | bna 128->256, bnb 256->256
  decoder  up(bnb) ++ e2b -> d2a 384->128, d2b 128->128 | up ++ e1b ->
           d1a 192->64, d1b 64->64 | up ++ e0b -> d0a 96->32, d0b 32->32
  head     1x1 conv 32->3 (no ReLU)
  every 3x3 conv: zero padding 1, bias, ReLU; pool = 2x2 average;
  up = 2x nearest; "++" = channel concatenation (upsampled first).

Tensors are NHWC.  Weights of a 3x3 conv are w[tap][cin][cout] with
tap = 3 * (dy + 1) + (dx + 1).  The GPU rounds weights and every stored
activation to bfloat16 (round to nearest even) and accumulates in float32;
`forward(..., bf16=True)` applies the same roundings.
"""

from __future__ import annotations

import numpy as np

LAYERS = [  # name, cin, cout
    ("e0a", 8, 32), ("e0b", 32, 32), ("e1a", 32, 64), ("e1b", 64, 64),
    ("e2a", 64, 128), ("e2b", 128, 128), ("bna", 128, 256),
    ("bnb", 256, 256), ("d2a", 384, 128), ("d2b", 128, 128),
    ("d1a", 192, 64), ("d1b", 64, 64), ("d0a", 96, 32), ("d0b", 32, 32),
]
HEAD = ("out", 32, 3)


def bf16(x: np.ndarray) -> np.ndarray:
  """Round float32 values to bfloat16 (nearest even), back as float."""
  x32 = np.ascontiguousarray(x, np.float32)
  u = x32.view(np.uint32).astype(np.uint64)
  u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
  return u.astype(np.uint32).view(np.float32).astype(x.dtype)


def conv3x3(x, w, b):
  """x [N,H,W,Ci], w [9,Ci,Co], b [Co] -> [N,H,W,Co] (zero padding)."""
  n, h, ww, _ = x.shape
  xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
  out = np.zeros((n, h, ww, w.shape[2]), np.float64)
  for tap in range(9):
    dy, dx = tap // 3 - 1, tap % 3 - 1
    out += np.einsum("nhwc,cd->nhwd",
                     xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + ww], w[tap])
  return out + b


def conv3x3_bwd(x, w, dy):
  """Gradients of conv3x3 w.r.t. x, w and b for upstream gradient dy."""
  n, h, ww, _ = x.shape
  xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
  dxp = np.zeros_like(xp, dtype=np.float64)
  dw = np.zeros_like(w, dtype=np.float64)
  for tap in range(9):
    oy, ox = tap // 3 - 1, tap % 3 - 1
    win = xp[:, 1 + oy:1 + oy + h, 1 + ox:1 + ox + ww]
    dw[tap] = np.einsum("nhwc,nhwd->cd", win, dy)
    dxp[:, 1 + oy:1 + oy + h, 1 + ox:1 + ox + ww] += np.einsum(
        "nhwd,cd->nhwc", dy, w[tap])
  return dxp[:, 1:-1, 1:-1], dw, dy.sum((0, 1, 2))


def pool(x):
  n, h, w, c = x.shape
  return x.reshape(n, h // 2, 2, w // 2, 2, c).mean((2, 4))


def pool_bwd(dy):
  return np.repeat(np.repeat(dy, 2, 1), 2, 2) / 4.0


def up(x):
  return np.repeat(np.repeat(x, 2, 1), 2, 2)


def up_bwd(dy):
  n, h, w, c = dy.shape
  return dy.reshape(n, h // 2, 2, w // 2, 2, c).sum((2, 4))


def init_params(seed: int = 0) -> dict:
  """He-normal weights (std sqrt(2 / (9 * cin))), zero biases."""
  rng = np.random.default_rng(seed)
  p = {}
  for name, ci, co in LAYERS:
    p[name + ".w"] = rng.normal(0, np.sqrt(2.0 / (9 * ci)), (9, ci, co))
    p[name + ".b"] = np.zeros(co)
  name, ci, co = HEAD
  p[name + ".w"] = rng.normal(0, np.sqrt(1.0 / ci), (ci, co))
  p[name + ".b"] = np.zeros(co)
  return p


def forward(p: dict, x: np.ndarray, bf16_round: bool = True):
  """Returns (output [N,H,W,3], cache of activations for backward)."""
  r = bf16 if bf16_round else (lambda a: a)
  act = {"in": r(x.astype(np.float64))}

  def conv(name, inp):
    y = conv3x3(inp, r(p[name + ".w"]), p[name + ".b"])
    y = r(np.maximum(y, 0.0))
    act[name] = y
    return y

  s0 = conv("e0b", conv("e0a", act["in"]))
  act["p0"] = r(pool(s0))
  s1 = conv("e1b", conv("e1a", act["p0"]))
  act["p1"] = r(pool(s1))
  s2 = conv("e2b", conv("e2a", act["p1"]))
  act["p2"] = r(pool(s2))
  bn = conv("bnb", conv("bna", act["p2"]))
  act["c2"] = np.concatenate([up(bn), s2], -1)
  d2 = conv("d2b", conv("d2a", act["c2"]))
  act["c1"] = np.concatenate([up(d2), s1], -1)
  d1 = conv("d1b", conv("d1a", act["c1"]))
  act["c0"] = np.concatenate([up(d1), s0], -1)
  d0 = conv("d0b", conv("d0a", act["c0"]))
  out = np.einsum("nhwc,cd->nhwd", d0, r(p["out.w"])) + p["out.b"]
  return out, act


def backward(p: dict, act: dict, dout: np.ndarray,
             bf16_round: bool = True) -> dict:
  """Gradients of every parameter for upstream gradient dout [N,H,W,3]."""
  r = bf16 if bf16_round else (lambda a: a)
  g = {}
  d0 = act["d0b"]
  g["out.w"] = np.einsum("nhwc,nhwd->cd", d0, dout)
  g["out.b"] = dout.sum((0, 1, 2))
  dy = np.einsum("nhwd,cd->nhwc", dout, r(p["out.w"]))

  def conv_b(name, inp_key_or_arr, dy):
    y = act[name]
    dy = dy * (y > 0)
    inp = act[inp_key_or_arr] if isinstance(inp_key_or_arr, str) else \
        inp_key_or_arr
    dx, dw, db = conv3x3_bwd(inp, r(p[name + ".w"]), dy)
    g[name + ".w"], g[name + ".b"] = dw, db
    return dx

  dy = conv_b("d0b", "d0a", dy)
  dc0 = conv_b("d0a", "c0", dy)
  c_up = act["d1b"].shape[-1]
  ds0 = dc0[..., c_up:]
  dy = up_bwd(dc0[..., :c_up])
  dy = conv_b("d1b", "d1a", dy)
  dc1 = conv_b("d1a", "c1", dy)
  c_up = act["d2b"].shape[-1]
  ds1 = dc1[..., c_up:]
  dy = up_bwd(dc1[..., :c_up])
  dy = conv_b("d2b", "d2a", dy)
  dc2 = conv_b("d2a", "c2", dy)
  c_up = act["bnb"].shape[-1]
  ds2 = dc2[..., c_up:]
  dy = up_bwd(dc2[..., :c_up])
  dy = conv_b("bnb", "bna", dy)
  dy = conv_b("bna", "p2", dy)
  dy = conv_b("e2b", "e2a", pool_bwd(dy) + ds2)
  dy = conv_b("e2a", "p1", dy)
  dy = conv_b("e1b", "e1a", pool_bwd(dy) + ds1)
  dy = conv_b("e1a", "p0", dy)
  dy = conv_b("e0b", "e0a", pool_bwd(dy) + ds0)
  conv_b("e0a", "in", dy)
  return g


def masked_mse(out, target, mask):
  """Loss of fpie_amd.unet: mean over solved pixels and channels of
  (out - target)^2; returns (loss, dloss/dout)."""
  m = mask[..., None].astype(np.float64)
  cnt = max(1.0, m.sum() * out.shape[-1])
  d = (out - target) * m
  return float((d * d).sum() / cnt), 2.0 * d / cnt
