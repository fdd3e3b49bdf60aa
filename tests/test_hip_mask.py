"""Mask preparation on the device (GPU).

pie_prepare_mask must give what the processors' host preparation gives
(fpie/process.py:209-221, 338-349: channel mean >= 128, outermost rows and
columns cleared, bounding box plus a 1-px margin): on the ten shipped
benchmark masks, whose published '# of vars' pin that host preparation
(tests/test_process.py), and on random masks built to hit the edge cases
(channel sums at 383 / 384, solved pixels on the image edge, one-pixel
masks, 1-, 3- and 4-channel and 2-D masks).
"""

import numpy as np
import pytest

from conftest import GOLDEN  # noqa: F401  (import path set-up)
from test_process import _benchmark_masks

pytestmark = pytest.mark.gpu


def _same(mask):
  from fpie_amd import core_hip, process
  want = process.prepare_mask(mask)
  got = core_hip.prepare_mask(mask)
  assert got[1:] == want[1:], (got[1:], want[1:])
  np.testing.assert_array_equal(got[0], want[0])
  assert got[0].dtype == np.int32
  return got


@pytest.mark.parametrize("name,img,equ_vars,grid_vars",
                         list(_benchmark_masks()),
                         ids=[c[0] for c in _benchmark_masks()])
def test_shipped_masks(name, img, equ_vars, grid_vars):
  m, *_ = _same(img)
  assert int(m.sum()) + 1 == equ_vars
  assert m.size == grid_vars


def test_random_masks():
  rng = np.random.default_rng(5)
  for k in range(12):
    h, w = int(rng.integers(3, 300)), int(rng.integers(3, 300))
    channels = [None, 1, 3, 4][k % 4]
    shape = (h, w) if channels is None else (h, w, channels)
    # values around the threshold: channel sums land on 128 * c - 1, 128 * c
    base = rng.integers(120, 137, shape).astype(np.uint8)
    keep = rng.random((h, w)) < rng.uniform(0.02, 0.9)
    mask = np.where(keep[..., None] if channels else keep, base, 0)
    mask = mask.astype(np.uint8)
    mask[0, :] = 255  # solved-looking edge pixels are cleared
    mask[:, -1] = 255
    if not (process_has_pixel(mask)):
      mask[h // 2, w // 2] = 255
    _same(mask)


def process_has_pixel(mask):
  m = mask.mean(-1) if mask.ndim == 3 else mask
  m = (m >= 128)
  return bool(m[1:-1, 1:-1].any())


def test_exact_threshold():
  """sum 384 over three channels is solved, 383 is not (the reference's
  float mean: 128.0 >= 128, 127.67 < 128)."""
  mask = np.zeros((5, 6, 3), np.uint8)
  mask[2, 2] = (128, 128, 128)  # 384
  mask[2, 3] = (128, 128, 127)  # 383
  mask[3, 3] = (255, 129, 0)    # 384
  m, x0, x1, y0, y1 = _same(mask)
  assert (x0, x1, y0, y1) == (1, 5, 1, 5)
  assert m.sum() == 2


def test_single_pixel_and_nothing_solved():
  from fpie_amd import core_hip
  mask = np.zeros((7, 9), np.uint8)
  mask[3, 4] = 200
  m, x0, x1, y0, y1 = _same(mask)
  assert m.shape == (3, 3) and m[1, 1] == 1 and m.sum() == 1
  edge_only = np.zeros((7, 9), np.uint8)
  edge_only[0, :] = 255
  with pytest.raises(ValueError):
    core_hip.prepare_mask(edge_only)


def test_processor_uses_the_device_preparation():
  """GridProcessor / EquProcessor on the hip backend give the same crop
  and variable count as the host preparation."""
  from fpie_amd import process
  rng = np.random.default_rng(9)
  src = rng.integers(0, 256, (90, 110, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (95, 120, 3), dtype=np.uint8)
  mask = np.zeros((90, 110, 3), np.uint8)
  mask[10:70, 20:95] = rng.integers(100, 256, (60, 75, 3))
  g = process.GridProcessor("max", "hip")
  n = g.reset(src, mask, tgt, (0, 0), (3, 5))
  m, x0, x1, y0, y1 = process.prepare_mask(mask)
  assert n == m.size
  assert (g.x0, g.x1, g.y0, g.y1) == (3 + x0, 3 + x1, 5 + y0, 5 + y1)
  e = process.EquProcessor("max", "hip")
  assert e.reset(src, mask, tgt, (0, 0), (3, 5)) == int(m.sum()) + 1
