"""GridSolver and EquSolver resets built on the device from uint8 crops (GPU).

`GridSolver.reset_images` must load exactly what GridProcessor.reset
computes on the host (fpie/process.py:352-386): the target crop as the state
and the mixed-gradient divergence as the guidance.  Every value is an
integer or a half, so the two paths agree bit for bit: checked through the
state after sweeps (one Jacobi sweep maps the guidance one to one onto the
state) and the residual sums, for every gradient mode, with offsets, on a
small random composite and on the 4096 x 4096 benchmark composite.

`EquSolver.reset_images` must build the system of EquProcessor.reset
(fpie/process.py:180-190, 227-270): ids row-major from 1, X = target,
B = divergence + target of unsolved neighbours, again exact values; checked
against the host-built system through the state, the output image and the
residual after sweeps of each schedule.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _both(src, mask, tgt, on_src, on_tgt, gradient, sweeps):
  from fpie_amd import core_hip, process
  proc = process.GridProcessor(gradient, "hip")
  assert hasattr(proc.core, "reset_images")
  proc.reset(src, mask, tgt, on_src, on_tgt)  # device path
  m, x0, x1, y0, y1 = process.prepare_mask(mask)
  sc = src[on_src[0] + x0:on_src[0] + x1,
           on_src[1] + y0:on_src[1] + y1].astype(np.float32)
  tc = tgt[on_tgt[0] + x0:on_tgt[0] + x1,
           on_tgt[1] + y0:on_tgt[1] + y1].astype(np.float32)
  grad = process.divergence(sc, tc, m, proc.mixgrad)
  host = core_hip.GridSolver()
  host.reset(0, m, tc, grad)  # host path
  out = []
  for core in (proc.core, host):
    assert np.array_equal(_bits(core.state()), _bits(tc))
    img, err = core.step(sweeps)
    out.append((core.state(), img, err))
  return out


@pytest.mark.parametrize("gradient", ["src", "avg", "max"])
def test_device_reset_matches_host_reset(gradient):
  rng = np.random.default_rng({"src": 1, "avg": 2, "max": 3}[gradient])
  src = rng.integers(0, 256, (90, 110, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (100, 120, 3), dtype=np.uint8)
  mask = (rng.random((90, 110)) < 0.6).astype(np.uint8) * 255
  (xd, id_, ed), (xh, ih, eh) = _both(src, mask, tgt, (0, 0), (5, 7),
                                      gradient, 3)
  assert np.array_equal(_bits(xd), _bits(xh))
  assert np.array_equal(id_, ih)
  assert np.array_equal(ed, eh)


def test_device_reset_on_the_benchmark_composite():
  from fpie_amd import synth
  w = synth.grid4096()
  (xd, _, ed), (xh, _, eh) = _both(w.src, w.mask, w.tgt, w.mask_on_src,
                                   w.mask_on_tgt, w.gradient, 8)
  assert np.array_equal(_bits(xd), _bits(xh))
  assert np.array_equal(ed, eh)


def test_device_reset_rejects_bad_input():
  from fpie_amd import core_hip
  s = core_hip.GridSolver()
  m = np.zeros((8, 8), np.int32)
  m[3, 3] = 1
  img = np.zeros((8, 8, 3), np.uint8)
  with pytest.raises(ValueError):
    s.reset_images(m, img, img, "mean")
  with pytest.raises(ValueError):
    s.reset_images(m, img[:, :4], img, "src")
  m[0, 0] = 1  # solved pixel on the border
  with pytest.raises(RuntimeError):
    s.reset_images(m, img, img, "src")


def _equ_both(src, mask, tgt, on_src, on_tgt, gradient, schedule, sweeps):
  from fpie_amd import process
  dev = process.EquProcessor(gradient, "hip", schedule=schedule)
  host = process.EquProcessor(gradient, "hip", schedule=schedule)
  n_dev = dev.reset(src, mask, tgt, on_src, on_tgt)  # uint8: device path
  # int16 copies of the same images take the host path
  n_host = host.reset(src.astype(np.int16), mask, tgt.astype(np.int16),
                      on_src, on_tgt)
  assert n_dev == n_host
  for a, b in zip(dev.tgt_index, host.tgt_index):
    assert np.array_equal(a, b)
  assert np.array_equal(_bits(dev.core.state()), _bits(host.core.state()))
  out = []
  for p in (dev, host):
    img, err = p.step(sweeps)
    out.append((p.core.state(), np.asarray(img, np.int16), err))
  return out


@pytest.mark.parametrize("schedule", ["jacobi", "gs_lex", "redblack"])
@pytest.mark.parametrize("gradient", ["src", "avg", "max"])
def test_equ_device_reset_matches_host_reset(gradient, schedule):
  rng = np.random.default_rng(10 + ["src", "avg", "max"].index(gradient))
  src = rng.integers(0, 256, (70, 90, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (80, 100, 3), dtype=np.uint8)
  mask = (rng.random((70, 90)) < 0.55).astype(np.uint8) * 255
  (xd, id_, ed), (xh, ih, eh) = _equ_both(src, mask, tgt, (0, 0), (4, 6),
                                          gradient, schedule, 5)
  assert np.array_equal(_bits(xd), _bits(xh))
  assert np.array_equal(id_, ih)
  assert np.array_equal(ed, eh)


def test_equ_device_reset_rejects_bad_input():
  from fpie_amd import core_hip
  s = core_hip.EquSolver()
  m = np.zeros((8, 8), np.int32)
  m[3, 3] = 1
  img = np.zeros((8, 8, 3), np.uint8)
  with pytest.raises(ValueError):
    s.reset_images(m, img, img, "mean")
  with pytest.raises(ValueError):
    s.reset_images(m, img[:, :4], img, "src")
  assert s.reset_images(m, img, img, "src") == 2
  crop, err = s.step_crop(1)
  assert crop.shape == (8, 8, 3) and err.shape == (3,)
  host = core_hip.EquSolver()
  host.reset(2, np.zeros((2, 4), np.int32), np.zeros((2, 3), np.float32),
             np.zeros((2, 3), np.float32))
  with pytest.raises(RuntimeError, match="pie_reset_equ_images"):
    host.step_crop(1)
  m[7, 0] = 1  # solved pixel on the border
  with pytest.raises(RuntimeError):
    s.reset_images(m, img, img, "src")
