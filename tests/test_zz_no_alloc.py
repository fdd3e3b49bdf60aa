"""The allocation-free GridSolver solve (GPU).

pie_grid_solve_noalloc runs reset + step + read on the caller's device
buffers with every intermediate in the library's static device array.  Two
calls in a row, the second on a smaller crop (the array then holds the first
call's contents), must give the uint8 image and the residual sums of
pie_reset_grid + pie_step on the same inputs bit for bit.  Device buffers
come from the HIP runtime the library itself links (no torch in this
process: two HIP runtimes must not share it).
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class _Hip:
  def __init__(self):
    self.rt = ctypes.CDLL("libamdhip64.so.7")
    self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p),
                                  ctypes.c_size_t]
    self.rt.hipFree.argtypes = [ctypes.c_void_p]
    self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_size_t, ctypes.c_int]
    self.held = []

  def put(self, a):
    a = np.ascontiguousarray(a)
    p = ctypes.c_void_p()
    assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)) == 0
    self.held.append(p)
    assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p

  def empty(self, nbytes):
    p = ctypes.c_void_p()
    assert self.rt.hipMalloc(ctypes.byref(p), max(nbytes, 1)) == 0
    self.held.append(p)
    return p

  def get(self, p, shape, dtype):
    out = np.empty(shape, dtype)
    assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
    return out

  def free(self):
    for p in self.held:
      self.rt.hipFree(p)
    self.held = []


def _problem(h, w, seed, gradient):
  from fpie_amd import synth
  mask, tgt, grad, _ = synth.composite((h, w), seed, gradient)
  return mask, tgt, grad


@pytest.fixture(scope="module")
def lib():
  from fpie_amd import native
  return native.load()


def test_noalloc_matches_reset_and_step(lib):
  from fpie_amd import core_hip, native
  hip = _Hip()
  try:
    cases = [((720, 960), 11, "max", "jacobi", 203),
             ((257, 300), 12, "src", "gs_lex", 61)]
    for (h, w), seed, grad_mode, sched, iters in cases:
      print(f"\nnoalloc: {h} x {w} seed {seed} {sched} {iters}", flush=True)
      mask, tgt, grad = _problem(h, w, seed, grad_mode)
      d_mask, d_tgt, d_grad = hip.put(mask), hip.put(tgt), hip.put(grad)
      d_img, d_err = hip.empty(h * w * 3), hip.empty(12)
      native.check(lib.pie_grid_solve_noalloc(
          h, w, d_mask, d_tgt, d_grad, iters, native.SCHEDULES[sched],
          native.PIE_ORDER_ULRD, d_img, d_err, None))
      img = hip.get(d_img, (h, w, 3), np.uint8)
      err = hip.get(d_err, (3,), np.float32)
      core = core_hip.GridSolver(schedule=sched, order="ulrd")
      core.reset(0, mask, tgt, grad)
      want_img, want_err = core.step(iters)
      np.testing.assert_array_equal(img, want_img)
      np.testing.assert_array_equal(err.view(np.uint32),
                                    np.asarray(want_err,
                                               np.float32).view(np.uint32))
  finally:
    hip.free()


def test_noalloc_rejects_a_crop_beyond_its_scratch(lib):
  from fpie_amd import native
  hip = _Hip()
  try:
    h, w = 1500, 1500
    d = hip.empty(16)
    rc = lib.pie_grid_solve_noalloc(h, w, d, d, d, 1, native.PIE_JACOBI,
                                    native.PIE_ORDER_ULRD, d, d, None)
    assert rc == native.PIE_EINVAL
    assert b"MiB of scratch" in lib.pie_last_error()
  finally:
    hip.free()
