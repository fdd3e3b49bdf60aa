"""The product processors against the reference processor's own output.

tests/golden/processor_cases.npz holds, per case, what fpie/process.py's
EquProcessor / GridProcessor (imported from the reference tree by
tests/golden/make_processor_golden.py) passed to their solver core: the
EquSolver system reset(N, A, X, B) and the GridSolver reset(N, mask, tgt,
grad), for non-zero source and target offsets, all three gradient modes,
2-D, 1- and 3-channel masks and a source offset whose indices wrap.  Here
the product processors, on their host path, must pass the same arrays bit
for bit (the GPU test of the device path is tests/test_hip_processor.py).
"""

import numpy as np
import pytest

from conftest import processor_cases

from fpie_amd import process

CASES = processor_cases()


class _Rec:
  def partition(self, mask):
    self.part_mask = np.array(mask, np.int32)
    flat = (np.asarray(mask) > 0).reshape(-1)
    return (np.cumsum(flat).reshape(mask.shape)).astype(np.int32)

  def reset(self, *a):
    self.a = a

  def sync(self):
    pass


@pytest.fixture(scope="module", autouse=True)
def recording_backend():
  process.register_backend("golden-rec", lambda **kw: _Rec(),
                           lambda **kw: _Rec())


def _bits(a):
  a = np.asarray(a)
  return a.view(np.uint32) if a.dtype == np.float32 else a


@pytest.mark.parametrize("name", sorted(CASES))
def test_equ_system_equals_reference(name):
  c = CASES[name]
  p = process.EquProcessor(str(c["mode"]), "golden-rec")
  n = p.reset(c["src"], c["mask"], c["tgt"], tuple(c["mask_on_src"]),
              tuple(c["mask_on_tgt"]))
  assert n == int(c["equ_n"])
  np.testing.assert_array_equal(p.core.part_mask, c["equ_mask"])
  N, A, X, B = p.core.a
  assert N == int(c["equ_n"])
  np.testing.assert_array_equal(np.asarray(A, np.int32), c["equ_A"])
  np.testing.assert_array_equal(_bits(np.asarray(X, np.float32)),
                                _bits(c["equ_X"]))
  np.testing.assert_array_equal(_bits(np.asarray(B, np.float32)),
                                _bits(c["equ_B"]))


@pytest.mark.parametrize("name", sorted(n for n in CASES
                                        if str(CASES[n]["what"]) == "both"))
def test_grid_inputs_equal_reference(name):
  c = CASES[name]
  p = process.GridProcessor(str(c["mode"]), "golden-rec")
  n = p.reset(c["src"], c["mask"], c["tgt"], tuple(c["mask_on_src"]),
              tuple(c["mask_on_tgt"]))
  assert n == int(c["grid_n"])
  _, mask, tgt, grad = p.core.a
  np.testing.assert_array_equal(np.asarray(mask, np.int32), c["grid_mask"])
  np.testing.assert_array_equal(_bits(np.asarray(tgt, np.float32)),
                                _bits(c["grid_tgt"]))
  np.testing.assert_array_equal(_bits(np.asarray(grad, np.float32)),
                                _bits(c["grid_grad"]))
