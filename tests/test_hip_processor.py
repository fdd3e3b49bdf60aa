"""The hip processors' device resets against the reference processor (GPU).

For every case of tests/golden/processor_cases.npz (recorded from the
reference's fpie/process.py by tests/golden/make_processor_golden.py) the
product processor on the hip backend builds its system on the device
(pie_reset_equ_images / pie_reset_grid_images, and pie_prepare_mask for the
mask; the host path where the offsets send indices out of the image, as in
the "wrap" case).  A few Gauss-Seidel sweeps of that device system must give
bit for bit what the oracle's gcc loops (po_equ_gs, po_grid_gs) give on the
reference processor's recorded system.
"""

import numpy as np
import pytest

from conftest import processor_cases

pytestmark = pytest.mark.gpu

CASES = processor_cases()
SWEEPS = 25


def _bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", sorted(CASES))
def test_equ_device_system(name, oracle):
  from fpie_amd import process
  c = CASES[name]
  print(f"\nequ {name}", flush=True)
  p = process.EquProcessor(str(c["mode"]), "hip", schedule="gs_lex")
  n = p.reset(c["src"], c["mask"], c["tgt"], tuple(c["mask_on_src"]),
              tuple(c["mask_on_tgt"]))
  assert n == int(c["equ_n"])
  p.core.iterate(SWEEPS)
  want = oracle.equ_gs(c["equ_A"], c["equ_X"], c["equ_B"], SWEEPS)
  np.testing.assert_array_equal(_bits(p.core.state()), _bits(want))


@pytest.mark.parametrize("name", sorted(n for n in CASES
                                        if str(CASES[n]["what"]) == "both"))
def test_grid_device_system(name, oracle):
  from fpie_amd import process
  c = CASES[name]
  print(f"\ngrid {name}", flush=True)
  p = process.GridProcessor(str(c["mode"]), "hip", schedule="gs_lex")
  n = p.reset(c["src"], c["mask"], c["tgt"], tuple(c["mask_on_src"]),
              tuple(c["mask_on_tgt"]))
  assert n == int(c["grid_n"])
  p.core.iterate(SWEEPS)
  want = oracle.grid_gs(c["grid_mask"], c["grid_tgt"], c["grid_grad"], SWEEPS)
  np.testing.assert_array_equal(_bits(p.core.state()), _bits(want))
