"""The learned warm start on the GPU.

* Fallback through the real GridSolver: a forced bad coarse correction
  makes the warm-started state worse than the target, and the target state
  (the cold start) comes back bit for bit.
* A small training run (768 samples of 64 x 64, composites drawn at
  256 x 256, labels from 4000 batched Jacobi sweeps) must halve the loss;
  on an unseen 256 x 256 composite (4x coarse grid) three cycles with the
  trained network must end at under half the error of three cycles with the
  untrained network or with plain conjugate gradients (no network, same
  iteration count), and the warm start plus plain
  Jacobi must reach the plain method's residual in under a third of its
  sweeps with no larger error against converged plain Jacobi.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws_mod():
  from fpie_amd import native, warmstart
  native.load()
  return warmstart


def _solver(mask, tgt, grad):
  from fpie_amd.core_hip import GridSolver
  s = GridSolver(schedule="jacobi")
  s.reset(0, mask, tgt, grad)
  return s


def test_fallback_restores_the_target_state(ws_mod, monkeypatch):
  from fpie_amd import synth, unet
  mask, tgt, grad, _ = synth.composite((96, 96), 3)
  ws = ws_mod.WarmStart(unet.init_params(0), coarse=24, cycles=2)
  rng = np.random.default_rng(0)
  monkeypatch.setattr(
      ws, "coarse_solve",
      lambda rc, mc, low=None: ((1e4 * rng.normal(0, 1, rc.shape) *
                                 mc[..., None])
                      .astype(np.float32), 1, np.ones(3)))
  s = _solver(mask, tgt, grad)
  info = ws.run(s, mask, tgt)
  assert not info["used"]
  assert np.array_equal(s.state(), tgt)


def _error_after(ws_mod, params, precond, mask, tgt, grad, x_ref, cycles):
  ws = ws_mod.WarmStart(params, coarse=64, cycles=cycles, stop=0.0,
                        precond=precond)
  s = _solver(mask, tgt, grad)
  ws.run(s, mask, tgt)
  return float(np.abs(s.state() - x_ref)[mask != 0].max())


def test_trained_network_gains(ws_mod):
  from fpie_amd import synth, unet
  print("training set", flush=True)
  M, X, Y = ws_mod.training_set(768, 64, 256, sweeps=4000, workers=8,
                                log=lambda *_: None)
  print("training", flush=True)
  net, tr = unet.train(M, X, Y, epochs=12, batch=32, lr=2e-3,
                       log=lambda *_: None)
  first, last = np.mean(tr["losses"][:5]), np.mean(tr["losses"][-5:])
  print(f"loss {first:.4f} -> {last:.4f}", flush=True)
  assert last < 0.5 * first, (first, last)
  params = net.params()
  del net

  mask, tgt, grad, _ = synth.composite((256, 256), 777_777)
  m = mask != 0
  x_ref, r_floor, _ = ws_mod.converged(mask, tgt, grad, lambda *_: None,
                                       chunk=20_000)
  tol = max(1e-3, 1.05 * r_floor)
  errs = {name: _error_after(ws_mod, p, pre, mask, tgt, grad, x_ref, 3)
          for name, p, pre in (("trained", params, "net"),
                               ("untrained", unet.init_params(0), "net"),
                               ("cg", params, "none"))}
  print("max error after three cycles", errs, flush=True)
  assert errs["trained"] < 0.5 * errs["untrained"], errs
  assert errs["trained"] < 0.5 * errs["cg"], errs

  def err(s):
    return float(np.abs(s.state() - x_ref)[m].max())

  s = _solver(mask, tgt, grad)
  n_cold, _ = s.solve_to_residual(tol, 64, 10**7)
  err_cold = err(s)
  ws = ws_mod.WarmStart(params, coarse=64)
  s = _solver(mask, tgt, grad)
  info = ws.run(s, mask, tgt)
  n_warm, _ = s.solve_to_residual(tol, 64, 10**7)
  print(f"sweeps cold {n_cold} warm {info['sweeps']} + {n_warm}; "
        f"error cold {err_cold:.4g} warm {err(s):.4g}; cycles "
        f"{info['cycles']}", flush=True)
  assert info["used"]
  assert info["sweeps"] + n_warm < n_cold / 3, (info, n_warm, n_cold)
  assert err(s) <= err_cold, (err(s), err_cold)
