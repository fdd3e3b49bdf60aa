"""The learned warm start on the GPU.

* Fallback through the real GridSolver: a forced bad coarse correction,
  taken with a forced huge step, makes the warm-started state worse than
  the target, and the target state (the cold start) comes back bit for
  bit.
* A small training run (768 samples of 64 x 64, composites drawn at
  256 x 256, labels from 4000 batched Jacobi sweeps) must halve the loss;
  on an unseen 256 x 256 composite (4x coarse grid) three cycles with the
  trained network must end at under half the error of three cycles with the
  untrained network and under a tenth of the error of the same cycles
  without a network; and the warm start plus plain
  Jacobi must reach the plain method's residual in under a third of its
  sweeps with no larger error against converged plain Jacobi.
* The packaged network (fpie_amd/data/warmstart_unet.npz) on two unseen
  512 x 512 composites (gradient modes max and src): the warm start reaches
  the plain method's residual in under a tenth of its sweeps and ends
  closer to the exact solution (sparse direct solve, float64) than plain
  Jacobi does, and ten times closer than the same cycles without the
  network.
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
      lambda solver, f, mc, cs=None: ((1e4 * rng.normal(0, 1, mc.shape + (3,))
                                       * mc[..., None]).astype(np.float32),
                                      {"coarse_sweeps": 0}))
  s = _solver(mask, tgt, grad)
  # and a forced huge step past the line search
  prolong = s.prolong

  def bad_apply(f, cc=None):
    c = (cc * 1e4).astype(np.float32)
    prolong(f, c)
    return np.full(3, 1e4), float(np.abs(c).max())

  s.warm_apply = bad_apply
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
                               ("none", params, "none"))}
  print("max error after three cycles", errs, flush=True)
  assert errs["trained"] < 0.5 * errs["untrained"], errs
  assert errs["trained"] < 0.1 * errs["none"], errs

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


def test_packaged_network_beats_plain_jacobi(ws_mod):
  import coarse_oracle as co
  from fpie_amd import synth
  for seed in (900_123, 900_001):  # gradient modes max and src
    mask, tgt, grad, mode = synth.composite((512, 512), seed)
    m = mask != 0
    print(f"seed {seed} ({mode}): exact solution (CPU)", flush=True)
    x_star = co.exact_solution(mask, tgt, grad)

    def err(s):
      return float(np.abs(s.state() - x_star)[m].max())

    s = _solver(mask, tgt, grad)
    n_cold, _ = s.solve_to_residual(1e-3, 256, 10**8)
    err_cold = err(s)
    params = ws_mod.load_params()
    runs = {}
    for pre in ("net", "none"):
      ws = ws_mod.WarmStart(params, precond=pre)
      s = _solver(mask, tgt, grad)
      info = ws.run(s, mask, tgt)
      n, _ = s.solve_to_residual(1e-3, 256, 10**8)
      runs[pre] = (info, n, err(s))
    info, n_warm, err_warm = runs["net"]
    print(f"seed {seed}: sweeps cold {n_cold}, warm {info['sweeps']} + "
          f"{n_warm}; max error vs exact: cold {err_cold:.4g} warm "
          f"{err_warm:.4g} without network {runs['none'][2]:.4g}",
          flush=True)
    assert info["used"]
    assert info["sweeps"] + n_warm < n_cold / 10, (info, n_warm, n_cold)
    assert err_warm < err_cold, (err_warm, err_cold)
    assert err_warm < runs["none"][2] / 10, (err_warm, runs["none"][2])


def test_packaged_network_at_1024(ws_mod):
  """The study's 1024^2 case (seed 900001, gradient src, 406,377 solved
  pixels): the packaged network's warm start ends below the tolerance in a
  tenth of the cold sweeps and closer to the exact solution (the oracle's
  sparse direct solve) than the cold run."""
  import coarse_oracle as co
  from fpie_amd import synth
  mask, tgt, grad, mode = synth.composite((1024, 1024), 900_001)
  m = mask != 0
  print(f"1024^2 ({mode}): exact solution (CPU)", flush=True)
  x_star = co.exact_solution(mask, tgt, grad)

  def err(s):
    return float(np.abs(s.state() - x_star)[m].max())

  s = _solver(mask, tgt, grad)
  n_cold, _ = s.solve_to_residual(1e-3, 256, 10**8)
  err_cold = err(s)
  ws = ws_mod.WarmStart(ws_mod.load_params())
  s = _solver(mask, tgt, grad)
  info = ws.run(s, mask, tgt)
  n_warm, _ = s.solve_to_residual(1e-3, 256, 10**8)
  err_warm = err(s)
  print(f"1024^2: sweeps cold {n_cold}, warm {info['sweeps']} + {n_warm}; "
        f"max error vs exact: cold {err_cold:.4g} warm {err_warm:.4g}",
        flush=True)
  assert info["used"]
  assert info["sweeps"] + n_warm < n_cold / 10, (info, n_warm, n_cold)
  assert err_warm < err_cold, (err_warm, err_cold)
