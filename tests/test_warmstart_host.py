"""Host logic of the learned warm start (CPU).

The warm start drives a GridSolver through restrict / prolong / iterate /
relax / solve_to_residual / set_state; here a stand-in built from the
coarse-grid oracle plays the solver, and the network is replaced:

* by a deliberately bad coarse correction: the fallback must return the
  exact target state (the cold start);
* by the exact coarse solve (sparse direct) as preconditioner: one network
  call solves the coarse problem and the cycles converge to the sparse
  direct solution of the fine problem;
* without a network (plain conjugate gradients) the coarse solve converges
  to the sparse direct coarse solution, and with a garbage preconditioner
  the coarse error's energy norm never grows from one iteration to the
  next.

The training-data residual (host numpy) is checked against the oracle's
restriction.
"""

import numpy as np
import pytest

import coarse_oracle as co
from fpie_amd import synth, warmstart


class CpuGrid:
  def __init__(self, mask, tgt, grad):
    self.m, self.g = mask, grad
    self.x = np.asarray(tgt, np.float32).copy()

  def restrict(self, f):
    return co.restrict(self.m, self.x, self.g, f)

  def prolong(self, f, c):
    self.x = co.prolong(self.m, self.x, c, f)

  def iterate(self, n):
    self.x = co.relax(self.m, self.x, self.g, n, 1.0)

  def relax(self, n, omega):
    self.x = co.relax(self.m, self.x, self.g, n, omega)

  def solve_to_residual(self, tol, every, max_sweeps):
    r = np.abs(co.restrict(self.m, self.x, self.g, 1)).sum()
    return 0, float(r) / (3 * max(1, int((self.m != 0).sum())))

  def set_state(self, x):
    self.x = np.asarray(x, np.float32).copy()

  def state(self):
    return self.x


def _coarse_exact(rc, mc):
  return co.exact_solution(mc, np.zeros(mc.shape + (3,)),
                           rc * mc[..., None]).astype(np.float32)


def test_training_residual_matches_oracle_restriction():
  mask, tgt, grad, _ = synth.composite((64, 72), 5)
  for f in (1, 4):
    host = warmstart.block_sum(warmstart.signed_residual(mask, tgt, grad), f)
    np.testing.assert_allclose(host, co.restrict(mask, tgt, grad, f),
                               rtol=1e-6, atol=1e-3)


def test_coarse_mask_majority_and_ring():
  m = np.zeros((16, 16), np.int32)
  m[2:14, 2:14] = 1
  m[4:6, 4:5] = 0  # 2 of the 16 pixels of block (1, 1): still a majority
  mc = warmstart.coarse_mask(m, 4)
  assert mc.shape == (4, 4)
  assert not mc[0].any() and not mc[-1].any()
  assert not mc[:, 0].any() and not mc[:, -1].any()
  assert mc[1, 1] and mc[2, 2]
  m[8:12, 8:12] = 0
  m[8, 8] = 1  # 1 of 16 solved: not a majority
  assert not warmstart.coarse_mask(m, 4)[2, 2]


def _coarse_problem(size=96, seed=17, f=4):
  mask, tgt, grad, _ = synth.composite((size, size), seed)
  mc = warmstart.coarse_mask(mask, f)
  rc = co.restrict(mask, tgt, grad, f) * mc[..., None]
  return mc, rc, _coarse_exact(rc, mc).astype(np.float64)


def test_coarse_solve_without_network_is_conjugate_gradients():
  mc, rc, e = _coarse_problem()
  ws = warmstart.WarmStart(np.zeros(1, np.float32), precond="none",
                           coarse_tol=1e-9, coarse_iters=400)
  c, calls, rel = ws.coarse_solve(rc, mc)
  assert (rel < 1e-9).all() and calls < 400
  np.testing.assert_allclose(c, e, rtol=0, atol=1e-4 * np.abs(e).max())


def test_coarse_solve_energy_never_grows(monkeypatch):
  mc, rc, e = _coarse_problem(seed=19)
  rng = np.random.default_rng(0)
  ws = warmstart.WarmStart(np.zeros(1, np.float32), coarse_tol=0.0)
  noise = [rng.normal(0, 1, rc.shape) for _ in range(6)]

  def energy(v):
    return warmstart._dot(v, warmstart.coarse_apply(mc, v))

  prev = energy(e)
  for k in range(1, 7):
    calls = iter(noise)
    monkeypatch.setattr(ws, "predict", lambda r, m: next(calls))
    ws.coarse_iters = k
    c, _, _ = ws.coarse_solve(rc, mc)
    cur = energy(e - c)
    assert (cur <= prev * (1 + 1e-9) + 1e-9).all(), (k, cur, prev)
    prev = cur


def test_exact_preconditioner_solves_in_one_call(monkeypatch):
  mc, rc, e = _coarse_problem(seed=21)
  ws = warmstart.WarmStart(np.zeros(1, np.float32), coarse_tol=1e-6,
                           lowpass_sweeps=0)
  monkeypatch.setattr(ws, "predict", _coarse_exact)
  c, calls, rel = ws.coarse_solve(rc, mc)
  assert calls == 1 and (rel < 1e-6).all()
  np.testing.assert_allclose(c, e, rtol=0, atol=1e-5 * np.abs(e).max())


def test_lowpass_keeps_smooth_fields_and_removes_checkerboards():
  n = 40
  mc = np.zeros((n, n), bool)
  mc[1:-1, 1:-1] = True
  i, j = np.mgrid[0:n, 0:n]
  smooth = np.sin(np.pi * i / (n - 1)) * np.sin(np.pi * j / (n - 1))
  smooth = np.repeat(smooth[..., None], 3, -1) * mc[..., None]
  checker = ((-1.0) ** (i + j))[..., None] * smooth
  out_s = warmstart.lowpass(mc, smooth, 4)
  out_c = warmstart.lowpass(mc, checker, 4)
  assert np.abs(out_s - smooth).max() < 0.02 * np.abs(smooth).max()
  assert np.abs(out_c).max() < 0.02 * np.abs(checker).max()


def test_coarsest_direction_is_the_exact_coarsest_solve():
  mc, rc, _ = _coarse_problem(size=256, seed=31)
  low = warmstart.CoarsestSolve(mc, 4)
  mcc = low.mcc
  rcc = co.restrict(mc.astype(np.int32), np.zeros(mc.shape + (3,)), rc, 4)
  ecc = co.exact_solution(mcc, np.zeros(mcc.shape + (3,)),
                          rcc * mcc[..., None])
  want = co.prolong(mc, np.zeros(mc.shape + (3,), np.float32),
                    ecc.astype(np.float32), 4)
  got = low.direction(rc, mc)
  np.testing.assert_allclose(got, want, rtol=0,
                             atol=1e-5 * np.abs(want).max())


def test_exact_reference_matches_sparse_direct_solve():
  from fpie_amd import exact_ref
  mask, tgt, grad, _ = synth.composite((48, 56), 41)
  want = co.exact_solution(mask, tgt, grad)
  got = exact_ref.solve(mask, tgt, grad, tgt, log=lambda *_: None)
  np.testing.assert_allclose(got, want, rtol=0, atol=1e-8)


def test_bad_prediction_falls_back_to_cold_start(monkeypatch):
  mask, tgt, grad, _ = synth.composite((64, 64), 3)
  ws = warmstart.WarmStart(np.zeros(1, np.float32), coarse=16, cycles=2)
  rng = np.random.default_rng(1)
  monkeypatch.setattr(
      ws, "coarse_solve",
      lambda rc, mc, low=None: ((1e4 * rng.normal(0, 1, rc.shape) *
                                 mc[..., None])
                      .astype(np.float32), 1, np.ones(3)))
  s = CpuGrid(mask, tgt, grad)
  info = ws.run(s, mask, tgt)
  assert not info["used"]
  assert info["start_residual_warm"] >= info["start_residual_cold"]
  assert np.array_equal(s.state(), tgt)


def test_cycles_with_exact_coarse_solve_converge(monkeypatch):
  mask, tgt, grad, _ = synth.composite((128, 128), 23)
  xs = co.exact_solution(mask, tgt, grad)
  m = mask != 0
  ws = warmstart.WarmStart(np.zeros(1, np.float32), coarse=32, cycles=4,
                           stop=0.0, lowpass_sweeps=0)
  monkeypatch.setattr(ws, "predict", _coarse_exact)
  s = CpuGrid(mask, tgt, grad)
  info = ws.run(s, mask, tgt)
  assert info["used"] and info["factor"] == 4 and info["cycles"] == 4
  assert all(st["coarse_iters"] == 1 for st in info["steps"])
  err0 = np.abs(tgt - xs)[m].max()
  err = np.abs(s.state() - xs)[m].max()
  assert err < err0 / 100, (err0, err)
