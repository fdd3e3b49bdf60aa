"""Host logic of the learned warm start (CPU).

The warm start drives a fine GridSolver through warm_net_start /
warm_apply / iterate / relax / solve_to_residual / set_state (and restrict
for the optional coarse polish, which solves the coarse problem with a
second GridSolver); here stand-ins built from the coarse-grid oracle play
both solvers, with the oracle's restatement of the learned coarse
correction, and the network is replaced:

* by a deliberately bad coarse correction taken with a forced huge step:
  the fallback must return the exact target state (the cold start);
* by the exact coarse solve (sparse direct): the learned start is then the
  coarse solution after one step, and the cycles converge to the sparse
  direct solution of the fine problem;
* by a garbage direction: the learned start never increases the coarse
  error's energy norm, and the coarse Jacobi still delivers the coarse
  solution to its tolerance.

The measurement reference (fpie_amd.exact_ref) is checked against the
sparse direct solve.
"""

import numpy as np
import pytest

import coarse_oracle as co
from fpie_amd import synth, warmstart


class CpuGrid:
  """GridSolver stand-in on the oracle (float32 Jacobi, exact residual)."""

  def __init__(self, mask=None, tgt=None, grad=None):
    if mask is not None:
      self.reset(0, mask, tgt, grad)

  def reset(self, n, mask, tgt, grad):
    self.m, self.g = mask, grad
    self.x = np.asarray(tgt, np.float32).copy()

  def restrict(self, f):
    return co.restrict(self.m, self.x, self.g, f)

  def prolong(self, f, c):
    self.x = co.prolong(self.m, self.x, c, f)

  def prolong_dots(self, f, c):
    return co.prolong_dots(self.m, self.x, self.g, c, f)

  def iterate(self, n):
    self.x = co.relax(self.m, self.x, self.g, n, 1.0)

  def relax(self, n, omega):
    self.x = co.relax(self.m, self.x, self.g, n, omega)

  def _mean(self):
    r = np.abs(co.restrict(self.m, self.x, self.g, 1)).sum()
    return float(r) / (3 * max(1, int((self.m != 0).sum())))

  def solve_to_residual(self, tol, every, max_sweeps):
    n, r = 0, self._mean()
    while r > tol and n < max_sweeps:
      k = min(every, max_sweeps - n)
      self.iterate(k)
      n += k
      r = self._mean()
    return n, r

  def set_state(self, x):
    self.x = np.asarray(x, np.float32).copy()

  def state(self):
    return self.x

  # the learned coarse correction, with `predict` standing in for the
  # network (set by each test)
  predict = None

  def warm_net_start(self, f, mc, net, net_steps, lowpass_sweeps,
                     want=False):
    rc = co.restrict(self.m, self.x, self.g, f) * mc[..., None]
    self._corr = co.net_start(rc, mc, self.predict, net_steps, lowpass_sweeps)
    return self._corr.copy() if want else None

  def warm_apply(self, f, corr=None):
    c = self._corr if corr is None else np.asarray(corr, np.float32)
    num, den = self.prolong_dots(f, c)
    alpha = np.maximum(np.where(den > 0, num / np.where(den > 0, den, 1.0),
                                0.0), 0.0)
    cc = (c * alpha).astype(np.float32)
    self.prolong(f, cc)
    return alpha, float(np.abs(cc).max()) if cc.size else 0.0


def _ws(**kw):
  return warmstart.WarmStart(np.zeros(1, np.float32), coarse_solver=CpuGrid,
                             **kw)


def _coarse_exact(rc, mc):
  return co.exact_solution(mc, np.zeros(mc.shape + (3,)),
                           rc * mc[..., None]).astype(np.float32)


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
  return mc, rc, _coarse_exact(rc, mc).astype(np.float64), CpuGrid(
      mask, tgt, grad)


def test_coarse_solve_without_network_is_coarse_jacobi():
  mc, rc, e, s = _coarse_problem()
  ws = _ws(precond="none", coarse_rtol=1e-4)
  c, info = ws.coarse_solve(s, 4, mc)
  assert info["coarse_rel_residual"] <= 1e-4 and info["coarse_sweeps"] > 0
  # the exact coarse solution within what a 1e-4 residual leaves
  assert np.abs(c - e).max() < 0.05 * np.abs(e).max()


def test_learned_start_energy_never_grows():
  mc, rc, e, s = _coarse_problem(seed=19)
  rng = np.random.default_rng(0)
  noise = [rng.normal(0, 1, rc.shape) for _ in range(6)]

  def energy(v):
    return co.dot(v, co.coarse_apply(mc, v))

  prev = energy(e)
  for k in range(1, 7):
    calls = iter(noise)
    c = co.net_start(rc, mc, lambda r, m: next(calls), k, 4)
    cur = energy(e - c)
    assert (cur <= prev * (1 + 1e-9) + 1e-9).all(), (k, cur, prev)
    prev = cur
  # and the coarse Jacobi still reaches its tolerance from that start
  calls = iter(noise)
  s.predict = lambda r, m: next(calls)
  ws = _ws(coarse_rtol=1e-3)
  c, info = ws.coarse_solve(s, 4, mc)
  assert info["coarse_rel_residual"] <= 1e-3


def test_exact_direction_gives_the_coarse_solution():
  mc, rc, e, s = _coarse_problem(seed=21)
  c = co.net_start(rc, mc, _coarse_exact, 1, 0)
  np.testing.assert_allclose(c, e, rtol=0, atol=1e-5 * np.abs(e).max())
  s.predict = _coarse_exact
  ws = _ws(lowpass_sweeps=0, net_steps=1, coarse_rtol=1e-3)
  _, info = ws.coarse_solve(s, 4, mc)
  assert info["coarse_sweeps"] == 0  # the polish finds it converged


def test_lowpass_keeps_smooth_fields_and_removes_checkerboards():
  n = 40
  mc = np.zeros((n, n), bool)
  mc[1:-1, 1:-1] = True
  i, j = np.mgrid[0:n, 0:n]
  smooth = np.sin(np.pi * i / (n - 1)) * np.sin(np.pi * j / (n - 1))
  smooth = np.repeat(smooth[..., None], 3, -1) * mc[..., None]
  checker = ((-1.0) ** (i + j))[..., None] * smooth
  out_s = co.lowpass(mc, smooth, 4)
  out_c = co.lowpass(mc, checker, 4)
  assert np.abs(out_s - smooth).max() < 0.02 * np.abs(smooth).max()
  assert np.abs(out_c).max() < 0.02 * np.abs(checker).max()


def test_exact_reference_matches_sparse_direct_solve():
  from fpie_amd import exact_ref
  mask, tgt, grad, _ = synth.composite((48, 56), 41)
  want = co.exact_solution(mask, tgt, grad)
  got = exact_ref.solve(mask, tgt, grad, tgt, log=lambda *_: None)
  np.testing.assert_allclose(got, want, rtol=0, atol=1e-8)


def test_bad_prediction_falls_back_to_cold_start(monkeypatch):
  mask, tgt, grad, _ = synth.composite((64, 64), 3)
  ws = _ws(coarse=16, cycles=2)
  rng = np.random.default_rng(1)
  monkeypatch.setattr(
      ws, "coarse_solve",
      lambda solver, f, mc, cs=None: ((1e4 * rng.normal(0, 1, mc.shape + (3,))
                                       * mc[..., None]).astype(np.float32),
                                      {"coarse_sweeps": 0}))
  s = CpuGrid(mask, tgt, grad)
  # and a forced huge step past the line search
  s.prolong_dots = lambda f, c: np.array([[1e4] * 3, [1.0] * 3])
  info = ws.run(s, mask, tgt)
  assert not info["used"]
  assert info["start_residual_warm"] >= info["start_residual_cold"]
  assert np.array_equal(s.state(), tgt)


def test_cycles_with_exact_coarse_solve_converge(monkeypatch):
  mask, tgt, grad, _ = synth.composite((128, 128), 23)
  xs = co.exact_solution(mask, tgt, grad)
  m = mask != 0
  ws = _ws(coarse=32, cycles=4, stop=0.0, lowpass_sweeps=0, net_steps=1)
  s = CpuGrid(mask, tgt, grad)
  s.predict = _coarse_exact
  info = ws.run(s, mask, tgt)
  assert info["used"] and info["factor"] == 4 and info["cycles"] == 4
  assert info["coarse_sweeps"] == 0  # the exact start needs no coarse sweep
  err0 = np.abs(tgt - xs)[m].max()
  err = np.abs(s.state() - xs)[m].max()
  assert err < err0 / 100, (err0, err)
