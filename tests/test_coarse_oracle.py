"""Pins of the coarse-grid correction oracle (CPU).

Each check is chosen so that a plausible slip (wrong sign, a missing
neighbour, an off-by-half in the cell-centred upsampling, a transposed
block index) fails it:

* restriction at factor 1 equals b - A x from a Laplacian assembled
  independently (Kronecker sums on the full grid);
* restriction at factor f equals the f x f block sums of factor 1, also on
  sizes that f does not divide;
* an exact fixed point (a linear ramp, discrete Laplacian 0) restricts to 0
  exactly; the sparse direct solution restricts to float32 noise only and
  is where the reference's Jacobi loop (C restatement) converges;
* upsampling a dyadic constant adds it exactly; a dyadic linear coarse field
  is reproduced exactly at the fine pixel centres between coarse centres;
* weighted Jacobi with omega 1 is the reference's Jacobi sweep bit for bit,
  omega 1/2 the mean of two consecutive sweeps;
* the line-search sums: for the exact error as the correction (factor 1),
  r . d equals d . A d (r = A e), and the energy-optimal multiple of a
  random correction never raises the error's energy norm;
* the two-grid cycle built from these pieces (exact coarse solve) cuts the
  error of a composite by a factor of 2.5 or more per cycle after the first.
"""

import numpy as np
import pytest
import scipy.sparse as sp

import coarse_oracle as co


def _problem(h, w, seed):
  from fpie_amd import synth
  return synth.composite((h, w), seed)


def _laplacian_residual(mask, x, grad):
  """g + sum(neighbours) - 4 x at solved pixels via a full-grid 5-point
  operator assembled as Kronecker sums (float64)."""
  h, w = mask.shape
  def lap1(n):
    return sp.diags([np.ones(n - 1), -2 * np.ones(n), np.ones(n - 1)],
                    [-1, 0, 1])
  L = sp.kron(lap1(h), sp.eye(w)) + sp.kron(sp.eye(h), lap1(w))
  out = np.zeros((h, w, 3))
  for c in range(3):
    out[..., c] = (L @ x[..., c].astype(np.float64).reshape(-1)).reshape(h, w)
  out += grad
  return out * (mask[..., None] != 0)


def test_restrict_factor1_is_the_residual():
  mask, tgt, grad, _ = _problem(40, 52, 3)
  rng = np.random.default_rng(0)
  x = (tgt + rng.normal(0, 20, tgt.shape)).astype(np.float32)
  got = co.restrict(mask, x, grad, 1)
  want = _laplacian_residual(mask, x, grad)
  np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("f", [2, 3, 4, 7])
def test_restrict_is_block_sum(f):
  mask, tgt, grad, _ = _problem(45, 61, 4)
  x = np.random.default_rng(1).uniform(0, 255, tgt.shape).astype(np.float32)
  fine = co.restrict(mask, x, grad, 1).astype(np.float64)
  got = co.restrict(mask, x, grad, f)
  hc, wc = 45 // f, 61 // f
  assert got.shape == (hc, wc, 3)
  for I in (0, hc // 2, hc - 1):
    for J in (0, wc // 3, wc - 1):
      want = fine[I * f:(I + 1) * f, J * f:(J + 1) * f].sum((0, 1))
      np.testing.assert_allclose(got[I, J], want, rtol=1e-5, atol=1e-3)


def test_restrict_of_exact_fixed_points():
  from fpie_amd import process
  h, w = 30, 34
  ii, jj = np.mgrid[0:h, 0:w]
  ramp = np.stack([2 * ii + jj, 3 * jj, 100 + ii - jj], -1).astype(np.float32)
  mask = np.zeros((h, w), np.int32)
  mask[3:-3, 4:-4] = 1
  grad = process.divergence(ramp, ramp, mask,
                            lambda a, b: process.mixgrad(a, b, "src"))
  assert not co.restrict(mask, ramp, grad, 1).any()
  assert not co.restrict(mask, ramp, grad, 5).any()
  m2, t2, g2, _ = _problem(36, 40, 9)
  xs = co.exact_solution(m2, t2, g2).astype(np.float32)
  assert np.abs(co.restrict(m2, xs, g2, 1)).max() < 2e-3


def test_exact_solution_is_where_jacobi_converges(oracle):
  mask, tgt, grad, _ = _problem(20, 24, 11)
  xs = co.exact_solution(mask, tgt, grad)
  xj = oracle.grid_jacobi(mask, tgt, grad, 4000, "ulrd")
  m = mask != 0
  assert np.abs(xj - xs)[m].max() < 1e-3
  assert np.array_equal(xs[~m], tgt[~m])


def test_prolong_constant_and_linear():
  h, w, f = 32, 40, 4
  mask = np.zeros((h, w), np.int32)
  mask[2:-2, 3:-3] = 1
  x = np.random.default_rng(2).uniform(0, 255, (h, w, 3)).astype(np.float32)
  x = np.round(x * 8) / 8  # dyadic, so every sum below is exact
  hc, wc = h // f, w // f
  const = np.full((hc, wc, 3), 2.5, np.float32)
  got = co.prolong(mask, x, const, f)
  m = mask != 0
  assert np.array_equal(got[m], x[m] + np.float32(2.5))
  assert np.array_equal(got[~m], x[~m])
  I, J = np.mgrid[0:hc, 0:wc]
  lin = np.stack([I * 2.0 - J, J * 0.5, I + J], -1).astype(np.float32)
  got = co.prolong(np.ones((h, w), np.int32), np.zeros_like(x), lin, f)
  i, j = np.mgrid[0:h, 0:w]
  ys, xs = (i + 0.5) / f - 0.5, (j + 0.5) / f - 0.5
  want = np.stack([ys * 2.0 - xs, xs * 0.5, ys + xs], -1)
  inner = (ys >= 0) & (ys <= hc - 1) & (xs >= 0) & (xs <= wc - 1)
  np.testing.assert_array_equal(got[inner], want[inner].astype(np.float32))


def test_relax_weights(oracle):
  mask, tgt, grad, _ = _problem(33, 47, 5)
  x = np.random.default_rng(3).uniform(0, 255, tgt.shape).astype(np.float32)
  one = co.relax(mask, x, grad, 1, 1.0)
  ref = oracle.grid_jacobi(mask, x, grad, 1, "ulrd")
  assert np.array_equal(one.view(np.uint32), ref.view(np.uint32))
  half = co.relax(mask, x, grad, 1, 0.5)
  np.testing.assert_allclose(half, 0.5 * (x.astype(np.float64) + ref),
                             rtol=2e-7, atol=1e-5)


def test_prolong_dots_line_search():
  mask, tgt, grad, _ = _problem(40, 44, 13)
  m = mask != 0
  xs = co.exact_solution(mask, tgt, grad)
  x = tgt.astype(np.float32)
  e = (xs - x).astype(np.float32) * m[..., None]
  rd, dad = co.prolong_dots(mask, x, grad, e, 1)
  np.testing.assert_allclose(rd, dad, rtol=1e-5)

  def energy(v):  # e^T A e from the exact solution, per channel
    err = (xs - v) * m[..., None]
    return co.prolong_dots(mask, np.zeros_like(x), np.zeros_like(grad),
                           err.astype(np.float32), 1)[1]

  rng = np.random.default_rng(4)
  for f in (1, 2, 4):
    corr = rng.normal(0, 20, (40 // f, 44 // f, 3)).astype(np.float32)
    rd, dad = co.prolong_dots(mask, x, grad, corr, f)
    a = rd / dad
    new = co.prolong(mask, x, (corr * a).astype(np.float32), f)
    assert (energy(new) <= energy(x) + 1e-6).all(), f


def _two_grid_errors(mask, tgt, grad, f, cycles, sweeps):
  xs = co.exact_solution(mask, tgt, grad)
  m = mask != 0
  hc, wc = mask.shape[0] // f, mask.shape[1] // f
  mc = (m[:hc * f, :wc * f].reshape(hc, f, wc, f).mean((1, 3)) > 0.5)
  mc[[0, -1], :] = False
  mc[:, [0, -1]] = False
  x = tgt.astype(np.float32).copy()
  errs = [np.abs(x - xs)[m].max()]
  for _ in range(cycles):
    rc = co.restrict(mask, x, grad, f)
    cc = co.exact_solution(mc, np.zeros((hc, wc, 3)), rc * mc[..., None])
    x = co.prolong(mask, x, cc.astype(np.float32), f)
    x = co.relax(mask, x, grad, sweeps, 1.0)
    x = co.relax(mask, x, grad, 1, 0.5)
    errs.append(np.abs(x - xs)[m].max())
  return errs


def test_two_grid_cycle_converges():
  mask, tgt, grad, _ = _problem(96, 96, 21)
  errs = _two_grid_errors(mask, tgt, grad, 4, 3, 40)
  assert errs[1] < errs[0] / 1.5, errs
  assert errs[2] < errs[1] / 2.5, errs
  assert errs[3] < errs[2] / 2.5, errs
  assert errs[3] < errs[0] / 15, errs


# -- the learned coarse correction's flexible CG (coarse_oracle.net_start) --
#
# * a perfect predictor (the exact coarse error A_c^-1 r, A_c assembled
#   independently as a Kronecker sum restricted to the solved cells) makes
#   one step return the exact correction, and a second step (zero residual,
#   zero direction) change nothing;
# * predicting the residual itself gives one steepest-descent step,
#   c = (r.r / r.A r) r;
# * with arbitrary predictions the final residual is orthogonal to every
#   direction taken (the Galerkin property of A-orthogonalised directions
#   with exact line searches), which a wrong sign or a missed
#   orthogonalisation breaks;
# * the low-pass smoother halves an isolated cell per sweep and averages a
#   pair as (v + v / 4) / 2.


def _coarse_case(seed, hc=12, wc=14):
  rng = np.random.default_rng(seed)
  mc = np.zeros((hc, wc), bool)
  mc[1:-1, 1:-1] = rng.random((hc - 2, wc - 2)) < 0.85
  rc = rng.normal(0.0, 1.0, (hc, wc, 3)) * mc[..., None]
  return rc, mc


def _coarse_matrix(mc):
  """A_c on the solved cells: 4 on the diagonal, -1 between 4-neighbours,
  unsolved cells (and the outside) held at zero."""
  hc, wc = mc.shape
  def t(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)],
                    [-1, 0, 1])
  L = (sp.kron(t(hc), sp.eye(wc)) + sp.kron(sp.eye(hc), t(wc))).tocsr()
  idx = np.flatnonzero(mc.reshape(-1))
  return L[idx][:, idx], idx


def _apply(mc, p):
  A, idx = _coarse_matrix(mc)
  out = np.zeros(p.shape)
  for c in range(3):
    out.reshape(-1, 3)[idx, c] = A @ p.reshape(-1, 3)[idx, c]
  return out


def _solve(mc, r):
  import scipy.sparse.linalg as spl
  A, idx = _coarse_matrix(mc)
  out = np.zeros(r.shape)
  for c in range(3):
    out.reshape(-1, 3)[idx, c] = spl.spsolve(A.tocsc(), r.reshape(-1, 3)[idx, c])
  return out


@pytest.mark.parametrize("seed", [1, 2])
def test_net_start_with_a_perfect_predictor_is_exact(seed):
  rc, mc = _coarse_case(seed)
  exact = _solve(mc, rc)
  one = co.net_start(rc, mc, lambda r, m: _solve(m, r), 1, 0)
  np.testing.assert_allclose(one, exact, rtol=1e-6, atol=1e-6 * np.abs(exact).max())
  two = co.net_start(rc, mc, lambda r, m: _solve(m, r), 2, 0)
  np.testing.assert_allclose(two, exact, rtol=1e-6, atol=1e-6 * np.abs(exact).max())


def test_net_start_steepest_descent_step():
  rc, mc = _coarse_case(3)
  got = co.net_start(rc, mc, lambda r, m: r, 1, 0)
  ar = _apply(mc, rc)
  a = np.einsum("ijc,ijc->c", rc, rc) / np.einsum("ijc,ijc->c", rc, ar)
  np.testing.assert_allclose(got, (rc * a).astype(np.float32), rtol=1e-6,
                             atol=1e-7)


@pytest.mark.parametrize("steps,sweeps", [(2, 0), (3, 2), (4, 4)])
def test_net_start_residual_orthogonal_to_its_directions(steps, sweeps):
  rc, mc = _coarse_case(10 + steps)
  rng = np.random.default_rng(steps)
  taken = []
  def predict(r, m):
    z = rng.normal(0.0, 1.0, r.shape)
    taken.append(z)
    return z
  c = co.net_start(rc, mc, predict, steps, sweeps).astype(np.float64)
  assert len(taken) == steps
  res = rc - _apply(mc, c)
  for z in taken:
    d = co.lowpass(mc, z, sweeps) * mc[..., None]
    num = np.abs(np.einsum("ijc,ijc->c", res, d))
    scale = np.linalg.norm(rc) * np.linalg.norm(d)
    assert (num <= 1e-5 * scale).all(), (num, scale)
  # and the correction lowers the energy of the error below the start's
  e0 = _solve(mc, rc)
  def energy(e):
    return np.einsum("ijc,ijc->c", e, _apply(mc, e))
  assert (energy(e0 - c) < energy(e0)).all()


def test_lowpass_isolated_cell_and_pair():
  mc = np.zeros((5, 6), bool)
  mc[2, 2] = True
  p = np.zeros((5, 6, 3))
  p[2, 2] = [1.0, -2.0, 4.0]
  for s in range(4):
    np.testing.assert_array_equal(co.lowpass(mc, p, s)[2, 2],
                                  p[2, 2] / 2.0 ** s)
  mc[2, 3] = True
  p[2, 3] = p[2, 2]
  q = co.lowpass(mc, p, 1)
  np.testing.assert_array_equal(q[2, 2], (p[2, 2] + p[2, 3] / 4) / 2)
  np.testing.assert_array_equal(q[2, 3], q[2, 2])
  assert (q[mc == 0] == 0).all()
