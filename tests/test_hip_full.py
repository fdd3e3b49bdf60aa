"""Parity at the benchmark sizes (GPU).

* 512 x 512 disc, EquSolver, 5000 sweeps: the whole run is checked bit for
  bit against the oracle's gcc Gauss-Seidel loop, red-black loop and numpy
  Jacobi loop (each a few seconds of CPU).
* 4096 x 4096 square, GridSolver: a few sweeps of every schedule are
  checked bit for bit on the full grid; the full 5000-sweep run is checked
  through exact size-independent properties: a constant image stays
  exactly constant, and translating the problem inside a larger canvas
  (which moves every tile and halo boundary) translates the result bit
  for bit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def hip():
  from fpie_amd import core_hip, native
  native.load()
  return core_hip


class _Grab:
  def reset(self, *a):
    self.a = a

  def partition(self, mask):
    flat = (mask.reshape(-1) > 0)
    return (np.cumsum(flat).reshape(mask.shape) * flat.reshape(
        mask.shape)).astype(np.int32)


def _inputs(w, method):
  from fpie_amd import process
  process.register_backend("grab", lambda **kw: _Grab(), lambda **kw: _Grab())
  cls = process.EquProcessor if method == "equ" else process.GridProcessor
  p = cls(w.gradient, "grab")
  p.reset(w.src, w.mask, w.tgt, w.mask_on_src, w.mask_on_tgt)
  return p.core.a


@pytest.fixture(scope="module")
def equ512():
  from fpie_amd import synth
  return _inputs(synth.equ512(), "equ")


@pytest.fixture(scope="module")
def grid4096():
  from fpie_amd import synth
  return _inputs(synth.grid4096(), "grid")


def _equ_core(hip, sched, N, A, X, B, mask_ids):
  core = hip.EquSolver(schedule=sched)
  core.partition((mask_ids > 0).astype(np.int32))
  core.reset(N, A, X, B)
  return core


def _mask_ids(A, N):
  """Rebuild the cropped id grid from the row-major A of the disc."""
  from fpie_amd import synth
  from fpie_amd.process import prepare_mask
  m, *_ = prepare_mask(synth.disc_mask(512, 512, 200, 256, 256))
  ids = np.zeros(m.shape, np.int32)
  ids[m > 0] = np.arange(1, N)
  return ids


def test_equ512_gs_lex_5000_is_gcc(hip, oracle, equ512):
  N, A, X, B = equ512
  assert N == 125630
  core = _equ_core(hip, "gs_lex", N, A, X, B, _mask_ids(A, N))
  u8, err = core.step(5000)
  want = oracle.equ_gs(A, X, B, 5000)
  np.testing.assert_array_equal(bits(core.state()), bits(want))
  np.testing.assert_array_equal(u8, oracle.clamp_u8(want))
  np.testing.assert_allclose(err, oracle.equ_err(A, want, B), rtol=1e-4)
  terms = oracle.equ_terms(A, want, B)
  np.testing.assert_allclose(err, terms.sum(0), rtol=1e-6)


def test_equ512_redblack_5000_is_openmp(hip, oracle, equ512):
  N, A, X, B = equ512
  ids = _mask_ids(A, N)
  core = _equ_core(hip, "redblack", N, A, X, B, ids)
  core.step(5000)
  # the openmp core numbers (i+j)-odd pixels first; values per pixel do not
  # depend on the numbering
  m = (ids > 0).astype(np.int32)
  rb_ids, n_mid = oracle.partition_redblack(m)
  rb_ids = np.where(m > 0, rb_ids, 0)
  perm = np.zeros(N, np.int64)
  perm[rb_ids[m > 0]] = ids[m > 0]
  inv = np.zeros(N, np.int64)
  inv[perm] = np.arange(N)
  A_rb = inv[A[perm]].astype(np.int32)
  want = oracle.equ_redblack(A_rb, X[perm], B[perm], n_mid, 5000, threads=16)
  np.testing.assert_array_equal(bits(core.state()[perm]), bits(want))


def test_equ512_jacobi_5000_is_numpy_order(hip, oracle, equ512):
  N, A, X, B = equ512
  core = _equ_core(hip, "jacobi", N, A, X, B, _mask_ids(A, N))
  core.step(5000)
  want = oracle.equ_jacobi(A, X, B, 5000)
  np.testing.assert_array_equal(bits(core.state()), bits(want))


@pytest.mark.parametrize("sched,iters", [("jacobi", 24), ("gs_lex", 3),
                                         ("redblack", 5)])
def test_grid4096_full_grid_matches_oracle(hip, oracle, grid4096, sched,
                                           iters):
  _, m, t, g = grid4096
  assert int(m.sum()) == 3800 * 3800
  core = hip.GridSolver(schedule=sched, order="ulrd")
  core.reset(0, m, t, g)
  u8, err = core.step(iters)
  if sched == "jacobi":
    want = oracle.grid_jacobi(m, t, g, iters, "ulrd", threads=16)
  elif sched == "gs_lex":
    want = oracle.grid_gs(m, t, g, iters)
  else:
    want = oracle.grid_colour(m, t, g, iters)
  got = core.state()
  assert np.array_equal(bits(got), bits(want)), (
      f"{int((bits(got) != bits(want)).sum())} values differ")
  np.testing.assert_array_equal(u8, oracle.clamp_u8(want))
  terms = oracle.grid_terms(m, want, g)
  np.testing.assert_allclose(err, terms.sum((0, 1)), rtol=1e-6)


def test_grid4096_constant_image_stays_exact(hip):
  h = w = 3802
  m = np.zeros((h, w), np.int32)
  m[1:-1, 1:-1] = 1
  t = np.full((h, w, 3), 255.0, np.float32)
  g = np.zeros_like(t)
  core = hip.GridSolver(schedule="jacobi")
  core.reset(0, m, t, g)
  u8, err = core.step(5000)
  assert (core.state() == 255.0).all()
  assert (u8 == 255).all()
  assert (err == 0).all()


def test_grid4096_5000_sweeps_translate_exactly(hip, grid4096):
  _, m, t, g = grid4096
  core = hip.GridSolver(schedule="jacobi")
  core.reset(0, m, t, g)
  core.step(5000)
  ref = core.state()
  dy, dx = 21, 45  # moves every stripe and row-range boundary
  pad = ((dy, 11), (dx, 3))
  m2 = np.pad(m, pad)
  t2 = np.pad(t, pad + ((0, 0),))
  g2 = np.pad(g, pad + ((0, 0),))
  core = hip.GridSolver(schedule="jacobi")
  core.reset(0, m2, t2, g2)
  core.step(5000)
  moved = core.state()[dy:dy + m.shape[0], dx:dx + m.shape[1]]
  np.testing.assert_array_equal(bits(moved), bits(ref))
