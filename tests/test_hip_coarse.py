"""Coarse-grid correction calls of the GridSolver against the oracle (GPU).

pie_restrict is checked to float32 rounding of the block sums (the kernel
and the oracle both add float64 terms, in different orders), pie_prolong
and pie_relax bit for bit (same float32 operations, no contraction).
"""

import numpy as np
import pytest

import coarse_oracle as co

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
  from fpie_amd import core_hip, native
  native.load()
  return core_hip


def _bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _loaded(hip, h, w, seed):
  from fpie_amd import synth
  mask, tgt, grad, _ = synth.composite((h, w), seed)
  x = (tgt + np.random.default_rng(seed).normal(0, 15, tgt.shape))
  x = np.where(mask[..., None] != 0, x, tgt).astype(np.float32)
  s = hip.GridSolver(schedule="jacobi")
  s.reset(0, mask, tgt, grad)
  s.set_state(x)
  return s, mask, tgt, grad, x


@pytest.mark.parametrize("h,w,f", [(64, 80, 1), (64, 80, 4), (67, 91, 3),
                                   (130, 200, 8), (300, 257, 16)])
def test_restrict_matches_oracle(hip, h, w, f):
  print(f"restrict {h}x{w} factor {f}", flush=True)
  s, mask, _, grad, x = _loaded(hip, h, w, 100 + f)
  got = s.restrict(f)
  want = co.restrict(mask, x, grad, f)
  assert got.shape == (h // f, w // f, 3)
  np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-4)


@pytest.mark.parametrize("h,w,f", [(64, 80, 4), (67, 91, 3), (300, 257, 16)])
def test_prolong_matches_oracle(hip, h, w, f):
  print(f"prolong {h}x{w} factor {f}", flush=True)
  s, mask, _, _, x = _loaded(hip, h, w, 200 + f)
  corr = np.random.default_rng(f).normal(0, 30, (h // f, w // f, 3))
  corr = corr.astype(np.float32)
  s.prolong(f, corr)
  assert np.array_equal(_bits(s.state()), _bits(co.prolong(mask, x, corr, f)))


@pytest.mark.parametrize("h,w,f", [(64, 80, 4), (67, 91, 3), (300, 257, 16),
                                   (64, 80, 1)])
def test_prolong_dots_match_oracle(hip, h, w, f):
  print(f"prolong_dots {h}x{w} factor {f}", flush=True)
  s, mask, _, grad, x = _loaded(hip, h, w, 300 + f)
  corr = np.random.default_rng(f).normal(0, 30, (h // f, w // f, 3))
  corr = corr.astype(np.float32)
  got = s.prolong_dots(f, corr)
  want = co.prolong_dots(mask, x, grad, corr, f)
  np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)
  assert np.array_equal(_bits(s.state()), _bits(x))  # state untouched


@pytest.mark.parametrize("omega,sweeps", [(0.5, 3), (1.0, 2), (0.8, 1)])
def test_relax_matches_oracle(hip, omega, sweeps):
  print(f"relax omega {omega} sweeps {sweeps}", flush=True)
  s, mask, _, grad, x = _loaded(hip, 70, 96, 7)
  s.relax(sweeps, omega)
  want = co.relax(mask, x, grad, sweeps, omega)
  assert np.array_equal(_bits(s.state()), _bits(want))


def test_relax_then_sweeps_continue_from_the_relaxed_state(hip):
  s, mask, _, grad, x = _loaded(hip, 70, 96, 8)
  s.relax(1, 0.5)
  s.iterate(9)  # the sweep kernel starts from the relaxed buffer
  want = co.relax(mask, co.relax(mask, x, grad, 1, 0.5), grad, 9, 1.0)
  assert np.array_equal(_bits(s.state()), _bits(want))


def test_coarse_calls_reject_bad_arguments(hip):
  s, *_ = _loaded(hip, 40, 48, 9)
  with pytest.raises(ValueError):
    s.restrict(0)
  with pytest.raises(RuntimeError):
    s.restrict(64)
  with pytest.raises(ValueError):
    s.prolong(4, np.zeros((3, 3, 3), np.float32))
  with pytest.raises(RuntimeError):
    s.relax(1, 0.0)
  with pytest.raises(RuntimeError):
    s.relax(1, 1.5)
  from fpie_amd import native
  e = hip.EquSolver(256, "jacobi")
  out = np.zeros(64 * 3, np.float32)
  with pytest.raises(RuntimeError):
    native.check(e._lib.pie_restrict(e._h, 2, out.ctypes.data))
