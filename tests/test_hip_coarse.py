"""Coarse-grid correction calls of the GridSolver against the oracle (GPU).

pie_restrict is checked to float32 rounding of the block sums (the kernel
and the oracle both add float64 terms, in different orders), pie_prolong
and pie_relax bit for bit (same float32 operations, no contraction).

pie_warm_net_start against the oracle's restatement driven by the same
network (the packaged weights, forward pass on the host path of the same
library): the two differ only where the per-channel rms of the residual,
summed in another order, moves a network input across a bf16 rounding
boundary, so the corrections agree to 2e-3 of their size; the device
correction never raises the coarse error's energy norm (exact coarse
solution by a sparse direct solve).  pie_warm_apply against the oracle's
line search and prolongation of the same correction (alpha from double
sums in another order: to 1e-6).
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


def _net_problem(hip, size=256, f=4, seed=300):
  from fpie_amd import synth, unet, warmstart
  mask, tgt, grad, _ = synth.composite((size, size), seed)
  s = hip.GridSolver(schedule="jacobi")
  s.reset(0, mask, tgt, grad)
  s.iterate(24 * f * f)
  s.relax(1, 0.5)
  mc = warmstart.coarse_mask(mask, f)
  hc, wc = mc.shape
  net = unet.UNet(1, (hc + 7) // 8 * 8, (wc + 7) // 8 * 8,
                  params=warmstart.load_params())
  return s, mask, grad, mc, net


def _predict(net, mc):
  hc, wc = mc.shape
  hp, wp = net.shape[1:]

  def predict(r, m):
    x = co.net_input(r, m, hp, wp)
    if x is None:
      return np.zeros((hc, wc, 3), np.float32)
    return net.forward(x)[0, :hc, :wc] * m[..., None]

  return predict


@pytest.mark.parametrize("steps,lowpass", [(1, 0), (2, 4), (3, 2)])
def test_warm_net_start_matches_oracle(hip, steps, lowpass):
  print(f"warm_net_start steps {steps} lowpass {lowpass}", flush=True)
  s, mask, grad, mc, net = _net_problem(hip)
  f = 4
  rc = s.restrict(f)
  rc[~mc] = 0.0
  want = co.net_start(rc, mc, _predict(net, mc), steps, lowpass)
  got = s.warm_net_start(f, mc, lambda: net, steps, lowpass, want=True)
  scale = float(np.abs(want).max())
  assert scale > 0
  err = float(np.abs(got - want).max())
  print(f"max |device - oracle| = {err:.3g} of {scale:.3g}", flush=True)
  assert err <= 2e-3 * scale
  assert not got[~mc].any()
  # the energy norm of the coarse error never grows
  e = co.exact_solution(mc, np.zeros(mc.shape + (3,)), rc)

  def energy(v):
    return co.dot(v, co.coarse_apply(mc, v))

  assert (energy(e - got) < energy(e)).all()


def test_warm_apply_matches_oracle(hip):
  print("warm_apply", flush=True)
  s, mask, grad, mc, net = _net_problem(hip, seed=301)
  f = 4
  x = s.state().copy()
  corr = s.warm_net_start(f, mc, lambda: net, 2, 4, want=True)
  alpha, size = s.warm_apply(f)  # the correction left on the device
  want_x, want_alpha, want_size = co.warm_apply(mask, x, grad, corr, f)
  np.testing.assert_allclose(alpha, want_alpha, rtol=1e-6)
  assert abs(size - want_size) <= 1e-6 * want_size
  np.testing.assert_allclose(s.state(), want_x, rtol=1e-6, atol=1e-4)
  # a host correction takes the same path
  s.set_state(x)
  alpha2, size2 = s.warm_apply(f, corr)
  np.testing.assert_array_equal(alpha2, alpha)
  assert size2 == size


def test_warm_calls_reject_bad_input(hip):
  from fpie_amd import unet
  s, mask, grad, mc, net = _net_problem(hip, size=128, seed=302)
  with pytest.raises(RuntimeError, match="no device correction"):
    s.warm_apply(4)  # nothing left by warm_net_start
  wrong = unet.UNet(1, 64, 64)
  with pytest.raises(RuntimeError, match="network shape"):
    s.warm_net_start(4, mc, lambda: wrong, 1, 0)
  with pytest.raises(RuntimeError, match="net_steps"):
    s.warm_net_start(4, mc, lambda: net, 9, 0)
  with pytest.raises(ValueError):
    s.warm_net_start(4, mc[:-1], lambda: net, 1, 0)
  s.warm_net_start(4, mc, lambda: net, 1, 0)
  with pytest.raises(RuntimeError, match="no device correction"):
    s.warm_apply(8)  # the device correction is of factor 4
