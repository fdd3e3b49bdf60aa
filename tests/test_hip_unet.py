"""Warm-start U-Net kernels (GPU) against the CPU oracle (oracle/unet_ref.py).

The GPU rounds weights and stored activations to bfloat16 and accumulates
in float32; the oracle applies the same roundings and accumulates in
float64, so an activation can land one bfloat16 step (2^-8 relative) away
and the difference travels through 15 layers.  Forward outputs are held to
an RMS error of 1 % and a largest error of 5 % of their spread; every parameter
gradient to cosine similarity 0.99 and norm within 5 % of the oracle's (the
GPU also rounds the intermediate gradients to bfloat16 before each MFMA).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def unet():
  from fpie_amd import native, unet
  native.load()
  return unet


def _problem(b, h, w, seed):
  rng = np.random.default_rng(seed)
  x = np.zeros((b, h, w, 8), np.float32)
  x[..., :6] = rng.normal(0, 1, (b, h, w, 6))
  m = (rng.random((b, h, w)) < 0.7).astype(np.uint8)
  x[..., 6] = m
  y = rng.normal(0, 0.5, (b, h, w, 3)).astype(np.float32)
  return x, y, m


def _oracle_params(unet, flat):
  return {k: v.astype(np.float64) for k, v in unet.unflatten(flat).items()}


def test_forward_matches_oracle(unet):
  import unet_ref as R
  x, _, _ = _problem(2, 32, 40, 1)
  p = unet.init_params(3)
  p[unet.unflatten(np.arange(p.size))["out.w"].reshape(-1)] *= 10  # larger head
  net = unet.UNet(2, 32, 40, params=p)
  got = net.forward(x)
  want, _ = R.forward(_oracle_params(unet, p), x.astype(np.float64))
  spread = want.std() + 1e-12
  rms = np.sqrt(np.mean((got - want) ** 2)) / spread
  worst = np.abs(got - want).max() / spread
  assert rms < 1e-2 and worst < 5e-2, (rms, worst)


def test_gradients_match_oracle(unet):
  import unet_ref as R
  x, y, m = _problem(2, 32, 32, 2)
  p = unet.init_params(5)
  net = unet.UNet(2, 32, 32, params=p)
  loss, g = net.loss_grad(x, y, m)
  po = _oracle_params(unet, p)
  out, act = R.forward(po, x.astype(np.float64))
  want_loss, dout = R.masked_mse(out, y, m)
  assert abs(loss - want_loss) <= 2e-2 * want_loss
  gw = R.backward(po, act, dout)
  gd = unet.unflatten(g)
  for name, _ in unet.param_shapes():
    a, b = gd[name].astype(np.float64).ravel(), gw[name].ravel()
    nb = np.linalg.norm(b)
    if nb == 0:
      assert np.linalg.norm(a) < 1e-6, name
      continue
    cos = a @ b / (np.linalg.norm(a) * nb)
    assert cos > 0.99, (name, cos)
    assert abs(np.linalg.norm(a) / nb - 1) < 5e-2, (name, np.linalg.norm(a) / nb)


def test_adam_steps_reduce_the_loss(unet):
  x, _, m = _problem(4, 32, 32, 4)
  y = (0.5 * x[..., 3:6] - 0.25 * x[..., 0:3]).astype(np.float32)  # learnable
  net = unet.UNet(4, 32, 32, seed=1)
  first = net.train_step(x, y, m, 1e-3)
  for _ in range(40):
    last = net.train_step(x, y, m, 1e-3)
  assert last < 0.5 * first, (first, last)
  st = net.stats()
  assert st["conv_flops"] > 0 and st["ms"] > 0
