"""Pins of the U-Net oracle (oracle/unet_ref.py) on the CPU.

The oracle is what the GPU kernels are held to, so it is checked against
what the mathematics fixes rather than against itself:

* every parameter gradient of the float64 network (no bf16 rounding) against
  central finite differences of the loss, for sampled entries of every
  layer, bias and the head;
* the 3x3 convolution against a direct per-pixel loop with the documented
  weight layout w[tap][cin][cout], tap = 3 (dy + 1) + (dx + 1), zero
  padding;
* pooling and upsampling backward as the adjoints of their forward maps;
* the bf16 rounding (round to nearest, ties to even) on known values.
"""

import numpy as np

import unet_ref as R


def _input(n, h, w, seed):
  rng = np.random.default_rng(seed)
  x = np.zeros((n, h, w, 8))
  x[..., :6] = rng.normal(0, 1, (n, h, w, 6))
  m = (rng.random((n, h, w)) < 0.7)
  x[..., 6] = m
  y = rng.normal(0, 0.5, (n, h, w, 3))
  return x, y, m


def test_gradients_match_finite_differences():
  x, y, m = _input(1, 16, 16, 0)
  p = R.init_params(1)
  p["out.w"] = p["out.w"] * 10.0  # a head large enough to matter

  def loss(params):
    out, _ = R.forward(params, x, bf16_round=False)
    return R.masked_mse(out, y, m)[0]

  out, act = R.forward(p, x, bf16_round=False)
  _, dout = R.masked_mse(out, y, m)
  g = R.backward(p, act, dout, bf16_round=False)
  rng = np.random.default_rng(2)
  eps = 1e-5
  for name in p:
    flat = p[name].reshape(-1)
    for k in rng.choice(flat.size, size=min(3, flat.size), replace=False):
      plus = {kk: vv.copy() for kk, vv in p.items()}
      minus = {kk: vv.copy() for kk, vv in p.items()}
      plus[name].reshape(-1)[k] += eps
      minus[name].reshape(-1)[k] -= eps
      fd = (loss(plus) - loss(minus)) / (2 * eps)
      an = g[name].reshape(-1)[k]
      assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (name, k, fd, an)


def test_conv3x3_is_the_documented_correlation():
  rng = np.random.default_rng(3)
  x = rng.normal(0, 1, (2, 5, 6, 3))
  w = rng.normal(0, 1, (9, 3, 4))
  b = rng.normal(0, 1, 4)
  got = R.conv3x3(x, w, b)
  want = np.zeros((2, 5, 6, 4))
  for n in range(2):
    for i in range(5):
      for j in range(6):
        acc = b.copy()
        for dy in (-1, 0, 1):
          for dx in (-1, 0, 1):
            ii, jj = i + dy, j + dx
            if 0 <= ii < 5 and 0 <= jj < 6:
              acc += x[n, ii, jj] @ w[3 * (dy + 1) + (dx + 1)]
        want[n, i, j] = acc
  np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_pool_and_up_backward_are_adjoints():
  rng = np.random.default_rng(4)
  x = rng.normal(0, 1, (2, 8, 6, 3))
  y = rng.normal(0, 1, (2, 4, 3, 3))
  assert np.isclose((R.pool(x) * y).sum(), (x * R.pool_bwd(y)).sum())
  assert np.isclose((R.up(y) * x).sum(), (y * R.up_bwd(x)).sum())


def test_bf16_rounding():
  # bf16 keeps 7 fraction bits: near 1 the step is 2^-7
  v = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, 1.0 + 2.0**-9,
                1.0 + 5 * 2.0**-9, -2.5, 2.0**-130])
  want = np.array([1.0, 1.0, 1.0 + 2.0**-6, 1.0, 1.0 + 2.0**-7, -2.5,
                   2.0**-130])
  np.testing.assert_array_equal(R.bf16(v), want)
