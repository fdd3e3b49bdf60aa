"""CPU oracle for the coarse-grid correction of the learned warm start.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product package never imports it.

The warm start has no counterpart in the reference (fpie 0.3.2), whose
GridSolver only iterates (fpie/core/gcc/grid.cc:21-80,
fpie/core/openmp/grid.cc:29-110).  What the device calls pie_restrict,
pie_prolong and pie_relax compute is restated here with numpy in the same
float32 operation order (the restriction in float64, as the kernel), plus an
exact sparse solve (scipy, float64) of the
GridSolver system that the tests use as the converged solution:

  4 x_p - sum_{q in N(p), q solved} x_q = g_p + sum_{q in N(p), q unsolved} t_q

for every solved pixel p (the fixed point of the update in
fpie/core/gcc/grid.cc:21-36).

The learned coarse correction (pie_warm_net_start / pie_warm_apply) is
restated in float64 numpy too: the network input, the damped homogeneous
smoothing of its directions, the 5-point coarse operator, and the flexible
conjugate-gradient steps with a caller-supplied `predict` (the tests pass
the device network's forward pass, an exact coarse solve or noise).
"""

from __future__ import annotations

import numpy as np

_OFFS = {"ulrd": ((-1, 0), (0, -1), (0, 1), (1, 0)),
         "udlr": ((-1, 0), (1, 0), (0, -1), (0, 1))}


def neighbour_sum(x: np.ndarray, grad: np.ndarray, order: str = "ulrd"):
  """float32 g + four neighbours, added left to right in the solver's order,
  at every interior pixel ([h, w, 3]; the outermost ring is left 0)."""
  x = np.asarray(x, np.float32)
  s = np.zeros_like(x)
  h, w = x.shape[:2]
  acc = np.asarray(grad, np.float32)[1:-1, 1:-1].copy()
  for di, dj in _OFFS[order]:
    acc = acc + x[1 + di:h - 1 + di, 1 + dj:w - 1 + dj]
  s[1:-1, 1:-1] = acc
  return s


def restrict(mask, x, grad, f: int) -> np.ndarray:
  """Block sums over f x f blocks of the signed residual g + neighbours - 4x
  at solved pixels, all in float64 (exact for float32 x and g).  Returns
  float32 [h // f, w // f, 3]."""
  m = np.asarray(mask) != 0
  x = np.asarray(x, np.float32).astype(np.float64)
  g = np.asarray(grad, np.float32).astype(np.float64)
  r = np.zeros_like(x)
  r[1:-1, 1:-1] = (g[1:-1, 1:-1] + x[:-2, 1:-1] + x[1:-1, :-2] +
                   x[1:-1, 2:] + x[2:, 1:-1] - 4.0 * x[1:-1, 1:-1])
  r = r * m[..., None]
  h, w = (m.shape[0] // f) * f, (m.shape[1] // f) * f
  blocks = r[:h, :w].reshape(h // f, f, w // f, f, 3)
  return blocks.sum((1, 3)).astype(np.float32)


def prolong(mask, x, corr, f: int) -> np.ndarray:
  """x + cell-centred bilinear upsampling of corr at solved pixels, in the
  kernel's float32 operation order."""
  m = np.asarray(mask) != 0
  x = np.asarray(x, np.float32)
  corr = np.asarray(corr, np.float32)
  hc, wc = corr.shape[:2]
  h, w = m.shape
  f32 = np.float32
  inv = f32(1.0) / f32(f)
  ys = (np.arange(h, dtype=f32) + f32(0.5)) * inv - f32(0.5)
  xs = (np.arange(w, dtype=f32) + f32(0.5)) * inv - f32(0.5)
  y0 = np.clip(np.floor(ys).astype(np.int64), 0, hc - 1)
  x0 = np.clip(np.floor(xs).astype(np.int64), 0, wc - 1)
  y1, x1 = np.minimum(y0 + 1, hc - 1), np.minimum(x0 + 1, wc - 1)
  fy = np.clip(ys - y0.astype(f32), f32(0), f32(1))[:, None, None]
  fx = np.clip(xs - x0.astype(f32), f32(0), f32(1))[None, :, None]
  one = f32(1.0)
  top = corr[y0][:, x0] * (one - fx) + corr[y0][:, x1] * fx
  bot = corr[y1][:, x0] * (one - fx) + corr[y1][:, x1] * fx
  v = top * (one - fy) + bot * fy
  return np.where(m[..., None], x + v, x).astype(np.float32)


def prolong_dots(mask, x, grad, corr, f: int) -> np.ndarray:
  """[2, 3] float64: per channel sum r . d and sum d . A d over the solved
  pixels, d the float32 correction prolong() adds (0 off the mask), r the
  exact residual of x, (A d)_p = 4 d_p - sum of the neighbours' d."""
  m = np.asarray(mask) != 0
  d = prolong(mask, np.zeros(m.shape + (3,), np.float32), corr, f)
  d = d.astype(np.float64) * m[..., None]
  z = np.pad(d, ((1, 1), (1, 1), (0, 0)))
  ad = 4.0 * d - (z[:-2, 1:-1] + z[2:, 1:-1] + z[1:-1, :-2] + z[1:-1, 2:])
  h, w = m.shape
  x64 = np.asarray(x, np.float32).astype(np.float64)
  g64 = np.asarray(grad, np.float32).astype(np.float64)
  r = np.zeros_like(x64)
  r[1:-1, 1:-1] = (g64[1:-1, 1:-1] + x64[:-2, 1:-1] + x64[1:-1, :-2] +
                   x64[1:-1, 2:] + x64[2:, 1:-1] - 4.0 * x64[1:-1, 1:-1])
  mm = m[..., None]
  return np.stack([(r * d * mm).sum((0, 1)), (d * ad * mm).sum((0, 1))])


def relax(mask, x, grad, sweeps: int, omega: float,
          order: str = "ulrd") -> np.ndarray:
  """Weighted Jacobi in the kernel's float32 order: j = s * 0.25;
  x + omega * (j - x) (exactly j when omega is 1)."""
  m = (np.asarray(mask) != 0)[..., None]
  x = np.asarray(x, np.float32).copy()
  om = np.float32(omega)
  for _ in range(sweeps):
    j = neighbour_sum(x, grad, order) * np.float32(0.25)
    new = j if omega == 1.0 else x + om * (j - x)
    x = np.where(m, new, x).astype(np.float32)
  return x


def exact_solution(mask, tgt, grad) -> np.ndarray:
  """float64 [h, w, 3]: the solved pixels from a sparse direct solve of the
  GridSolver system, the other pixels from tgt."""
  import scipy.sparse as sp
  import scipy.sparse.linalg as spl
  m = np.asarray(mask) != 0
  ii, jj = np.nonzero(m)
  n = ii.size
  ids = -np.ones(m.shape, np.int64)
  ids[ii, jj] = np.arange(n)
  rows, cols, vals = [np.arange(n)], [np.arange(n)], [np.full(n, 4.0)]
  t = np.asarray(tgt, np.float64)
  b = np.asarray(grad, np.float64)[ii, jj].copy()
  for di, dj in _OFFS["ulrd"]:
    nid = ids[ii + di, jj + dj]
    inside = nid >= 0
    rows.append(np.arange(n)[inside])
    cols.append(nid[inside])
    vals.append(-np.ones(int(inside.sum())))
    b += np.where(inside[:, None], 0.0, t[ii + di, jj + dj])
  A = sp.csc_matrix((np.concatenate(vals),
                     (np.concatenate(rows), np.concatenate(cols))),
                    shape=(n, n))
  x = t.copy()
  lu = spl.splu(A)
  for c in range(3):
    x[ii, jj, c] = lu.solve(b[:, c])
  return x


# ------------------------------------------- learned coarse correction ---


def coarse_apply(mc: np.ndarray, p: np.ndarray) -> np.ndarray:
  """A_c p: 4 p - (up + down + left + right) on the solved cells."""
  q = np.where(mc[..., None], p, 0.0).astype(np.float64)
  z = np.pad(q, ((1, 1), (1, 1), (0, 0)))
  out = 4.0 * q - (z[:-2, 1:-1] + z[2:, 1:-1] + z[1:-1, :-2] + z[1:-1, 2:])
  return out * mc[..., None]


def lowpass(mc: np.ndarray, p: np.ndarray, sweeps: int = 4) -> np.ndarray:
  """`sweeps` damped sweeps q <- (q + (up + down + left + right) / 4) / 2 of
  the homogeneous coarse problem (q = 0 off the solved cells)."""
  m = mc[..., None]
  q = np.where(m, p, 0.0).astype(np.float64)
  for _ in range(sweeps):
    z = np.pad(q, ((1, 1), (1, 1), (0, 0)))
    avg = 0.25 * (z[:-2, 1:-1] + z[2:, 1:-1] + z[1:-1, :-2] + z[1:-1, 2:])
    q = np.where(m, 0.5 * (q + avg), 0.0)
  return q


def dot(a: np.ndarray, b: np.ndarray) -> np.ndarray:
  """Per-channel inner products of [h, w, 3] fields (float64)."""
  return np.einsum("ijc,ijc->c", a, b, dtype=np.float64)


def net_input(r: np.ndarray, mc: np.ndarray, hp: int, wp: int):
  """Network input [1, hp, wp, 8] float32 for the coarse residual r:
  channels 0..2 = r over its per-channel rms on the solved cells, channel 6
  = the mask; None when r vanishes."""
  hc, wc = mc.shape
  sig = np.sqrt(np.mean(r[mc] ** 2, axis=0)) if mc.any() else np.zeros(3)
  if not (sig > 0).any():
    return None
  x = np.zeros((1, hp, wp, 8), np.float32)
  x[0, :hc, :wc, 0:3] = r / np.where(sig > 0, sig, 1.0)
  x[0, :hc, :wc, 6] = mc
  return x


def net_start(rc: np.ndarray, mc: np.ndarray, predict, net_steps: int = 2,
              lowpass_sweeps: int = 4) -> np.ndarray:
  """The learned coarse correction (float32 [hc, wc, 3]) for the coarse
  residual rc: `net_steps` flexible CG steps whose directions are
  predict(r, mc) low-passed, A_c-orthogonalised one after the other against
  the earlier ones, and taken with the energy-optimal step."""
  m = mc[..., None]
  r = np.asarray(rc, np.float64) * m
  c = np.zeros_like(r)
  dirs = []  # (p, A p, <p, A p>)
  for _ in range(net_steps):
    z = lowpass(mc, predict(r, mc), lowpass_sweeps) * m
    for p, ap, pap in dirs:
      z -= (dot(z, ap) / pap) * p
    az = coarse_apply(mc, z)
    pap = dot(z, az)
    ok = pap > 1e-30 * np.maximum(dot(z, z), 1e-300)
    a = np.where(ok, dot(r, z) / np.where(ok, pap, 1.0), 0.0)
    c += a * z
    r -= a * az
    dirs.append((z, az, np.where(ok, pap, 1.0)))
  return c.astype(np.float32)


def warm_apply(mask, x, grad, corr, f: int):
  """x + alpha prolong(corr), alpha = max(0, <r, d> / <d, A d>) per
  channel; returns (new x, alpha, max |alpha corr|)."""
  num, den = prolong_dots(mask, x, grad, corr, f)
  alpha = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
  alpha = np.maximum(alpha, 0.0)
  cc = (np.asarray(corr, np.float32) * alpha).astype(np.float32)
  size = float(np.abs(cc).max()) if cc.size else 0.0
  return prolong(mask, x, cc, f), alpha, size
