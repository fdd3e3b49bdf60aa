"""Exact float64 solution of a GridSolver system, for measurements only.

The accuracy reference of fpie_amd.warmstart's evaluation: conjugate
gradients in float64 (torch tensors, on the GPU when there is one) on

  4 x_p - sum_{q in N(p), q solved} x_q = g_p + sum_{q in N(p), q unsolved} t_q

for every solved pixel p -- the fixed point of the reference's update
(fpie/core/gcc/grid.cc:21-36) in exact arithmetic.  Not part of the solver
path; it runs in its own process (torch's HIP runtime and the solver
library's do not share one):

  python -m fpie_amd.exact_ref problem.npz out.npy

problem.npz holds mask, tgt, grad and the start state x0; out.npy gets the
float64 solution [h, w, 3].
"""

from __future__ import annotations

import sys

import numpy as np


def solve(mask, tgt, grad, x0, rtol: float = 1e-13,
          max_iters: int = 100_000, log=print) -> np.ndarray:
  """float64 solution by conjugate gradients from x0 (the unsolved pixels
  keep tgt); the residual is reduced to `rtol` times the target state's."""
  import torch
  dev = torch.device("cuda" if torch.cuda.is_available() else "cpu")
  m = torch.from_numpy(np.asarray(mask) != 0).to(dev)
  t = torch.from_numpy(np.asarray(tgt, np.float64)).to(dev).permute(2, 0, 1)
  g = torch.from_numpy(np.asarray(grad, np.float64)).to(dev).permute(2, 0, 1)
  x = torch.from_numpy(np.asarray(x0, np.float64)).to(dev).permute(2, 0, 1)
  x = torch.where(m, x, t).contiguous()

  def lap(v):  # sum of the four neighbours - 4 v (zero outside the grid)
    out = -4.0 * v
    out[:, 1:, :] += v[:, :-1, :]
    out[:, :-1, :] += v[:, 1:, :]
    out[:, :, 1:] += v[:, :, :-1]
    out[:, :, :-1] += v[:, :, 1:]
    return out

  def resid(v):
    return torch.where(m, g + lap(v), torch.zeros_like(v))

  def dot(a, b):
    return (a * b).sum((1, 2))

  r0 = torch.sqrt(dot(resid(t), resid(t)))
  r = resid(x)
  p = r.clone()
  rr = dot(r, r)
  k = 0
  for k in range(1, max_iters + 1):
    ap = torch.where(m, -lap(p), torch.zeros_like(p))
    a = rr / dot(p, ap).clamp_min(1e-300)
    x += a[:, None, None] * p
    r -= a[:, None, None] * ap
    rr_new = dot(r, r)
    if bool((torch.sqrt(rr_new) <= rtol * r0).all()):
      break
    p = r + (rr_new / rr.clamp_min(1e-300))[:, None, None] * p
    rr = rr_new
    if k % 2000 == 0:
      log(f"#   exact solution: {k} iterations, relative residual "
          f"{float((torch.sqrt(rr) / r0).max()):.3g}")
  rel = float((torch.sqrt(dot(resid(x), resid(x))) / r0).max())
  log(f"#   exact solution: {k} iterations, relative residual {rel:.3g}")
  return x.permute(1, 2, 0).cpu().numpy()


def main(argv) -> None:
  src, dst = argv
  with np.load(src, allow_pickle=False) as f:
    x = solve(f["mask"], f["tgt"], f["grad"], f["x0"],
              log=lambda m: print(m, flush=True))
  np.save(dst, x)


if __name__ == "__main__":
  main(sys.argv[1:])
