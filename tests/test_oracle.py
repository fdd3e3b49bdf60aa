"""Pins of the CPU oracle (CPU only).

The oracle restates the reference solver cores; these tests check it
against (a) golden vectors written from the reference itself (compiled gcc
and openmp cores, the numpy solver), (b) the compiled reference cores
directly when oracle/_ref is built, and (c) what the mathematics fixes:
fixed points, one-pixel masks, a dense linear solve, and the equivalence of
the GPU schedules with the sequential loops.
"""

import math

import numpy as np
import pytest

from conftest import golden_cases, load_golden


def bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


# ------------------------------------------------------ golden vectors ---


@pytest.mark.parametrize("name", golden_cases())
def test_equ_gcc_matches_golden(oracle, name):
  g = load_golden(name)
  X = oracle.equ_gs(g["equ_A"], g["equ_X"], g["equ_B"], int(g["iters"]))
  np.testing.assert_array_equal(oracle.clamp_u8(X), g["equ_gcc_u8"])
  np.testing.assert_array_equal(bits(oracle.equ_err(g["equ_A"], X,
                                                    g["equ_B"])),
                                bits(g["equ_gcc_err"]))


@pytest.mark.parametrize("name", golden_cases())
def test_equ_openmp_matches_golden(oracle, name):
  g = load_golden(name)
  _, n_mid = oracle.partition_redblack(g["equ_mask"])
  X = oracle.equ_redblack(g["equ_rb_A"], g["equ_rb_X"], g["equ_rb_B"], n_mid,
                          int(g["iters"]), threads=3)
  np.testing.assert_array_equal(oracle.clamp_u8(X), g["equ_openmp_u8"])
  np.testing.assert_array_equal(
      bits(oracle.equ_err(g["equ_rb_A"], X, g["equ_rb_B"], "openmp")),
      bits(g["equ_openmp_err"]))


@pytest.mark.parametrize("name", golden_cases())
def test_equ_numpy_jacobi_bits_match_golden(oracle, name):
  g = load_golden(name)
  X = oracle.equ_jacobi(g["equ_A"], g["equ_X"], g["equ_B"], int(g["iters"]))
  np.testing.assert_array_equal(bits(X), bits(g["equ_numpy_X"]))
  # numpy sums the float32 terms in its own order: tolerance of
  # tests/test_smoke.py:66 against the exactly rounded sum
  exact = [math.fsum(c) for c in oracle.equ_terms(g["equ_A"], X,
                                                  g["equ_B"]).T]
  np.testing.assert_allclose(g["equ_numpy_err"], exact, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("name", golden_cases())
def test_grid_gcc_matches_golden(oracle, name):
  g = load_golden(name)
  T = oracle.grid_gs(g["grid_mask"], g["grid_tgt"], g["grid_grad"],
                     int(g["iters"]), 8, 8)
  np.testing.assert_array_equal(oracle.clamp_u8(T), g["grid_gcc_u8"])
  np.testing.assert_array_equal(
      bits(oracle.grid_err(g["grid_mask"], T, g["grid_grad"])),
      bits(g["grid_gcc_err"]))


@pytest.mark.parametrize("name", golden_cases())
def test_grid_openmp_matches_golden(oracle, name):
  g = load_golden(name)
  T = oracle.grid_jacobi(g["grid_mask"], g["grid_tgt"], g["grid_grad"],
                         int(g["iters"]), "ulrd", threads=3)
  np.testing.assert_array_equal(oracle.clamp_u8(T), g["grid_openmp_u8"])
  np.testing.assert_array_equal(
      bits(oracle.grid_err(g["grid_mask"], T, g["grid_grad"], "openmp")),
      bits(g["grid_openmp_err"]))


@pytest.mark.parametrize("name", golden_cases())
def test_grid_numpy_jacobi_bits_match_golden(oracle, name):
  g = load_golden(name)
  T = oracle.grid_jacobi(g["grid_mask"], g["grid_tgt"], g["grid_grad"],
                         int(g["iters"]), "udlr")
  np.testing.assert_array_equal(bits(T), bits(g["grid_numpy_tgt"]))


@pytest.mark.parametrize("name", golden_cases())
def test_system_assembly_is_stable(oracle, name):
  """The pixel-loop restatement of fpie/process.py still builds the
  systems the golden vectors were made from."""
  g = load_golden(name)
  args = (g["src"], g["mask_in"], g["tgt_in"], tuple(g["mask_on_src"]),
          tuple(g["mask_on_tgt"]), str(g["mode"]))
  e = oracle.build_equ(*args)
  for k in ("A", "X", "B"):
    np.testing.assert_array_equal(e[k], g[f"equ_{k}"])
  gr = oracle.build_grid(*args)
  for k in ("mask", "tgt", "grad"):
    np.testing.assert_array_equal(gr[k], g[f"grid_{k}"])


# ----------------------------------------- compiled reference (optional) ---


def _ref(oracle, name):
  mod = oracle.ref_core(name)
  if mod is None:
    pytest.skip("oracle/_ref not built (needs the reference checkout)")
  return mod


@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_compiled_reference(oracle, seed):
  gcc, omp = _ref(oracle, "core_gcc"), _ref(oracle, "core_openmp")
  rng = np.random.default_rng(100 + seed)
  h, w = rng.integers(8, 48, 2)
  src = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  mask = (rng.random((h, w)) > 0.25).astype(np.uint8) * 255
  mode = ("max", "src", "avg", "max")[seed]
  iters = int(rng.integers(1, 60))
  e = oracle.build_equ(src, mask, tgt, (0, 0), (0, 0), mode)
  c = gcc.EquSolver()
  c.partition(e["mask"])
  c.reset(e["max_id"], e["A"], e["X"], e["B"])
  x, err = c.step(iters)
  X = oracle.equ_gs(e["A"], e["X"], e["B"], iters)
  np.testing.assert_array_equal(oracle.clamp_u8(X), x)
  np.testing.assert_array_equal(bits(oracle.equ_err(e["A"], X, e["B"])),
                                bits(err))
  gr = oracle.build_grid(src, mask, tgt, (0, 0), (0, 0), mode)
  gx, gy = int(rng.integers(1, 9)), int(rng.integers(1, 9))
  c = gcc.GridSolver(gx, gy)
  c.reset(0, gr["mask"], gr["tgt"], gr["grad"])
  x, err = c.step(iters)
  T = oracle.grid_gs(gr["mask"], gr["tgt"], gr["grad"], iters, gx, gy)
  np.testing.assert_array_equal(oracle.clamp_u8(T), x)
  np.testing.assert_array_equal(
      bits(oracle.grid_err(gr["mask"], T, gr["grad"])), bits(err))
  c = omp.GridSolver(gx, gy, 4)
  c.reset(0, gr["mask"], gr["tgt"], gr["grad"])
  x, err = c.step(iters)
  T = oracle.grid_jacobi(gr["mask"], gr["tgt"], gr["grad"], iters, "ulrd", 4)
  np.testing.assert_array_equal(oracle.clamp_u8(T), x)
  np.testing.assert_array_equal(
      bits(oracle.grid_err(gr["mask"], T, gr["grad"], "openmp")), bits(err))


# ------------------------------------------------------ mathematics ---


def _problem(h=24, w=30, seed=0, density=0.7):
  rng = np.random.default_rng(seed)
  src = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  mask = (rng.random((h, w)) < density).astype(np.uint8) * 255
  return src, mask, tgt


def test_gs_tile_order_does_not_change_bits(oracle):
  """Any gcc grid tile shape gives lexicographic GS (same new/old split)."""
  src, mask, tgt = _problem()
  g = oracle.build_grid(src, mask, tgt, (0, 0), (0, 0), "max")
  ref = oracle.grid_gs(g["mask"], g["tgt"], g["grad"], 9, 1, 10**6)
  for gx, gy in [(8, 8), (2, 128), (3, 5), (1, 1), (64, 2)]:
    np.testing.assert_array_equal(
        bits(oracle.grid_gs(g["mask"], g["tgt"], g["grad"], 9, gx, gy)),
        bits(ref))


@pytest.mark.parametrize("order", ["ulrd", "udlr"])
def test_wavefront_schedule_equals_sequential_gs(oracle, order):
  """The skewed two-colour schedule the GPU runs is lexicographic GS."""
  src, mask, tgt = _problem(19, 23, 3)
  g = oracle.build_grid(src, mask, tgt, (0, 0), (0, 0), "max")
  wf = oracle.grid_wavefront(g["mask"], g["tgt"], g["grad"], 13, order)
  if order == "ulrd":
    seq = oracle.grid_gs(g["mask"], g["tgt"], g["grad"], 13)
  else:  # the Equ system in grid form: B + U + D + L + R, 0 off the mask
    e = oracle.build_equ(src, mask, tgt, (0, 0), (0, 0), "max")
    X = oracle.equ_gs(e["A"], e["X"], e["B"], 13)
    m = e["mask"]
    grid_x = np.zeros(m.shape + (3,), np.float32)
    grid_b = np.zeros_like(grid_x)
    ids = e["ids"]
    grid_x[m > 0] = e["X"][ids[m > 0]]
    grid_b[m > 0] = e["B"][ids[m > 0]]
    wf = oracle.grid_wavefront(m, grid_x, grid_b, 13, "udlr")[m > 0]
    seq = X[ids[m > 0]]
  np.testing.assert_array_equal(bits(wf), bits(seq))


def test_colour_schedule_equals_openmp_redblack(oracle):
  src, mask, tgt = _problem(21, 17, 4)
  e = oracle.build_equ(src, mask, tgt, (0, 0), (0, 0), "avg",
                       partition=oracle.partition_redblack)
  _, n_mid = oracle.partition_redblack(e["mask"])
  X = oracle.equ_redblack(e["A"], e["X"], e["B"], n_mid, 11)
  m, ids = e["mask"], e["ids"]
  gx = np.zeros(m.shape + (3,), np.float32)
  gb = np.zeros_like(gx)
  gx[m > 0] = e["X"][ids[m > 0]]
  gb[m > 0] = e["B"][ids[m > 0]]
  col = oracle.grid_colour(m, gx, gb, 11, "udlr")
  np.testing.assert_array_equal(bits(col[m > 0]), bits(X[ids[m > 0]]))


def test_schedules_and_orders_are_distinguishable(oracle):
  """A kernel that ran the wrong schedule or neighbour order would fail
  the parity tests: all four variants give different bits here.  (After
  a few sweeps every value is a short dyadic fraction and all orders round
  alike; 40 sweeps leave enough low bits for the order to show.)"""
  src, mask, tgt = _problem(20, 20, 5)
  g = oracle.build_grid(src, mask, tgt, (0, 0), (0, 0), "max")
  args = (g["mask"], g["tgt"], g["grad"], 40)
  outs = [oracle.grid_gs(*args), oracle.grid_jacobi(*args, "ulrd"),
          oracle.grid_jacobi(*args, "udlr"),
          oracle.grid_colour(*args, "ulrd")]
  for i in range(len(outs)):
    for j in range(i + 1, len(outs)):
      assert not np.array_equal(bits(outs[i]), bits(outs[j])), (i, j)


@pytest.mark.parametrize("mode", ["max", "src", "avg"])
def test_constant_image_is_a_fixed_point(oracle, mode):
  img = np.full((30, 30, 3), 200, np.uint8)
  mask = np.zeros((30, 30), np.uint8)
  mask[5:25, 4:26] = 255
  e = oracle.build_equ(img, mask, img, (0, 0), (0, 0), mode)
  X = oracle.equ_gs(e["A"], e["X"], e["B"], 25)
  assert (X[1:] == 200).all()
  assert (oracle.equ_err(e["A"], X, e["B"]) == 0).all()
  g = oracle.build_grid(img, mask, img, (0, 0), (0, 0), mode)
  for T in (oracle.grid_gs(g["mask"], g["tgt"], g["grad"], 25),
            oracle.grid_jacobi(g["mask"], g["tgt"], g["grad"], 25)):
    assert (T == 200).all()
    assert (oracle.grid_err(g["mask"], T, g["grad"]) == 0).all()


def test_linear_ramp_is_a_fixed_point(oracle):
  """A linear image has zero discrete Laplacian: with source = target =
  ramp and gradient 'src' the target is the exact solution."""
  ii, jj = np.mgrid[0:40, 0:36]
  ramp = np.stack([2 * ii + 3 * jj + 10, ii + jj, 5 * ii + 20], -1)
  ramp = ramp.astype(np.uint8)
  rng = np.random.default_rng(2)
  mask = (rng.random((40, 36)) < 0.8).astype(np.uint8) * 255
  e = oracle.build_equ(ramp, mask, ramp, (0, 0), (0, 0), "src")
  X = oracle.equ_gs(e["A"], e["X"], e["B"], 10)
  np.testing.assert_array_equal(X, e["X"])
  assert (oracle.equ_err(e["A"], X, e["B"]) == 0).all()
  g = oracle.build_grid(ramp, mask, ramp, (0, 0), (0, 0), "src")
  T = oracle.grid_jacobi(g["mask"], g["tgt"], g["grad"], 10, "udlr")
  np.testing.assert_array_equal(T, g["tgt"])


def test_single_pixel_mask_exact_value(oracle):
  """One unknown: x = (4 s_c - sum s_n + sum t_n) / 4 ('src' gradient)."""
  rng = np.random.default_rng(9)
  src = rng.integers(0, 256, (7, 8, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (7, 8, 3), dtype=np.uint8)
  mask = np.zeros((7, 8), np.uint8)
  mask[3, 4] = 255
  s, t = src.astype(np.float64), tgt.astype(np.float64)
  nb = [(2, 4), (4, 4), (3, 3), (3, 5)]
  want = (4 * s[3, 4] - sum(s[p] for p in nb) + sum(t[p] for p in nb)) / 4
  e = oracle.build_equ(src, mask, tgt, (0, 0), (0, 0), "src")
  assert e["max_id"] == 2
  X = oracle.equ_gs(e["A"], e["X"], e["B"], 1)
  np.testing.assert_array_equal(X[1], want.astype(np.float32))
  g = oracle.build_grid(src, mask, tgt, (0, 0), (0, 0), "src")
  T = oracle.grid_gs(g["mask"], g["tgt"], g["grad"], 1)
  np.testing.assert_array_equal(T[1, 1], want.astype(np.float32))


def test_converges_to_dense_solve(oracle):
  """Many sweeps of every scheme reach the solution of (4I - N) x = b
  assembled independently from the definition of Poisson blending."""
  rng = np.random.default_rng(11)
  h, w = 9, 11
  src = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  mask = (rng.random((h, w)) < 0.8).astype(np.uint8) * 255
  e = oracle.build_equ(src, mask, tgt, (0, 0), (0, 0), "avg")
  m = e["mask"]
  pts = list(zip(*np.nonzero(m)))
  idx = {p: k for k, p in enumerate(pts)}
  n = len(pts)
  M = np.zeros((n, n))
  b = np.zeros((n, 3))
  x0, y0 = np.argwhere(mask[1:-1, 1:-1] >= 128).min(0)  # crop offset
  S, T = src.astype(np.float64), tgt.astype(np.float64)
  for k, (i, j) in enumerate(pts):
    M[k, k] = 4
    gi, gj = i + x0, j + y0
    for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
      gs = S[gi, gj] - S[gi + di, gj + dj]
      gt = T[gi, gj] - T[gi + di, gj + dj]
      b[k] += (gs + gt) / 2
      q = (i + di, j + dj)
      if q in idx:
        M[k, idx[q]] = -1
      else:
        b[k] += T[gi + di, gj + dj]
  want = np.linalg.solve(M, b)
  order = np.argsort([e["ids"][p] for p in pts])
  for X in (oracle.equ_gs(e["A"], e["X"], e["B"], 3000),
            oracle.equ_jacobi(e["A"], e["X"], e["B"], 6000)):
    np.testing.assert_allclose(X[1:], want[order], atol=2e-3)
