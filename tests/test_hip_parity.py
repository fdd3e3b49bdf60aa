"""Parity of the hip backend with the oracle and the reference (GPU).

Bar: float32 state bit-identical to the sequential CPU loops for every
schedule (gs_lex = gcc backend, redblack = openmp EquSolver, jacobi =
openmp GridSolver / numpy), uint8 output identical, err within rtol 1e-5 of
the reference's float-accumulated err (tests/test_smoke.py:66) and within
rtol 1e-6 of the exactly rounded sum of the oracle's double terms (the
device sums in double).
"""

import math

import numpy as np
import pytest

from conftest import golden_cases, load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
  from fpie_amd import core_hip, native
  native.load()
  return core_hip


def bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


def exact_err(terms):
  return np.array([math.fsum(terms[..., c].ravel().tolist())
                   for c in range(3)])


def check_err(err, ref_err, terms):
  np.testing.assert_allclose(err, exact_err(terms), rtol=1e-6, atol=1e-4)
  if ref_err is not None:
    np.testing.assert_allclose(err, ref_err, rtol=1e-5, atol=1e-5)


# ------------------------------------------------------ golden vectors ---


@pytest.mark.parametrize("name", golden_cases())
def test_equ_gs_lex_is_gcc(hip, oracle, name):
  g = load_golden(name)
  it = int(g["iters"])
  core = hip.EquSolver(schedule="gs_lex")
  ids = core.partition(g["equ_mask"])
  np.testing.assert_array_equal(np.where(g["equ_mask"] > 0, ids, 0),
                                g["equ_ids"])
  core.reset(len(g["equ_X"]), g["equ_A"], g["equ_X"], g["equ_B"])
  u8, err = core.step(it)
  np.testing.assert_array_equal(u8, g["equ_gcc_u8"])
  X = oracle.equ_gs(g["equ_A"], g["equ_X"], g["equ_B"], it)
  np.testing.assert_array_equal(bits(core.state()), bits(X))
  check_err(err, g["equ_gcc_err"], oracle.equ_terms(g["equ_A"], X,
                                                    g["equ_B"]))


@pytest.mark.parametrize("name", golden_cases())
def test_equ_jacobi_is_numpy(hip, oracle, name):
  g = load_golden(name)
  core = hip.EquSolver(schedule="jacobi")
  core.partition(g["equ_mask"])
  core.reset(len(g["equ_X"]), g["equ_A"], g["equ_X"], g["equ_B"])
  u8, err = core.step(int(g["iters"]))
  np.testing.assert_array_equal(bits(core.state()), bits(g["equ_numpy_X"]))
  np.testing.assert_array_equal(u8, oracle.clamp_u8(g["equ_numpy_X"]))
  check_err(err, g["equ_numpy_err"],
            oracle.equ_terms(g["equ_A"], g["equ_numpy_X"], g["equ_B"]))


@pytest.mark.parametrize("name", golden_cases())
def test_equ_redblack_is_openmp(hip, oracle, name):
  """Red-black values do not depend on the id numbering: our row-major
  system gives, pixel for pixel, the openmp core's results."""
  g = load_golden(name)
  it = int(g["iters"])
  core = hip.EquSolver(schedule="redblack")
  core.partition(g["equ_mask"])
  core.reset(len(g["equ_X"]), g["equ_A"], g["equ_X"], g["equ_B"])
  u8, err = core.step(it)
  m = g["equ_mask"] > 0
  rm, rb = g["equ_ids"][m], g["equ_rb_ids"][m]
  np.testing.assert_array_equal(u8[rm], g["equ_openmp_u8"][rb])
  _, n_mid = oracle.partition_redblack(g["equ_mask"])
  X = oracle.equ_redblack(g["equ_rb_A"], g["equ_rb_X"], g["equ_rb_B"], n_mid,
                          it)
  np.testing.assert_array_equal(bits(core.state()[rm]), bits(X[rb]))
  check_err(err, g["equ_openmp_err"],
            oracle.equ_terms(g["equ_rb_A"], X, g["equ_rb_B"]))


@pytest.mark.parametrize("name", golden_cases())
def test_grid_gs_lex_is_gcc(hip, oracle, name):
  g = load_golden(name)
  it = int(g["iters"])
  core = hip.GridSolver(8, 8, schedule="gs_lex", order="ulrd")
  core.reset(0, g["grid_mask"], g["grid_tgt"], g["grid_grad"])
  u8, err = core.step(it)
  np.testing.assert_array_equal(u8, g["grid_gcc_u8"])
  T = oracle.grid_gs(g["grid_mask"], g["grid_tgt"], g["grid_grad"], it)
  np.testing.assert_array_equal(bits(core.state()), bits(T))
  check_err(err, g["grid_gcc_err"],
            oracle.grid_terms(g["grid_mask"], T, g["grid_grad"]))


@pytest.mark.parametrize("name", golden_cases())
def test_grid_jacobi_is_openmp(hip, oracle, name):
  g = load_golden(name)
  it = int(g["iters"])
  core = hip.GridSolver(2, 3, schedule="jacobi", order="ulrd")
  core.reset(0, g["grid_mask"], g["grid_tgt"], g["grid_grad"])
  u8, err = core.step(it)
  np.testing.assert_array_equal(u8, g["grid_openmp_u8"])
  T = oracle.grid_jacobi(g["grid_mask"], g["grid_tgt"], g["grid_grad"], it)
  np.testing.assert_array_equal(bits(core.state()), bits(T))
  check_err(err, g["grid_openmp_err"],
            oracle.grid_terms(g["grid_mask"], T, g["grid_grad"]))


@pytest.mark.parametrize("name", golden_cases())
def test_grid_jacobi_udlr_is_numpy(hip, name):
  g = load_golden(name)
  core = hip.GridSolver(schedule="jacobi", order="udlr")
  core.reset(0, g["grid_mask"], g["grid_tgt"], g["grid_grad"])
  core.step(int(g["iters"]))
  np.testing.assert_array_equal(bits(core.state()), bits(g["grid_numpy_tgt"]))


# --------------------------------------------- every tile, every schedule ---


class _Grab:
  """Stand-in core that keeps what the processor hands to reset()."""

  def reset(self, *a):
    self.a = a


def _grid_inputs(h, w, seed, density=0.75, mode="max"):
  """Cropped mask, float target and divergence as the processor builds
  them for a random composite."""
  from fpie_amd import process
  process.register_backend("grab", lambda **kw: _Grab(),
                           lambda **kw: _Grab())
  rng = np.random.default_rng(seed)
  src = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  tgt = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
  mask = (rng.random((h, w)) < density).astype(np.uint8) * 255
  p = process.GridProcessor(mode, "grab")
  p.reset(src, mask, tgt, (0, 0), (0, 0))
  _, m, t, gr = p.core.a
  return m, t, gr


def _tiles():
  import re
  import os
  from conftest import PKG
  text = open(os.path.join(PKG, "csrc", "kernels.h")).read()
  line = re.search(r"#define PIE_TILES\(X\)((?:.*\\\n)*.*)", text).group(1)
  return [(int(a), int(b), int(c))
          for a, b, c in re.findall(r"X\((\d+), (\d+), (\d+)\)", line)]


@pytest.mark.parametrize("sched,order", [("jacobi", "ulrd"),
                                         ("jacobi", "udlr"),
                                         ("gs_lex", "ulrd"),
                                         ("redblack", "ulrd")])
def test_every_tile_matches_oracle(hip, oracle, sched, order):
  m, t, gr = _grid_inputs(150, 171, 21)
  it = 37
  if sched == "jacobi":
    want = oracle.grid_jacobi(m, t, gr, it, order)
  elif sched == "gs_lex":
    want = oracle.grid_gs(m, t, gr, it)
  else:
    want = oracle.grid_colour(m, t, gr, it, order)
  for S, C, N in _tiles():
    for rows in (16, 48, 256):
      core = hip.GridSolver(schedule=sched, order=order)
      core.set_tile(S, C, rows, channels=N)
      assert core.tile() == (S, C, rows) and core.tile_channels() == N
      core.reset(0, m, t, gr)
      core.step(it)
      got = core.state()
      assert np.array_equal(bits(got), bits(want)), (
          f"tile S={S} C={C} N={N} rows={rows}: "
          f"{int((bits(got) != bits(want)).sum())} values differ")


@pytest.mark.parametrize("sched", ["jacobi", "gs_lex", "redblack"])
def test_state_persists_across_steps(hip, oracle, sched):
  m, t, gr = _grid_inputs(70, 90, 5)
  core = hip.GridSolver(schedule=sched)
  core.reset(0, m, t, gr)
  core.step(7)
  u8, _ = core.step(11)
  if sched == "jacobi":
    want = oracle.grid_jacobi(m, t, gr, 18)
  elif sched == "gs_lex":
    want = oracle.grid_gs(m, t, gr, 18)
  else:
    want = oracle.grid_colour(m, t, gr, 18)
  np.testing.assert_array_equal(bits(core.state()), bits(want))
  np.testing.assert_array_equal(u8, oracle.clamp_u8(want))


def test_repeat_runs_are_bitwise_identical(hip):
  m, t, gr = _grid_inputs(200, 260, 8)
  outs = []
  for _ in range(2):
    core = hip.GridSolver(schedule="jacobi")
    core.reset(0, m, t, gr)
    u8, err = core.step(101)
    outs.append((bits(core.state()).copy(), u8, err.copy()))
  np.testing.assert_array_equal(outs[0][0], outs[1][0])
  np.testing.assert_array_equal(outs[0][1], outs[1][1])
  np.testing.assert_array_equal(bits(outs[0][2]), bits(outs[1][2]))


def test_equ_arbitrary_A_uses_gather_path(hip, oracle):
  """A system whose ids are not the row-major numbering of partition()
  (here: the red-black ids) is solved by the gather kernel; Jacobi bits
  equal the numpy-order oracle."""
  g = load_golden("disc48")
  core = hip.EquSolver(schedule="jacobi")
  core.reset(len(g["equ_rb_X"]), g["equ_rb_A"], g["equ_rb_X"], g["equ_rb_B"])
  u8, err = core.step(57)
  X = oracle.equ_jacobi(g["equ_rb_A"], g["equ_rb_X"], g["equ_rb_B"], 57)
  np.testing.assert_array_equal(bits(core.state()), bits(X))
  np.testing.assert_array_equal(u8, oracle.clamp_u8(X))
  check_err(err, None, oracle.equ_terms(g["equ_rb_A"], X, g["equ_rb_B"]))


# ---------------------------------------------------------- errors ---


def test_errors_raise_runtime_error(hip):
  g = load_golden("rand0")
  core = hip.EquSolver(schedule="gs_lex")
  with pytest.raises(RuntimeError, match="partition"):
    core.reset(len(g["equ_rb_X"]), g["equ_rb_A"], g["equ_rb_X"],
               g["equ_rb_B"])
  bad = g["equ_A"].copy()
  bad[3, 1] = len(bad) + 5
  core = hip.EquSolver()
  with pytest.raises(RuntimeError, match="outside"):
    core.reset(len(bad), bad, g["equ_X"], g["equ_B"])
  with pytest.raises(RuntimeError, match="before reset"):
    hip.GridSolver().step(3)
  m = g["grid_mask"].copy()
  m[0, 2] = 1
  core = hip.GridSolver()
  with pytest.raises(RuntimeError, match="outermost"):
    core.reset(0, m, g["grid_tgt"], g["grid_grad"])


# -------------------------------------------------------- processor ---


def test_processor_hip_matches_reference_processor(hip):
  """EquProcessor/GridProcessor on the hip backend give the images of the
  reference gcc and openmp cores (golden vectors), end to end."""
  from fpie_amd import process
  for name in ("disc48", "rand2"):
    g = load_golden(name)
    args = (g["src"], g["mask_in"], g["tgt_in"], tuple(g["mask_on_src"]),
            tuple(g["mask_on_tgt"]))
    it, mode = int(g["iters"]), str(g["mode"])
    p = process.EquProcessor(mode, "hip", schedule="gs_lex")
    assert p.reset(*args) == len(g["equ_X"])
    out, _ = p.step(it)
    want = g["tgt_in"].copy()
    want[p.tgt_index] = g["equ_gcc_u8"][1:]
    np.testing.assert_array_equal(out, want)
    for sched, key in (("gs_lex", "grid_gcc_u8"),
                       ("jacobi", "grid_openmp_u8")):
      p = process.GridProcessor(mode, "hip", schedule=sched)
      p.reset(*args)
      out, _ = p.step(it)
      want = g["tgt_in"].copy()
      x0, x1, y0, y1 = g["grid_window"]
      want[x0:x1, y0:y1] = g[key]
      np.testing.assert_array_equal(out, want)
