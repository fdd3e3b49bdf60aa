"""Batched GridSolver (GPU): one persistent Jacobi kernel over many items.

Every item must end bit-identical to the oracle's Jacobi loop run for the
number of sweeps the item reports, and the stop rule must have fired at the
first launch whose last sweep moved no value by `tol` or more: the oracle
recomputes that last-sweep change exactly.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


class _Grab:
  def reset(self, *a):
    self.a = a


def _items(n, seed, side=(24, 160)):
  from fpie_amd import process, synth
  process.register_backend("grab", lambda **kw: _Grab(), lambda **kw: _Grab())
  out = []
  for w in synth.batch(n, seed=seed, side=side):
    p = process.GridProcessor(w.gradient, "grab")
    p.reset(w.src, w.mask, w.tgt, (0, 0), (0, 0))
    _, m, t, g = p.core.a
    out.append((m, t, g))
  return out


@pytest.fixture(scope="module")
def hip():
  from fpie_amd import core_hip, native
  native.load()
  return core_hip


def test_fixed_budget_items_match_oracle(hip, oracle):
  """tol = 0 never stops early: every item runs exactly its budget."""
  items = _items(12, 3)
  budgets = [0, 1, 7, 8, 9, 16, 37, 64, 100, 5, 23, 50]
  b = hip.BatchSolver(sweeps_per_launch=8)
  for (m, t, g), k in zip(items, budgets):
    b.add(m, t, g, max_sweeps=k, tol=0.0)
  b.solve(rounds_per_graph=4)
  for i, ((m, t, g), k) in enumerate(zip(items, budgets)):
    info = b.item(i)
    assert info["sweeps"] == k
    img, err, st = b.read(i, state=True)
    want = oracle.grid_jacobi(m, t, g, k, "ulrd")
    assert np.array_equal(bits(st), bits(want)), (i, k)
    np.testing.assert_array_equal(img, oracle.clamp_u8(want))
    np.testing.assert_allclose(err, oracle.grid_terms(m, want, g).sum((0, 1)),
                               rtol=1e-6, atol=1e-4)


def test_stop_rule_fires_at_first_quiet_launch(hip, oracle):
  items = _items(10, 7)
  tol = 0.05
  b = hip.BatchSolver(sweeps_per_launch=8)
  for m, t, g in items:
    b.add(m, t, g, max_sweeps=2000, tol=tol)
  b.solve(rounds_per_graph=8)
  for i, (m, t, g) in enumerate(items):
    info = b.item(i)
    k = info["sweeps"]
    _, _, st = b.read(i, state=True)
    want = oracle.grid_jacobi(m, t, g, k, "ulrd")
    assert np.array_equal(bits(st), bits(want)), i
    if k < 2000:
      assert k % 8 == 0
      before = oracle.grid_jacobi(m, t, g, k - 1, "ulrd")
      change = float(np.abs(want - before).max())
      assert info["last_dmax"] == change < tol
      if k > 8:  # the previous launch's last sweep was still moving
        a = oracle.grid_jacobi(m, t, g, k - 9, "ulrd")
        bb = oracle.grid_jacobi(m, t, g, k - 8, "ulrd")
        assert float(np.abs(bb - a).max()) >= tol
    assert info["latency_s"] >= 0.0


def test_solve_restarts_from_the_initial_state(hip):
  items = _items(4, 11)
  b = hip.BatchSolver()
  for m, t, g in items:
    b.add(m, t, g, max_sweeps=40, tol=0.0)
  b.solve()
  first = [bits(b.read(i, state=True)[2]).copy() for i in range(4)]
  b.solve()
  for i in range(4):
    np.testing.assert_array_equal(bits(b.read(i, state=True)[2]), first[i])


@pytest.mark.parametrize("tile", [(8, 1, 16, 1), (8, 2, 32, 0), (4, 1, 64, 3)])
def test_channel_split_tiles_match_oracle(hip, oracle, tile):
  """One channel a wave task (or the two-column tile's default) changes only
  how the work is split over waves: every item still equals the oracle."""
  items = _items(6, 17)
  budgets = [3, 8, 21, 40, 64, 13]
  s, c, rows, ch = tile
  b = hip.BatchSolver(sweeps_per_launch=s, cols_per_lane=c,
                      rows_per_task=rows, channels=ch)
  for (m, t, g), k in zip(items, budgets):
    b.add(m, t, g, max_sweeps=k, tol=0.0)
  b.solve(rounds_per_graph=4)
  for i, ((m, t, g), k) in enumerate(zip(items, budgets)):
    assert b.item(i)["sweeps"] == k
    want = oracle.grid_jacobi(m, t, g, k, "ulrd")
    assert np.array_equal(bits(b.read(i, state=True)[2]), bits(want)), (i, k)


def test_batch_rejects_bad_channels(hip):
  b = hip.BatchSolver()
  with pytest.raises(RuntimeError, match="channels per task"):
    b.set_tile_channels(2)


def test_batch_rejects_bad_items(hip):
  b = hip.BatchSolver()
  m = np.zeros((10, 10), np.int32)
  m[0, 3] = 1
  t = np.zeros((10, 10, 3), np.float32)
  with pytest.raises(RuntimeError, match="outermost"):
    b.add(m, t, t)
  with pytest.raises(RuntimeError, match="empty batch"):
    b.solve()


def test_budget_stop_reports_the_last_sweeps_change(hip, oracle):
  """A budget that is not a multiple of the sweeps per launch: the last
  launch is aligned to end on the budget, and last_dmax is the change of
  the item's last sweep (not 0)."""
  items = _items(5, 23)
  budgets = [13, 37, 5, 44, 9]
  b = hip.BatchSolver(sweeps_per_launch=8)
  for (m, t, g), k in zip(items, budgets):
    b.add(m, t, g, max_sweeps=k, tol=1e-12)
  b.solve(rounds_per_graph=4)
  for i, ((m, t, g), k) in enumerate(zip(items, budgets)):
    info = b.item(i)
    assert info["sweeps"] == k
    want = oracle.grid_jacobi(m, t, g, k, "ulrd")
    assert np.array_equal(bits(b.read(i, state=True)[2]), bits(want)), i
    before = oracle.grid_jacobi(m, t, g, k - 1, "ulrd")
    assert info["last_dmax"] == float(np.abs(want - before).max()) > 0.0


def test_solo_and_shared_items_match_oracle(hip, oracle):
  """Items small enough to be run by one work-group (several launches a
  round) and items shared over the tile queue, in one batch: every item
  equals the oracle at its budget, and the small items finish first."""
  items = _items(8, 29, side=(40, 700))
  b = hip.BatchSolver(sweeps_per_launch=8)
  budgets = [200, 120, 333, 64, 250, 17, 96, 300]
  for (m, t, g), k in zip(items, budgets):
    b.add(m, t, g, max_sweeps=k, tol=0.0)
  b.solve(rounds_per_graph=4)
  sizes = [int(np.count_nonzero(m)) for m, _, _ in items]
  for i, ((m, t, g), k) in enumerate(zip(items, budgets)):
    assert b.item(i)["sweeps"] == k
    want = oracle.grid_jacobi(m, t, g, k, "ulrd", threads=16)
    assert np.array_equal(bits(b.read(i, state=True)[2]), bits(want)), (i, k)
  lat = [b.item(i)["latency_s"] for i in range(len(items))]
  assert all(x >= 0.0 for x in lat)
  print(f"\nsizes {sizes}\nlatencies {lat}")


def test_item_without_solved_pixels(hip, oracle):
  """An item whose mask solves nothing has no wave task: it keeps its
  initial state, stops at its budget (tol 0) or after one launch (tol > 0),
  and the other items are unaffected."""
  items = _items(3, 41)
  m0, t0, g0 = items[0]
  empty = np.zeros_like(m0)
  b = hip.BatchSolver(sweeps_per_launch=8)
  b.add(empty, t0, g0, max_sweeps=40, tol=0.0)
  b.add(empty, t0, g0, max_sweeps=40, tol=1e-3)
  for m, t, g in items[1:]:
    b.add(m, t, g, max_sweeps=24, tol=0.0)
  b.solve(rounds_per_graph=4)
  want0 = np.ascontiguousarray(t0, np.float32)
  assert b.item(0)["sweeps"] == 40 and b.item(1)["sweeps"] == 8
  for i in (0, 1):
    assert np.array_equal(bits(b.read(i, state=True)[2]), bits(want0))
  for i, (m, t, g) in enumerate(items[1:], start=2):
    want = oracle.grid_jacobi(m, t, g, 24, "ulrd")
    assert np.array_equal(bits(b.read(i, state=True)[2]), bits(want)), i
