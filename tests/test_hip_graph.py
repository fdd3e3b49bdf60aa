"""Captured sweep graphs (GPU).

On grids below 2^22 pixels, step(n) calls of 16 or more launches replay a
hipGraph of their launches, cached per call shape.  Each replay must equal
the same sweeps issued launch by launch, also with the program's own eager
work (weighted sweeps, state reads) between the replays, and after
switching between call shapes (several cached graphs).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
  return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _solver(sched, mask, tgt, grad):
  from fpie_amd import core_hip
  s = core_hip.GridSolver(schedule=sched)
  s.reset(0, mask, tgt, grad)
  s.set_tile(8, 1, 16, 1)  # the small-grid tile: 1 launch = 8 sweeps
  return s


def test_jacobi_replays_equal_eager_launches():
  from fpie_amd import synth
  mask, tgt, grad, _ = synth.composite((300, 280), 41)
  g = _solver("jacobi", mask, tgt, grad)  # 200 sweeps = 25 launches: graph
  e = _solver("jacobi", mask, tgt, grad)  # 8 sweeps = 1 launch: eager
  for r in range(4):
    g.iterate(200)
    for _ in range(25):
      e.iterate(8)
    np.testing.assert_array_equal(_bits(g.state()), _bits(e.state()),
                                  err_msg=f"replay {r}")
    # eager work between replays, on both
    g.relax(1, 0.5)
    e.relax(1, 0.5)
    # a second call shape (another cached graph), then back
    g.iterate(136)
    for _ in range(17):
      e.iterate(8)
    np.testing.assert_array_equal(_bits(g.state()), _bits(e.state()),
                                  err_msg=f"replay {r}, second shape")


def test_gs_lex_replays_equal_oracle(oracle):
  from fpie_amd import synth
  mask, tgt, grad, _ = synth.composite((120, 150), 42)
  g = _solver("gs_lex", mask, tgt, grad)
  want = tgt
  for r in range(3):
    g.iterate(20)  # 2*19 + 119 + 149 + 1 steps: 39 launches, a graph
    want = oracle.grid_gs(mask, want, grad, 20)
    np.testing.assert_array_equal(_bits(g.state()), _bits(want),
                                  err_msg=f"replay {r}")
