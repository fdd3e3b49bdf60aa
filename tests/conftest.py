"""Test configuration: import paths, the ``gpu`` marker, shared fixtures."""

import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fast-poisson-image-editing_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, os.path.join(ROOT, "oracle")):
  if p not in sys.path:
    sys.path.insert(0, p)


def pytest_configure(config):
  config.addinivalue_line(
      "markers", "gpu: needs an AMD GPU (runs the HIP kernels)")


def golden_cases():
  """Names of the solver fixtures written by tests/golden/make_golden.py."""
  return sorted(os.path.basename(p)[:-4]
                for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                if not p.endswith("benchmark_masks.npz"))


def load_golden(name):
  with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as f:
    return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def oracle():
  import oracle as O
  O.lib()
  return O
