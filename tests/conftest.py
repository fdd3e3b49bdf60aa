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
                if not p.endswith(("benchmark_masks.npz",
                                   "processor_cases.npz")))


def processor_cases():
  """Cases recorded from the reference processor by
  tests/golden/make_processor_golden.py: name -> dict of arrays."""
  path = os.path.join(GOLDEN, "processor_cases.npz")
  with np.load(path, allow_pickle=False) as f:
    d = {k: f[k] for k in f.files}
  out = {}
  for name in d["names"]:
    name = str(name)
    out[name] = {k[len(name) + 1:]: v for k, v in d.items()
                 if k.startswith(name + "_")}
  return out


def load_golden(name):
  with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as f:
    return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def oracle():
  import oracle as O
  O.lib()
  return O
