"""The C ABI library (CPU only: no kernel is launched here).

libfpie_hip.so must load and export every entry point include/fpie_hip.h
declares; the command line must list the hip backend.
"""

import os
import re
import subprocess
import sys

import pytest

from conftest import PKG, ROOT

from fpie_amd import native

HEADER = os.path.join(ROOT, "include", "fpie_hip.h")


def declared_symbols():
  text = open(HEADER).read()
  text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
  return sorted(set(re.findall(r"\b(pie_[a-z_]+)\s*\(", text)))


def test_header_declares_the_reference_core_interface():
  names = declared_symbols()
  for need in ("pie_create", "pie_destroy", "pie_partition", "pie_reset_equ",
               "pie_reset_grid", "pie_sync", "pie_step", "pie_last_error"):
    assert need in names


@pytest.mark.skipif(not native.available(), reason="libfpie_hip.so not built")
def test_library_exports_every_declared_symbol():
  lib = native.load()
  missing = [n for n in declared_symbols() if not hasattr(lib, n)]
  assert not missing, missing
  out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH],
                       capture_output=True, text=True, check=True).stdout
  exported = set(re.findall(r"\bT (pie_\w+)", out))
  assert set(declared_symbols()) <= exported


@pytest.mark.skipif(not native.available(), reason="libfpie_hip.so not built")
def test_library_is_built_for_gfx950():
  out = subprocess.run(["strings", native.LIB_PATH], capture_output=True,
                       text=True, check=True).stdout
  assert "gfx950" in out


@pytest.mark.skipif(not native.available(), reason="libfpie_hip.so not built")
def test_last_error_starts_empty():
  assert native.load().pie_last_error() == b""


@pytest.mark.skipif(not native.available(), reason="libfpie_hip.so not built")
def test_tuning_calls_validate_before_any_device_work():
  """Host-side argument checks of the tile controls (no HIP call made)."""
  lib = native.load()
  assert lib.pie_batch_set_tile_channels(None, 1) < 0
  assert b"NULL batch" in lib.pie_last_error()
  assert lib.pie_unet_set_conv_tiles(-2) < 0
  assert b"below -1" in lib.pie_last_error()
  assert lib.pie_unet_set_conv_tiles(512) == 0  # the default, restored


@pytest.mark.skipif(not native.available(), reason="libfpie_hip.so not built")
def test_noalloc_solve_checks_arguments_before_device_work():
  """The allocation-free solve rejects NULL buffers, bad shapes and crops
  beyond its scratch array on the host, before any HIP call (the pointers
  below are never dereferenced)."""
  import ctypes
  lib = native.load()
  p = ctypes.c_void_p(0x1000)
  jac, ulrd = native.PIE_JACOBI, native.PIE_ORDER_ULRD
  assert lib.pie_grid_solve_noalloc(64, 64, None, p, p, 10, jac, ulrd, p, p,
                                    None) == native.PIE_EINVAL
  assert b"NULL argument" in lib.pie_last_error()
  assert lib.pie_grid_solve_noalloc(2, 64, p, p, p, 10, jac, ulrd, p, p,
                                    None) == native.PIE_EINVAL
  assert lib.pie_grid_solve_noalloc(64, 64, p, p, p, 10, 7, ulrd, p, p,
                                    None) == native.PIE_EINVAL
  assert lib.pie_grid_solve_noalloc(2000, 2000, p, p, p, 10, jac, ulrd, p, p,
                                    None) == native.PIE_EINVAL
  assert b"MiB of scratch" in lib.pie_last_error()


def test_cli_check_backend_lists_hip():
  code = ("import sys; sys.path.insert(0, %r); "
          "from fpie_amd.cli import main; main(['--check-backend'])" % PKG)
  r = subprocess.run([sys.executable, "-c", code], capture_output=True,
                     text=True)
  assert r.returncode == 0, r.stderr
  if native.available():
    assert "hip" in r.stdout
