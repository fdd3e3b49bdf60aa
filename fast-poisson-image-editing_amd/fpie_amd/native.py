"""ctypes binding of ``libfpie_hip.so`` (the C ABI in include/fpie_hip.h).

The library is built in-tree by ``build.py`` next to this file.  Loading it
fails loudly (``OSError``) when it is missing: there is no CPU fallback.
ctypes releases the GIL around every call, unlike the reference's pybind11
cores, which hold it for the whole ``step`` (fpie/core/gcc/solver.cc).
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfpie_hip.so")

PIE_EQU, PIE_GRID = 0, 1
PIE_JACOBI, PIE_GS_LEX, PIE_REDBLACK = 0, 1, 2
PIE_ORDER_UDLR, PIE_ORDER_ULRD = 0, 1
PIE_OK, PIE_EINVAL, PIE_EHIP, PIE_EUNSUPPORTED, PIE_ESTATE = 0, -1, -2, -3, -4

SCHEDULES = {"jacobi": PIE_JACOBI, "gs_lex": PIE_GS_LEX,
             "redblack": PIE_REDBLACK}
ORDERS = {"udlr": PIE_ORDER_UDLR, "ulrd": PIE_ORDER_ULRD}

_vp = ctypes.c_void_p
_ip = ctypes.c_int
_SIGNATURES = {
    "pie_create": ([_ip, _ip, _ip, _ip, _ip, ctypes.POINTER(_vp)], _ip),
    "pie_destroy": ([_vp], _ip),
    "pie_partition": ([_vp, _vp, _ip, _ip, _vp], _ip),
    "pie_reset_equ": ([_vp, _ip, _vp, _vp, _vp], _ip),
    "pie_reset_grid": ([_vp, _ip, _ip, _vp, _vp, _vp], _ip),
    "pie_sync": ([_vp], _ip),
    "pie_step": ([_vp, _ip, _vp, _vp], _ip),
    "pie_last_error": ([], ctypes.c_char_p),
    "pie_iterate": ([_vp, _ip], _ip),
    "pie_read": ([_vp, _vp, _vp], _ip),
    "pie_get_state": ([_vp, _vp], _ip),
    "pie_set_tile": ([_vp, _ip, _ip, _ip], _ip),
    "pie_set_tile_channels": ([_vp, _ip], _ip),
    "pie_get_tile_channels": ([_vp, ctypes.POINTER(_ip)], _ip),
    "pie_get_tile": ([_vp, ctypes.POINTER(_ip), ctypes.POINTER(_ip),
                      ctypes.POINTER(_ip)], _ip),
    "pie_autotune": ([_vp, _ip], _ip),
    "pie_set_timing": ([_vp, _ip], _ip),
    "pie_get_timing": ([_vp, ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(ctypes.c_longlong),
                        ctypes.POINTER(ctypes.c_double)], _ip),
    "pie_shape": ([_vp, ctypes.POINTER(_ip), ctypes.POINTER(_ip),
                   ctypes.POINTER(_ip), ctypes.POINTER(ctypes.c_longlong)],
                  _ip),
    "pie_device_info": ([ctypes.c_char_p, _ip], _ip),
    "pie_set_state": ([_vp, _vp], _ip),
    "pie_solve_to_residual": ([_vp, ctypes.c_double, _ip, ctypes.c_longlong,
                               ctypes.POINTER(ctypes.c_longlong),
                               ctypes.POINTER(ctypes.c_double)], _ip),
    "pie_reset_grid_images": ([_vp, _ip, _ip, _vp, _vp, _vp, _ip], _ip),
"pie_reset_equ_images": ([_vp, _ip, _ip, _vp, _vp, _vp, _ip], _ip),
    "pie_read_crop": ([_vp, _vp, _vp], _ip),
    "pie_restrict": ([_vp, _ip, _vp], _ip),
    "pie_prolong": ([_vp, _ip, _ip, _ip, _vp], _ip),
    "pie_relax": ([_vp, _ip, ctypes.c_float], _ip),
    "pie_prolong_dots": ([_vp, _ip, _ip, _ip, _vp, _vp], _ip),
    "pie_warm_net_start": ([_vp, _vp, _ip, _vp, _ip, _ip, _vp], _ip),
    "pie_warm_apply": ([_vp, _ip, _vp, _vp, _vp], _ip),
    "pie_unet_create": ([_ip, _ip, _ip, ctypes.POINTER(_vp)], _ip),
    "pie_unet_destroy": ([_vp], _ip),
    "pie_unet_param_count": ([_vp, ctypes.POINTER(ctypes.c_longlong)], _ip),
    "pie_unet_set_params": ([_vp, _vp], _ip),
    "pie_unet_get_params": ([_vp, _vp], _ip),
    "pie_unet_forward": ([_vp, _vp, _vp], _ip),
    "pie_unet_forward_device": ([_vp, _ip, _ip, _vp, _vp], _ip),
    "pie_unet_loss_grad": ([_vp, _vp, _vp, _vp,
                            ctypes.POINTER(ctypes.c_float), _vp], _ip),
    "pie_unet_train_step": ([_vp, _vp, _vp, _vp, ctypes.c_float,
                             ctypes.POINTER(ctypes.c_float)], _ip),
    "pie_unet_staging": ([_vp, _ip, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                          ctypes.POINTER(_vp)], _ip),
    "pie_unet_train_async": ([_vp, _ip, ctypes.c_float], _ip),
    "pie_unet_finish": ([_vp, _ip, ctypes.POINTER(ctypes.c_float)], _ip),
    "pie_unet_stats": ([_vp, ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(ctypes.c_double)], _ip),
    "pie_unet_set_conv_tiles": ([_ip], _ip),
    "pie_prepare_mask": ([_vp, _ip, _ip, _ip, _vp, _vp, ctypes.c_longlong],
                         _ip),
    "pie_grid_solve_noalloc": ([_ip, _ip, _vp, _vp, _vp, _ip, _ip, _ip, _vp,
                                _vp, _vp], _ip),
    "pie_batch_create": ([_ip, _ip, _ip, _ip, ctypes.POINTER(_vp)], _ip),
    "pie_batch_destroy": ([_vp], _ip),
    "pie_batch_autotune": ([_vp, _ip], _ip),
    "pie_batch_tile": ([_vp, ctypes.POINTER(_ip), ctypes.POINTER(_ip),
                        ctypes.POINTER(_ip), ctypes.POINTER(_ip)], _ip),
    "pie_batch_set_tile_channels": ([_vp, _ip], _ip),
    "pie_batch_add": ([_vp, _ip, _ip, _vp, _vp, _vp, _ip, ctypes.c_float,
                       ctypes.POINTER(_ip)], _ip),
    "pie_batch_solve": ([_vp, _ip, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(_ip)], _ip),
    "pie_batch_item": ([_vp, _ip, ctypes.POINTER(_ip),
                        ctypes.POINTER(ctypes.c_float),
                        ctypes.POINTER(ctypes.c_double)], _ip),
    "pie_batch_read": ([_vp, _ip, _vp, _vp, _vp], _ip),
    "pie_batch_size": ([_vp, ctypes.POINTER(_ip), ctypes.POINTER(_ip),
                        ctypes.POINTER(ctypes.c_longlong)], _ip),
}

_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
  """Load the library once; raises OSError if it is not built."""
  global _lib
  if _lib is None:
    if not os.path.exists(LIB_PATH):
      raise OSError(f"{LIB_PATH} is not built: run "
                    "`python fast-poisson-image-editing_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
      fn = getattr(lib, name)
      fn.argtypes = args
      fn.restype = res
    _lib = lib
  return _lib


def available() -> bool:
  try:
    load()
    return True
  except OSError:
    return False


def check(status: int) -> None:
  """Raise RuntimeError with the library's message on a failed call, as the
  reference's pybind11 cores turn std::runtime_error into RuntimeError
  (fpie/core/base_solver.h:63-74)."""
  if status != PIE_OK:
    msg = load().pie_last_error().decode(errors="replace")
    raise RuntimeError(f"fpie_hip error {status}: {msg}")


def device_info() -> str:
  buf = ctypes.create_string_buffer(512)
  check(load().pie_device_info(buf, 512))
  return buf.value.decode()
