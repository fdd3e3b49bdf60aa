"""Command-line flags.

The flag names, defaults and choices are the reference's command line
(fpie/args.py:10-142), so a user switching backends keeps the same command;
the help text is this package's own.  ``--schedule`` and ``--order`` are
hip-only additions.  MPI is not offered, so ``mpi_sync_interval`` is 0.
"""

from __future__ import annotations

import argparse
import os
import sys

from . import __version__
from .process import ALL_BACKEND, CPU_COUNT, DEFAULT_BACKEND

# (flags, keyword arguments) in the order --help lists them
_FLAGS = [
    (("-v", "--version"),
     dict(action="store_true", help="print the package version and quit")),
    (("--check-backend",),
     dict(action="store_true",
          help="list the solver backends this installation can run")),
    (("-b", "--backend"),
     dict(type=str, choices=ALL_BACKEND or None, default=DEFAULT_BACKEND,
          help="solver backend to run the iterations on")),
    (("-c", "--cpu"),
     dict(type=int, default=CPU_COUNT,
          help="thread count of the CPU backends")),
    (("-z", "--block-size"),
     dict(type=int, default=1024,
          help="threads per block of the cuda backend (kept for "
          "compatibility; hip picks its tile itself)")),
    (("--method",),
     dict(type=str, choices=["equ", "grid"], default="equ",
          help="solver layout: 'equ' numbers the masked pixels as a list "
          "of equations, 'grid' iterates on the cropped image")),
    (("-s", "--source"), dict(type=str, help="path of the source image")),
    (("-m", "--mask"),
     dict(type=str, default="",
          help="path of the mask image; without one every source pixel "
          "is blended")),
    (("-t", "--target"), dict(type=str, help="path of the target image")),
    (("-o", "--output"), dict(type=str, help="path the result is written to")),
    (("-h0",), dict(type=int, default=0,
                    help="row of the source image the mask's top edge "
                    "sits on")),
    (("-w0",), dict(type=int, default=0,
                    help="column of the source image the mask's left edge "
                    "sits on")),
    (("-h1",), dict(type=int, default=0,
                    help="row of the target image the blended region "
                    "starts at")),
    (("-w1",), dict(type=int, default=0,
                    help="column of the target image the blended region "
                    "starts at")),
    (("-g", "--gradient"),
     dict(type=str, choices=["max", "src", "avg"], default="max",
          help="guidance field: the source gradient, the mean of source "
          "and target, or whichever of the two is larger")),
    (("-n",), dict(type=int, default=5000,
                   help="number of solver iterations (more iterations, "
                   "closer to the exact blend)")),
    (("-p",), dict(type=int, default=0,
                   help="also write an intermediate image every P "
                   "iterations (0: only the final one)")),
    (("--grid-x",), dict(type=int, default=8,
                         help="tile height of the CPU grid solvers")),
    (("--grid-y",), dict(type=int, default=8,
                         help="tile width of the CPU grid solvers")),
    (("--schedule",),
     dict(type=str, default="jacobi", choices=["jacobi", "gs_lex", "redblack"],
          help="hip update schedule: jacobi, gs_lex (the gcc backend's "
          "Gauss-Seidel result) or redblack (openmp EquSolver)")),
    (("--order",),
     dict(type=str, default="ulrd", choices=["ulrd", "udlr"],
          help="GridSolver neighbour order: ulrd (gcc/openmp) or udlr "
          "(numpy)")),
]


def _parser() -> argparse.ArgumentParser:
  p = argparse.ArgumentParser()
  for flags, kw in _FLAGS:
    p.add_argument(*flags, **kw)
  return p


def get_args(gen_type: str = "cli", argv=None) -> argparse.Namespace:
  args = _parser().parse_args(argv)
  if args.version:
    print(__version__)
    sys.exit(0)
  if args.check_backend:
    print(ALL_BACKEND)
    sys.exit(0)
  # the reference's messages for missing inputs (fpie/args.py:134-139)
  for what, path in (("Source", args.source), ("Target", args.target)):
    if not path or not os.path.exists(path):
      print(f"{what} image {path} not found.")
      sys.exit(1)
  args.mpi_sync_interval = 0
  return args
