"""Command-line flags (fpie/args.py:10-142) plus the hip-only flags
``--schedule`` and ``--order``."""

from __future__ import annotations

import argparse
import os
import sys

from . import __version__
from .process import ALL_BACKEND, CPU_COUNT, DEFAULT_BACKEND


def get_args(gen_type: str = "cli", argv=None) -> argparse.Namespace:
  p = argparse.ArgumentParser()
  p.add_argument("-v", "--version", action="store_true",
                 help="show the version and exit")
  p.add_argument("--check-backend", action="store_true",
                 help="print all available backends")
  p.add_argument("-b", "--backend", type=str, choices=ALL_BACKEND or None,
                 default=DEFAULT_BACKEND, help="backend choice")
  p.add_argument("-c", "--cpu", type=int, default=CPU_COUNT,
                 help="number of CPU used")
  p.add_argument("-z", "--block-size", type=int, default=1024,
                 help="cuda block size (accepted; the hip tile is autotuned)")
  p.add_argument("--method", type=str, choices=["equ", "grid"],
                 default="equ", help="how to parallelize computation")
  p.add_argument("-s", "--source", type=str, help="source image filename")
  p.add_argument("-m", "--mask", type=str, default="",
                 help="mask image filename (default is to use the whole "
                 "source image)")
  p.add_argument("-t", "--target", type=str, help="target image filename")
  p.add_argument("-o", "--output", type=str, help="output image filename")
  p.add_argument("-h0", type=int, default=0,
                 help="mask position (height) on source image")
  p.add_argument("-w0", type=int, default=0,
                 help="mask position (width) on source image")
  p.add_argument("-h1", type=int, default=0,
                 help="mask position (height) on target image")
  p.add_argument("-w1", type=int, default=0,
                 help="mask position (width) on target image")
  p.add_argument("-g", "--gradient", type=str, choices=["max", "src", "avg"],
                 default="max", help="how to calculate gradient for PIE")
  p.add_argument("-n", type=int, default=5000,
                 help="how many iteration would you perfer, the more the "
                 "better")
  p.add_argument("-p", type=int, default=0,
                 help="output result every P iteration")
  p.add_argument("--grid-x", type=int, default=8,
                 help="x axis stride for grid solver")
  p.add_argument("--grid-y", type=int, default=8,
                 help="y axis stride for grid solver")
  p.add_argument("--schedule", type=str, default="jacobi",
                 choices=["jacobi", "gs_lex", "redblack"],
                 help="hip update schedule: jacobi, gs_lex (the gcc "
                 "backend's Gauss-Seidel result) or redblack (openmp "
                 "EquSolver)")
  p.add_argument("--order", type=str, default="ulrd",
                 choices=["ulrd", "udlr"],
                 help="GridSolver neighbour order: ulrd (gcc/openmp) or "
                 "udlr (numpy)")
  args = p.parse_args(argv)
  if args.version:
    print(__version__)
    sys.exit(0)
  if args.check_backend:
    print(ALL_BACKEND)
    sys.exit(0)
  if not args.source or not os.path.exists(args.source):
    print(f"Source image {args.source} not found.")
    sys.exit(1)
  if not args.target or not os.path.exists(args.target):
    print(f"Target image {args.target} not found.")
    sys.exit(1)
  args.mpi_sync_interval = 0
  return args
