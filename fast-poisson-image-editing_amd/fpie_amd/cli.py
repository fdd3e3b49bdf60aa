"""``python -m fpie_amd.cli``: the reference command line (fpie/cli.py:10-63)
on the hip backend."""

from __future__ import annotations

import time

from .args import get_args
from .io import read_images, write_image
from .process import EquProcessor, GridProcessor


def main(argv=None) -> None:
  args = get_args("cli", argv)
  if args.method == "equ":
    proc = EquProcessor(args.gradient, args.backend, args.cpu,
                        args.mpi_sync_interval, args.block_size,
                        schedule=args.schedule)
  else:
    proc = GridProcessor(args.gradient, args.backend, args.cpu,
                         args.mpi_sync_interval, args.block_size,
                         args.grid_x, args.grid_y, schedule=args.schedule,
                         order=args.order)
  print(f"Successfully initialize PIE {args.method} solver "
        f"with {args.backend} backend")
  src, mask, tgt = read_images(args.source, args.mask, args.target)
  n = proc.reset(src, mask, tgt, (args.h0, args.w0), (args.h1, args.w1))
  print(f"# of vars: {n}")
  proc.sync()
  result = tgt
  t = time.time()
  p = args.p or args.n
  for i in range(0, args.n, p):
    result, err = proc.step(p)
    print(f"Iter {i + p}, abs error {err}")
    if i + p < args.n:
      write_image(f"iter{i + p:05d}.png", result)
  dt = time.time() - t
  print(f"Time elapsed: {dt:.4f}s")
  write_image(args.output, result)
  print(f"Successfully write image to {args.output}")


if __name__ == "__main__":
  main()
