"""Build recipe of the gfx950 solver library ``fpie_amd/libfpie_hip.so``.

hipcc only (no cmake): every translation unit under ``csrc/`` is compiled for
gfx950 in parallel, the sweep kernel once per compiled tile variant, then
linked into one shared library that ``fpie_amd.native`` loads with ctypes.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "fpie_amd", "libfpie_hip.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # every add and multiply rounds on its own (bit-exact to the CPU loops)
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    # no packed FP32 instructions (v_pk_*_f32) in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def tiles() -> list[tuple[int, int, int]]:
  """(sweeps, columns per lane, channels per task) of every compiled
  variant of the sweep kernel (PIE_TILES in csrc/kernels.h)."""
  text = open(os.path.join(CSRC, "kernels.h")).read()
  line = re.search(r"#define PIE_TILES\(X\)((?:.*\\\n)*.*)", text).group(1)
  return [(int(a), int(b), int(c))
          for a, b, c in re.findall(r"X\((\d+), (\d+), (\d+)\)", line)]


def _compile(src: str, obj: str, defines: list[str]) -> str:
  # PIE_DEFINES: extra -D flags for tuning builds (e.g. the sweep kernel's
  # PIE_RING_MAX / PIE_WAVES_PER_EU in csrc/march.h, PIE_FINE_SHARE in
  # csrc/solver.hip)
  defines = defines + os.environ.get("PIE_DEFINES", "").split()
  cmd = [HIPCC, *FLAGS, *defines, "-c", src, "-o", obj]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n"
                       f"{' '.join(cmd)}\n{r.stdout}{r.stderr}")
  return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
  os.makedirs(OBJ, exist_ok=True)
  units = [(os.path.join(CSRC, f), os.path.join(OBJ, f[:-4] + ".o"), [])
           for f in sorted(os.listdir(CSRC))
           if f.endswith(".hip") and f != "march_inst.hip"]
  for s, c, n in tiles():
    units.append((os.path.join(CSRC, "march_inst.hip"),
                  os.path.join(OBJ, f"march_{s}x{c}x{n}.o"),
                  [f"-DPIE_S={s}", f"-DPIE_C={c}", f"-DPIE_N={n}"]))
  jobs = jobs or min(len(units), max(1, (os.cpu_count() or 2)))
  with cf.ThreadPoolExecutor(jobs) as ex:
    objs = list(ex.map(lambda u: _compile(*u), units))
  tmp = OUT + ".tmp"
  cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
  os.replace(tmp, OUT)
  if verbose:
    print(f"built {OUT}")
  return OUT


if __name__ == "__main__":
  build(int(sys.argv[1]) if len(sys.argv) > 1 else None)
