#!/bin/sh
# Compile the reference's own gcc and openmp solver cores (pybind11 modules)
# straight from their sources where they lie under the reference checkout,
# with the flags of fpie/core/{gcc,openmp}/CMakeLists.txt:1.  Nothing is
# copied; the only outputs are oracle/_ref/core_gcc*.so and core_openmp*.so.
# pybind11 is the header-only package installed in this image.
set -e
REF="${FPIE_REFERENCE:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF/fpie/core/gcc" ]; then
  echo "build_ref.sh: no reference checkout at $REF; skipped" >&2
  exit 0
fi
mkdir -p "$OUT"
INC="$(python3 -m pybind11 --includes)"
EXT="$(python3 -c 'import sysconfig; print(sysconfig.get_config_var("EXT_SUFFIX"))')"
CORE="$REF/fpie/core"
g++ -O3 -std=c++11 -shared -fPIC $INC -I"$CORE" \
  "$CORE/gcc/solver.cc" "$CORE/gcc/equ.cc" "$CORE/gcc/grid.cc" \
  -o "$OUT/core_gcc$EXT"
g++ -O3 -std=c++11 -fopenmp -shared -fPIC $INC -I"$CORE" \
  "$CORE/openmp/solver.cc" "$CORE/openmp/equ.cc" "$CORE/openmp/grid.cc" \
  -o "$OUT/core_openmp$EXT"
echo "build_ref.sh: built $OUT/core_gcc$EXT $OUT/core_openmp$EXT"
