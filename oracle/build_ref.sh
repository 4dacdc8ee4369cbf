#!/bin/bash
# Compiles the reference's own hot-path sources (read in place under
# /root/reference, never copied) into oracle/_ref/libparatt_ref.so.
#
# Toolchain: the image's MPICH 3.3.2 mpicc (gcc back end).  LAPACK/BLAS:
# the netlib LAPACK inside scipy's bundled OpenBLAS (LP64), whose symbols
# carry a "scipy_" prefix; the reference objects' undefined LAPACK/BLAS
# references are renamed to those symbols with objcopy.  (MKL is not used:
# its dgesvdq faults on the reference's NULL-pointer workspace query,
# pattac.c:370-391, which netlib LAPACK accepts.)
#
# The result is a checker only: tests/golden/gen_golden.sh links it to write
# golden vectors and bench.py may time it as the CPU baseline.
set -euo pipefail
REF=${PARATT_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
CONDA=${CONDA_PREFIX_FOR_REF:-/opt/conda}
if [ ! -d "$REF/source" ]; then
  echo "build_ref: reference sources not present at $REF; skipping" >&2
  exit 0
fi
OPENBLAS=$(python3 "$HERE/find_openblas.py")
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
export MPICH_CC=gcc
for s in source/paratt.c source/paratt_ode.c \
         source/mpi_attac_src/mpi_attac/pattac.c \
         source/mpi_attac_src/mpi_attac/tsqr.c \
         source/mpi_attac_src/serial_attac/attac.c \
         source/mpi_attac_src/serial_attac/utilities.c; do
  o="$TMP/$(basename "${s%.c}").o"
  "$CONDA/bin/mpicc" -O2 -fPIC -w -c "$REF/$s" -o "$o"
  # rename every undefined Fortran-ABI LAPACK/BLAS symbol (d*_ / lsame_)
  for sym in $(nm -u "$o" | awk '{print $2}' | grep -E '^(d[a-z0-9]+_|lsame_)$' || true); do
    objcopy --redefine-sym "$sym=scipy_$sym" "$o"
  done
done
"$CONDA/bin/mpicc" -shared -o "$OUT/libparatt_ref.so" "$TMP"/*.o \
  "$OPENBLAS" -Wl,-rpath,"$(dirname "$OPENBLAS")" -lm
echo "build_ref: wrote $OUT/libparatt_ref.so (LAPACK: $OPENBLAS)"
