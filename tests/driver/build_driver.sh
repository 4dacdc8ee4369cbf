#!/bin/bash
# Builds the reference's CDF driver (mpi_pattac_rxn_burgers_rank_max_read_file.c,
# read in place under the reference tree, never copied) twice, into
# tests/driver/_bin/ (git-ignored; the binaries travel with the tree):
#   cdf_driver_amd.x  against this repository's headers and libparatt_amd.so
#                     (the drop-in: the driver source is unchanged apart from
#                     the two initialisations below);
#   cdf_driver_ref.x  against the compiled reference (oracle/_ref), the checker.
# The driver reads `dt` while initialising trunc_eps_old before `dt` is set
# and increments `nrmlz_idx` without initialising it; both builds get the
# same fix on the compiler's input stream (dt starts at the driver's INIT_DT,
# nrmlz_idx at 0) so that their runs are defined and comparable.
# Quote includes of a stream resolve against the working directory, so each
# build runs in a scratch directory whose source/ holds the headers to use.
set -euo pipefail
REF=${PARATT_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
CONDA=${CONDA_PREFIX_FOR_REF:-/opt/conda}
DRIVER="$REF/mpi_pattac_rxn_burgers_rank_max_read_file.c"
if [ ! -f "$DRIVER" ] || [ ! -x "$CONDA/bin/mpicc" ]; then
  echo "build_driver: reference driver or MPICH not present; skipping" >&2
  exit 0
fi
OUT="$HERE/_bin"
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
export MPICH_CC=gcc
fix() {
  sed -e 's/^\(    double dt\);/\1 = INIT_DT;/' \
      -e 's/\(nrmlz_period = 500, nrmlz_idx\);/\1 = 0;/' "$DRIVER"
}
# the drop-in build: this repository's headers as ./source/*.h
mkdir -p "$TMP/amd/source"
for h in "$ROOT"/include/*.h; do ln -s "$h" "$TMP/amd/source/"; done
(cd "$TMP/amd" && fix | "$CONDA/bin/mpicc" -std=gnu99 -O2 -w -x c -c - -o driver.o)
# libamdhip64 needs the system libstdc++, newer than MPICH's own directory holds
gcc "$TMP/amd/driver.o" -o "$OUT/cdf_driver_amd.x" \
  -L"$ROOT/paratt_amd" -lparatt_amd "$CONDA/lib/libmpi.so" \
  /usr/lib/x86_64-linux-gnu/libstdc++.so.6 \
  -Wl,-rpath,'$ORIGIN/../../../paratt_amd':/usr/lib/x86_64-linux-gnu:"$CONDA/lib" \
  -Wl,-rpath-link,/usr/lib/x86_64-linux-gnu:"$CONDA/lib" -lm
# the checker build: the reference's own headers and compiled sources
if [ -f "$ROOT/oracle/_ref/libparatt_ref.so" ]; then
  mkdir -p "$TMP/ref"
  ln -s "$REF/source" "$TMP/ref/source"
  (cd "$TMP/ref" && fix | "$CONDA/bin/mpicc" -std=gnu99 -O2 -w -x c - \
     -o "$OUT/cdf_driver_ref.x" -L"$ROOT/oracle/_ref" \
     -Wl,-rpath,"$ROOT/oracle/_ref" -lparatt_ref -lm)
fi
echo "build_driver: built $(ls "$OUT" | tr '\n' ' ')"
