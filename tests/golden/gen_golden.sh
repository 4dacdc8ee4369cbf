#!/bin/bash
# Regenerates tests/golden/golden_ref.json from the compiled reference.
# Needs /root/reference (read in place) and the image's MPICH + MKL; run it
# where those are present.  The tests only read the fixtures it writes.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
REF=${PARATT_REFERENCE:-/root/reference}
CONDA=${CONDA_PREFIX_FOR_REF:-/opt/conda}
"$ROOT/oracle/build_ref.sh"
TMP=$(mktemp -d)
export MPICH_CC=gcc
"$CONDA/bin/mpicc" -O2 -w -I"$REF/source" -o "$TMP/gen_golden" "$HERE/gen_golden.c" \
  -L"$ROOT/oracle/_ref" -Wl,-rpath,"$ROOT/oracle/_ref" -lparatt_ref -lm
"$TMP/gen_golden" > "$HERE/golden_ref.json"
rm -rf "$TMP"
python3 -c "import json,sys; d=json.load(open('$HERE/golden_ref.json')); print('golden keys:', len(d))"
