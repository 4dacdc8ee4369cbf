"""Print the path of scipy's bundled LP64 OpenBLAS (netlib LAPACK inside).

Used by oracle/Makefile and oracle/build_ref.sh to link the CPU checkers."""
import glob
import os

import scipy

libdir = os.path.dirname(scipy.__file__) + ".libs"
cands = sorted(glob.glob(os.path.join(libdir, "libscipy_openblas-*.so")))
if not cands:
    raise SystemExit("scipy OpenBLAS not found in " + libdir)
print(cands[0])
